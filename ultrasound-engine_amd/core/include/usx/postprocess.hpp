// usx/postprocess.hpp -- display helpers downstream of the frame.
//
// Interface of the reference's core/include/usx/postprocess.hpp:27-59
// (envelope, apply_tgc, log_compress, analytic_along_axis1); host code, used
// by ImagingPipeline::to_display and usx.to_bmode.
#pragma once

#include <vector>

#include "usx/types.hpp"

namespace usx {

std::vector<real> envelope(const Frame& frame);
std::vector<real> apply_tgc(const std::vector<real>& env, const ScanGrid& grid,
                            real attenuation_db_cm_mhz, real center_frequency_hz);
std::vector<real> log_compress(const std::vector<real>& env, real dynamic_range_db = 60.0f,
                               real gain_db = 0.0f, real reference = -1.0f);
std::vector<cplx> analytic_along_axis1(const std::vector<real>& data, int n_axis0, int n_axis1);

}  // namespace usx
