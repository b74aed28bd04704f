// gpu.hpp -- host-side bridge from the usx:: C++ API to the C ABI
// (include/usx_gpu.h): contexts, status mapping, descriptor construction.
// Internal to the host layer.
#pragma once

#include <memory>
#include <vector>

#include "usx/beamformer.hpp"
#include "usx/demodulate.hpp"
#include "usx/pipeline.hpp"
#include "usx/types.hpp"
#include "usx_gpu.h"

namespace usx {

// Owns one usx_ctx (created on first use). Device: $USX_DEVICE or 0.
class GpuContext {
 public:
  GpuContext() = default;
  ~GpuContext();
  GpuContext(const GpuContext&) = delete;
  GpuContext& operator=(const GpuContext&) = delete;
  usx_ctx* get();

 private:
  usx_ctx* ctx_ = nullptr;
};

// The calling thread's context for the free functions (demodulate,
// Beamformer::beamform), which are reentrant in the reference.
GpuContext& thread_context();

// Throws std::invalid_argument for USX_ERR_INVALID, std::runtime_error otherwise.
void check(usx_ctx* ctx, usx_status st, const char* what);

struct DemodPlan {
  std::vector<real> taps;
  std::vector<cplx> rotator;
  real fdem = 0;
  real fs_out = 0;
  int n_out = 0;
  usx_demod_desc desc{};
};
DemodPlan make_demod_plan(const ChannelData& rf, const DemodConfig& cfg);

struct BeamformPlan {
  std::vector<real> element_x, axis0, axis1, s0, c0;
  std::vector<usx_tx_desc> tx;
  usx_bf_desc desc{};
};
// `samples`, `fs`, `is_iq`, `fdem` describe the data the plan will beamform
// (they differ from data.acq when the data is demodulated on the device).
BeamformPlan make_bf_plan(const ChannelData& data, int samples, real fs, bool is_iq, real fdem,
                          const ScanGrid& grid, const BeamformerConfig& cfg, int ev0, int ev1);
// Only the grid fields (the display tail needs no channel data).
BeamformPlan make_grid_plan(const ScanGrid& grid);
usx_display_desc make_display_desc(const PipelineConfig& cfg, const Medium& medium, real fdem);

}  // namespace usx
