// usx/demodulate.hpp -- RF to baseband IQ.
//
// Interface of the reference's core/include/usx/demodulate.hpp:23-49. The
// work runs on the GPU (usx_demodulate in include/usx_gpu.h); the filter
// taps and mixing rotator are built here on the host with the reference's
// expressions.
#pragma once

#include <vector>

#include "usx/types.hpp"

namespace usx {

struct DemodConfig {
  real demodulation_frequency = 0.0f;  // 0: probe centre frequency
  real cutoff_fraction = 0.5f;         // low-pass cutoff / demodulation frequency
  int filter_taps = 33;
  int decimation = 1;
};

ChannelData demodulate(const ChannelData& rf, DemodConfig config = {});

// Hamming-windowed sinc low-pass, odd length, unity DC gain.
std::vector<real> design_lowpass(int n_taps, real normalized_cutoff);

}  // namespace usx
