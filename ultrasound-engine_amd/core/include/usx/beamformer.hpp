// usx/beamformer.hpp -- delay, apodize and combine.
//
// Interface of the reference's core/include/usx/beamformer.hpp:40-138; the
// work runs on the GPU (usx_beamform in include/usx_gpu.h).
#pragma once

#include <vector>

#include "usx/types.hpp"

namespace usx {

// Learned: the learned-apodization network (no reference counterpart), with
// minimum variance as its fallback when no network is loaded.
enum class Combiner : std::uint8_t { DAS = 0, DMAS = 1, MV = 2, Learned = 3 };
enum class Compounding : std::uint8_t { Coherent = 0, Incoherent = 1, Select = 2 };

// Covariance input precision of the MV combine (no reference counterpart;
// FP32 is the reference's arithmetic and the default).
enum class MvPrecision : std::uint8_t { FP32 = 0, BF16 = 1, FP16 = 2 };

struct BeamformerConfig {
  real f_number = 1.75f;
  Window rx_window = Window::Hann;
  real tukey_alpha = 0.5f;
  Combiner combiner = Combiner::DAS;
  Compounding compounding = Compounding::Coherent;
  bool transmit_weighting = true;
  bool coherence_factor = false;
  int mv_subarray = 16;
  real mv_diagonal_loading = 1e-2f;
  int mv_range_average = 2;
  // Additions without a reference counterpart; the defaults keep the
  // reference's arithmetic.
  MvPrecision mv_precision = MvPrecision::FP32;
  bool mv_refine = false;  // one step of iterative refinement, FP64 residual (not in the reference)
  // Launch-shape tunables (0 = default): transmit events per beamforming
  // launch (1..16) and persistent MV workgroups per CU; usx.autotune picks them.
  int tile_slots = 0;
  int tile_mv_occupancy = 0;

  void validate() const;
};

class Beamformer {
 public:
  explicit Beamformer(BeamformerConfig config = {}) : config_(config) { config_.validate(); }

  // Beamform events [event_begin, event_end) (end < 0: to the last) into one frame.
  Frame beamform(const ChannelData& data, const ScanGrid& grid, int event_begin = 0,
                 int event_end = -1) const;

  // Consecutive groups of events_per_frame transmits, one frame each.
  std::vector<Frame> beamform_sequence(const ChannelData& data, const ScanGrid& grid,
                                       int events_per_frame) const;

  const BeamformerConfig& config() const { return config_; }
  BeamformerConfig& config() { return config_; }

 private:
  BeamformerConfig config_;
};

// Illumination weight of a transmit at a point (Tukey taper over the
// insonified band, widened by the diffraction-limited beam width).
real transmit_sensitivity(const Probe& probe, const Transmit& tx, const Vec3& point,
                          real wavelength);

}  // namespace usx
