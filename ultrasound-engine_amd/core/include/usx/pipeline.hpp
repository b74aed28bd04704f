// usx/pipeline.hpp -- the streaming front end.
//
// Interface of the reference's core/include/usx/pipeline.hpp:23-75. One
// ImagingPipeline owns one GPU context (stream + pooled device buffers), so
// a steady-state frame allocates no device memory.
#pragma once

#include <cstdint>
#include <memory>
#include <vector>

#include "usx/beamformer.hpp"
#include "usx/demodulate.hpp"
#include "usx/types.hpp"

namespace usx {

struct PipelineConfig {
  DemodConfig demod;
  BeamformerConfig beamformer;
  real dynamic_range_db = 60.0f;
  real gain_db = 0.0f;
  bool apply_tgc = true;
  real reference_smoothing = 0.85f;
};

struct PipelineStats {
  double demodulate_ms = 0;
  double beamform_ms = 0;
  double postprocess_ms = 0;
  double total_ms = 0;
  double frames_per_second() const { return total_ms > 0 ? 1000.0 / total_ms : 0.0; }
};

class GpuContext;

class ImagingPipeline {
 public:
  ImagingPipeline(ScanGrid grid, PipelineConfig config = {});
  ~ImagingPipeline();
  ImagingPipeline(ImagingPipeline&&) noexcept;
  ImagingPipeline& operator=(ImagingPipeline&&) noexcept;

  // Raw channel data (RF or IQ) to a complex beamformed frame.
  Frame process_frame(const ChannelData& data);
  // Raw channel data to a log-compressed display image in [0, 1].
  std::vector<real> process_display(const ChannelData& data);
  // A beamformed frame to a display image, with the smoothed reference level.
  std::vector<real> to_display(const Frame& frame, const Medium& medium);

  const ScanGrid& grid() const { return grid_; }
  const PipelineStats& stats() const { return stats_; }
  PipelineConfig& config() { return config_; }
  const PipelineConfig& config() const { return config_; }
  std::int64_t frame_count() const { return frame_count_; }
  void reset();

  // Learned apodization (no reference counterpart): the network's FP32
  // parameters (USX_APOD_PARAMS floats, include/usx_gpu.h). With
  // Combiner::Learned and no network loaded, frames are minimum variance.
  void set_apodization(const std::vector<float>& params);
  bool has_apodization() const { return has_apod_; }

 private:
  ScanGrid grid_;
  PipelineConfig config_;
  PipelineStats stats_;
  real reference_ = 0.0f;
  std::int64_t frame_count_ = 0;
  std::unique_ptr<GpuContext> gpu_;
  bool has_apod_ = false;
};

}  // namespace usx
