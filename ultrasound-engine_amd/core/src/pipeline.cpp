// pipeline.cpp -- ImagingPipeline over the C ABI.
// Reference: core/src/pipeline.cpp:14-79. RF input takes one fused
// usx_process_frame call (demodulate + beamform on the device, IQ never
// leaves HBM); IQ input takes usx_beamform. Stage times come from HIP events.
#include "usx/pipeline.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <stdexcept>
#include <string>

#include "gpu.hpp"

namespace usx {
namespace {
using Clock = std::chrono::steady_clock;
double ms_since(Clock::time_point t) {
  return std::chrono::duration<double, std::milli>(Clock::now() - t).count();
}
}  // namespace

ImagingPipeline::ImagingPipeline(ScanGrid grid, PipelineConfig config)
    : grid_(std::move(grid)), config_(config), gpu_(new GpuContext()) {
  config_.beamformer.validate();
  grid_.validate();
}

ImagingPipeline::~ImagingPipeline() = default;
ImagingPipeline::ImagingPipeline(ImagingPipeline&&) noexcept = default;
ImagingPipeline& ImagingPipeline::operator=(ImagingPipeline&&) noexcept = default;

void ImagingPipeline::reset() {
  reference_ = 0.0f;
  frame_count_ = 0;
  stats_ = PipelineStats{};
}

void ImagingPipeline::set_apodization(const std::vector<float>& params) {
  if (params.size() != static_cast<std::size_t>(USX_APOD_PARAMS))
    throw std::invalid_argument("set_apodization: expected " + std::to_string(USX_APOD_PARAMS) +
                                " parameters");
  usx_ctx* ctx = gpu_->get();
  check(ctx, usx_apod_set_params(ctx, params.data()), "apod_set_params");
  has_apod_ = true;
}

Frame ImagingPipeline::process_frame(const ChannelData& data) {
  const auto t_start = Clock::now();
  stats_ = PipelineStats{};
  data.validate();
  config_.beamformer.validate();
  usx_ctx* ctx = gpu_->get();
  BeamformerConfig bcfg = config_.beamformer;
  if (bcfg.combiner == Combiner::Learned && !has_apod_) bcfg.combiner = Combiner::MV;

  Frame frame;
  frame.grid = grid_;
  const int n_ev = data.n_events();
  if (!data.is_iq) {
    if (data.rf.empty()) throw std::invalid_argument("demodulate: input has no RF samples");
    const DemodPlan dp = make_demod_plan(data, config_.demod);
    frame.demodulation_frequency = dp.fdem;
    frame.allocate();
    const BeamformPlan bp = make_bf_plan(data, dp.n_out, dp.fs_out, true, dp.fdem, grid_,
                                         bcfg, 0, n_ev);
    usx_stage_times t{};
    check(ctx, usx_process_frame(ctx, &dp.desc, &bp.desc, data.rf.data(),
                                 reinterpret_cast<usx_c32*>(frame.data.data()), USX_HOST, &t),
          "process_frame");
    stats_.demodulate_ms = t.demodulate_ms;
    stats_.beamform_ms = t.beamform_ms;
  } else {
    if (data.iq.size() != data.size())
      throw std::invalid_argument("beamform: channel data buffer is empty");
    frame.demodulation_frequency = data.demodulation_frequency;
    frame.allocate();
    const BeamformPlan bp = make_bf_plan(data, data.n_samples(), data.acq.sampling_frequency, true,
                                         data.demodulation_frequency, grid_, bcfg, 0, n_ev);
    const auto t = Clock::now();
    check(ctx, usx_beamform(ctx, &bp.desc, data.iq.data(),
                            reinterpret_cast<usx_c32*>(frame.data.data()), USX_HOST),
          "beamform");
    stats_.beamform_ms = ms_since(t);
  }
  frame.sequence = frame_count_++;
  stats_.total_ms = ms_since(t_start);
  return frame;
}

std::vector<real> ImagingPipeline::to_display(const Frame& frame, const Medium& medium) {
  const auto t = Clock::now();
  std::vector<real> env = envelope(frame);
  if (config_.apply_tgc) {
    real f0 = frame.demodulation_frequency;
    if (f0 <= 0.0f) f0 = 5.0e6f;
    env = apply_tgc(env, frame.grid, medium.attenuation, f0);
  }
  real peak = 0.0f;
  for (real v : env) peak = std::max(peak, v);
  if (peak > 0.0f) {
    const real a = std::clamp(config_.reference_smoothing, 0.0f, 0.999f);
    reference_ = (reference_ > 0.0f) ? (a * reference_ + (1.0f - a) * peak) : peak;
  }
  std::vector<real> img = log_compress(env, config_.dynamic_range_db, config_.gain_db,
                                       reference_ > 0.0f ? reference_ : -1.0f);
  stats_.postprocess_ms += ms_since(t);
  stats_.total_ms += stats_.postprocess_ms;
  return img;
}

std::vector<real> ImagingPipeline::process_display(const ChannelData& data) {
  const Frame frame = process_frame(data);
  return to_display(frame, data.medium);
}

}  // namespace usx
