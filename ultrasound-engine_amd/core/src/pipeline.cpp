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
  const BeamformPlan gp = make_grid_plan(frame.grid);
  const usx_display_desc dd = make_display_desc(config_, medium, frame.demodulation_frequency);
  std::vector<real> img(static_cast<std::size_t>(frame.grid.n_axis0) * frame.grid.n_axis1);
  if (frame.data.size() != img.size()) throw std::invalid_argument("to_display: frame data does not match its grid");
  usx_ctx* ctx = gpu_->get();
  check(ctx, usx_display(ctx, &gp.desc, &dd, reinterpret_cast<const usx_c32*>(frame.data.data()),
                         frame.demodulation_frequency > 0.0f ? 0 : 1, img.data(), &reference_, USX_HOST),
        "display");
  stats_.postprocess_ms += ms_since(t);
  stats_.total_ms += stats_.postprocess_ms;
  return img;
}

std::vector<real> ImagingPipeline::process_display(const ChannelData& data) {
  if (data.is_iq) {
    const Frame frame = process_frame(data);
    return to_display(frame, data.medium);
  }
  // RF: demodulate, beamform and display on the device; the frame stays in HBM
  const auto t_start = Clock::now();
  stats_ = PipelineStats{};
  data.validate();
  if (data.rf.empty()) throw std::invalid_argument("demodulate: input has no RF samples");
  usx_ctx* ctx = gpu_->get();
  BeamformerConfig bcfg = config_.beamformer;
  if (bcfg.combiner == Combiner::Learned && !has_apod_) bcfg.combiner = Combiner::MV;
  const DemodPlan dp = make_demod_plan(data, config_.demod);
  const BeamformPlan bp = make_bf_plan(data, dp.n_out, dp.fs_out, true, dp.fdem, grid_, bcfg, 0,
                                       data.n_events());
  const usx_display_desc dd = make_display_desc(config_, data.medium, dp.fdem);
  std::vector<real> img(static_cast<std::size_t>(grid_.n_axis0) * grid_.n_axis1);
  usx_stage_times t{};
  check(ctx, usx_process_display(ctx, &dp.desc, &bp.desc, &dd, data.rf.data(), img.data(), &reference_,
                                 USX_HOST, &t),
        "process_display");
  stats_.demodulate_ms = t.demodulate_ms;
  stats_.beamform_ms = t.beamform_ms;
  stats_.postprocess_ms = t.postprocess_ms;
  ++frame_count_;
  stats_.total_ms = ms_since(t_start);
  return img;
}

}  // namespace usx
