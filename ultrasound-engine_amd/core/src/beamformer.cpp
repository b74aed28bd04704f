// beamformer.cpp -- Beamformer on the GPU path.
// Reference: core/src/beamformer.cpp:218-295 (validation, transmit
// sensitivity), 297-304 (event range), 322-325 (frame metadata), 551-569
// (sequences). The per-pixel work is usx_beamform (include/usx_gpu.h).
#include "usx/beamformer.hpp"

#include <algorithm>
#include <cmath>

#include "gpu.hpp"
#include "usx/geometry.hpp"

namespace usx {

void BeamformerConfig::validate() const {
  if (f_number <= 0) throw std::invalid_argument("BeamformerConfig.f_number must be > 0");
  if (mv_subarray < 2) throw std::invalid_argument("BeamformerConfig.mv_subarray must be >= 2");
  if (mv_diagonal_loading < 0)
    throw std::invalid_argument("BeamformerConfig.mv_diagonal_loading must be >= 0");
  if (mv_range_average < 0)
    throw std::invalid_argument("BeamformerConfig.mv_range_average must be >= 0");
}

real transmit_sensitivity(const Probe& probe, const Transmit& tx, const Vec3& point,
                          real wavelength) {
  // Firing-aperture extent (all elements when there is no apodization).
  bool valid = false;
  real lo = 0, hi = 0;
  for (int i = 0; i < probe.n_elements; ++i) {
    if (!tx.apodization.empty() && tx.apodization[i] == 0.0f) continue;
    const real x = probe.element_position(i).x;
    lo = valid ? std::min(lo, x) : x;
    hi = valid ? std::max(hi, x) : x;
    valid = true;
  }
  if (!valid || point.z <= 0.0f) return 0.0f;
  real a, b;
  if (tx.type == TransmitType::PlaneWave) {
    const real shift = point.z * std::tan(tx.angle);
    a = lo + shift;
    b = hi + shift;
  } else {
    const real zf = tx.focus.z;
    if (std::abs(zf) < 1e-9f) return 1.0f;
    a = lo + (tx.focus.x - lo) * point.z / zf;
    b = hi + (tx.focus.x - hi) * point.z / zf;
    if (a > b) std::swap(a, b);
  }
  const real aperture = std::max(hi - lo, probe.pitch);
  const real diffraction = std::clamp(point.z / aperture, 0.5f, 20.0f) * wavelength;
  a -= diffraction;
  b += diffraction;
  const real centre = 0.5f * (a + b);
  const real half = 0.5f * (b - a);
  if (half <= 0.0f) return 0.0f;
  return window_value(Window::Tukey, (point.x - centre) / half, 0.4f);
}

Frame Beamformer::beamform(const ChannelData& data, const ScanGrid& grid, int event_begin,
                           int event_end) const {
  data.validate();
  grid.validate();
  const int n_ev = data.n_events();
  if (event_end < 0) event_end = n_ev;
  event_begin = std::clamp(event_begin, 0, n_ev);
  event_end = std::clamp(event_end, event_begin, n_ev);
  const real fdem = data.is_iq ? data.demodulation_frequency : 0.0f;

  Frame frame;
  frame.grid = grid;
  frame.demodulation_frequency = fdem;
  frame.allocate();
  const void* src = data.is_iq ? static_cast<const void*>(data.iq.data())
                               : static_cast<const void*>(data.rf.data());
  if ((data.is_iq ? data.iq.size() : data.rf.size()) != data.size())
    throw std::invalid_argument("beamform: channel data buffer is empty");
  const BeamformPlan plan = make_bf_plan(data, data.n_samples(), data.acq.sampling_frequency,
                                         data.is_iq, fdem, grid, config_, event_begin, event_end);
  usx_ctx* ctx = thread_context().get();
  check(ctx, usx_beamform(ctx, &plan.desc, src, reinterpret_cast<usx_c32*>(frame.data.data()),
                          USX_HOST),
        "beamform");
  return frame;
}

std::vector<Frame> Beamformer::beamform_sequence(const ChannelData& data, const ScanGrid& grid,
                                                 int events_per_frame) const {
  if (events_per_frame <= 0)
    throw std::invalid_argument("beamform_sequence: events_per_frame must be > 0");
  const int n_frames = data.n_events() / events_per_frame;
  if (n_frames == 0)
    throw std::invalid_argument("beamform_sequence: fewer events than events_per_frame");
  std::vector<Frame> frames;
  frames.reserve(static_cast<std::size_t>(n_frames));
  for (int f = 0; f < n_frames; ++f) {
    Frame fr = beamform(data, grid, f * events_per_frame, (f + 1) * events_per_frame);
    fr.sequence = f;
    fr.timestamp = static_cast<double>(f) * events_per_frame / data.acq.prf;
    frames.push_back(std::move(fr));
  }
  return frames;
}

}  // namespace usx
