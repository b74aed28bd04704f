// gpu.cpp -- descriptor construction and context management.
//
// Everything the kernels take from the host is computed here with the
// reference's own float expressions (compiled like the reference: -O3, no
// -march, no fast-math), so the device sees bitwise the same constants:
//   FIR taps and mixing rotator       demodulate.cpp:13-34, 72-76
//   element positions                 types.hpp:77-80
//   per-transmit sin/cos/tan, focus
//   axis, firing-aperture extent      geometry.cpp:82-109, beamformer.cpp:238-288
//   sector sin/cos of theta           types.cpp:110-111
#include "gpu.hpp"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <mutex>
#include <stdexcept>
#include <string>
#include <thread>
#include <unordered_map>

namespace usx {

GpuContext::~GpuContext() {
  if (ctx_) usx_ctx_destroy(ctx_);
}

usx_ctx* GpuContext::get() {
  if (!ctx_) {
    int dev = 0;
    if (const char* e = std::getenv("USX_DEVICE")) dev = std::atoi(e);
    const usx_status st = usx_ctx_create(dev, &ctx_);
    if (st != USX_OK) {
      ctx_ = nullptr;
      throw std::runtime_error(
          "usx: no usable gfx950 GPU (device " + std::to_string(dev) +
          "); the imaging path runs only on MI355X-class hardware (status " +
          std::to_string(static_cast<int>(st)) + ")");
    }
  }
  return ctx_;
}

GpuContext& thread_context() {
  // One context per thread, kept for the life of the process: the free
  // functions are reentrant, and tearing HIP objects down from thread-exit
  // handlers races with runtime shutdown.
  static std::mutex mu;
  static std::unordered_map<std::thread::id, GpuContext*>* map =
      new std::unordered_map<std::thread::id, GpuContext*>();
  std::lock_guard<std::mutex> lk(mu);
  GpuContext*& g = (*map)[std::this_thread::get_id()];
  if (!g) g = new GpuContext();
  return *g;
}

void check(usx_ctx* ctx, usx_status st, const char* what) {
  if (st == USX_OK) return;
  const std::string msg = ctx ? usx_ctx_last_error(ctx) : std::string("no context");
  if (st == USX_ERR_INVALID) throw std::invalid_argument(msg);
  throw std::runtime_error(std::string(what) + ": " + msg);
}

std::vector<real> design_lowpass(int n_taps, real normalized_cutoff) {
  if (n_taps < 3) n_taps = 3;
  if (n_taps % 2 == 0) ++n_taps;
  const real fc = std::clamp(normalized_cutoff, 1e-4f, 0.499f);
  std::vector<real> h(static_cast<std::size_t>(n_taps));
  const int m = n_taps / 2;
  real sum = 0.0f;
  for (int i = 0; i < n_taps; ++i) {
    const int k = i - m;
    const real s = (k == 0) ? 2.0f * fc : std::sin(2.0f * kPi * fc * k) / (kPi * k);
    const real w = 0.54f - 0.46f * std::cos(2.0f * kPi * i / (n_taps - 1));
    h[i] = s * w;
    sum += s * w;
  }
  if (sum != 0.0f)
    for (real& v : h) v /= sum;
  return h;
}

DemodPlan make_demod_plan(const ChannelData& rf, const DemodConfig& cfg) {
  DemodPlan p;
  p.fdem = cfg.demodulation_frequency > 0.0f ? cfg.demodulation_frequency
                                             : rf.probe.center_frequency;
  const int dec = std::max(1, cfg.decimation);
  const real fs = rf.acq.sampling_frequency;
  p.taps = design_lowpass(cfg.filter_taps, cfg.cutoff_fraction * p.fdem / fs);
  const int n_in = rf.n_samples();
  p.rotator.resize(static_cast<std::size_t>(n_in));
  for (int n = 0; n < n_in; ++n) {
    const real ph = -2.0f * kPi * p.fdem * rf.acq.sample_time(n);
    p.rotator[n] = cplx(std::cos(ph), std::sin(ph));
  }
  p.fs_out = fs / dec;
  p.n_out = (n_in + dec - 1) / dec;
  p.desc.n_traces = rf.n_events() * rf.n_channels();
  p.desc.n_in = n_in;
  p.desc.n_out = p.n_out;
  p.desc.decimation = dec;
  p.desc.n_taps = static_cast<int>(p.taps.size());
  p.desc.taps = p.taps.data();
  p.desc.rotator = reinterpret_cast<const usx_c32*>(p.rotator.data());
  return p;
}

BeamformPlan make_bf_plan(const ChannelData& data, int samples, real fs, bool is_iq, real fdem,
                          const ScanGrid& grid, const BeamformerConfig& cfg, int ev0, int ev1) {
  BeamformPlan p;
  const int n_ch = data.n_channels();
  const real c = data.medium.speed_of_sound;
  const real pitch = data.probe.pitch;
  p.element_x.resize(n_ch);
  for (int i = 0; i < n_ch; ++i) p.element_x[i] = data.probe.element_position(i).x;

  p.tx.resize(data.transmits.size());
  for (std::size_t e = 0; e < data.transmits.size(); ++e) {
    const Transmit& tx = data.transmits[e];
    usx_tx_desc& d = p.tx[e];
    d = usx_tx_desc{};
    d.type = static_cast<int>(tx.type);
    d.sin_a = std::sin(tx.angle);
    d.cos_a = std::cos(tx.angle);
    d.tan_a = std::tan(tx.angle);
    const Vec3 O(tx.origin.x, 0.0f, 0.0f);
    d.origin_x = O.x;
    d.focus[0] = tx.focus.x;
    d.focus[1] = tx.focus.y;
    d.focus[2] = tx.focus.z;
    const Vec3 axis = tx.focus - O;
    d.axis_len = axis.norm();
    if (d.axis_len >= 1e-9f) {
      d.axis_dir[0] = axis.x / d.axis_len;
      d.axis_dir[1] = axis.y / d.axis_len;
      d.axis_dir[2] = axis.z / d.axis_len;
    }
    d.origin_focus = distance(O, tx.focus);
    // Lateral extent of the firing aperture (all elements if no apodization).
    bool valid = false;
    real lo = 0, hi = 0;
    for (int i = 0; i < n_ch; ++i) {
      if (!tx.apodization.empty() && tx.apodization[i] == 0.0f) continue;
      const real x = data.probe.element_position(i).x;
      lo = valid ? std::min(lo, x) : x;
      hi = valid ? std::max(hi, x) : x;
      valid = true;
    }
    d.extent_valid = valid ? 1 : 0;
    d.x_lo = lo;
    d.x_hi = hi;
    d.aperture = std::max(hi - lo, pitch);
  }

  p.axis0 = grid.axis0;
  p.axis1 = grid.axis1;
  if (grid.type == GridType::Sector) {
    p.s0.resize(grid.n_axis0);
    p.c0.resize(grid.n_axis0);
    for (int i = 0; i < grid.n_axis0; ++i) {
      p.s0[i] = std::sin(grid.axis0[i]);
      p.c0[i] = std::cos(grid.axis0[i]);
    }
  }

  usx_bf_desc& b = p.desc;
  b.n_events = data.n_events();
  b.n_channels = n_ch;
  b.n_samples = samples;
  b.is_iq = is_iq ? 1 : 0;
  b.sampling_frequency = fs;
  b.t0 = data.acq.t0;
  b.speed_of_sound = c;
  b.demodulation_frequency = is_iq ? fdem : 0.0f;
  b.wavelength = data.probe.wavelength(c);
  b.pitch = pitch;
  b.element_x = p.element_x.data();
  b.tx = p.tx.data();
  b.event_begin = ev0;
  b.event_end = ev1;
  b.grid_type = static_cast<int>(grid.type);
  b.n_axis0 = grid.n_axis0;
  b.n_axis1 = grid.n_axis1;
  b.axis0 = p.axis0.data();
  b.axis1 = p.axis1.data();
  b.axis0_sin = p.s0.empty() ? nullptr : p.s0.data();
  b.axis0_cos = p.c0.empty() ? nullptr : p.c0.data();
  b.apex_x = grid.apex.x;
  b.apex_z = grid.apex.z;
  b.f_number = cfg.f_number;
  b.rx_window = static_cast<int>(cfg.rx_window);
  b.tukey_alpha = cfg.tukey_alpha;
  b.combiner = static_cast<int>(cfg.combiner);
  b.compounding = static_cast<int>(cfg.compounding);
  b.transmit_weighting = cfg.transmit_weighting ? 1 : 0;
  b.coherence_factor = cfg.coherence_factor ? 1 : 0;
  b.mv_subarray = cfg.mv_subarray;
  b.mv_diagonal_loading = cfg.mv_diagonal_loading;
  b.mv_range_average = cfg.mv_range_average;
  b.mv_precision = static_cast<int>(cfg.mv_precision);
  b.mv_refine = cfg.mv_refine ? 1 : 0;
  b.tile_slots = cfg.tile_slots;
  b.tile_mv_occupancy = cfg.tile_mv_occupancy;
  return p;
}

BeamformPlan make_grid_plan(const ScanGrid& grid) {
  BeamformPlan p;
  p.axis0 = grid.axis0;
  p.axis1 = grid.axis1;
  if (grid.type == GridType::Sector) {
    p.s0.resize(grid.n_axis0);
    p.c0.resize(grid.n_axis0);
    for (int i = 0; i < grid.n_axis0; ++i) {
      p.s0[i] = std::sin(grid.axis0[i]);
      p.c0[i] = std::cos(grid.axis0[i]);
    }
  }
  usx_bf_desc& b = p.desc;
  b.grid_type = static_cast<int>(grid.type);
  b.n_axis0 = grid.n_axis0;
  b.n_axis1 = grid.n_axis1;
  b.axis0 = p.axis0.data();
  b.axis1 = p.axis1.data();
  b.axis0_sin = p.s0.empty() ? nullptr : p.s0.data();
  b.axis0_cos = p.c0.empty() ? nullptr : p.c0.data();
  b.apex_x = grid.apex.x;
  b.apex_z = grid.apex.z;
  return p;
}

// TGC frequency: the frame's demodulation frequency, else 5 MHz (pipeline.cpp:52-55).
usx_display_desc make_display_desc(const PipelineConfig& cfg, const Medium& medium, real fdem) {
  usx_display_desc d{};
  d.apply_tgc = cfg.apply_tgc ? 1 : 0;
  d.attenuation_db_cm_mhz = medium.attenuation;
  d.center_frequency_hz = fdem > 0.0f ? fdem : 5.0e6f;
  d.dynamic_range_db = cfg.dynamic_range_db;
  d.gain_db = cfg.gain_db;
  d.reference_smoothing = cfg.reference_smoothing;
  return d;
}

ChannelData demodulate(const ChannelData& rf, DemodConfig config) {
  rf.validate();
  if (rf.is_iq) return rf;
  if (rf.rf.empty()) throw std::invalid_argument("demodulate: input has no RF samples");
  const DemodPlan plan = make_demod_plan(rf, config);
  ChannelData out;
  out.probe = rf.probe;
  out.medium = rf.medium;
  out.transmits = rf.transmits;
  out.acq = rf.acq;
  out.acq.sampling_frequency = plan.fs_out;
  out.acq.n_samples = plan.n_out;
  out.is_iq = true;
  out.demodulation_frequency = plan.fdem;
  out.allocate();
  usx_ctx* ctx = thread_context().get();
  check(ctx, usx_demodulate(ctx, &plan.desc, rf.rf.data(), reinterpret_cast<usx_c32*>(out.iq.data()),
                            USX_HOST),
        "demodulate");
  return out;
}

}  // namespace usx
