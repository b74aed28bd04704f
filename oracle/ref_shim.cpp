// ref_shim.cpp -- C entry points over the reference's own compiled core.
//
// TEST INFRASTRUCTURE ONLY (see oracle_abi.h). oracle/Makefile compiles this
// file together with the reference's unmodified core/src/*.cpp, read in place
// from /root/reference, into oracle/_ref/libusx_ref.so. The shim only
// marshals the plain-C descriptors into the reference's own types and calls
// its public API; no reference source is copied here.
#include <chrono>
#include <complex>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "usx/beamformer.hpp"
#include "usx/demodulate.hpp"
#include "usx/doppler.hpp"
#include "usx/geometry.hpp"
#include "usx/pipeline.hpp"
#include "usx/simulator.hpp"
#include "usx/types.hpp"

#include "oracle_abi.h"

namespace {

thread_local std::string g_err;

usx::Probe to_probe(const oc_setup* s) {
  usx::Probe p;
  p.n_elements = s->n_elements;
  p.pitch = s->pitch;
  p.element_width = s->element_width;
  p.elevation_height = s->elevation_height;
  p.center_frequency = s->center_frequency;
  p.bandwidth = s->bandwidth;
  return p;
}
usx::Medium to_medium(const oc_setup* s) {
  usx::Medium m;
  m.speed_of_sound = s->speed_of_sound;
  m.attenuation = s->attenuation;
  m.density = s->density;
  return m;
}
usx::Acquisition to_acq(const oc_setup* s) {
  usx::Acquisition a;
  a.sampling_frequency = s->sampling_frequency;
  a.n_samples = s->n_samples;
  a.t0 = s->t0;
  a.prf = s->prf;
  return a;
}
usx::Transmit to_tx(const oc_transmit& t) {
  usx::Transmit tx;
  tx.type = static_cast<usx::TransmitType>(t.type);
  tx.angle = t.angle;
  tx.focus = usx::Vec3(t.focus[0], t.focus[1], t.focus[2]);
  tx.origin = usx::Vec3(t.origin[0], t.origin[1], t.origin[2]);
  if (t.n_delays > 0) tx.delays.assign(t.delays, t.delays + t.n_delays);
  if (t.n_apod > 0) tx.apodization.assign(t.apodization, t.apodization + t.n_apod);
  return tx;
}
std::vector<usx::Transmit> to_txs(const oc_transmit* txs, int n) {
  std::vector<usx::Transmit> v;
  for (int i = 0; i < n; ++i) v.push_back(to_tx(txs[i]));
  return v;
}
usx::ScanGrid full_grid(const oc_grid* g) {
  if (g->type == 0)
    return usx::ScanGrid::cartesian(g->a0_min, g->a0_max, g->n0, g->a1_min, g->a1_max, g->n1);
  return usx::ScanGrid::sector(g->a0_min, g->a0_max, g->n0, g->a1_min, g->a1_max, g->n1,
                               usx::Vec3(g->apex[0], g->apex[1], g->apex[2]));
}
// A line subset keeps the full grid's axis values (the reference's ScanGrid is
// an explicit point list, types.hpp:228-231, so any subset is a valid grid).
usx::ScanGrid to_grid(const oc_grid* g) {
  usx::ScanGrid G = full_grid(g);
  if (g->line_step <= 0) return G;
  if (g->line_first < 0 || g->line_first >= G.n_axis0)
    throw std::invalid_argument("oc_grid: line_first outside the grid");
  std::vector<float> a0;
  for (int i = g->line_first; i < G.n_axis0; i += g->line_step) a0.push_back(G.axis0[i]);
  G.axis0 = a0;
  G.n_axis0 = static_cast<int>(a0.size());
  G.build_points();
  G.validate();
  return G;
}
usx::DemodConfig to_demod(const oc_demod* d) {
  usx::DemodConfig c;
  c.demodulation_frequency = d->demodulation_frequency;
  c.cutoff_fraction = d->cutoff_fraction;
  c.filter_taps = d->filter_taps;
  c.decimation = d->decimation;
  return c;
}
usx::BeamformerConfig to_bf(const oc_bf* b) {
  usx::BeamformerConfig c;
  c.f_number = b->f_number;
  c.rx_window = static_cast<usx::Window>(b->rx_window);
  c.tukey_alpha = b->tukey_alpha;
  c.combiner = static_cast<usx::Combiner>(b->combiner);
  c.compounding = static_cast<usx::Compounding>(b->compounding);
  c.transmit_weighting = b->transmit_weighting != 0;
  c.coherence_factor = b->coherence_factor != 0;
  c.mv_subarray = b->mv_subarray;
  c.mv_diagonal_loading = b->mv_diagonal_loading;
  c.mv_range_average = b->mv_range_average;
  return c;
}
usx::SimulatorConfig to_sim(const oc_simcfg* s) {
  usx::SimulatorConfig c;
  c.pulse_cycles = s->pulse_cycles;
  c.tx_window = static_cast<usx::Window>(s->tx_window);
  c.tx_f_number = s->tx_f_number;
  c.oversampling = s->oversampling;
  c.full_transmit_diffraction = s->full_transmit_diffraction != 0;
  c.incident_trim_db = s->incident_trim_db;
  c.max_incident_window_us = s->max_incident_window_us;
  c.apply_attenuation = s->apply_attenuation != 0;
  c.apply_directivity = s->apply_directivity != 0;
  c.apply_spreading = s->apply_spreading != 0;
  c.noise_db = s->noise_db;
  c.noise_seed = s->noise_seed;
  c.enable_motion = s->enable_motion != 0;
  return c;
}
usx::ChannelData make_data(const oc_setup* s, const oc_transmit* txs, int n_tx, int is_iq,
                           const float* data, float fdem) {
  usx::ChannelData cd;
  cd.probe = to_probe(s);
  cd.medium = to_medium(s);
  cd.acq = to_acq(s);
  cd.transmits = to_txs(txs, n_tx);
  cd.is_iq = is_iq != 0;
  cd.demodulation_frequency = fdem;
  cd.allocate();
  const std::size_t n = cd.size();
  if (is_iq)
    std::memcpy(static_cast<void*>(cd.iq.data()), data, n * sizeof(std::complex<float>));
  else
    std::memcpy(cd.rf.data(), data, n * sizeof(float));
  return cd;
}

// Frames of grid g from interleaved complex data (n_frames after each other).
std::vector<usx::Frame> to_frames(const oc_grid* g, int n_frames, const float* data, float fdem) {
  const usx::ScanGrid grid = to_grid(g);
  const std::size_t n_pix = static_cast<std::size_t>(grid.n_pixels());
  std::vector<usx::Frame> out(static_cast<std::size_t>(n_frames));
  for (int f = 0; f < n_frames; ++f) {
    usx::Frame& fr = out[static_cast<std::size_t>(f)];
    fr.grid = grid;
    fr.demodulation_frequency = fdem;
    fr.allocate();
    std::memcpy(static_cast<void*>(fr.data.data()), data + 2 * n_pix * static_cast<std::size_t>(f), n_pix * 8);
  }
  return out;
}
usx::WallFilter to_wall(int t) {
  return t == 1 ? usx::WallFilter::PolynomialRegression
                : (t == 2 ? usx::WallFilter::MeanSubtraction : usx::WallFilter::None);
}

template <typename F>
int guarded(F&& f) {
  try {
    f();
    return 0;
  } catch (const std::invalid_argument& e) {
    g_err = e.what();
    return 1;
  } catch (const std::exception& e) {
    g_err = e.what();
    return 2;
  }
}

}  // namespace

extern "C" {

const char* oc_last_error(void) { return g_err.c_str(); }
const char* oc_flavour(void) { return "reference"; }

int oc_grid_build(const oc_grid* g, float* axis0, float* axis1, float* points) {
  return guarded([&] {
    const usx::ScanGrid grid = to_grid(g);
    std::memcpy(axis0, grid.axis0.data(), grid.axis0.size() * sizeof(float));
    std::memcpy(axis1, grid.axis1.data(), grid.axis1.size() * sizeof(float));
    for (std::size_t k = 0; k < grid.points.size(); ++k) {
      points[3 * k + 0] = grid.points[k].x;
      points[3 * k + 1] = grid.points[k].y;
      points[3 * k + 2] = grid.points[k].z;
    }
  });
}

int oc_design_lowpass(int n_taps, float fc, float* taps_out, int* n_out) {
  return guarded([&] {
    const std::vector<float> h = usx::design_lowpass(n_taps, fc);
    if (taps_out) std::memcpy(taps_out, h.data(), h.size() * sizeof(float));
    *n_out = static_cast<int>(h.size());
  });
}

float oc_window_value(int window, float u, float tukey_alpha) {
  return usx::window_value(static_cast<usx::Window>(window), u, tukey_alpha);
}

int oc_make_transmits(const oc_setup* s, const oc_simcfg* sc, int kind, int n,
                      const float* params, float extra, float* delays_out, float* apod_out,
                      float* focus_out, float* origin_out, float* angle_out) {
  return guarded([&] {
    usx::Simulator sim(to_probe(s), to_medium(s), to_acq(s), to_sim(sc));
    std::vector<usx::Transmit> txs;
    if (kind == 1) {
      txs = sim.make_plane_waves(std::vector<float>(params, params + n));
    } else if (kind == 0) {
      txs = sim.make_focused_scan(n, params[0], params[1], params[2], params[3]);
    } else {
      txs = sim.make_diverging_waves(std::vector<float>(params, params + n), extra);
    }
    const int ne = s->n_elements;
    for (std::size_t i = 0; i < txs.size(); ++i) {
      std::memcpy(delays_out + i * ne, txs[i].delays.data(), ne * sizeof(float));
      std::memcpy(apod_out + i * ne, txs[i].apodization.data(), ne * sizeof(float));
      focus_out[3 * i + 0] = txs[i].focus.x;
      focus_out[3 * i + 1] = txs[i].focus.y;
      focus_out[3 * i + 2] = txs[i].focus.z;
      origin_out[3 * i + 0] = txs[i].origin.x;
      origin_out[3 * i + 1] = txs[i].origin.y;
      origin_out[3 * i + 2] = txs[i].origin.z;
      angle_out[i] = txs[i].angle;
    }
  });
}

float oc_transmit_arrival_time(const oc_transmit* tx, const float* p, float c) {
  return usx::transmit_arrival_time(to_tx(*tx), usx::Vec3(p[0], p[1], p[2]), c);
}

float oc_transmit_sensitivity(const oc_setup* s, const oc_transmit* tx, const float* p,
                              float wavelength) {
  return usx::transmit_sensitivity(to_probe(s), to_tx(*tx), usx::Vec3(p[0], p[1], p[2]),
                                   wavelength);
}

int oc_speckle_box(float x_min, float x_max, float z_min, float z_max, float y_half, int n,
                   uint32_t seed, float sigma, float* pos_out, float* amp_out) {
  return guarded([&] {
    const usx::Phantom ph =
        usx::Phantom::speckle_box(x_min, x_max, z_min, z_max, y_half, n, seed, sigma);
    for (std::size_t i = 0; i < ph.scatterers.size(); ++i) {
      pos_out[3 * i + 0] = ph.scatterers[i].position.x;
      pos_out[3 * i + 1] = ph.scatterers[i].position.y;
      pos_out[3 * i + 2] = ph.scatterers[i].position.z;
      amp_out[i] = ph.scatterers[i].amplitude;
    }
  });
}

int oc_simulate(const oc_setup* s, const oc_simcfg* sc, const oc_transmit* txs, int n_tx,
                int n_scat, const float* pos, const float* amp, const float* vel,
                float* rf_out) {
  return guarded([&] {
    usx::Simulator sim(to_probe(s), to_medium(s), to_acq(s), to_sim(sc));
    usx::Phantom ph;
    for (int i = 0; i < n_scat; ++i) {
      usx::Scatterer sct;
      sct.position = usx::Vec3(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]);
      sct.amplitude = amp[i];
      if (vel) sct.velocity = usx::Vec3(vel[3 * i], vel[3 * i + 1], vel[3 * i + 2]);
      ph.add(sct);
    }
    const usx::ChannelData cd = sim.simulate(to_txs(txs, n_tx), ph);
    std::memcpy(rf_out, cd.rf.data(), cd.rf.size() * sizeof(float));
  });
}

int oc_demodulate(const oc_setup* s, const oc_transmit* txs, int n_tx, const float* rf,
                  const oc_demod* d, float* iq_out, int* n_out, float* fs_out,
                  float* fdem_out) {
  return guarded([&] {
    if (!iq_out) {
      const int dec = d->decimation < 1 ? 1 : d->decimation;
      *n_out = (s->n_samples + dec - 1) / dec;
      return;
    }
    const usx::ChannelData cd = make_data(s, txs, n_tx, 0, rf, 0.0f);
    const usx::ChannelData iq = usx::demodulate(cd, to_demod(d));
    std::memcpy(static_cast<void*>(iq_out), static_cast<const void*>(iq.iq.data()), iq.iq.size() * sizeof(std::complex<float>));
    *n_out = iq.acq.n_samples;
    *fs_out = iq.acq.sampling_frequency;
    *fdem_out = iq.demodulation_frequency;
  });
}

// Minimum variance with a negative range average makes the reference size a
// vector of 2 r + 1 entries inside its OpenMP region (beamformer.cpp:357),
// which ends the process; refuse it here instead of calling in.
static void refuse_unrunnable(const oc_bf* bf) {
  if (bf->combiner == 2 && bf->mv_range_average < 0)
    throw std::invalid_argument("BeamformerConfig.mv_range_average must be >= 0");
}

// The constructor's check (beamformer.hpp:110).
int oc_validate_bf(const oc_bf* bf) {
  return guarded([&] { const usx::Beamformer b(to_bf(bf)); (void)b; });
}

int oc_beamform(const oc_setup* s, const oc_transmit* txs, int n_tx, int is_iq,
                const float* data, float fdem, const oc_grid* g, const oc_bf* bf,
                int event_begin, int event_end, float* frame_out) {
  return guarded([&] {
    const usx::ChannelData cd = make_data(s, txs, n_tx, is_iq, data, fdem);
    // Beamformer.config assigned after construction (pymodule.cpp:430-431):
    // validated only by the constructor, so an out-of-range field reaches
    // beamform and its clamps.
    refuse_unrunnable(bf);
    usx::Beamformer b;
    b.config() = to_bf(bf);
    const usx::Frame f = b.beamform(cd, to_grid(g), event_begin, event_end);
    std::memcpy(static_cast<void*>(frame_out), static_cast<const void*>(f.data.data()), f.data.size() * sizeof(std::complex<float>));
  });
}

int oc_process_frame(const oc_setup* s, const oc_transmit* txs, int n_tx, const float* rf,
                     const oc_grid* g, const oc_demod* d, const oc_bf* bf, float* frame_out,
                     double* stats_out) {
  return guarded([&] {
    const usx::ChannelData cd = make_data(s, txs, n_tx, 0, rf, 0.0f);
    usx::PipelineConfig pc;
    pc.demod = to_demod(d);
    pc.beamformer = to_bf(bf);
    // ImagingPipeline.config assigned after construction (pymodule.cpp:617-619).
    refuse_unrunnable(bf);
    usx::ImagingPipeline pipe(to_grid(g), usx::PipelineConfig{});
    pipe.config() = pc;
    const usx::Frame f = pipe.process_frame(cd);
    std::memcpy(static_cast<void*>(frame_out), static_cast<const void*>(f.data.data()), f.data.size() * sizeof(std::complex<float>));
    if (stats_out) {
      stats_out[0] = pipe.stats().demodulate_ms;
      stats_out[1] = pipe.stats().beamform_ms;
      stats_out[2] = pipe.stats().total_ms;
    }
  });
}

int oc_display(const oc_setup* s, const oc_grid* g, const oc_display_cfg* p, int n_frames,
               const float* frames, float fdem, float* images_out) {
  return guarded([&] {
    usx::PipelineConfig pc;
    pc.apply_tgc = p->apply_tgc != 0;
    pc.dynamic_range_db = p->dynamic_range_db;
    pc.gain_db = p->gain_db;
    pc.reference_smoothing = p->reference_smoothing;
    const usx::ScanGrid grid = to_grid(g);
    usx::ImagingPipeline pipe(grid, pc);
    const usx::Medium med = to_medium(s);
    const std::size_t n = static_cast<std::size_t>(grid.n_axis0) * grid.n_axis1;
    for (int f = 0; f < n_frames; ++f) {
      usx::Frame fr;
      fr.grid = grid;
      fr.demodulation_frequency = fdem;
      fr.data.resize(n);
      std::memcpy(static_cast<void*>(fr.data.data()), frames + 2 * n * f, n * sizeof(std::complex<float>));
      const std::vector<usx::real> img = pipe.to_display(fr, med);
      std::memcpy(images_out + n * f, img.data(), n * sizeof(float));
    }
  });
}

int oc_wall_filter(const oc_grid* g, int n_frames, const float* ensemble, int type, int order,
                   float* out) {
  return guarded([&] {
    const std::vector<usx::Frame> f =
        usx::apply_wall_filter(to_frames(g, n_frames, ensemble, 0.0f), to_wall(type), order);
    for (std::size_t i = 0; i < f.size(); ++i)
      std::memcpy(out + 2 * f[i].data.size() * i, f[i].data.data(), f[i].data.size() * 8);
  });
}

int oc_power_doppler(const oc_grid* g, int n_frames, const float* ensemble, int type, int order,
                     float* power_out) {
  return guarded([&] {
    const std::vector<float> p =
        usx::power_doppler(to_frames(g, n_frames, ensemble, 0.0f), to_wall(type), order);
    std::memcpy(power_out, p.data(), p.size() * sizeof(float));
  });
}

int oc_color_flow(const oc_grid* g, int n_frames, const float* ensemble, float fdem, float prf,
                  float c, const oc_doppler* d, float* velocity, float* power, float* variance,
                  uint8_t* valid, float* nyquist) {
  return guarded([&] {
    usx::DopplerConfig dc;
    dc.wall_filter = to_wall(d->wall_filter);
    dc.polynomial_order = d->polynomial_order;
    dc.loupas = d->loupas != 0;
    dc.kernel_axial = d->kernel_axial;
    dc.kernel_lateral = d->kernel_lateral;
    dc.power_threshold_db = d->power_threshold_db;
    dc.coherence_threshold = d->coherence_threshold;
    const usx::ColorFlowMap m = usx::color_flow(to_frames(g, n_frames, ensemble, fdem), prf, c, dc);
    const std::size_t n = m.velocity.size();
    std::memcpy(velocity, m.velocity.data(), n * sizeof(float));
    std::memcpy(power, m.power.data(), n * sizeof(float));
    std::memcpy(variance, m.variance.data(), n * sizeof(float));
    std::memcpy(valid, m.valid.data(), n);
    *nyquist = m.nyquist_velocity;
  });
}

int oc_axial_displacement(const oc_grid* g, const float* before, const float* after, float fdem,
                          float c, int kernel_axial, int kernel_lateral, int max_lag, float* axial,
                          float* correlation) {
  return guarded([&] {
    const std::vector<usx::Frame> a = to_frames(g, 1, before, fdem);
    const std::vector<usx::Frame> b = to_frames(g, 1, after, fdem);
    const usx::DisplacementMap m =
        usx::estimate_axial_displacement(a[0], b[0], c, kernel_axial, kernel_lateral, max_lag);
    std::memcpy(axial, m.axial.data(), m.axial.size() * sizeof(float));
    std::memcpy(correlation, m.correlation.data(), m.correlation.size() * sizeof(float));
  });
}

int oc_axial_strain(const oc_grid* g, const float* axial, int fit_length, float* strain) {
  return guarded([&] {
    usx::DisplacementMap m;
    m.grid = to_grid(g);
    const std::size_t n = static_cast<std::size_t>(m.grid.n_pixels());
    m.axial.assign(axial, axial + n);
    m.correlation.assign(n, 0.0f);
    const std::vector<float> s = usx::axial_strain(m, fit_length);
    std::memcpy(strain, s.data(), n * sizeof(float));
  });
}

}  // extern "C"
