// simulator.cpp -- point-scatterer forward model (host RF generator).
//
// Follows the reference's core/src/simulator.cpp:21-419: phantom factories,
// Gaussian two-way pulse, transmit factories, and the two-stage model
// (stage 1: the diffracted incident waveform at each scatterer, summed over
// firing elements and trimmed at incident_trim_db; stage 2: its delayed,
// spread, direction-weighted and attenuated copy at every receive element),
// then per-event receive noise. The float expressions are the reference's.
//
// Parallel structure (this file's own): stage 1 runs in parallel over
// scatterers into per-scatterer buffers; stage 2 runs in parallel over
// receive channels, each channel visiting the scatterers in index order. So
// every RF sample accumulates its contributions in scatterer order whatever
// the thread count -- the reference's single-threaded result, bit for bit
// -- where the reference's own parallel reduction (simulator.cpp:397-402)
// adds per-thread partial sums in arbitrary order.
#include "usx/simulator.hpp"

#include <algorithm>
#include <cmath>
#include <random>
#include <stdexcept>

#include "usx/geometry.hpp"
#include "gpu.hpp"

#ifdef _OPENMP
#include <omp.h>
#endif

namespace usx {
namespace {

real gaussian_sigma(real f0, real bandwidth) {
  const real ln2 = 0.69314718f;
  return std::sqrt(2.0f * ln2) / (kPi * bandwidth * f0);
}

real db_to_amp(real db) { return std::pow(10.0f, db / 20.0f); }

}  // namespace

Phantom Phantom::point_targets(const std::vector<Vec3>& positions, real amplitude) {
  Phantom p;
  for (const Vec3& v : positions) {
    Scatterer s;
    s.position = v;
    s.amplitude = amplitude;
    p.scatterers.push_back(s);
  }
  return p;
}

Phantom Phantom::speckle_box(real x_min, real x_max, real z_min, real z_max, real y_half,
                             int n, std::uint32_t seed, real sigma) {
  Phantom p;
  p.scatterers.reserve(static_cast<std::size_t>(std::max(0, n)));
  std::mt19937 rng(seed);
  std::uniform_real_distribution<real> ux(x_min, x_max);
  std::uniform_real_distribution<real> uz(z_min, z_max);
  std::uniform_real_distribution<real> uy(-y_half, y_half);
  std::normal_distribution<real> amp(0.0f, sigma);
  for (int i = 0; i < n; ++i) {
    // The reference draws Vec3(ux, uy, uz) as constructor arguments, which
    // GCC on x86-64 evaluates last-to-first: z, then y, then x.
    const real z = uz(rng);
    const real y = uy(rng);
    const real x = ux(rng);
    Scatterer s;
    s.position = Vec3(x, y, z);
    s.amplitude = amp(rng);
    p.scatterers.push_back(s);
  }
  return p;
}

int Phantom::recommended_scatterer_count(const Probe& probe, const Medium& medium, real x_span,
                                         real z_span, real y_span, real f_number,
                                         real per_cell) {
  const real lambda = probe.wavelength(medium.speed_of_sound);
  const real dx = std::max(f_number * lambda, 1e-5f);
  const real dz = std::max(lambda / std::max(probe.bandwidth, 0.1f), 1e-5f);
  const real dy = std::max(probe.elevation_height, 1e-4f);
  const real cells = (x_span / dx) * (z_span / dz) * std::max(y_span / dy, 1.0f);
  const double n = static_cast<double>(cells) * per_cell;
  return static_cast<int>(std::min(n, 4.0e6));
}

void Phantom::scale_amplitude_in_circle(const Vec3& center, real radius, real scale) {
  const real r2 = radius * radius;
  for (Scatterer& s : scatterers) {
    const real dx = s.position.x - center.x;
    const real dz = s.position.z - center.z;
    if (dx * dx + dz * dz <= r2) s.amplitude *= scale;
  }
}

void Phantom::add_flow_tube(const Vec3& center, real radius, real length, real axis_angle,
                            real peak_velocity, int n, std::uint32_t seed, real amplitude) {
  std::mt19937 rng(seed);
  std::uniform_real_distribution<real> along(-0.5f * length, 0.5f * length);
  std::uniform_real_distribution<real> u01(0.0f, 1.0f);
  std::normal_distribution<real> amp(0.0f, amplitude);
  const real ax = std::cos(axis_angle), az = std::sin(axis_angle);
  const real nx = -std::sin(axis_angle), nz = std::cos(axis_angle);
  for (int i = 0; i < n; ++i) {
    // Uniform over the disc cross-section: radius sqrt(u) R, angle 2 pi u.
    const real rr = radius * std::sqrt(u01(rng));
    const real phi = 2.0f * kPi * u01(rng);
    const real r_off = rr * std::cos(phi);
    const real y_off = rr * std::sin(phi);
    const real s_along = along(rng);
    Scatterer s;
    s.position = Vec3(center.x + ax * s_along + nx * r_off, y_off, center.z + az * s_along + nz * r_off);
    s.amplitude = amp(rng);
    const real v = peak_velocity * (1.0f - (rr / radius) * (rr / radius));
    s.velocity = Vec3(ax * v, 0.0f, az * v);
    scatterers.push_back(s);
  }
}

Simulator::Simulator(Probe probe, Medium medium, Acquisition acq, SimulatorConfig config)
    : probe_(probe), medium_(medium), acq_(acq), config_(config) {
  probe_.validate();
  medium_.validate();
  acq_.validate();
  if (config_.oversampling < 1) config_.oversampling = 1;
}

std::vector<real> Simulator::pulse(real* half_duration) const {
  const real f0 = probe_.center_frequency;
  const real sigma = gaussian_sigma(f0, probe_.bandwidth);
  const real fs_os = acq_.sampling_frequency * config_.oversampling;
  const int n_half = static_cast<int>(std::ceil(3.0f * sigma * fs_os));
  std::vector<real> p(static_cast<std::size_t>(2 * n_half + 1));
  for (int i = -n_half; i <= n_half; ++i) {
    const real t = static_cast<real>(i) / fs_os;
    p[i + n_half] = std::exp(-0.5f * (t * t) / (sigma * sigma)) * std::cos(2.0f * kPi * f0 * t);
  }
  if (half_duration) *half_duration = static_cast<real>(n_half) / fs_os;
  return p;
}

std::vector<Transmit> Simulator::make_plane_waves(const std::vector<real>& angles) const {
  std::vector<Transmit> txs;
  for (real a : angles) {
    Transmit tx;
    tx.type = TransmitType::PlaneWave;
    tx.angle = a;
    tx.origin = Vec3(0, 0, 0);
    compute_transmit_delays(probe_, medium_, tx, config_.tx_window, config_.tx_f_number, nullptr);
    txs.push_back(std::move(tx));
  }
  return txs;
}

std::vector<Transmit> Simulator::make_focused_scan(int n_lines, real x_min, real x_max,
                                                   real focal_depth, real angle) const {
  std::vector<Transmit> txs;
  for (int i = 0; i < n_lines; ++i) {
    const real x = (n_lines == 1) ? x_min : x_min + (x_max - x_min) * i / (n_lines - 1);
    Transmit tx;
    tx.type = TransmitType::Focused;
    tx.angle = angle;
    tx.origin = Vec3(x, 0, 0);
    tx.focus = Vec3(x + focal_depth * std::tan(angle), 0, focal_depth);
    compute_transmit_delays(probe_, medium_, tx, config_.tx_window, config_.tx_f_number, nullptr);
    txs.push_back(std::move(tx));
  }
  return txs;
}

std::vector<Transmit> Simulator::make_diverging_waves(const std::vector<real>& angles,
                                                      real depth) const {
  std::vector<Transmit> txs;
  for (real a : angles) {
    Transmit tx;
    tx.type = TransmitType::Diverging;
    tx.angle = a;
    tx.origin = Vec3(0, 0, 0);
    tx.focus = Vec3(-depth * std::tan(a), 0, -depth);
    compute_transmit_delays(probe_, medium_, tx, config_.tx_window, config_.tx_f_number, nullptr);
    txs.push_back(std::move(tx));
  }
  return txs;
}

std::vector<Transmit> Simulator::repeat(const std::vector<Transmit>& txs, int n_ensemble) {
  std::vector<Transmit> out;
  for (int e = 0; e < n_ensemble; ++e) out.insert(out.end(), txs.begin(), txs.end());
  return out;
}

namespace {

// Stage-1 result for one scatterer: the trimmed incident waveform.
struct Incident {
  bool live = false;
  Vec3 pos;
  real amplitude = 0.0f;
  real t0 = 0.0f;           // time of its first kept sample
  std::size_t off = 0;      // offset into the owning thread's sample pool
  int n = 0;                // kept samples
  int pool = 0;             // owning thread
};

}  // namespace

ChannelData Simulator::simulate(const std::vector<Transmit>& transmits,
                                const Phantom& phantom) const {
  ChannelData cd;
  cd.probe = probe_;
  cd.medium = medium_;
  cd.acq = acq_;
  cd.transmits = transmits;
  cd.is_iq = false;
  cd.allocate();

  const int n_ch = probe_.n_elements;
  const int n_smp = acq_.n_samples;
  const real c = medium_.speed_of_sound;
  const real fs_os = acq_.sampling_frequency * config_.oversampling;
  const real lambda = probe_.wavelength(c);
  real pulse_half = 0.0f;
  const std::vector<real> pw = pulse(&pulse_half);
  const int pulse_n = static_cast<int>(pw.size());
  const real f0_mhz = probe_.center_frequency * 1e-6f;
  const real atten = config_.apply_attenuation ? medium_.attenuation * f0_mhz * 100.0f : 0.0f;
  const real max_window = config_.max_incident_window_us * 1e-6f;
  const real trim_amp = db_to_amp(config_.incident_trim_db);
  std::vector<Vec3> elem(n_ch);
  for (int i = 0; i < n_ch; ++i) elem[i] = probe_.element_position(i);

  // Amplitude of the path between a scatterer and element pe at range r.
  auto path_amp = [&](real a, const Vec3& pos, const Vec3& pe, real r) {
    if (config_.apply_spreading) a /= r;
    if (config_.apply_directivity)
      a *= element_directivity(probe_.element_width, lambda, (pos.x - pe.x) / r);
    if (atten > 0.0f) a *= db_to_amp(-atten * r);
    return a;
  };

  const std::ptrdiff_t n_scat = static_cast<std::ptrdiff_t>(phantom.scatterers.size());
  std::vector<Incident> inc(static_cast<std::size_t>(n_scat));
  int n_threads = 1;
#ifdef _OPENMP
  n_threads = omp_get_max_threads();
#endif
  std::vector<std::vector<real>> pools(static_cast<std::size_t>(n_threads));

  for (int ev = 0; ev < static_cast<int>(transmits.size()); ++ev) {
    const Transmit& tx = transmits[ev];
    const real event_time = static_cast<real>(ev) / acq_.prf;
    real* out_event = cd.rf.data() + cd.index(ev, 0, 0);
    std::vector<int> firing;
    for (int i = 0; i < n_ch; ++i)
      if (!tx.apodization.empty() && tx.apodization[i] != 0.0f) firing.push_back(i);
    if (firing.empty())
      for (int i = 0; i < n_ch; ++i) firing.push_back(i);
    for (auto& p : pools) p.clear();

    // ---- stage 1: incident waveform at every scatterer (parallel) --------
#pragma omp parallel num_threads(n_threads)
    {
      int tid = 0;
#ifdef _OPENMP
      tid = omp_get_thread_num();
#endif
      std::vector<real> buf;
      std::vector<real>& pool = pools[static_cast<std::size_t>(tid)];
#pragma omp for schedule(static)
      for (std::ptrdiff_t si = 0; si < n_scat; ++si) {
        Incident& I = inc[static_cast<std::size_t>(si)];
        I.live = false;
        const Scatterer& sc = phantom.scatterers[static_cast<std::size_t>(si)];
        if (sc.amplitude == 0.0f) continue;
        Vec3 pos = sc.position;
        if (config_.enable_motion) pos = pos + sc.velocity * event_time;
        if (pos.z <= 0.0f) continue;

        real t_min = 0, t_max = 0;
        bool first = true;
        for (int i : firing) {
          const real t = tx.delays[i] + distance(pos, elem[i]) / c;
          t_min = first ? t : std::min(t_min, t);
          t_max = first ? t : std::max(t_max, t);
          first = false;
        }
        real span = (t_max - t_min) + 2.0f * pulse_half;
        span = std::min(span, max_window + 2.0f * pulse_half);
        const real inc_t0 = t_min - pulse_half;
        const int inc_n = static_cast<int>(std::ceil(span * fs_os)) + 2;
        buf.assign(static_cast<std::size_t>(inc_n), 0.0f);
        for (int i : firing) {
          const Vec3& pe = elem[i];
          const real r = distance(pos, pe);
          if (r < 1e-6f) continue;
          const real amp = path_amp(tx.apodization.empty() ? 1.0f : tx.apodization[i], pos, pe, r);
          if (amp == 0.0f) continue;
          const real centre = (tx.delays[i] + r / c - inc_t0) * fs_os;
          const int base = static_cast<int>(std::floor(centre)) - pulse_n / 2;
          const real frac = centre - std::floor(centre);
          for (int k = 0; k < pulse_n; ++k) {
            const int idx = base + k;
            if (idx < 0 || idx + 1 >= inc_n) continue;
            const real v = amp * pw[k];
            buf[idx] += v * (1.0f - frac);
            buf[idx + 1] += v * frac;
          }
          if (!config_.full_transmit_diffraction) break;
        }
        real peak = 0.0f;
        for (real v : buf) peak = std::max(peak, std::abs(v));
        if (peak <= 0.0f) continue;
        const real thr = peak * trim_amp;
        int lo = 0, hi = inc_n - 1;
        while (lo < inc_n && std::abs(buf[lo]) < thr) ++lo;
        while (hi > lo && std::abs(buf[hi]) < thr) --hi;
        if (lo >= hi) continue;
        I.live = true;
        I.pos = pos;
        I.amplitude = sc.amplitude;
        I.t0 = inc_t0 + static_cast<real>(lo) / fs_os;
        I.n = hi - lo + 1;
        I.pool = tid;
        I.off = pool.size();
        pool.insert(pool.end(), buf.begin() + lo, buf.begin() + hi + 1);
      }
    }

    // ---- stage 2: echoes at every receive element (parallel over channels,
    // scatterers in index order within each channel) ----------------------
#pragma omp parallel for schedule(dynamic, 1) num_threads(n_threads)
    for (int j = 0; j < n_ch; ++j) {
      const Vec3& pe = elem[j];
      real* dst = out_event + static_cast<std::size_t>(j) * n_smp;
      for (std::ptrdiff_t si = 0; si < n_scat; ++si) {
        const Incident& I = inc[static_cast<std::size_t>(si)];
        if (!I.live) continue;
        const real r = distance(I.pos, pe);
        if (r < 1e-6f) continue;
        const real amp = path_amp(I.amplitude, I.pos, pe, r);
        if (amp == 0.0f) continue;
        const real delay = r / c;
        const real t_start = I.t0 + delay;
        const real t_end = t_start + static_cast<real>(I.n - 1) / fs_os;
        const int n0 = std::max(static_cast<int>(std::ceil(acq_.time_to_sample(t_start))), 0);
        const int n1 = std::min(static_cast<int>(std::floor(acq_.time_to_sample(t_end))), n_smp - 1);
        const real* w = pools[static_cast<std::size_t>(I.pool)].data() + I.off;
        for (int n = n0; n <= n1; ++n) {
          const real t = acq_.sample_time(n) - delay;
          const real u = (t - I.t0) * fs_os;
          const int k = static_cast<int>(u);
          if (k < 0 || k + 1 >= I.n) continue;
          const real f = u - static_cast<real>(k);
          const real a = w[k], b = w[k + 1];
          dst[n] += amp * (a + (b - a) * f);
        }
      }
    }

    // ---- receive noise ----------------------------------------------------
    if (config_.noise_db > -200.0f) {
      const std::size_t n_total = static_cast<std::size_t>(n_ch) * n_smp;
      real peak = 0.0f;
      for (std::size_t k = 0; k < n_total; ++k) peak = std::max(peak, std::abs(out_event[k]));
      if (peak > 0.0f) {
        std::mt19937 rng(config_.noise_seed + static_cast<std::uint32_t>(ev));
        std::normal_distribution<real> nd(0.0f, peak * db_to_amp(config_.noise_db));
        for (std::size_t k = 0; k < n_total; ++k) out_event[k] += nd(rng);
      }
    }
  }
  return cd;
}

ChannelData Simulator::simulate_gpu(const std::vector<Transmit>& transmits,
                                    const Phantom& phantom) const {
  ChannelData cd;
  cd.probe = probe_;
  cd.medium = medium_;
  cd.acq = acq_;
  cd.transmits = transmits;
  cd.is_iq = false;
  cd.allocate();
  const int n_ch = probe_.n_elements;
  const int n_smp = acq_.n_samples;
  const int n_ev = static_cast<int>(transmits.size());
  real pulse_half = 0.0f;
  const std::vector<real> pw = pulse(&pulse_half);
  std::vector<real> ex(n_ch), delays(static_cast<std::size_t>(n_ev) * n_ch, 0.0f),
      apod(static_cast<std::size_t>(n_ev) * n_ch, 0.0f);
  std::vector<int> has_apod(std::max(n_ev, 1), 0);
  for (int i = 0; i < n_ch; ++i) ex[i] = probe_.element_position(i).x;
  for (int e = 0; e < n_ev; ++e) {
    const Transmit& tx = transmits[e];
    if (static_cast<int>(tx.delays.size()) != n_ch)
      throw std::invalid_argument("simulate: transmit delays must have one entry per element");
    std::copy(tx.delays.begin(), tx.delays.end(), delays.begin() + static_cast<std::ptrdiff_t>(e) * n_ch);
    if (!tx.apodization.empty()) {
      if (static_cast<int>(tx.apodization.size()) != n_ch)
        throw std::invalid_argument("simulate: transmit apodization must have one entry per element");
      has_apod[e] = 1;
      std::copy(tx.apodization.begin(), tx.apodization.end(),
                apod.begin() + static_cast<std::ptrdiff_t>(e) * n_ch);
    }
  }
  const real f0_mhz = probe_.center_frequency * 1e-6f;
  const bool noise = config_.noise_db > -200.0f;
  const std::size_t n_tot = static_cast<std::size_t>(n_ch) * n_smp;
  if (noise && (!noise_ || noise_->seed != config_.noise_seed || noise_->n_events < n_ev ||
                noise_->n_per_event != n_tot)) {
    auto t = std::make_shared<NoiseTable>();
    t->seed = config_.noise_seed;
    t->n_events = n_ev;
    t->n_per_event = n_tot;
    t->z.resize(n_tot * static_cast<std::size_t>(n_ev));
    for (int e = 0; e < n_ev; ++e) {
      std::mt19937 rng(config_.noise_seed + static_cast<std::uint32_t>(e));
      std::normal_distribution<real> nd(0.0f, 1.0f);
      real* z = t->z.data() + n_tot * static_cast<std::size_t>(e);
      for (std::size_t k = 0; k < n_tot; ++k) z[k] = nd(rng);
    }
    noise_ = t;
  }
  std::vector<real> sc(7 * phantom.scatterers.size());
  for (std::size_t i = 0; i < phantom.scatterers.size(); ++i) {
    const Scatterer& p = phantom.scatterers[i];
    real* d = sc.data() + 7 * i;
    d[0] = p.position.x;
    d[1] = p.position.y;
    d[2] = p.position.z;
    d[3] = p.amplitude;
    d[4] = p.velocity.x;
    d[5] = p.velocity.y;
    d[6] = p.velocity.z;
  }
  usx_sim_desc d{};
  d.n_elements = n_ch;
  d.n_samples = n_smp;
  d.n_events = n_ev;
  d.sampling_frequency = acq_.sampling_frequency;
  d.t0 = acq_.t0;
  d.prf = acq_.prf;
  d.speed_of_sound = medium_.speed_of_sound;
  d.wavelength = probe_.wavelength(medium_.speed_of_sound);
  d.element_width = probe_.element_width;
  d.oversampling = config_.oversampling;
  d.element_x = ex.data();
  d.delays = delays.data();
  d.apodization = apod.data();
  d.has_apodization = has_apod.data();
  d.pulse_n = static_cast<int>(pw.size());
  d.pulse = pw.data();
  d.pulse_half = pulse_half;
  d.attenuation = config_.apply_attenuation ? medium_.attenuation * f0_mhz * 100.0f : 0.0f;
  d.max_window = config_.max_incident_window_us * 1e-6f;
  d.trim_amp = db_to_amp(config_.incident_trim_db);
  d.spreading = config_.apply_spreading;
  d.directivity = config_.apply_directivity;
  d.motion = config_.enable_motion;
  d.full_diffraction = config_.full_transmit_diffraction;
  d.noise_amp = noise ? db_to_amp(config_.noise_db) : 0.0f;
  d.noise = noise ? noise_->z.data() : nullptr;
  usx_ctx* ctx = thread_context().get();
  check(ctx, usx_simulate(ctx, &d, static_cast<int>(phantom.scatterers.size()), sc.data(), cd.rf.data(),
                          USX_HOST),
        "simulate_gpu");
  return cd;
}

}  // namespace usx
