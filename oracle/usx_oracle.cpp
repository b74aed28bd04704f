// usx_oracle.cpp -- CPU restatement of the reference imaging path.
//
// TEST INFRASTRUCTURE ONLY: the checker for the HIP path. Only tests/,
// __graft_entry__.smoke() and the cpu_baseline leg of bench.py may load the
// library built from this file, and only as the checker. The product never
// links or calls it.
//
// What it restates (reference file:line in each function's comment):
//   ScanGrid factories              core/src/types.cpp:66-115
//   window_value                    core/src/types.cpp:126-148
//   design_lowpass, demodulate      core/src/demodulate.cpp:13-121
//   transmit_arrival_time           core/src/geometry.cpp:82-109
//   dynamic_aperture_weight         core/src/geometry.cpp:123-130
//   tx extent / tx_sensitivity      core/src/beamformer.cpp:238-288
//   beamform (geometry, delay-interpolate-rotate-accumulate, combiners,
//   coherence factor, compounding)  core/src/beamformer.cpp:15-214, 297-549
//   process_frame                   core/src/pipeline.cpp:25-46
//   wall filter, colour flow, power Doppler, axial displacement and strain
//                                   core/src/doppler.cpp:15-307
//
// Two numeric modes:
//   FP32 exact order  -- every float expression is evaluated in the
//     reference's operation order with std::complex<float>, built with the
//     reference's flags (-O3, no -march, no fast-math), so results are
//     bitwise those of the reference on the same inputs;
//   FP64 MV           -- channel data Y identical to the FP32 mode, but the
//     MV covariance, loading, Cholesky, solve and weighting in double. This
//     is the accuracy yardstick for the GPU precision variants.
//
// The loop structure is this file's own (one lateral line at a time, pixel
// outer within the combine); the per-element arithmetic is the reference's,
// because bitwise parity requires it.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

#include "oracle_abi.h"

namespace {

using cf32 = std::complex<float>;
using cf64 = std::complex<double>;
constexpr float kPiF = static_cast<float>(3.14159265358979323846);

thread_local std::string g_err;

struct Invalid : std::invalid_argument {
  using std::invalid_argument::invalid_argument;
};

// ---- geometry helpers ------------------------------------------------------

// Probe::element_position (types.hpp:77-80).
inline float element_x(int i, int n_elements, float pitch) {
  return (static_cast<float>(i) - 0.5f * (n_elements - 1)) * pitch;
}

// window_value (types.cpp:126-148).
float window_value(int w, float u, float tukey_alpha) {
  if (u < -1.0f || u > 1.0f) return 0.0f;
  const float t = 0.5f * (u + 1.0f);
  switch (w) {
    case 0:
      return 1.0f;
    case 1:
      return 0.5f * (1.0f - std::cos(2.0f * kPiF * t));
    case 2:
      return 0.54f - 0.46f * std::cos(2.0f * kPiF * t);
    case 4:
      return 0.42f - 0.5f * std::cos(2.0f * kPiF * t) + 0.08f * std::cos(4.0f * kPiF * t);
    case 3: {
      const float a = std::min(std::max(tukey_alpha, 0.0f), 1.0f);
      if (a <= 0.0f) return 1.0f;
      if (t < a * 0.5f) return 0.5f * (1.0f + std::cos(2.0f * kPiF / a * (t - a * 0.5f)));
      if (t > 1.0f - a * 0.5f)
        return 0.5f * (1.0f + std::cos(2.0f * kPiF / a * (t - 1.0f + a * 0.5f)));
      return 1.0f;
    }
  }
  return 1.0f;
}

struct P3 {
  float x, y, z;
};

inline float norm3(float x, float y, float z) { return std::sqrt(x * x + y * y + z * z); }

// Vec3 distance a - b (types.hpp:48-56).
inline float dist(const P3& a, const P3& b) { return norm3(a.x - b.x, a.y - b.y, a.z - b.z); }

// transmit_arrival_time (geometry.cpp:82-109).
float arrival_time(const oc_transmit& tx, const P3& p, float c) {
  const P3 O{tx.origin[0], 0.0f, 0.0f};
  if (tx.type == 1) {
    const float s = std::sin(tx.angle), k = std::cos(tx.angle);
    return ((p.x - O.x) * s + (p.z - O.z) * k) / c;
  }
  const P3 F{tx.focus[0], tx.focus[1], tx.focus[2]};
  if (tx.type == 0) {
    const P3 ax{F.x - O.x, F.y - O.y, F.z - O.z};
    const float len = norm3(ax.x, ax.y, ax.z);
    if (len < 1e-9f) return norm3(p.x, p.y, p.z) / c;
    const P3 d{ax.x / len, ax.y / len, ax.z / len};
    const P3 rel{p.x - F.x, p.y - F.y, p.z - F.z};
    const float proj = rel.x * d.x + rel.y * d.y + rel.z * d.z;
    const float sign = (proj >= 0.0f) ? 1.0f : -1.0f;
    return (len + sign * norm3(rel.x, rel.y, rel.z)) / c;
  }
  if (tx.type == 2) return (dist(p, F) - dist(O, F)) / c;
  return 0.0f;
}

struct Extent {
  float lo = 0, hi = 0;
  bool valid = false;
};

// compute_tx_extent (beamformer.cpp:238-247).
Extent tx_extent(const oc_setup& s, const oc_transmit& tx) {
  Extent e;
  for (int i = 0; i < s.n_elements; ++i) {
    if (tx.n_apod > 0 && tx.apodization[i] == 0.0f) continue;
    const float x = element_x(i, s.n_elements, s.pitch);
    if (!e.valid) {
      e.lo = e.hi = x;
      e.valid = true;
    } else {
      e.lo = std::min(e.lo, x);
      e.hi = std::max(e.hi, x);
    }
  }
  return e;
}

// tx_sensitivity (beamformer.cpp:249-288).
float tx_weight(const Extent& e, const oc_transmit& tx, const P3& p, float lambda, float pitch) {
  if (!e.valid || p.z <= 0.0f) return 0.0f;
  float a, b;
  if (tx.type == 1) {
    const float shift = p.z * std::tan(tx.angle);
    a = e.lo + shift;
    b = e.hi + shift;
  } else {
    const float zf = tx.focus[2];
    if (std::abs(zf) < 1e-9f) return 1.0f;
    a = e.lo + (tx.focus[0] - e.lo) * p.z / zf;
    b = e.hi + (tx.focus[0] - e.hi) * p.z / zf;
    if (a > b) std::swap(a, b);
  }
  const float aperture = std::max(e.hi - e.lo, pitch);
  const float fne = std::min(std::max(p.z / aperture, 0.5f), 20.0f);
  const float diffraction = fne * lambda;
  a -= diffraction;
  b += diffraction;
  const float centre = 0.5f * (a + b);
  const float half = 0.5f * (b - a);
  if (half <= 0.0f) return 0.0f;
  return window_value(3, (p.x - centre) / half, 0.4f);
}

// ---- validation (types.cpp:8-64, beamformer.cpp:218-225) -------------------

void validate_setup(const oc_setup& s, const oc_transmit* txs, int n_tx) {
  if (s.n_elements <= 0) throw Invalid("Probe.n_elements must be > 0");
  if (s.pitch <= 0) throw Invalid("Probe.pitch must be > 0");
  if (s.element_width <= 0) throw Invalid("Probe.element_width must be > 0");
  if (s.element_width > s.pitch * 1.001f)
    throw Invalid("Probe.element_width exceeds pitch (elements would overlap)");
  if (s.center_frequency <= 0) throw Invalid("Probe.center_frequency must be > 0");
  if (s.bandwidth <= 0) throw Invalid("Probe.bandwidth must be > 0");
  if (s.speed_of_sound <= 0) throw Invalid("Medium.speed_of_sound must be > 0");
  if (s.attenuation < 0) throw Invalid("Medium.attenuation must be >= 0");
  if (s.sampling_frequency <= 0) throw Invalid("Acquisition.sampling_frequency must be > 0");
  if (s.n_samples <= 0) throw Invalid("Acquisition.n_samples must be > 0");
  if (s.prf <= 0) throw Invalid("Acquisition.prf must be > 0");
  if (n_tx <= 0) throw Invalid("ChannelData has no transmit events");
  for (int i = 0; i < n_tx; ++i) {
    if (txs[i].n_delays != 0 && txs[i].n_delays != s.n_elements)
      throw Invalid("Transmit.delays length != probe.n_elements");
    if (txs[i].n_apod != 0 && txs[i].n_apod != s.n_elements)
      throw Invalid("Transmit.apodization length != probe.n_elements");
  }
}

// BeamformerConfig::validate (beamformer.cpp:218-225). The reference runs it
// when a Beamformer is constructed (beamformer.hpp:110), never inside
// beamform (beamformer.cpp:297-299): a config assigned after construction is
// beamformed with the clamps of geometry.cpp:127 and beamformer.cpp:147, 187.
// The beamform entries below follow that; oc_validate_bf is the constructor's
// check on its own.
void validate_bf(const oc_bf& b) {
  if (b.f_number <= 0) throw Invalid("BeamformerConfig.f_number must be > 0");
  if (b.mv_subarray < 2) throw Invalid("BeamformerConfig.mv_subarray must be >= 2");
  if (b.mv_diagonal_loading < 0)
    throw Invalid("BeamformerConfig.mv_diagonal_loading must be >= 0");
  if (b.mv_range_average < 0) throw Invalid("BeamformerConfig.mv_range_average must be >= 0");
}

// What beamform itself cannot run: minimum variance with a negative range
// average sizes a scratch vector of 2 r + 1 entries (beamformer.cpp:357),
// which fails for r < 0; every other out-of-range field has a clamp.
void check_bf_runnable(const oc_bf& b) {
  if (b.combiner == 2 && b.mv_range_average < 0)
    throw Invalid("BeamformerConfig.mv_range_average must be >= 0");
}

// ---- grid (types.cpp:66-115) ------------------------------------------------

struct Grid {
  int type = 0, n0 = 0, n1 = 0;
  std::vector<float> axis0, axis1;
  std::vector<P3> pts;
};

// Keep lines line_first, line_first + line_step, ... (line_step 0: all) with
// their axis values and points as the full grid has them.
Grid subset_lines(Grid G, const oc_grid& g) {
  if (g.line_step <= 0) return G;
  if (g.line_first < 0 || g.line_first >= G.n0) throw Invalid("oc_grid: line_first outside the grid");
  Grid S;
  S.type = G.type;
  S.n1 = G.n1;
  S.axis1 = G.axis1;
  for (int i = g.line_first; i < G.n0; i += g.line_step) {
    S.axis0.push_back(G.axis0[i]);
    S.pts.insert(S.pts.end(), G.pts.begin() + static_cast<std::ptrdiff_t>(i) * G.n1,
                 G.pts.begin() + static_cast<std::ptrdiff_t>(i + 1) * G.n1);
  }
  S.n0 = static_cast<int>(S.axis0.size());
  return S;
}

Grid build_grid(const oc_grid& g) {
  Grid G;
  G.type = g.type;
  G.n0 = g.n0;
  G.n1 = g.n1;
  if (g.n0 <= 0 || g.n1 <= 0) throw Invalid("ScanGrid dimensions must be > 0");
  G.axis0.resize(g.n0);
  G.axis1.resize(g.n1);
  for (int i = 0; i < g.n0; ++i)
    G.axis0[i] = g.n0 == 1 ? g.a0_min : g.a0_min + (g.a0_max - g.a0_min) * i / (g.n0 - 1);
  for (int j = 0; j < g.n1; ++j)
    G.axis1[j] = g.n1 == 1 ? g.a1_min : g.a1_min + (g.a1_max - g.a1_min) * j / (g.n1 - 1);
  G.pts.resize(static_cast<std::size_t>(g.n0) * g.n1);
  for (int i = 0; i < g.n0; ++i)
    for (int j = 0; j < g.n1; ++j) {
      P3& p = G.pts[static_cast<std::size_t>(i) * g.n1 + j];
      if (g.type == 0) {
        p = {G.axis0[i], 0.0f, G.axis1[j]};
      } else {
        const float th = G.axis0[i], r = G.axis1[j];
        p = {g.apex[0] + r * std::sin(th), 0.0f, g.apex[2] + r * std::cos(th)};
      }
    }
  return subset_lines(G, g);
}

// ---- demodulation (demodulate.cpp:13-121) ----------------------------------

std::vector<float> lowpass(int n_taps, float fc_in) {
  if (n_taps < 3) n_taps = 3;
  if (n_taps % 2 == 0) ++n_taps;
  const float fc = std::min(std::max(fc_in, 1e-4f), 0.499f);
  std::vector<float> h(static_cast<std::size_t>(n_taps));
  const int m = n_taps / 2;
  float sum = 0.0f;
  for (int i = 0; i < n_taps; ++i) {
    const int k = i - m;
    const float sinc = (k == 0) ? 2.0f * fc : std::sin(2.0f * kPiF * fc * k) / (kPiF * k);
    const float w = 0.54f - 0.46f * std::cos(2.0f * kPiF * i / (n_taps - 1));
    h[i] = sinc * w;
    sum += sinc * w;
  }
  if (sum != 0.0f)
    for (float& v : h) v /= sum;
  return h;
}

struct IQ {
  std::vector<cf32> data;
  int n = 0;
  float fs = 0, fdem = 0;
};

IQ demodulate(const oc_setup& s, int n_tx, const float* rf, const oc_demod& d) {
  const float fdem = d.demodulation_frequency > 0.0f ? d.demodulation_frequency
                                                     : s.center_frequency;
  const int dec = std::max(1, d.decimation);
  const float fs = s.sampling_frequency;
  const std::vector<float> h = lowpass(d.filter_taps, d.cutoff_fraction * fdem / fs);
  const int n_taps = static_cast<int>(h.size());
  const int gd = n_taps / 2;
  const int n_in = s.n_samples;
  IQ out;
  out.fs = fs / dec;
  out.n = (n_in + dec - 1) / dec;
  out.fdem = fdem;
  const std::size_t n_traces = static_cast<std::size_t>(n_tx) * s.n_elements;
  out.data.assign(n_traces * out.n, cf32(0, 0));

  // Mixing phasor per input sample, from absolute acquisition time.
  std::vector<cf32> mix(static_cast<std::size_t>(n_in));
  for (int n = 0; n < n_in; ++n) {
    const float t = s.t0 + static_cast<float>(n) / fs;
    const float ph = -2.0f * kPiF * fdem * t;
    mix[n] = cf32(std::cos(ph), std::sin(ph));
  }

#pragma omp parallel
  {
    std::vector<cf32> m(static_cast<std::size_t>(n_in));
#pragma omp for schedule(static)
    for (std::ptrdiff_t tr = 0; tr < static_cast<std::ptrdiff_t>(n_traces); ++tr) {
      const float* x = rf + static_cast<std::size_t>(tr) * n_in;
      for (int n = 0; n < n_in; ++n) m[n] = mix[n] * x[n];
      cf32* y = out.data.data() + static_cast<std::size_t>(tr) * out.n;
      for (int k = 0; k < out.n; ++k) {
        const int c = k * dec + gd;
        const int j0 = std::max(0, c - n_in + 1);
        const int j1 = std::min(n_taps - 1, c);
        cf32 acc(0, 0);
        for (int j = j0; j <= j1; ++j) acc += m[c - j] * h[j];
        y[k] = acc * 2.0f;
      }
    }
  }
  return out;
}

// ---- combiners (beamformer.cpp:47-214) --------------------------------------

// Row-oriented complex Cholesky and triangular solves (beamformer.cpp:47-79).
template <class T>
bool chol_solve(std::vector<std::complex<T>>& A, int L, std::vector<std::complex<T>>& b) {
  using C = std::complex<T>;
  for (int i = 0; i < L; ++i) {
    for (int j = 0; j <= i; ++j) {
      C sum = A[i * L + j];
      for (int k = 0; k < j; ++k) sum -= A[i * L + k] * std::conj(A[j * L + k]);
      if (i == j) {
        const T d = sum.real();
        if (!(d > T(0))) return false;
        A[i * L + j] = C(std::sqrt(d), T(0));
      } else {
        A[i * L + j] = sum / A[j * L + j];
      }
    }
  }
  for (int i = 0; i < L; ++i) {
    C sum = b[i];
    for (int k = 0; k < i; ++k) sum -= A[i * L + k] * b[k];
    b[i] = sum / A[i * L + i];
  }
  for (int i = L - 1; i >= 0; --i) {
    C sum = b[i];
    for (int k = i + 1; k < L; ++k) sum -= std::conj(A[k * L + i]) * b[k];
    b[i] = sum / A[i * L + i];
  }
  return true;
}

struct MvInfo {
  int n_active = 0, L = 0, fallback = 0;  // fallback: 0 none, 1 few channels, 2 solve failed
};

// combine_mv_windowed (beamformer.cpp:137-214). Y(k, ch) returns the
// line-scratch value at depth row k. `g` (optional, n_ch) receives the
// effective per-channel weights: out == sum_ch g[ch] * Y(kw, ch).
template <class T, class Get>
cf32 mv_combine(Get Y, int kw, int k_lo, int k_hi, const int* act, int na, int subarray,
                float loading, int ra, MvInfo* info, cf32* g) {
  using C = std::complex<T>;
  auto das = [&] {
    cf32 acc(0, 0);
    for (int a = 0; a < na; ++a) acc += Y(kw, act[a]);
    if (g)
      for (int a = 0; a < na; ++a) g[act[a]] = cf32(1, 0);
    return acc;
  };
  if (info) info->n_active = na;
  if (na < 4) {
    if (info) info->fallback = 1;
    return das();
  }
  const int L = std::clamp(subarray, 2, std::max(2, na / 2));
  const int ns = na - L + 1;
  if (info) info->L = L;
  std::vector<C> R(static_cast<std::size_t>(L) * L, C(0, 0)), sub(L);
  int n_snap = 0;
  for (int d = -ra; d <= ra; ++d) {
    const int k = kw + d;
    if (k < k_lo || k >= k_hi) continue;
    for (int s0 = 0; s0 < ns; ++s0) {
      for (int a = 0; a < L; ++a) {
        const cf32 v = Y(k, act[s0 + a]);
        sub[a] = C(v.real(), v.imag());
      }
      for (int a = 0; a < L; ++a) {
        const C va = sub[a];
        for (int b = a; b < L; ++b) R[a * L + b] += va * std::conj(sub[b]);
      }
      ++n_snap;
    }
  }
  if (n_snap == 0) return cf32(0, 0);
  const T inv = T(1) / static_cast<T>(n_snap);
  T trace = 0;
  for (int a = 0; a < L; ++a) {
    for (int b = a; b < L; ++b) {
      R[a * L + b] *= inv;
      if (b != a) R[b * L + a] = std::conj(R[a * L + b]);
    }
    trace += R[a * L + a].real();
  }
  const T eps = std::max(static_cast<T>(loading) * trace / static_cast<T>(L), T(1e-20f));
  for (int a = 0; a < L; ++a) R[a * L + a] += C(eps, T(0));
  std::vector<C> w(static_cast<std::size_t>(L), C(1, 0));
  if (!chol_solve<T>(R, L, w)) {
    if (info) info->fallback = 2;
    return das();
  }
  C den(0, 0);
  for (int a = 0; a < L; ++a) den += w[a];
  if (std::abs(den) < T(1e-30f)) {
    if (info) info->fallback = 2;
    return das();
  }
  C out(0, 0);
  for (int a = 0; a < L; ++a) {
    C avg(0, 0);
    for (int s0 = 0; s0 < ns; ++s0) {
      const cf32 v = Y(kw, act[s0 + a]);
      avg += C(v.real(), v.imag());
    }
    avg /= static_cast<T>(ns);
    out += std::conj(w[a]) * avg;
  }
  const C res = out / den * static_cast<T>(na);
  if (g) {
    // g_c = n_active / (n_sub * sum w) * sum_{a covering c} conj(w_a)
    for (int c = 0; c < na; ++c) {
      C acc(0, 0);
      for (int a = std::max(0, c - ns + 1); a <= std::min(L - 1, c); ++a) acc += std::conj(w[a]);
      const C gc = acc / den * (static_cast<T>(na) / static_cast<T>(ns));
      g[act[c]] = cf32(static_cast<float>(gc.real()), static_cast<float>(gc.imag()));
    }
  }
  return cf32(static_cast<float>(res.real()), static_cast<float>(res.imag()));
}

// combine_dmas (beamformer.cpp:89-130), over the active channels in order.
cf32 dmas(const cf32* y, const int* act, int na) {
  cf32 s(0, 0), q(0, 0);
  for (int a = 0; a < na; ++a) {
    const cf32 yi = y[act[a]];
    const float m = std::abs(yi);
    if (m <= 0.0f) continue;
    const cf32 v = yi / std::sqrt(m);
    s += v;
    q += v * v;
  }
  const cf32 acc = (s * s - q) * 0.5f;
  const float nrm = 0.5f * std::max(1.0f, static_cast<float>(na - 1));
  return acc / nrm;
}

// ---- beamformer (beamformer.cpp:297-549) -------------------------------------

struct BfExtra {
  int mv_fp64 = 0;
  float* weights = nullptr;  // n_pix * n_ch complex (interleaved), MV only
  float* ych = nullptr;      // n_pix * n_ch complex: the channel vector Y of each pixel
  int* info = nullptr;       // n_pix * 3: n_active, L, fallback
};

void beamform(const oc_setup& s, const oc_transmit* txs, int n_tx, bool is_iq, const void* data,
              float fdem_in, const Grid& G, const oc_bf& bf, int ev0, int ev1, cf32* out,
              const BfExtra& ex) {
  if (ev1 < 0) ev1 = n_tx;
  ev0 = std::clamp(ev0, 0, n_tx);
  ev1 = std::clamp(ev1, ev0, n_tx);
  const int n_ch = s.n_elements, n_smp = s.n_samples, n1 = G.n1;
  const float c = s.speed_of_sound, fs = s.sampling_frequency, t0 = s.t0;
  const float fdem = is_iq ? fdem_in : 0.0f;
  const float lambda = c / s.center_frequency;
  const bool incoh = bf.compounding == 1, select = bf.compounding == 2;
  const bool is_mv = bf.combiner == 2;
  const int ra = is_mv ? bf.mv_range_average : 0;

  std::vector<P3> el(n_ch);
  for (int i = 0; i < n_ch; ++i) el[i] = {element_x(i, n_ch, s.pitch), 0.0f, 0.0f};
  std::vector<Extent> ext(n_tx);
  for (int e = ev0; e < ev1; ++e) ext[e] = tx_extent(s, txs[e]);
  const float* rf = static_cast<const float*>(data);
  const cf32* iq = static_cast<const cf32*>(data);

#pragma omp parallel
  {
    const std::size_t line = static_cast<std::size_t>(n1) * n_ch;
    std::vector<float> apod(line), rtau(line), rcos(line), rsin(line);
    std::vector<cf32> Y(line);
    std::vector<float> ttau(n1), tw(n1), tcos(n1), tsin(n1), acc(n1);
    std::vector<int> best_ev(n1), act(n_ch);
    std::vector<float> best_w(n1);
    std::vector<cf32> ybuf(n_ch);

#pragma omp for schedule(dynamic, 1)
    for (int i0 = 0; i0 < G.n0; ++i0) {
      const P3* pts = G.pts.data() + static_cast<std::size_t>(i0) * n1;
      cf32* o = out + static_cast<std::size_t>(i0) * n1;

      // Receive geometry per (depth, channel), reused across events.
      for (int k = 0; k < n1; ++k) {
        for (int ch = 0; ch < n_ch; ++ch) {
          const std::size_t q = static_cast<std::size_t>(k) * n_ch + ch;
          const P3& p = pts[k];
          float a = 0.0f;
          const float depth = p.z - el[ch].z;
          if (depth > 0.0f) {
            const float half = 0.5f * depth / std::max(bf.f_number, 1e-3f);
            const float u = (el[ch].x - p.x) / std::max(half, 1e-9f);
            a = window_value(bf.rx_window, u, bf.tukey_alpha);
          }
          apod[q] = is_mv ? (a > 0.0f ? 1.0f : 0.0f) : a;
          if (a <= 0.0f) continue;
          const float tau = dist(el[ch], p) / c;
          rtau[q] = tau;
          if (fdem > 0.0f) {
            float sn, cs;
            sincosf(2.0f * kPiF * fdem * tau, &sn, &cs);
            rsin[q] = sn;
            rcos[q] = cs;
          }
        }
      }

      if (select) {
        for (int k = 0; k < n1; ++k) {
          best_ev[k] = -1;
          best_w[k] = 0.0f;
        }
        for (int e = ev0; e < ev1; ++e)
          for (int k = 0; k < n1; ++k) {
            const float w = tx_weight(ext[e], txs[e], pts[k], lambda, s.pitch);
            if (w > best_w[k]) {
              best_w[k] = w;
              best_ev[k] = e;
            }
          }
      }

      const int n_pass = incoh ? (ev1 - ev0) : 1;
      std::fill(acc.begin(), acc.end(), 0.0f);
      for (int pass = 0; pass < std::max(1, n_pass); ++pass) {
        std::fill(Y.begin(), Y.end(), cf32(0, 0));
        const int e_lo = incoh ? ev0 + pass : ev0;
        const int e_hi = incoh ? e_lo + 1 : ev1;
        for (int e = e_lo; e < e_hi; ++e) {
          const oc_transmit& tx = txs[e];
          for (int k = 0; k < n1; ++k) {
            float w;
            if (select) {
              w = (best_ev[k] == e) ? best_w[k] : 0.0f;
              if (w > 0.0f) w = 1.0f;
            } else if (bf.transmit_weighting) {
              w = tx_weight(ext[e], tx, pts[k], lambda, s.pitch);
            } else {
              w = 1.0f;
            }
            tw[k] = w;
            if (w <= 0.0f) continue;
            const float tau = arrival_time(tx, pts[k], c);
            ttau[k] = tau;
            if (fdem > 0.0f) {
              float sn, cs;
              sincosf(2.0f * kPiF * fdem * tau, &sn, &cs);
              tsin[k] = sn;
              tcos[k] = cs;
            } else {
              tcos[k] = 1.0f;
              tsin[k] = 0.0f;
            }
          }
          for (int k = 0; k < n1; ++k) {
            if (tw[k] <= 0.0f) continue;
            for (int ch = 0; ch < n_ch; ++ch) {
              const std::size_t q = static_cast<std::size_t>(k) * n_ch + ch;
              const float a = apod[q];
              if (a <= 0.0f) continue;
              const float idx = (ttau[k] + rtau[q] - t0) * fs;
              const int kk = static_cast<int>(std::floor(idx));
              const std::size_t base = (static_cast<std::size_t>(e) * n_ch + ch) * n_smp;
              if (is_iq) {
                cf32 v(0, 0);
                if (!(kk < 0 || kk + 1 >= n_smp)) {
                  const float f = idx - static_cast<float>(kk);
                  v = iq[base + kk] * (1.0f - f) + iq[base + kk + 1] * f;
                }
                const float tc = tcos[k], ts = tsin[k], rc = rcos[q], rs = rsin[q];
                Y[q] += v * cf32(tc * rc - ts * rs, ts * rc + tc * rs) * (a * tw[k]);
              } else {
                cf32 v(0, 0);
                if (!(kk < 0 || kk + 1 >= n_smp)) {
                  const float f = idx - static_cast<float>(kk);
                  v = cf32(rf[base + kk] * (1.0f - f) + rf[base + kk + 1] * f, 0.0f);
                }
                Y[q] += v * (a * tw[k]);
              }
            }
          }
        }

        // Combine each depth's aligned channels.
        for (int k = 0; k < n1; ++k) {
          const std::size_t pix = static_cast<std::size_t>(i0) * n1 + k;
          int na = 0;
          for (int ch = 0; ch < n_ch; ++ch) {
            const std::size_t q = static_cast<std::size_t>(k) * n_ch + ch;
            if (apod[q] <= 0.0f) {
              ybuf[ch] = cf32(0, 0);
              continue;
            }
            act[na++] = ch;
            ybuf[ch] = Y[q];
          }
          if (ex.ych)
            std::memcpy(ex.ych + 2 * pix * n_ch, ybuf.data(), sizeof(float) * 2 * n_ch);
          if (na == 0) {
            if (!incoh) o[k] = cf32(0, 0);
            if (ex.info) {
              ex.info[3 * pix] = 0;
              ex.info[3 * pix + 1] = 0;
              ex.info[3 * pix + 2] = 1;
            }
            continue;
          }
          if (ex.info) {
            ex.info[3 * pix] = na;
            ex.info[3 * pix + 1] = 0;
            ex.info[3 * pix + 2] = 0;
          }
          cf32 r(0, 0);
          if (bf.combiner == 0) {
            for (int ch = 0; ch < n_ch; ++ch) r += ybuf[ch];
          } else if (bf.combiner == 1) {
            r = dmas(ybuf.data(), act.data(), na);
          } else {
            auto get = [&](int kr, int ch) { return Y[static_cast<std::size_t>(kr) * n_ch + ch]; };
            MvInfo info;
            cf32* gw = nullptr;
            if (ex.weights) {
              gw = reinterpret_cast<cf32*>(ex.weights) + pix * n_ch;
              std::fill(gw, gw + n_ch, cf32(0, 0));
            }
            if (ex.mv_fp64)
              r = mv_combine<double>(get, k, 0, n1, act.data(), na, bf.mv_subarray,
                                     bf.mv_diagonal_loading, ra, &info, gw);
            else
              r = mv_combine<float>(get, k, 0, n1, act.data(), na, bf.mv_subarray,
                                    bf.mv_diagonal_loading, ra, &info, gw);
            if (ex.info) {
              ex.info[3 * pix] = info.n_active;
              ex.info[3 * pix + 1] = info.L;
              ex.info[3 * pix + 2] = info.fallback;
            }
          }
          if (bf.coherence_factor) {
            cf32 coh(0, 0);
            float inc = 0.0f;
            for (int a = 0; a < na; ++a) {
              const cf32& v = ybuf[act[a]];
              coh += v;
              inc += std::norm(v);
            }
            r *= (inc > 0.0f) ? std::norm(coh) / (static_cast<float>(na) * inc) : 0.0f;
          }
          if (incoh)
            acc[k] += std::norm(r);
          else
            o[k] = r;
        }
      }
      if (incoh)
        for (int k = 0; k < n1; ++k) o[k] = cf32(std::sqrt(acc[k]), 0.0f);
    }
  }
}

// ---- display tail (postprocess.cpp:12-117, pipeline.cpp:48-70) -------------

// Iterative radix-2 FFT with the twiddle carried by complex multiplication,
// in float, as the reference computes it.
void fft(std::vector<cf32>& a, bool inverse) {
  const std::size_t n = a.size();
  if (n < 2) return;
  for (std::size_t i = 1, j = 0; i < n; ++i) {
    std::size_t bit = n >> 1;
    for (; j & bit; bit >>= 1) j ^= bit;
    j ^= bit;
    if (i < j) std::swap(a[i], a[j]);
  }
  for (std::size_t len = 2; len <= n; len <<= 1) {
    const float ang = 2.0f * kPiF / static_cast<float>(len) * (inverse ? 1.0f : -1.0f);
    const cf32 step(std::cos(ang), std::sin(ang));
    for (std::size_t i = 0; i < n; i += len) {
      cf32 w(1, 0);
      for (std::size_t k = 0; k < len / 2; ++k) {
        const cf32 u = a[i + k];
        const cf32 v = a[i + k + len / 2] * w;
        a[i + k] = u + v;
        a[i + k + len / 2] = u - v;
        w *= step;
      }
    }
  }
  if (inverse)
    for (cf32& v : a) v /= static_cast<float>(n);
}

// envelope + TGC of one frame (postprocess.cpp:40-98)
std::vector<float> display_envelope(const Grid& G, const cf32* frame, float fdem, int apply_tgc,
                                    float atten) {
  const std::size_t n = static_cast<std::size_t>(G.n0) * G.n1;
  std::vector<float> env(n);
  if (fdem > 0.0f) {
    for (std::size_t k = 0; k < n; ++k) env[k] = std::abs(frame[k]);
  } else {
    std::size_t nf = 1;
    while (nf < static_cast<std::size_t>(G.n1)) nf <<= 1;
    std::vector<cf32> buf(nf);
    for (int i = 0; i < G.n0; ++i) {
      std::fill(buf.begin(), buf.end(), cf32(0, 0));
      for (int j = 0; j < G.n1; ++j) buf[j] = cf32(frame[static_cast<std::size_t>(i) * G.n1 + j].real(), 0.0f);
      fft(buf, false);
      for (std::size_t k = 1; k < nf / 2; ++k) buf[k] *= 2.0f;
      for (std::size_t k = nf / 2 + 1; k < nf; ++k) buf[k] = cf32(0, 0);
      fft(buf, true);
      for (int j = 0; j < G.n1; ++j) env[static_cast<std::size_t>(i) * G.n1 + j] = std::abs(buf[j]);
    }
  }
  if (apply_tgc) {
    const float f0 = fdem > 0.0f ? fdem : 5.0e6f;
    const float f_mhz = f0 * 1e-6f;
    for (std::size_t k = 0; k < n; ++k) {
      const float depth_cm = G.pts[k].z * 100.0f;
      const float gain_db = 2.0f * atten * f_mhz * std::max(depth_cm, 0.0f);
      env[k] = env[k] * std::pow(10.0f, gain_db / 20.0f);
    }
  }
  return env;
}

// ---- Doppler and speckle tracking (doppler.cpp:15-307) ----------------------

// The (I - P) slow-time projector of the polynomial wall filter
// (doppler.cpp:15-49): orthonormalised monomials, subtracted outer products.
std::vector<float> reject_matrix(int n, int order) {
  order = std::clamp(order, 0, n - 1);
  std::vector<std::vector<float>> basis;
  for (int d = 0; d <= order; ++d) {
    std::vector<float> v(static_cast<std::size_t>(n));
    for (int k = 0; k < n; ++k) {
      const float x = (n == 1) ? 0.0f : 2.0f * k / (n - 1) - 1.0f;
      v[static_cast<std::size_t>(k)] = std::pow(x, d);
    }
    for (const auto& b : basis) {
      float dot = 0;
      for (int k = 0; k < n; ++k) dot += v[static_cast<std::size_t>(k)] * b[static_cast<std::size_t>(k)];
      for (int k = 0; k < n; ++k) v[static_cast<std::size_t>(k)] -= dot * b[static_cast<std::size_t>(k)];
    }
    float nrm = 0;
    for (float x : v) nrm += x * x;
    nrm = std::sqrt(nrm);
    if (nrm < 1e-8f) continue;
    for (float& x : v) x /= nrm;
    basis.push_back(std::move(v));
  }
  std::vector<float> M(static_cast<std::size_t>(n) * n, 0.0f);
  for (int i = 0; i < n; ++i) M[static_cast<std::size_t>(i) * n + i] = 1.0f;
  for (const auto& b : basis)
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j)
        M[static_cast<std::size_t>(i) * n + j] -= b[static_cast<std::size_t>(i)] * b[static_cast<std::size_t>(j)];
  return M;
}

// Filtered ensemble (doppler.cpp:51-75); type 0 copies.
std::vector<cf32> wall_filtered(const float* ens, int n, std::size_t n_pix, int type, int order) {
  const cf32* z = reinterpret_cast<const cf32*>(ens);
  std::vector<cf32> out(z, z + n_pix * static_cast<std::size_t>(n));
  if (type == 0 || n == 0) return out;
  const std::vector<float> M = reject_matrix(n, type == 2 ? 0 : order);
  for (std::size_t k = 0; k < n_pix; ++k)
    for (int i = 0; i < n; ++i) {
      cf32 acc(0, 0);
      for (int j = 0; j < n; ++j)
        acc += z[static_cast<std::size_t>(j) * n_pix + k] * M[static_cast<std::size_t>(i) * n + j];
      out[static_cast<std::size_t>(i) * n_pix + k] = acc;
    }
  return out;
}

float axial_step(const Grid& G) { return G.n1 > 1 ? std::abs(G.axis1[1] - G.axis1[0]) : 1e-4f; }

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

void check_data(const oc_setup& s, const oc_transmit* txs, int n_tx, bool is_iq, float fdem) {
  validate_setup(s, txs, n_tx);
  if (is_iq && fdem <= 0)
    throw Invalid("IQ ChannelData needs a positive demodulation_frequency");
}

}  // namespace

extern "C" {

const char* oc_last_error(void) { return g_err.c_str(); }
const char* oc_flavour(void) { return "port"; }

int oc_grid_build(const oc_grid* g, float* axis0, float* axis1, float* points) {
  return guarded([&] {
    const Grid G = build_grid(*g);
    std::memcpy(axis0, G.axis0.data(), G.axis0.size() * sizeof(float));
    std::memcpy(axis1, G.axis1.data(), G.axis1.size() * sizeof(float));
    std::memcpy(points, G.pts.data(), G.pts.size() * sizeof(P3));
  });
}

int oc_design_lowpass(int n_taps, float fc, float* taps_out, int* n_out) {
  return guarded([&] {
    const std::vector<float> h = lowpass(n_taps, fc);
    if (taps_out) std::memcpy(taps_out, h.data(), h.size() * sizeof(float));
    *n_out = static_cast<int>(h.size());
  });
}

float oc_window_value(int window, float u, float tukey_alpha) {
  return window_value(window, u, tukey_alpha);
}

int oc_make_transmits(const oc_setup*, const oc_simcfg*, int, int, const float*, float, float*,
                      float*, float*, float*, float*) {
  g_err = "the restatement does not include the simulator's transmit factories";
  return 2;
}

float oc_transmit_arrival_time(const oc_transmit* tx, const float* p, float c) {
  return arrival_time(*tx, P3{p[0], p[1], p[2]}, c);
}

float oc_transmit_sensitivity(const oc_setup* s, const oc_transmit* tx, const float* p,
                              float wavelength) {
  return tx_weight(tx_extent(*s, *tx), *tx, P3{p[0], p[1], p[2]}, wavelength, s->pitch);
}

int oc_speckle_box(float, float, float, float, float, int, uint32_t, float, float*, float*) {
  g_err = "the restatement does not include the phantom factories";
  return 2;
}

int oc_simulate(const oc_setup*, const oc_simcfg*, const oc_transmit*, int, int, const float*,
                const float*, const float*, float*) {
  g_err = "the restatement does not include the simulator";
  return 2;
}

int oc_demodulate(const oc_setup* s, const oc_transmit* txs, int n_tx, const float* rf,
                  const oc_demod* d, float* iq_out, int* n_out, float* fs_out, float* fdem_out) {
  return guarded([&] {
    validate_setup(*s, txs, n_tx);
    if (!iq_out) {
      const int dec = std::max(1, d->decimation);
      *n_out = (s->n_samples + dec - 1) / dec;
      return;
    }
    const IQ r = demodulate(*s, n_tx, rf, *d);
    std::memcpy(iq_out, r.data.data(), r.data.size() * sizeof(cf32));
    *n_out = r.n;
    *fs_out = r.fs;
    *fdem_out = r.fdem;
  });
}

int oc_validate_bf(const oc_bf* bf) {
  return guarded([&] { validate_bf(*bf); });
}

// Port-only extension: beamform with optional per-pixel outputs.
// mv_fp64: 1 runs the MV combine in double (same FP32 channel data).
int oc_beamform_ex(const oc_setup* s, const oc_transmit* txs, int n_tx, int is_iq,
                   const float* data, float fdem, const oc_grid* g, const oc_bf* bf,
                   int event_begin, int event_end, float* frame_out, int mv_fp64,
                   float* weights_out, float* ych_out, int* info_out) {
  return guarded([&] {
    check_data(*s, txs, n_tx, is_iq != 0, fdem);
    check_bf_runnable(*bf);
    const Grid G = build_grid(*g);
    BfExtra ex;
    ex.mv_fp64 = mv_fp64;
    ex.weights = weights_out;
    ex.ych = ych_out;
    ex.info = info_out;
    beamform(*s, txs, n_tx, is_iq != 0, data, fdem, G, *bf, event_begin, event_end,
             reinterpret_cast<cf32*>(frame_out), ex);
  });
}

int oc_beamform(const oc_setup* s, const oc_transmit* txs, int n_tx, int is_iq,
                const float* data, float fdem, const oc_grid* g, const oc_bf* bf,
                int event_begin, int event_end, float* frame_out) {
  return oc_beamform_ex(s, txs, n_tx, is_iq, data, fdem, g, bf, event_begin, event_end,
                        frame_out, 0, nullptr, nullptr, nullptr);
}

int oc_display(const oc_setup* s, const oc_grid* g, const oc_display_cfg* p, int n_frames,
               const float* frames, float fdem, float* images_out) {
  return guarded([&] {
    const Grid G = build_grid(*g);
    const std::size_t n = static_cast<std::size_t>(G.n0) * G.n1;
    float reference = 0.0f;
    for (int f = 0; f < n_frames; ++f) {
      const cf32* fr = reinterpret_cast<const cf32*>(frames) + n * f;
      const std::vector<float> env = display_envelope(G, fr, fdem, p->apply_tgc, s->attenuation);
      float peak = 0.0f;
      for (float v : env) peak = std::max(peak, v);
      if (peak > 0.0f) {
        const float a = std::clamp(p->reference_smoothing, 0.0f, 0.999f);
        reference = (reference > 0.0f) ? (a * reference + (1.0f - a) * peak) : peak;
      }
      // log_compress(env, dr, gain, reference > 0 ? reference : -1)
      float ref = reference > 0.0f ? reference : -1.0f;
      if (ref <= 0.0f) {
        ref = 0.0f;
        for (float v : env) ref = std::max(ref, v);
      }
      float* out = images_out + n * f;
      if (ref <= 0.0f) {
        std::fill(out, out + n, 0.0f);
        continue;
      }
      const float dr = std::max(p->dynamic_range_db, 1.0f);
      for (std::size_t k = 0; k < n; ++k) {
        const float v = env[k] / ref;
        const float db = (v > 0.0f) ? 20.0f * std::log10(v) : -1e6f;
        out[k] = std::clamp((db + p->gain_db + dr) / dr, 0.0f, 1.0f);
      }
    }
  });
}

int oc_process_frame(const oc_setup* s, const oc_transmit* txs, int n_tx, const float* rf,
                     const oc_grid* g, const oc_demod* d, const oc_bf* bf, float* frame_out,
                     double* stats_out) {
  using clk = std::chrono::steady_clock;
  return guarded([&] {
    const auto t_start = clk::now();
    check_bf_runnable(*bf);
    const Grid G = build_grid(*g);
    check_data(*s, txs, n_tx, false, 0.0f);
    const auto t_d = clk::now();
    const IQ iq = demodulate(*s, n_tx, rf, *d);
    const double dm = std::chrono::duration<double, std::milli>(clk::now() - t_d).count();
    oc_setup s2 = *s;
    s2.sampling_frequency = iq.fs;
    s2.n_samples = iq.n;
    const auto t_b = clk::now();
    beamform(s2, txs, n_tx, true, iq.data.data(), iq.fdem, G, *bf, 0, -1,
             reinterpret_cast<cf32*>(frame_out), BfExtra{});
    const double bm = std::chrono::duration<double, std::milli>(clk::now() - t_b).count();
    if (stats_out) {
      stats_out[0] = dm;
      stats_out[1] = bm;
      stats_out[2] = std::chrono::duration<double, std::milli>(clk::now() - t_start).count();
    }
  });
}

int oc_wall_filter(const oc_grid* g, int n_frames, const float* ensemble, int type, int order,
                   float* out) {
  return guarded([&] {
    const Grid G = build_grid(*g);
    const std::size_t n_pix = static_cast<std::size_t>(G.n0) * G.n1;
    const std::vector<cf32> f = wall_filtered(ensemble, n_frames, n_pix, type, order);
    std::memcpy(out, f.data(), f.size() * sizeof(cf32));
  });
}

int oc_power_doppler(const oc_grid* g, int n_frames, const float* ensemble, int type, int order,
                     float* power_out) {
  return guarded([&] {
    const Grid G = build_grid(*g);
    const std::size_t n_pix = static_cast<std::size_t>(G.n0) * G.n1;
    const std::vector<cf32> f = wall_filtered(ensemble, n_frames, n_pix, type, order);
    for (std::size_t k = 0; k < n_pix; ++k) power_out[k] = 0.0f;
    for (int e = 0; e < n_frames; ++e)
      for (std::size_t k = 0; k < n_pix; ++k) power_out[k] += std::norm(f[static_cast<std::size_t>(e) * n_pix + k]);
  });
}

// color_flow (doppler.cpp:77-176): per pixel the window (lateral outer, axial
// inner) and, inside, the ensemble, in one running sum each.
int oc_color_flow(const oc_grid* g, int n_frames, const float* ensemble, float fdem, float prf,
                  float c, const oc_doppler* d, float* velocity, float* power, float* variance,
                  uint8_t* valid, float* nyquist) {
  return guarded([&] {
    if (n_frames < 2) throw Invalid("color_flow: need at least 2 frames in the ensemble");
    const Grid G = build_grid(*g);
    const int n0 = G.n0, n1 = G.n1;
    const std::size_t n_pix = static_cast<std::size_t>(n0) * n1;
    const std::vector<cf32> f =
        wall_filtered(ensemble, n_frames, n_pix, d->wall_filter, d->polynomial_order);
    const float dz = axial_step(G);
    *nyquist = (fdem > 0.0f) ? c * prf / (4.0f * fdem) : 0.0f;
    const int ka = std::max(0, d->kernel_axial), kl = std::max(0, d->kernel_lateral);
    auto at = [&](int e, std::size_t kk) { return f[static_cast<std::size_t>(e) * n_pix + kk]; };
    for (int i = 0; i < n0; ++i)
      for (int j = 0; j < n1; ++j) {
        const std::size_t k = static_cast<std::size_t>(i) * n1 + j;
        velocity[k] = power[k] = variance[k] = 0.0f;
        valid[k] = 0;
        cf32 r01(0, 0), r10(0, 0);
        float r00 = 0.0f;
        for (int ii = i - kl; ii <= i + kl; ++ii) {
          if (ii < 0 || ii >= n0) continue;
          for (int jj = j - ka; jj <= j + ka; ++jj) {
            if (jj < 0 || jj >= n1) continue;
            const std::size_t kk = static_cast<std::size_t>(ii) * n1 + jj;
            for (int e = 0; e < n_frames; ++e) {
              const cf32 z = at(e, kk);
              r00 += std::norm(z);
              if (e + 1 < n_frames) r01 += std::conj(z) * at(e + 1, kk);
              if (jj + 1 < n1) r10 += std::conj(z) * at(e, kk + 1);
            }
          }
        }
        power[k] = r00;
        if (r00 <= 0.0f) continue;
        const float coh = std::abs(r01) / r00;
        variance[k] = std::clamp(2.0f * (1.0f - coh), 0.0f, 1.0f);
        if (coh < d->coherence_threshold) continue;
        float fc = fdem;
        if (d->loupas) {
          const float fe = std::arg(r10) * c / (4.0f * kPiF * dz);
          if (fe > 0.25f * fdem && fe < 2.0f * fdem) fc = fe;
        }
        if (fc <= 0.0f) continue;
        velocity[k] = -std::arg(r01) * c * prf / (4.0f * kPiF * fc);
        valid[k] = 1;
      }
    float pmax = 0.0f;
    for (std::size_t k = 0; k < n_pix; ++k) pmax = std::max(pmax, power[k]);
    if (pmax > 0.0f) {
      const float thr = pmax * std::pow(10.0f, d->power_threshold_db / 10.0f);
      for (std::size_t k = 0; k < n_pix; ++k)
        if (power[k] < thr) valid[k] = 0;
    }
  });
}

// estimate_axial_displacement (doppler.cpp:188-273).
int oc_axial_displacement(const oc_grid* g, const float* before, const float* after, float fdem,
                          float c, int kernel_axial, int kernel_lateral, int max_lag, float* axial,
                          float* correlation) {
  return guarded([&] {
    const Grid G = build_grid(*g);
    const int n0 = G.n0, n1 = G.n1;
    const cf32* a = reinterpret_cast<const cf32*>(before);
    const cf32* b = reinterpret_cast<const cf32*>(after);
    const float dz = axial_step(G);
    const int ka = std::max(1, kernel_axial), kl = std::max(0, kernel_lateral);
    const int lmax = std::max(0, max_lag);
    for (int i = 0; i < n0; ++i)
      for (int j = 0; j < n1; ++j) {
        const std::size_t k = static_cast<std::size_t>(i) * n1 + j;
        axial[k] = correlation[k] = 0.0f;
        cf32 r10(0, 0);
        for (int ii = i - kl; ii <= i + kl; ++ii) {
          if (ii < 0 || ii >= n0) continue;
          for (int jj = j - ka; jj <= j + ka; ++jj) {
            if (jj < 0 || jj + 1 >= n1) continue;
            const std::size_t kk = static_cast<std::size_t>(ii) * n1 + jj;
            r10 += std::conj(a[kk]) * a[kk + 1];
          }
        }
        float fc = fdem;
        const float fe = std::arg(r10) * c / (4.0f * kPiF * dz);
        if (fe > 0.25f * fdem && fe < 2.0f * fdem) fc = fe;
        if (fc <= 0.0f) continue;
        cf32 best(0, 0);
        float best_corr = -1.0f;
        int best_lag = 0;
        for (int lag = -lmax; lag <= lmax; ++lag) {
          cf32 r(0, 0);
          float ea = 0.0f, eb = 0.0f;
          for (int ii = i - kl; ii <= i + kl; ++ii) {
            if (ii < 0 || ii >= n0) continue;
            for (int jj = j - ka; jj <= j + ka; ++jj) {
              const int jl = jj + lag;
              if (jj < 0 || jj >= n1 || jl < 0 || jl >= n1) continue;
              const cf32 x = a[static_cast<std::size_t>(ii) * n1 + jj];
              const cf32 y = b[static_cast<std::size_t>(ii) * n1 + jl];
              r += std::conj(x) * y;
              ea += std::norm(x);
              eb += std::norm(y);
            }
          }
          if (ea <= 0.0f || eb <= 0.0f) continue;
          const float corr = std::abs(r) / std::sqrt(ea * eb);
          if (corr > best_corr) {
            best_corr = corr;
            best = r;
            best_lag = lag;
          }
        }
        if (best_corr < 0.0f) continue;
        correlation[k] = std::clamp(best_corr, 0.0f, 1.0f);
        axial[k] = static_cast<float>(best_lag) * dz + (-std::arg(best) * c / (4.0f * kPiF * fc));
      }
  });
}

// axial_strain (doppler.cpp:275-307): least-squares slope over the window.
int oc_axial_strain(const oc_grid* g, const float* axial, int fit_length, float* strain) {
  return guarded([&] {
    const Grid G = build_grid(*g);
    const int n0 = G.n0, n1 = G.n1;
    const float dz = axial_step(G);
    const int half = std::max(1, fit_length / 2);
    for (int i = 0; i < n0; ++i)
      for (int j = 0; j < n1; ++j) {
        float sx = 0, sy = 0, sxx = 0, sxy = 0;
        int n = 0;
        for (int dd = -half; dd <= half; ++dd) {
          const int jj = j + dd;
          if (jj < 0 || jj >= n1) continue;
          const float x = static_cast<float>(dd) * dz;
          const float y = axial[static_cast<std::size_t>(i) * n1 + jj];
          sx += x;
          sy += y;
          sxx += x * x;
          sxy += x * y;
          ++n;
        }
        float s = 0.0f;
        if (n >= 2) {
          const float den = n * sxx - sx * sx;
          if (!(std::abs(den) < 1e-20f)) s = (n * sxy - sx * sy) / den;
        }
        strain[static_cast<std::size_t>(i) * n1 + j] = s;
      }
  });
}

}  // extern "C"
