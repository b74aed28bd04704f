// postprocess.cpp -- envelope, TGC and log compression (host).
// Reference: core/src/postprocess.cpp:12-117.
#include "usx/postprocess.hpp"

#include <algorithm>
#include <cmath>

namespace usx {
namespace {

// In-place iterative radix-2 FFT (n a power of two).
void fft_inplace(std::vector<cplx>& a, bool inverse) {
  const std::size_t n = a.size();
  if (n < 2) return;
  for (std::size_t i = 1, j = 0; i < n; ++i) {
    std::size_t bit = n >> 1;
    for (; j & bit; bit >>= 1) j ^= bit;
    j ^= bit;
    if (i < j) std::swap(a[i], a[j]);
  }
  for (std::size_t len = 2; len <= n; len <<= 1) {
    const real ang = 2.0f * kPi / static_cast<real>(len) * (inverse ? 1.0f : -1.0f);
    const cplx step(std::cos(ang), std::sin(ang));
    for (std::size_t i = 0; i < n; i += len) {
      cplx w(1, 0);
      for (std::size_t k = 0; k < len / 2; ++k) {
        const cplx u = a[i + k];
        const cplx v = a[i + k + len / 2] * w;
        a[i + k] = u + v;
        a[i + k + len / 2] = u - v;
        w *= step;
      }
    }
  }
  if (inverse)
    for (cplx& v : a) v /= static_cast<real>(n);
}

}  // namespace

std::vector<cplx> analytic_along_axis1(const std::vector<real>& data, int n_axis0, int n_axis1) {
  std::vector<cplx> out(static_cast<std::size_t>(n_axis0) * n_axis1);
  std::size_t n_fft = 1;
  while (n_fft < static_cast<std::size_t>(n_axis1)) n_fft <<= 1;
  std::vector<cplx> buf(n_fft);
  for (int i = 0; i < n_axis0; ++i) {
    std::fill(buf.begin(), buf.end(), cplx(0, 0));
    for (int j = 0; j < n_axis1; ++j) buf[j] = cplx(data[static_cast<std::size_t>(i) * n_axis1 + j], 0.0f);
    fft_inplace(buf, false);
    for (std::size_t k = 1; k < n_fft / 2; ++k) buf[k] *= 2.0f;
    for (std::size_t k = n_fft / 2 + 1; k < n_fft; ++k) buf[k] = cplx(0, 0);
    fft_inplace(buf, true);
    for (int j = 0; j < n_axis1; ++j) out[static_cast<std::size_t>(i) * n_axis1 + j] = buf[j];
  }
  return out;
}

std::vector<real> envelope(const Frame& frame) {
  const std::size_t n = static_cast<std::size_t>(frame.grid.n_axis0) * frame.grid.n_axis1;
  std::vector<real> env(n);
  if (frame.demodulation_frequency > 0.0f) {
    for (std::size_t k = 0; k < n; ++k) env[k] = std::abs(frame.data[k]);
    return env;
  }
  std::vector<real> re(n);
  for (std::size_t k = 0; k < n; ++k) re[k] = frame.data[k].real();
  const std::vector<cplx> an = analytic_along_axis1(re, frame.grid.n_axis0, frame.grid.n_axis1);
  for (std::size_t k = 0; k < n; ++k) env[k] = std::abs(an[k]);
  return env;
}

std::vector<real> apply_tgc(const std::vector<real>& env, const ScanGrid& grid,
                            real attenuation_db_cm_mhz, real center_frequency_hz) {
  std::vector<real> out(env.size());
  const real f_mhz = center_frequency_hz * 1e-6f;
  for (int i = 0; i < grid.n_axis0; ++i)
    for (int j = 0; j < grid.n_axis1; ++j) {
      const std::size_t k = static_cast<std::size_t>(i) * grid.n_axis1 + j;
      const real depth_cm = grid.points[k].z * 100.0f;
      const real gain_db = 2.0f * attenuation_db_cm_mhz * f_mhz * std::max(depth_cm, 0.0f);
      out[k] = env[k] * std::pow(10.0f, gain_db / 20.0f);
    }
  return out;
}

std::vector<real> log_compress(const std::vector<real>& env, real dynamic_range_db, real gain_db,
                               real reference) {
  real ref = reference;
  if (ref <= 0.0f) {
    ref = 0.0f;
    for (real v : env) ref = std::max(ref, v);
  }
  if (ref <= 0.0f) return std::vector<real>(env.size(), 0.0f);
  std::vector<real> out(env.size());
  const real dr = std::max(dynamic_range_db, 1.0f);
  for (std::size_t k = 0; k < env.size(); ++k) {
    const real v = env[k] / ref;
    const real db = (v > 0.0f) ? 20.0f * std::log10(v) : -1e6f;
    out[k] = std::clamp((db + gain_db + dr) / dr, 0.0f, 1.0f);
  }
  return out;
}

}  // namespace usx
