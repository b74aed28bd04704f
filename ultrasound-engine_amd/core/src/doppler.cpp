// doppler.cpp -- colour / power Doppler and axial speckle tracking on the GPU.
// Reference: core/src/doppler.cpp (polynomial_reject_matrix 15-49,
// apply_wall_filter 51-75, color_flow 77-176, power_doppler 178-186,
// estimate_axial_displacement 188-273, axial_strain 275-307). The host
// validates, builds the wall-filter projector and the per-call constants with
// the reference's float expressions, and packs the frames; the per-pixel work
// is usx_wall_filter / usx_color_flow / usx_power_doppler /
// usx_axial_displacement / usx_axial_strain (include/usx_gpu.h).
#include "usx/doppler.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <stdexcept>

#include "gpu.hpp"

namespace usx {

std::vector<real> polynomial_reject_matrix(int n, int order) {
  if (n < 1) throw std::invalid_argument("polynomial_reject_matrix: n must be >= 1");
  const int degree = std::clamp(order, 0, n - 1);
  const std::size_t un = static_cast<std::size_t>(n);
  // Orthonormal slow-time basis: monomials x^d on x in [-1, 1] (x = 0 for a
  // single frame), Gram-Schmidt against the vectors kept so far; a vector
  // that vanishes is dropped.
  std::vector<std::vector<real>> q;
  for (int d = 0; d <= degree; ++d) {
    std::vector<real> v(un);
    for (int t = 0; t < n; ++t) {
      const real x = (n == 1) ? 0.0f : 2.0f * t / (n - 1) - 1.0f;
      v[static_cast<std::size_t>(t)] = std::pow(x, d);
    }
    for (const std::vector<real>& u : q) {
      real dot = 0;
      for (std::size_t t = 0; t < un; ++t) dot += v[t] * u[t];
      for (std::size_t t = 0; t < un; ++t) v[t] -= dot * u[t];
    }
    real len = 0;
    for (real x : v) len += x * x;
    len = std::sqrt(len);
    if (len < 1e-8f) continue;
    for (real& x : v) x /= len;
    q.push_back(std::move(v));
  }
  // I - sum_u u u^T, subtracting the basis vectors' outer products in order.
  std::vector<real> m(un * un, 0.0f);
  for (std::size_t i = 0; i < un; ++i) m[i * un + i] = 1.0f;
  for (const std::vector<real>& u : q)
    for (std::size_t i = 0; i < un; ++i)
      for (std::size_t j = 0; j < un; ++j) m[i * un + j] -= u[i] * u[j];
  return m;
}

namespace {

real axial_spacing(const ScanGrid& g) {
  real dz = 1e-4f;
  if (g.n_axis1 > 1 && g.axis1.size() > 1) dz = std::abs(g.axis1[1] - g.axis1[0]);
  return dz;
}

// Frames laid one after the other, after checking they share one size.
std::vector<cplx> pack(const std::vector<Frame>& ens, const char* who) {
  const std::size_t n_pix = ens.front().data.size();
  if (n_pix != static_cast<std::size_t>(ens.front().grid.n_pixels()))
    throw std::invalid_argument(std::string(who) + ": frame data does not match its grid");
  for (const Frame& f : ens)
    if (f.data.size() != n_pix)
      throw std::invalid_argument(std::string(who) + ": frames have different sizes");
  std::vector<cplx> buf(n_pix * ens.size());
  for (std::size_t e = 0; e < ens.size(); ++e)
    std::memcpy(buf.data() + e * n_pix, ens[e].data.data(), n_pix * sizeof(cplx));
  return buf;
}

std::vector<real> reject_for(WallFilter type, int n, int order) {
  if (type == WallFilter::None) return {};
  return polynomial_reject_matrix(n, type == WallFilter::MeanSubtraction ? 0 : order);
}

usx_ensemble_desc ensemble_desc(const std::vector<Frame>& ens, const std::vector<real>& reject) {
  usx_ensemble_desc d{};
  d.n_frames = static_cast<int>(ens.size());
  d.n_axis0 = ens.front().grid.n_axis0;
  d.n_axis1 = ens.front().grid.n_axis1;
  d.reject = reject.empty() ? nullptr : reject.data();
  return d;
}

}  // namespace

std::vector<Frame> apply_wall_filter(const std::vector<Frame>& ensemble, WallFilter type,
                                     int polynomial_order) {
  if (ensemble.empty()) return {};
  if (type == WallFilter::None) return ensemble;
  const std::vector<cplx> in = pack(ensemble, "apply_wall_filter");
  const std::vector<real> reject =
      reject_for(type, static_cast<int>(ensemble.size()), polynomial_order);
  const usx_ensemble_desc d = ensemble_desc(ensemble, reject);
  std::vector<cplx> out(in.size());
  usx_ctx* ctx = thread_context().get();
  check(ctx,
        usx_wall_filter(ctx, &d, reinterpret_cast<const usx_c32*>(in.data()),
                        reinterpret_cast<usx_c32*>(out.data()), USX_HOST),
        "apply_wall_filter");
  std::vector<Frame> res = ensemble;
  const std::size_t n_pix = ensemble.front().data.size();
  for (std::size_t e = 0; e < res.size(); ++e)
    std::memcpy(res[e].data.data(), out.data() + e * n_pix, n_pix * sizeof(cplx));
  return res;
}

ColorFlowMap color_flow(const std::vector<Frame>& ensemble, real prf, real speed_of_sound,
                        DopplerConfig config) {
  if (ensemble.size() < 2)
    throw std::invalid_argument("color_flow: need at least 2 frames in the ensemble");
  const std::vector<cplx> in =
      pack(ensemble, config.wall_filter == WallFilter::None ? "color_flow" : "apply_wall_filter");
  const std::vector<real> reject =
      reject_for(config.wall_filter, static_cast<int>(ensemble.size()), config.polynomial_order);
  const usx_ensemble_desc d = ensemble_desc(ensemble, reject);
  const ScanGrid& grid = ensemble.front().grid;
  const real fdem = ensemble.front().demodulation_frequency;

  usx_color_flow_desc f{};
  f.prf = prf;
  f.speed_of_sound = speed_of_sound;
  f.demodulation_frequency = fdem;
  f.dz = axial_spacing(grid);
  f.loupas = config.loupas ? 1 : 0;
  f.kernel_axial = config.kernel_axial;
  f.kernel_lateral = config.kernel_lateral;
  f.coherence_threshold = config.coherence_threshold;
  f.power_factor = std::pow(10.0f, config.power_threshold_db / 10.0f);

  ColorFlowMap out;
  out.grid = grid;
  const std::size_t n_pix = ensemble.front().data.size();
  out.velocity.assign(n_pix, 0.0f);
  out.power.assign(n_pix, 0.0f);
  out.variance.assign(n_pix, 0.0f);
  out.valid.assign(n_pix, 0);
  out.nyquist_velocity = (fdem > 0.0f) ? speed_of_sound * prf / (4.0f * fdem) : 0.0f;
  usx_ctx* ctx = thread_context().get();
  check(ctx,
        usx_color_flow(ctx, &d, &f, reinterpret_cast<const usx_c32*>(in.data()),
                       out.velocity.data(), out.power.data(), out.variance.data(),
                       out.valid.data(), USX_HOST),
        "color_flow");
  return out;
}

std::vector<real> power_doppler(const std::vector<Frame>& ensemble, WallFilter type,
                                int polynomial_order) {
  if (ensemble.empty()) return {};
  const std::vector<cplx> in = pack(ensemble, "power_doppler");
  const std::vector<real> reject =
      reject_for(type, static_cast<int>(ensemble.size()), polynomial_order);
  const usx_ensemble_desc d = ensemble_desc(ensemble, reject);
  std::vector<real> p(ensemble.front().data.size(), 0.0f);
  usx_ctx* ctx = thread_context().get();
  check(ctx,
        usx_power_doppler(ctx, &d, reinterpret_cast<const usx_c32*>(in.data()), p.data(),
                          USX_HOST),
        "power_doppler");
  return p;
}

DisplacementMap estimate_axial_displacement(const Frame& before, const Frame& after,
                                            real speed_of_sound, int kernel_axial,
                                            int kernel_lateral, int max_lag) {
  if (before.data.size() != after.data.size())
    throw std::invalid_argument("estimate_axial_displacement: frame size mismatch");
  const ScanGrid& grid = before.grid;
  if (before.data.size() != static_cast<std::size_t>(grid.n_pixels()))
    throw std::invalid_argument("estimate_axial_displacement: frame data does not match its grid");
  DisplacementMap out;
  out.grid = grid;
  const std::size_t n_pix = before.data.size();
  out.axial.assign(n_pix, 0.0f);
  out.correlation.assign(n_pix, 0.0f);
  usx_displacement_desc d{};
  d.n_axis0 = grid.n_axis0;
  d.n_axis1 = grid.n_axis1;
  d.speed_of_sound = speed_of_sound;
  d.demodulation_frequency = before.demodulation_frequency;
  d.dz = axial_spacing(grid);
  d.kernel_axial = kernel_axial;
  d.kernel_lateral = kernel_lateral;
  d.max_lag = max_lag;
  usx_ctx* ctx = thread_context().get();
  check(ctx,
        usx_axial_displacement(ctx, &d, reinterpret_cast<const usx_c32*>(before.data.data()),
                               reinterpret_cast<const usx_c32*>(after.data.data()),
                               out.axial.data(), out.correlation.data(), USX_HOST),
        "estimate_axial_displacement");
  return out;
}

std::vector<real> axial_strain(const DisplacementMap& disp, int fit_length) {
  const ScanGrid& g = disp.grid;
  const std::size_t n_pix = static_cast<std::size_t>(g.n_pixels());
  if (disp.axial.size() != n_pix)
    throw std::invalid_argument("axial_strain: displacement does not match its grid");
  std::vector<real> strain(n_pix, 0.0f);
  usx_ctx* ctx = thread_context().get();
  check(ctx,
        usx_axial_strain(ctx, g.n_axis0, g.n_axis1, axial_spacing(g), fit_length,
                         disp.axial.data(), strain.data(), USX_HOST),
        "axial_strain");
  return strain;
}

}  // namespace usx
