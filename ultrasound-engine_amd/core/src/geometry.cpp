// geometry.cpp -- delay laws and arrival times (host).
// Follows the reference's core/src/geometry.cpp:17-130; float expressions
// kept in the reference's order because the simulator's RF and the
// beamformer's delays must agree bitwise with it.
#include "usx/geometry.hpp"

#include <algorithm>
#include <cmath>

namespace usx {

void compute_transmit_delays(const Probe& probe, const Medium& medium, Transmit& tx,
                             Window apod_window, real f_number, real* t0_offset) {
  const real c = medium.speed_of_sound;
  const int n = probe.n_elements;
  tx.delays.assign(n, 0.0f);
  tx.apodization.assign(n, 0.0f);
  const Vec3 O(tx.origin.x, 0.0f, 0.0f);

  // Firing half-aperture: depth / F# for a focused beam, the full array
  // otherwise, never wider than the array.
  real half = 0.5f * probe.aperture_width();
  if (tx.type == TransmitType::Focused)
    half = 0.5f * std::abs(tx.focus.z - O.z) / std::max(f_number, 0.1f);
  half = std::min(half, 0.5f * probe.aperture_width());

  bool any = false;
  real tmin = 0.0f;
  for (int i = 0; i < n; ++i) {
    const Vec3 p = probe.element_position(i);
    const real w = window_value(apod_window, (p.x - O.x) / std::max(half, 1e-9f));
    if (w <= 0.0f) continue;
    tx.apodization[i] = w;
    real d = 0.0f;
    if (tx.type == TransmitType::PlaneWave)
      d = (p.x - O.x) * std::sin(tx.angle) / c;
    else if (tx.type == TransmitType::Focused)
      d = (distance(tx.focus, O) - distance(tx.focus, p)) / c;
    else
      d = (distance(p, tx.focus) - distance(O, tx.focus)) / c;
    tx.delays[i] = d;
  }
  for (int i = 0; i < n; ++i) {
    if (tx.apodization[i] == 0.0f) continue;
    tmin = any ? std::min(tmin, tx.delays[i]) : tx.delays[i];
    any = true;
  }
  if (t0_offset) *t0_offset = any ? tmin : 0.0f;
}

real transmit_arrival_time(const Transmit& tx, const Vec3& point, real c) {
  const Vec3 O(tx.origin.x, 0.0f, 0.0f);
  switch (tx.type) {
    case TransmitType::PlaneWave: {
      const real s = std::sin(tx.angle), k = std::cos(tx.angle);
      return ((point.x - O.x) * s + (point.z - O.z) * k) / c;
    }
    case TransmitType::Focused: {
      const Vec3 axis = tx.focus - O;
      const real len = axis.norm();
      if (len < 1e-9f) return point.norm() / c;
      const Vec3 d(axis.x / len, axis.y / len, axis.z / len);
      const Vec3 rel = point - tx.focus;
      const real proj = rel.x * d.x + rel.y * d.y + rel.z * d.z;
      return (len + ((proj >= 0.0f) ? 1.0f : -1.0f) * rel.norm()) / c;
    }
    case TransmitType::Diverging:
      return (distance(point, tx.focus) - distance(O, tx.focus)) / c;
  }
  return 0.0f;
}

real element_directivity(real element_width, real wavelength, real sin_theta) {
  const real st = std::clamp(sin_theta, -1.0f, 1.0f);
  const real ct = std::sqrt(std::max(0.0f, 1.0f - st * st));
  const real x = kPi * element_width * st / wavelength;
  const real sinc = (std::abs(x) < 1e-6f) ? 1.0f - x * x / 6.0f : std::sin(x) / x;
  return ct * sinc;
}

real dynamic_aperture_weight(const Vec3& element, const Vec3& point, real f_number, Window w,
                             real tukey_alpha) {
  const real depth = point.z - element.z;
  if (depth <= 0.0f) return 0.0f;
  const real half = 0.5f * depth / std::max(f_number, 1e-3f);
  return window_value(w, (element.x - point.x) / std::max(half, 1e-9f), tukey_alpha);
}

}  // namespace usx
