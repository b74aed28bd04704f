// device_math.hpp -- per-pixel geometry on gfx950, bit-faithful to the
// reference's FP32 expressions.
//
// Every file in this library is compiled with -ffp-contract=off and with
// correctly rounded FP32 division and square root, so an expression written
// here in the reference's operation order rounds exactly as the reference's
// x86-64 build does (-O3, no -march, hence no FMA contraction). That matters
// most for the delay and phase arithmetic: 2*pi*fdem*tau reaches ~2e3 rad,
// where one float ulp is ~2e-4 rad, so a single differently rounded step
// would move the rotated sample by ~1e-4 relative.
//
// Transcendentals: phases use the device sincosf (value error ~1 ulp, harmless).
// Window weights use cos_window(): the device cosf, replaced by a correctly
// rounded cos (evaluated in double) whenever the weight is near zero. Near
// zero the weight is a difference of nearly equal terms, and whether it
// comes out exactly 0 decides whether a channel is "active" (beamformer.cpp
// :385, 496-500), which fixes the MV active set and DMAS/CF normalisation.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

namespace usxd {

constexpr float kPi = static_cast<float>(3.14159265358979323846);
constexpr float kTwoPi = 2.0f * kPi;  // the reference's (2.0f * kPi), exact

struct Pt {
  float x, y, z;
};

// std::max / std::min semantics (the first argument wins unordered compares).
__device__ __forceinline__ float smax(float a, float b) { return (a < b) ? b : a; }
__device__ __forceinline__ float smin(float a, float b) { return (b < a) ? b : a; }

// Correctly rounded float cosine (double evaluation, one final rounding).
__device__ __forceinline__ float cos_cr(float x) {
  return static_cast<float>(cos(static_cast<double>(x)));
}

// window_value (types.cpp:126-148) with cosine `C` for the cos evaluations.
template <bool kExact>
__device__ __forceinline__ float window_eval(int w, float u, float tukey_alpha) {
  auto C = [](float x) { return kExact ? cos_cr(x) : cosf(x); };
  if (u < -1.0f || u > 1.0f) return 0.0f;
  const float t = 0.5f * (u + 1.0f);
  switch (w) {
    case 0:
      return 1.0f;
    case 1:
      return 0.5f * (1.0f - C(kTwoPi * t));
    case 2:
      return 0.54f - 0.46f * C(kTwoPi * t);
    case 4:
      return 0.42f - 0.5f * C(kTwoPi * t) + 0.08f * C((4.0f * kPi) * t);
    case 3: {
      const float a = smin(smax(tukey_alpha, 0.0f), 1.0f);
      if (a <= 0.0f) return 1.0f;
      if (t < a * 0.5f) return 0.5f * (1.0f + C(kTwoPi / a * (t - a * 0.5f)));
      if (t > 1.0f - a * 0.5f) return 0.5f * (1.0f + C(kTwoPi / a * (t - 1.0f + a * 0.5f)));
      return 1.0f;
    }
  }
  return 1.0f;
}

// Window weight whose zero/non-zero decision and small values match the
// reference's libm evaluation: fast cosf first, exact cos where the result
// is small enough for cancellation to matter.
__device__ __forceinline__ float window_value(int w, float u, float tukey_alpha) {
  const float fast = window_eval<false>(w, u, tukey_alpha);
  if (fast < 0.02f) return window_eval<true>(w, u, tukey_alpha);
  return fast;
}

// dynamic_aperture_weight (geometry.cpp:123-130) for an element at (ex, 0, 0).
__device__ __forceinline__ float aperture_weight(float ex, const Pt& p, float f_number, int w,
                                                 float tukey_alpha) {
  const float depth = p.z - 0.0f;
  if (depth <= 0.0f) return 0.0f;
  const float half = 0.5f * depth / smax(f_number, 1e-3f);
  const float u = (ex - p.x) / smax(half, 1e-9f);
  return window_value(w, u, tukey_alpha);
}

// receive_arrival_time (geometry.hpp:39-41): |element - p| / c, element on
// the face (y = z = 0), in Vec3::norm's operation order (types.hpp:52).
__device__ __forceinline__ float rx_time(float ex, const Pt& p, float c) {
  const float dx = ex - p.x;
  const float dy = 0.0f - p.y;
  const float dz = 0.0f - p.z;
  return sqrtf(dx * dx + dy * dy + dz * dz) / c;
}

}  // namespace usxd

#include "usx_gpu.h"

namespace usxd {

// ScanGrid::build_points (types.cpp:102-115): Cartesian (x, 0, z); sector
// apex + r (sin th, 0, cos th) with sin/cos tabulated on the host.
struct GridView {
  int type, n0, n1;
  const float* axis0;
  const float* axis1;
  const float* s0;
  const float* c0;
  float apex_x, apex_z;
};

__device__ __forceinline__ Pt grid_point(const GridView& g, int i0, int i1) {
  const float a = g.axis0[i0], r = g.axis1[i1];
  if (g.type == 0) return Pt{a, 0.0f, r};
  return Pt{g.apex_x + r * g.s0[i0], 0.0f, g.apex_z + r * g.c0[i0]};
}

// transmit_arrival_time (geometry.cpp:82-109) with the pixel-independent
// terms taken from the host descriptor.
__device__ __forceinline__ float tx_time(const usx_tx_desc& t, const Pt& p, float c) {
  if (t.type == 1) {
    return ((p.x - t.origin_x) * t.sin_a + (p.z - 0.0f) * t.cos_a) / c;
  }
  if (t.type == 0) {
    if (t.axis_len < 1e-9f) return sqrtf(p.x * p.x + p.y * p.y + p.z * p.z) / c;
    const float rx = p.x - t.focus[0], ry = p.y - t.focus[1], rz = p.z - t.focus[2];
    const float proj = rx * t.axis_dir[0] + ry * t.axis_dir[1] + rz * t.axis_dir[2];
    const float sign = (proj >= 0.0f) ? 1.0f : -1.0f;
    return (t.axis_len + sign * sqrtf(rx * rx + ry * ry + rz * rz)) / c;
  }
  if (t.type == 2) {
    const float dx = p.x - t.focus[0], dy = p.y - t.focus[1], dz = p.z - t.focus[2];
    return (sqrtf(dx * dx + dy * dy + dz * dz) - t.origin_focus) / c;
  }
  return 0.0f;
}

// tx_sensitivity (beamformer.cpp:249-288): Tukey(0.4) taper over the
// illuminated band widened by the diffraction-limited beam width. Always
// uses the correctly rounded cosine: Select compounding takes the argmax of
// this weight over transmits, so its exact value can matter.
__device__ __forceinline__ float tx_weight(const usx_tx_desc& t, const Pt& p, float lambda) {
  if (!t.extent_valid || p.z <= 0.0f) return 0.0f;
  float a, b;
  if (t.type == 1) {
    const float shift = p.z * t.tan_a;
    a = t.x_lo + shift;
    b = t.x_hi + shift;
  } else {
    const float zf = t.focus[2];
    if (fabsf(zf) < 1e-9f) return 1.0f;
    a = t.x_lo + (t.focus[0] - t.x_lo) * p.z / zf;
    b = t.x_hi + (t.focus[0] - t.x_hi) * p.z / zf;
    if (a > b) {
      const float s = a;
      a = b;
      b = s;
    }
  }
  const float fne = smin(smax(p.z / t.aperture, 0.5f), 20.0f);
  const float diffraction = fne * lambda;
  a -= diffraction;
  b += diffraction;
  const float centre = 0.5f * (a + b);
  const float half = 0.5f * (b - a);
  if (half <= 0.0f) return 0.0f;
  return window_eval<true>(3, (p.x - centre) / half, 0.4f);
}

}  // namespace usxd
