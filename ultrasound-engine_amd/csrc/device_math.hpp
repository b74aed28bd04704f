// device_math.hpp -- per-pixel geometry on gfx950, faithful to the
// reference's FP32 expressions.
//
// Every file in this library is compiled with -ffp-contract=off and with
// correctly rounded FP32 division and square root, so an expression written
// here in the reference's operation order rounds exactly as the reference's
// x86-64 build does (-O3, no -march, hence no FMA contraction). That matters
// most for the delay arithmetic: an interpolation index or a phase
// 2*pi*fdem*tau (~2e3 rad, where one float ulp is ~2e-4 rad) rounded
// differently would move the rotated sample by ~1e-4 relative.
//
// Divisions by a divisor that many quotients share (the speed of sound, a
// pixel's half aperture) use Markstein's correction step, q = x r,
// e = fma(-q, d, x), q' = fma(e, r, q) with r = 1/d correctly rounded, which
// returns the correctly rounded quotient x/d (P. Markstein, IBM J. Res. Dev.
// 34(1), 1990); checked against IEEE division on every float distance over
// c = 1540 and on 4e8 random pairs, with no difference.
//
// Transcendentals: phases are reduced exactly by 2 pi (Cody-Waite, two
// terms, fma) and evaluated by the hardware sine and cosine (error ~1e-7
// absolute, harmless for the 1e-5 bar). Window weights: what must be exact
// is whether a weight is 0, because that decides whether a channel is
// "active" (beamformer.cpp:385, 496-500), which fixes the MV active set and
// the DMAS/CF normalisation. For the Hann window (and rect, Hamming) that is
// decided exactly from thresholds on the cosine's argument, measured on the
// host's own cosf (where it returns exactly 1.0f); Tukey and Blackman keep the
// device cosf with a correctly rounded (double) cosine near zero.
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

// x / d, correctly rounded, from r = 1/d correctly rounded (Markstein's
// correction step; d > 0, no underflow).
__device__ __forceinline__ float div_r(float x, float d, float r) {
  const float q = x * r;
  const float e = fmaf(-q, d, x);
  return fmaf(e, r, q);
}

// cos and sin of a float phase. The phase is reduced by 2 pi with the
// whole-turn count k: phase - k P1 is a multiple of 2^-21 below 4 in
// magnitude (P1 is, and so is a float phase of magnitude >= 2), so one fma
// forms it exactly; the second term corrects P1 to 2 pi, and the hardware
// sin/cos take the remaining fraction of a turn.
__device__ __forceinline__ void rot_of(float phase, float* c, float* s) {
  constexpr float kInv2Pi = 0.15915494309189535f;
  constexpr float kP1 = 6.28318548202514648f;      // RN(2 pi)
  constexpr float kP2 = -1.7484556000744487e-07f;  // RN(2 pi - P1)
  const float k = rintf(phase * kInv2Pi);
  float r = fmaf(-k, kP1, phase);
  r = fmaf(-k, kP2, r);
  const float turns = r * kInv2Pi;
  *c = __builtin_amdgcn_cosf(turns);
  *s = __builtin_amdgcn_sinf(turns);
}

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

// window_value kept out of line: the Tukey and Blackman receive windows are
// off the common path, and inlined, the constants of its double-precision
// cosine would hold registers through the callers' channel loops.
__device__ __attribute__((noinline)) float window_value_slow(int w, float u, float tukey_alpha) {
  return window_value(w, u, tukey_alpha);
}

// Receive apodization of an element at lateral offset dx = x_e - x_p from
// the pixel, given d = max(half aperture, 1e-9) and r = 1/d:
// dynamic_aperture_weight (geometry.cpp:123-130) -> window_value
// (types.cpp:126-148). *on is whether the reference's weight is non-zero,
// exactly (a NaN weight counts as non-zero, as `a <= 0` is false for it);
// the value is the reference's within a few 1e-8. Hann's 0.5 (1 - cos x) is
// 0 exactly where the C library's cosf(x) returns 1.0f, i.e. for
// x <= cos1_lo or x >= cos1_hi; rect and Hamming are non-zero on the whole
// aperture; Tukey and Blackman take window_value. HANN: the window is known
// to be Hann (w == 1), so no call to the out-of-line window_value_slow is
// compiled in -- a call in the channel loop makes the compiler keep the
// loop's scalar values in VGPR lanes (a readlane per use).
template <bool HANN = false>
__device__ __forceinline__ float rx_weight(int w, float tukey_alpha, float dx, float d, float r,
                                           float cos1_lo, float cos1_hi, bool* on) {
  if (HANN) w = 1;
  const float u = div_r(dx, d, r);
  if (u < -1.0f || u > 1.0f) {
    *on = false;
    return 0.0f;
  }
  const float t = 0.5f * (u + 1.0f);
  if (w == 0) {
    *on = true;
    return 1.0f;
  }
  if (w == 1 || w == 2) {
    const float cs = __builtin_amdgcn_cosf(t);  // cos(2 pi t)
    if (w == 2) {
      *on = true;
      return 0.54f - 0.46f * cs;
    }
    const float x = kTwoPi * t;
    *on = !(x <= cos1_lo || x >= cos1_hi);
    return 0.5f * (1.0f - cs);
  }
  if constexpr (HANN) {
    return 0.0f;  // unreachable: w == 1
  } else {
    const float a = window_value_slow(w, u, tukey_alpha);
    *on = !(a <= 0.0f);
    return a;
  }
}

// Correctly rounded square root of a positive normal float, by the
// correction the compiler itself emits for sqrtf (the hardware root, then the
// neighbouring floats tested by their fma residuals) without its rescaling of
// inputs below 2^-96 and its infinity/zero tests.
__device__ __forceinline__ float sqrt_normal(float x) {
  const float s = __builtin_amdgcn_sqrtf(x);
  const float lo = __int_as_float(__float_as_int(s) - 1);
  const float hi = __int_as_float(__float_as_int(s) + 1);
  float r = (fmaf(-lo, s, x) <= 0.0f) ? lo : s;
  r = (fmaf(-hi, s, x) > 0.0f) ? hi : r;
  return r;
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
__device__ __forceinline__ float rx_time(float ex, const Pt& p, float c, float inv_c) {
  const float dx = ex - p.x;
  const float dy = 0.0f - p.y;
  const float dz = 0.0f - p.z;
  return div_r(sqrtf(dx * dx + dy * dy + dz * dz), c, inv_c);
}
// The same for pixels deeper than 1e-12 m, where the squared distance is a
// positive normal float.
__device__ __forceinline__ float rx_time_normal(float ex, const Pt& p, float c, float inv_c) {
  const float dx = ex - p.x;
  const float dy = 0.0f - p.y;
  const float dz = 0.0f - p.z;
  return div_r(sqrt_normal(dx * dx + dy * dy + dz * dz), c, inv_c);
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
// (inv_c = 1/c correctly rounded: the divisions by c are div_r's.)
__device__ __forceinline__ float tx_time(const usx_tx_desc& t, const Pt& p, float c, float inv_c) {
  if (t.type == 1) {
    return div_r((p.x - t.origin_x) * t.sin_a + (p.z - 0.0f) * t.cos_a, c, inv_c);
  }
  if (t.type == 0) {
    if (t.axis_len < 1e-9f) return div_r(sqrtf(p.x * p.x + p.y * p.y + p.z * p.z), c, inv_c);
    const float rx = p.x - t.focus[0], ry = p.y - t.focus[1], rz = p.z - t.focus[2];
    const float proj = rx * t.axis_dir[0] + ry * t.axis_dir[1] + rz * t.axis_dir[2];
    const float sign = (proj >= 0.0f) ? 1.0f : -1.0f;
    return div_r(t.axis_len + sign * sqrtf(rx * rx + ry * ry + rz * rz), c, inv_c);
  }
  if (t.type == 2) {
    const float dx = p.x - t.focus[0], dy = p.y - t.focus[1], dz = p.z - t.focus[2];
    return div_r(sqrtf(dx * dx + dy * dy + dz * dz) - t.origin_focus, c, inv_c);
  }
  return 0.0f;
}

// tx_sensitivity (beamformer.cpp:249-288): Tukey(0.4) taper over the
// illuminated band widened by the diffraction-limited beam width. EXACT:
// the correctly rounded cosine, for Select compounding, which takes the
// argmax of this weight over transmits, so its exact value can matter;
// otherwise the weight only scales a transmit's contribution, and the device
// cosf (an ulp or two) is used.
template <bool EXACT>
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
  return window_eval<EXACT>(3, (p.x - centre) / half, 0.4f);
}

}  // namespace usxd
