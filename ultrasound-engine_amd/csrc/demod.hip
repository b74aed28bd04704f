// demod.hip -- RF to baseband IQ: mix, FIR low-pass, decimate (gfx950).
//
// Reference: demodulate, core/src/demodulate.cpp:36-121.
//
// One workgroup per (event, channel) trace. The trace is read once from HBM
// (coalesced), mixed with the host-tabulated rotator exp(-j 2 pi fdem t)
// (demodulate.cpp:72-76, 98-99) into LDS, and every output sample k is then
// the FIR sum over taps j = j_lo..j_hi in ascending order
//   iq[k] = 2 * sum_j mixed[k*dec + gd - j] * h[j]   (demodulate.cpp:106-116)
// with the reference's edge clipping and the same multiply/add sequence, so
// the output is bitwise the reference's. Traces longer than the LDS budget
// take the same arithmetic straight from global memory.
#include <hip/hip_runtime.h>

#include "kernels.hpp"

namespace usxd {

namespace {

constexpr int kDemodLdsMax = 8192;  // complex samples staged per trace (64 KiB)

__device__ __forceinline__ float2 fir_tap(float2 acc, float2 m, float h) {
  acc.x += m.x * h;
  acc.y += m.y * h;
  return acc;
}

}  // namespace

__global__ __launch_bounds__(256) void demod_lds_kernel(int n_in, int n_out, int dec, int n_taps,
                                                        const float* __restrict__ taps,
                                                        const float2* __restrict__ rot,
                                                        const float* __restrict__ rf,
                                                        float2* __restrict__ iq) {
  extern __shared__ float2 mixed[];
  const size_t tr = blockIdx.x;
  const float* x = rf + tr * n_in;
  for (int n = threadIdx.x; n < n_in; n += blockDim.x) {
    const float v = x[n];
    const float2 r = rot[n];
    mixed[n] = make_float2(r.x * v, r.y * v);
  }
  __syncthreads();
  const int gd = n_taps / 2;
  float2* y = iq + tr * n_out;
  for (int k = threadIdx.x; k < n_out; k += blockDim.x) {
    const int c = k * dec + gd;
    const int j0 = max(0, c - n_in + 1);
    const int j1 = min(n_taps - 1, c);
    float2 acc = make_float2(0.0f, 0.0f);
    for (int j = j0; j <= j1; ++j) acc = fir_tap(acc, mixed[c - j], taps[j]);
    y[k] = make_float2(acc.x * 2.0f, acc.y * 2.0f);
  }
}

__global__ __launch_bounds__(256) void demod_global_kernel(int n_traces, int n_in, int n_out,
                                                           int dec, int n_taps,
                                                           const float* __restrict__ taps,
                                                           const float2* __restrict__ rot,
                                                           const float* __restrict__ rf,
                                                           float2* __restrict__ iq) {
  const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= static_cast<size_t>(n_traces) * n_out) return;
  const size_t tr = i / n_out;
  const int k = static_cast<int>(i - tr * n_out);
  const float* x = rf + tr * n_in;
  const int gd = n_taps / 2;
  const int c = k * dec + gd;
  const int j0 = max(0, c - n_in + 1);
  const int j1 = min(n_taps - 1, c);
  float2 acc = make_float2(0.0f, 0.0f);
  for (int j = j0; j <= j1; ++j) {
    const int n = c - j;
    const float v = x[n];
    const float2 r = rot[n];
    acc = fir_tap(acc, make_float2(r.x * v, r.y * v), taps[j]);
  }
  iq[i] = make_float2(acc.x * 2.0f, acc.y * 2.0f);
}

void launch_demod(int n_traces, int n_in, int n_out, int dec, int n_taps, const float* taps,
                  const float2* rotator, const float* rf, float2* iq, hipStream_t st) {
  if (n_in <= kDemodLdsMax) {
    hipLaunchKernelGGL(demod_lds_kernel, dim3(n_traces), dim3(256),
                       static_cast<size_t>(n_in) * sizeof(float2), st, n_in, n_out, dec, n_taps,
                       taps, rotator, rf, iq);
  } else {
    const size_t n = static_cast<size_t>(n_traces) * n_out;
    hipLaunchKernelGGL(demod_global_kernel, dim3(static_cast<unsigned>((n + 255) / 256)),
                       dim3(256), 0, st, n_traces, n_in, n_out, dec, n_taps, taps, rotator, rf,
                       iq);
  }
}

}  // namespace usxd
