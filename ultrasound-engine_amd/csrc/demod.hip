// demod.hip -- RF to baseband IQ: mix, FIR low-pass, decimate (gfx950).
//
// Reference: demodulate, core/src/demodulate.cpp:36-121.
//
// Every output sample k of a trace is
//   iq[k] = 2 * sum_j mixed[k*dec + gd - j] * h[j],  j = j_lo..j_hi ascending
// (demodulate.cpp:106-116), mixed[n] = rotator[n] * rf[n] with the
// host-tabulated rotator exp(-j 2 pi fdem t) (demodulate.cpp:72-76, 98-99),
// with the reference's edge clipping and the same multiply/add sequence, so
// the output is bitwise the reference's.
//
// demod_tile_kernel (33 taps, decimation 1, 2 or 4: the configurations'
// filters): one wavefront per 64 consecutive outputs of one trace, four per
// workgroup (so the CUs' workgroup slots do not cap residency). It mixes
// the input window those outputs read (64 dec + 32 samples, coalesced) into
// LDS split by decimation phase -- sample i of the window at
// [i mod dec][i / dec] -- so that for each tap j all 64 lanes read
// consecutive entries of one phase row (no bank conflicts) at an offset the
// compiler knows. Interior tiles run the 33 taps without edge tests.
// demod_lds_kernel / demod_global_kernel: any other filter length or
// decimation (one workgroup per trace, or straight from global memory for
// traces longer than the LDS budget).
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

template <int DEC, int NT>
__global__ __launch_bounds__(256) void demod_tile_kernel(int n_in, int n_out, int n_tiles,
                                                         int tiles_per_trace,
                                                         const float* __restrict__ taps,
                                                         const float2* __restrict__ rot,
                                                         const float* __restrict__ rf,
                                                         float2* __restrict__ iq) {
  constexpr int GD = NT / 2;
  constexpr int W = 63 * DEC + NT;            // input samples the tile reads
  constexpr int S = (W + DEC - 1) / DEC;      // entries per phase row
  __shared__ float2 wins[4][DEC * S];
  // four tiles per workgroup, one per wavefront (tiles numbered trace-major)
  const int wv = threadIdx.x >> 6;
  // a wavefront past the last tile repeats it and stores nothing (every
  // wavefront reaches the barrier)
  const int tile = min(blockIdx.x * 4 + wv, n_tiles - 1);
  const bool store = blockIdx.x * 4 + wv < n_tiles;
  float2* win = wins[wv];
  const int lane = threadIdx.x & 63;
  const size_t tr = static_cast<size_t>(tile / tiles_per_trace);
  const int k0 = (tile - static_cast<int>(tr) * tiles_per_trace) * 64;
  const int n_lo = k0 * DEC + GD - (NT - 1);  // input index of window sample 0
  const float* x = rf + tr * n_in;
  // every load of the window issued before any is used: one memory round trip
  constexpr int NL = (W + 63) / 64;
  float v[NL];
  float2 r[NL];
#pragma unroll
  for (int q = 0; q < NL; ++q) {
    const int i = lane + 64 * q;
    const int n = n_lo + i;
    const bool in = i < W && n >= 0 && n < n_in;
    v[q] = in ? x[n] : 0.0f;
    r[q] = in ? rot[n] : make_float2(0.0f, 0.0f);
  }
#pragma unroll
  for (int q = 0; q < NL; ++q) {
    const int i = lane + 64 * q;
    if (i < W) win[(i % DEC) * S + i / DEC] = make_float2(r[q].x * v[q], r[q].y * v[q]);
  }
  __syncthreads();
  const int k = k0 + lane;
  // output k, tap j reads window sample lane*DEC + (NT - 1 - j)
  float2 acc = make_float2(0.0f, 0.0f);
  const bool interior = n_lo >= 0 && n_lo + W <= n_in;  // every lane's j range is 0..NT-1
  if (interior) {
#pragma unroll
    for (int j = 0; j < NT; ++j) {
      const int i = NT - 1 - j;
      const float2 m = win[(i % DEC) * S + i / DEC + lane];
      const float h = taps[j];
      acc.x += m.x * h;
      acc.y += m.y * h;
    }
  } else {
    const int c = k * DEC + GD;
    const int j0 = max(0, c - n_in + 1);
    const int j1 = min(NT - 1, c);
#pragma unroll
    for (int j = 0; j < NT; ++j) {
      const int i = NT - 1 - j;
      const float2 m = win[(i % DEC) * S + i / DEC + lane];
      const float h = taps[j];
      if (j >= j0 && j <= j1) {
        acc.x += m.x * h;
        acc.y += m.y * h;
      }
    }
  }
  if (store && k < n_out) iq[tr * n_out + k] = make_float2(acc.x * 2.0f, acc.y * 2.0f);
}

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
  if (n_taps == 33 && (dec == 1 || dec == 2 || dec == 4)) {
    const int per = (n_out + 63) / 64;
    const int n_tiles = n_traces * per;
    const dim3 grid((n_tiles + 3) / 4);
#define USX_DEMOD(D)                                                                          \
  hipLaunchKernelGGL((demod_tile_kernel<D, 33>), grid, dim3(256), 0, st, n_in, n_out, n_tiles, \
                     per, taps, rotator, rf, iq)
    if (dec == 1) USX_DEMOD(1);
    else if (dec == 2) USX_DEMOD(2);
    else USX_DEMOD(4);
#undef USX_DEMOD
    return;
  }
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
