// display.hip -- the display tail on gfx950: envelope, time-gain
// compensation and log compression of a beamformed frame.
//
// Reference: ImagingPipeline::to_display / process_display
// (core/src/pipeline.cpp:48-79) over envelope, analytic_along_axis1,
// apply_tgc and log_compress (core/src/postprocess.cpp:12-117).
//
// display_env_kernel: one thread per pixel, envelope |frame| of an IQ frame
//   (demodulation_frequency > 0), times the TGC gain 10^(g/20) with
//   g = 2 alpha f depth_cm; per-block maxima for the reference level.
// display_analytic_kernel: RF frames (no demodulation frequency): one
//   workgroup per lateral line, the analytic signal along depth by FFT in
//   LDS (radix 2, zero padding to a power of two, positive frequencies
//   doubled, negative ones zeroed, inverse FFT), then its magnitude and TGC.
// display_compress_kernel: clamp((20 log10(env / ref) + gain + dr) / dr, 0, 1),
//   with the frame's peak when no reference is given.
#include <hip/hip_runtime.h>

#include "kernels.hpp"

namespace usxd {

namespace {

__device__ __forceinline__ float tgc_gain(const DisplayArgs& D, const Pt& p) {
  if (!D.apply_tgc) return 1.0f;
  const float depth_cm = p.z * 100.0f;
  const float gain_db = 2.0f * D.attenuation * D.f_mhz * fmaxf(depth_cm, 0.0f);
  return powf(10.0f, gain_db / 20.0f);
}

__device__ __forceinline__ void block_max(float m, float* blockmax) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o, 64));
  __shared__ float wm[16];
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) wm[w] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    float r = 0.0f;
    for (int q = 0; q < static_cast<int>(blockDim.x + 63) / 64; ++q) r = fmaxf(r, wm[q]);
    blockmax[blockIdx.x] = r;
  }
}

}  // namespace

__global__ __launch_bounds__(256) void display_env_kernel(DisplayArgs D, const float2* __restrict__ frame,
                                                          float* __restrict__ env, float* blockmax) {
  const int n_pix = D.g.n0 * D.g.n1;
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  float v = 0.0f;
  if (k < n_pix) {
    const int i0 = k / D.g.n1, j = k - i0 * D.g.n1;
    const float2 y = frame[k];
    v = hypotf(y.x, y.y) * tgc_gain(D, grid_point(D.g, i0, j));
    env[k] = v;
  }
  block_max(v, blockmax);
}

// n_fft = 2^log2n <= 8192: the line in LDS as complex, bit-reversed on load.
__global__ __launch_bounds__(256) void display_analytic_kernel(DisplayArgs D, const float2* __restrict__ frame,
                                                               float* __restrict__ env, float* blockmax,
                                                               int log2n) {
  extern __shared__ float2 buf[];
  const int n = 1 << log2n;
  const int i0 = blockIdx.x;
  const int n1 = D.g.n1;
  const float2* line = frame + static_cast<size_t>(i0) * n1;
  for (int q = threadIdx.x; q < n; q += blockDim.x) {
    const int r = static_cast<int>(__brev(static_cast<unsigned>(q)) >> (32 - log2n));
    buf[r] = make_float2(q < n1 ? line[q].x : 0.0f, 0.0f);
  }
  __syncthreads();
  // forward transform, e^{-2 pi i k / len}
  for (int len = 2; len <= n; len <<= 1) {
    const int h = len >> 1;
    for (int q = threadIdx.x; q < n / 2; q += blockDim.x) {
      const int grp = q / h, k = q - grp * h;
      const int a = grp * len + k, b = a + h;
      float sn, cs;
      sincospif(-2.0f * static_cast<float>(k) / static_cast<float>(len), &sn, &cs);
      const float2 u = buf[a], v = buf[b];
      const float2 t = make_float2(v.x * cs - v.y * sn, v.x * sn + v.y * cs);
      buf[a] = make_float2(u.x + t.x, u.y + t.y);
      buf[b] = make_float2(u.x - t.x, u.y - t.y);
    }
    __syncthreads();
  }
  // analytic spectrum: k in [1, n/2) doubled, (n/2, n) zeroed; then bit-reverse
  // for the inverse transform
  for (int q = threadIdx.x; q < n; q += blockDim.x) {
    float2 v = buf[q];
    if (q >= 1 && q < n / 2) v = make_float2(2.0f * v.x, 2.0f * v.y);
    if (q > n / 2) v = make_float2(0.0f, 0.0f);
    buf[q] = v;
  }
  __syncthreads();
  for (int q = threadIdx.x; q < n; q += blockDim.x) {
    const int r = static_cast<int>(__brev(static_cast<unsigned>(q)) >> (32 - log2n));
    if (q < r) {
      const float2 t = buf[q];
      buf[q] = buf[r];
      buf[r] = t;
    }
  }
  __syncthreads();
  for (int len = 2; len <= n; len <<= 1) {  // inverse, e^{+2 pi i k / len}
    const int h = len >> 1;
    for (int q = threadIdx.x; q < n / 2; q += blockDim.x) {
      const int grp = q / h, k = q - grp * h;
      const int a = grp * len + k, b = a + h;
      float sn, cs;
      sincospif(2.0f * static_cast<float>(k) / static_cast<float>(len), &sn, &cs);
      const float2 u = buf[a], v = buf[b];
      const float2 t = make_float2(v.x * cs - v.y * sn, v.x * sn + v.y * cs);
      buf[a] = make_float2(u.x + t.x, u.y + t.y);
      buf[b] = make_float2(u.x - t.x, u.y - t.y);
    }
    __syncthreads();
  }
  float m = 0.0f;
  const float inv = 1.0f / static_cast<float>(n);
  for (int q = threadIdx.x; q < n1; q += blockDim.x) {
    const float2 v = buf[q];
    const float e = hypotf(v.x * inv, v.y * inv) * tgc_gain(D, grid_point(D.g, i0, q));
    env[static_cast<size_t>(i0) * n1 + q] = e;
    m = fmaxf(m, e);
  }
  block_max(m, blockmax);
}

__global__ __launch_bounds__(256) void display_compress_kernel(DisplayArgs D, int n_pix,
                                                               const float* __restrict__ env,
                                                               const float* __restrict__ blockmax,
                                                               int n_blocks, const float* ref_in,
                                                               float* __restrict__ image) {
  __shared__ float ref_s;
  if (threadIdx.x < 64) {
    float r = *ref_in;
    if (!(r > 0.0f)) {  // no reference: the frame's own peak
      float m = 0.0f;
      for (int b = threadIdx.x; b < n_blocks; b += 64) m = fmaxf(m, blockmax[b]);
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o, 64));
      r = m;
    }
    if (threadIdx.x == 0) ref_s = r;
  }
  __syncthreads();
  const float ref = ref_s;
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n_pix) return;
  float out = 0.0f;
  if (ref > 0.0f) {
    const float dr = fmaxf(D.dynamic_range_db, 1.0f);
    const float v = env[k] / ref;
    const float db = (v > 0.0f) ? 20.0f * log10f(v) : -1e6f;
    out = fminf(fmaxf((db + D.gain_db + dr) / dr, 0.0f), 1.0f);
  }
  image[k] = out;
}

int display_blocks(const DisplayArgs& D, bool rf) { return rf ? D.g.n0 : (D.g.n0 * D.g.n1 + 255) / 256; }

void launch_display_envelope(const DisplayArgs& D, const float2* frame, bool rf, float* env,
                             float* blockmax, hipStream_t st) {
  if (!rf) {
    hipLaunchKernelGGL(display_env_kernel, dim3(display_blocks(D, false)), dim3(256), 0, st, D, frame,
                       env, blockmax);
    return;
  }
  int log2n = 0;
  while ((1 << log2n) < D.g.n1) ++log2n;
  hipLaunchKernelGGL(display_analytic_kernel, dim3(D.g.n0), dim3(256),
                     sizeof(float2) * (static_cast<size_t>(1) << log2n), st, D, frame, env, blockmax,
                     log2n);
}

void launch_display_compress(const DisplayArgs& D, const float* env, const float* blockmax,
                             int n_blocks, const float* ref, float* image, hipStream_t st) {
  const int n_pix = D.g.n0 * D.g.n1;
  hipLaunchKernelGGL(display_compress_kernel, dim3((n_pix + 255) / 256), dim3(256), 0, st, D, n_pix,
                     env, blockmax, n_blocks, ref, image);
}

}  // namespace usxd
