// sim.hip -- the point-scatterer forward model on gfx950.
//
// Reference: Simulator::simulate, core/src/simulator.cpp:213-419, with the
// path amplitude of element_directivity (geometry.cpp:111-121). Per event:
//
// sim_incident_kernel (one wavefront per scatterer): the diffracted incident
//   waveform at the scatterer, i.e. the pulse delayed and weighted per
//   firing element and summed, in an LDS buffer of the reference's length
//   (span of the arrival times plus the pulse, capped by the incident
//   window), then trimmed at incident_trim_db of its peak and stored. The
//   firing elements are added in order; a lane owns one buffer sample per
//   element, so every sample adds exactly the reference's two terms in the
//   reference's order (the frac term of step k-1, then the 1 - frac term of
//   step k). Each lane first evaluates the geometry of one of 64 elements;
//   the element loop then broadcasts it with readlane.
// sim_echo_kernel (one wavefront per (channel, chunk of scatterers)): each
//   scatterer's echo at the channel (spread, directivity, attenuation,
//   delayed and linearly interpolated from the incident waveform) added to
//   an LDS trace in scatterer order; the trace is written as the chunk's
//   partial sum.
// sim_reduce_kernel: RF = the partial traces summed in chunk order, and the
//   block maxima of |RF|; sim_noise_kernel: RF += z * (peak * noise
//   amplitude) with z the reference's standard-normal stream for the event
//   (drawn on the host, std::mt19937(noise_seed + event)).
//
// Floating-point expressions keep the reference's operation order
// (contraction is off), so the sample indices and trimming decisions match
// it; amplitudes go through device powf/sinf, whose last-bit differences
// from the host's libm make the RF match the host simulator to ~1e-6
// relative rather than bit for bit. The reduction order (chunks of
// kSimChunk scatterers) is fixed, so the output is deterministic.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "kernels.hpp"

namespace usxd {

namespace {

struct V3 {
  float x, y, z;
};

__device__ __forceinline__ float dist(V3 a, V3 b) {
  const float dx = a.x - b.x, dy = a.y - b.y, dz = a.z - b.z;
  return sqrtf(dx * dx + dy * dy + dz * dz);
}

// element_directivity (geometry.cpp:111-121)
__device__ __forceinline__ float directivity(float width, float lambda, float sin_theta) {
  const float st = fminf(fmaxf(sin_theta, -1.0f), 1.0f);
  const float ct = sqrtf(fmaxf(0.0f, 1.0f - st * st));
  const float x = kPi * width * st / lambda;
  const float sinc = (fabsf(x) < 1e-6f) ? 1.0f - x * x / 6.0f : sinf(x) / x;
  return ct * sinc;
}

// The path amplitude between a scatterer and an element at range r
// (simulator.cpp:255-261).
__device__ __forceinline__ float path_amp(const SimArgs& S, float a, V3 pos, V3 pe, float r) {
  if (S.spreading) a /= r;
  if (S.directivity) a *= directivity(S.element_width, S.wavelength, (pos.x - pe.x) / r);
  if (S.atten > 0.0f) a *= powf(10.0f, (-S.atten * r) / 20.0f);
  return a;
}

__device__ __forceinline__ float rl(float v, int k) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), k));
}
__device__ __forceinline__ int rl(int v, int k) { return __builtin_amdgcn_readlane(v, k); }

}  // namespace

// ---------------------------------------------------------------------------
// Stage 1: incident waveform per scatterer.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void sim_incident_kernel(SimArgs S, int ev, const float* __restrict__ scat,
                                                           int n_scat, SimIncident I) {
  extern __shared__ float sbuf[];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int si = blockIdx.x * 4 + wv;
  if (si >= n_scat) return;
  float* buf = sbuf + wv * S.inc_cap;
  const float* sc = scat + 7 * static_cast<size_t>(si);
  I.live[si] = 0;
  const float amp0 = sc[3];
  if (amp0 == 0.0f) return;
  V3 pos{sc[0], sc[1], sc[2]};
  if (S.motion) {
    const float et = static_cast<float>(ev) / S.prf;
    pos = V3{pos.x + sc[4] * et, pos.y + sc[5] * et, pos.z + sc[6] * et};
  }
  if (pos.z <= 0.0f) return;
  const int nf = S.n_firing[ev];
  const int* fire = S.firing + static_cast<size_t>(ev) * S.n_elements;
  const float* fw = S.fire_w + static_cast<size_t>(ev) * S.n_elements;
  const float* delays = S.delays + static_cast<size_t>(ev) * S.n_elements;

  // arrival time span over the firing elements
  float tmin = INFINITY, tmax = -INFINITY;
  for (int q = lane; q < nf; q += 64) {
    const int i = fire[q];
    const float t = delays[i] + dist(pos, V3{S.ex[i], 0.0f, 0.0f}) / S.c;
    tmin = fminf(tmin, t);
    tmax = fmaxf(tmax, t);
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    tmin = fminf(tmin, __shfl_xor(tmin, o, 64));
    tmax = fmaxf(tmax, __shfl_xor(tmax, o, 64));
  }
  float span = (tmax - tmin) + 2.0f * S.pulse_half;
  span = fminf(span, S.max_window + 2.0f * S.pulse_half);
  const float inc_t0 = tmin - S.pulse_half;
  // (the host sizes inc_cap for the capped span; the clamp only guards
  // against non-finite positions)
  const int inc_n = min(max(static_cast<int>(ceilf(span * S.fs_os)) + 2, 0), S.inc_cap);
  for (int q = lane; q < inc_n; q += 64) buf[q] = 0.0f;

  const int pn = S.pulse_n;
  const int half = pn / 2;
  for (int q0 = 0; q0 < nf; q0 += 64) {
    // lane: geometry of firing element q0 + lane
    int base = 0;
    float amp = 0.0f, frac = 0.0f;
    if (q0 + lane < nf) {
      const int i = fire[q0 + lane];
      const V3 pe{S.ex[i], 0.0f, 0.0f};
      const float r = dist(pos, pe);
      if (r >= 1e-6f) {
        amp = path_amp(S, fw[q0 + lane], pos, pe, r);
        const float centre = (delays[i] + r / S.c - inc_t0) * S.fs_os;
        const float fl = floorf(centre);
        base = static_cast<int>(fl) - half;
        frac = centre - fl;
      }
    }
    const int nq = min(64, nf - q0);
    for (int j = 0; j < nq; ++j) {
      const float a = rl(amp, j);
      if (a == 0.0f) continue;
      const int b = rl(base, j);
      const float fr = rl(frac, j);
      const float g = 1.0f - fr;
      // sample s = b + k gets (step k-1: v_{k-1} frac) then (step k: v_k (1 - frac));
      // step k is skipped when b + k < 0 or b + k + 1 >= inc_n
      for (int k = lane; k <= pn; k += 64) {
        const int s = b + k;
        if (s < 0 || s >= inc_n) continue;
        float v = buf[s];
        if (k >= 1 && s - 1 >= 0) v += (a * S.pulse[k - 1]) * fr;
        if (k < pn && s + 1 < inc_n) v += (a * S.pulse[k]) * g;
        buf[s] = v;
      }
      if (!S.full_diffraction) break;
    }
    if (!S.full_diffraction) break;
  }

  // trim at trim_amp of the peak
  float peak = 0.0f;
  for (int q = lane; q < inc_n; q += 64) peak = fmaxf(peak, fabsf(buf[q]));
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) peak = fmaxf(peak, __shfl_xor(peak, o, 64));
  if (!(peak > 0.0f)) return;
  const float thr = peak * S.trim_amp;
  int lo = inc_n, hi = -1;
  for (int q = lane; q < inc_n; q += 64) {
    if (fabsf(buf[q]) >= thr) {
      lo = min(lo, q);
      hi = max(hi, q);
    }
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    lo = min(lo, __shfl_xor(lo, o, 64));
    hi = max(hi, __shfl_xor(hi, o, 64));
  }
  if (hi < 0 || lo >= hi) return;
  const int n = hi - lo + 1;
  float* dst = I.wave + static_cast<size_t>(si) * S.inc_cap;
  for (int q = lane; q < n; q += 64) dst[q] = buf[lo + q];
  if (lane == 0) {
    I.live[si] = 1;
    I.t0[si] = inc_t0 + static_cast<float>(lo) / S.fs_os;
    I.n[si] = n;
    I.pos[si] = make_float4(pos.x, pos.y, pos.z, amp0);
  }
}

// ---------------------------------------------------------------------------
// Stage 2: echoes. Grid (n_chunks, ceil(n_elements / 4)); wave = one
// channel, one chunk of kSimChunk scatterers, trace in LDS.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void sim_echo_kernel(SimArgs S, int n_scat, SimIncident I,
                                                       float* __restrict__ partial) {
  extern __shared__ float trace_all[];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int ch = blockIdx.y * 4 + wv;
  const int chunk = blockIdx.x;
  if (ch >= S.n_elements) return;
  float* tr = trace_all + wv * S.n_samples;
  for (int q = lane; q < S.n_samples; q += 64) tr[q] = 0.0f;
  const V3 pe{S.ex[ch], 0.0f, 0.0f};
  const int s_lo = chunk * kSimChunk, s_hi = min(n_scat, s_lo + kSimChunk);
  for (int g0 = s_lo; g0 < s_hi; g0 += 64) {
    // lane: the (channel, scatterer g0 + lane) path
    const int si = g0 + lane;
    int n0 = 1, n1 = 0;
    float amp = 0.0f, delay = 0.0f, t0 = 0.0f;
    int nn = 0;
    if (si < s_hi && I.live[si]) {
      const float4 p = I.pos[si];
      const V3 pos{p.x, p.y, p.z};
      const float r = dist(pos, pe);
      if (r >= 1e-6f) {
        amp = path_amp(S, p.w, pos, pe, r);
        if (amp != 0.0f) {
          delay = r / S.c;
          t0 = I.t0[si];
          nn = I.n[si];
          const float t_start = t0 + delay;
          const float t_end = t_start + static_cast<float>(nn - 1) / S.fs_os;
          n0 = max(static_cast<int>(ceilf((t_start - S.t0) * S.fs)), 0);
          n1 = min(static_cast<int>(floorf((t_end - S.t0) * S.fs)), S.n_samples - 1);
        }
      }
    }
    const int nq = min(64, s_hi - g0);
    for (int j = 0; j < nq; ++j) {
      const int a0 = rl(n0, j), a1 = rl(n1, j);
      if (a0 > a1) continue;
      const float am = rl(amp, j), dl = rl(delay, j), tt0 = rl(t0, j);
      const int n = rl(nn, j);
      const float* w = I.wave + static_cast<size_t>(g0 + j) * S.inc_cap;
      for (int q = a0 + lane; q <= a1; q += 64) {
        const float t = (S.t0 + static_cast<float>(q) / S.fs) - dl;
        const float u = (t - tt0) * S.fs_os;
        const int k = static_cast<int>(u);
        if (k < 0 || k + 1 >= n) continue;
        const float f = u - static_cast<float>(k);
        const float va = w[k], vb = w[k + 1];
        tr[q] += am * (va + (vb - va) * f);
      }
    }
  }
  float* dst = partial + (static_cast<size_t>(chunk) * S.n_elements + ch) * S.n_samples;
  for (int q = lane; q < S.n_samples; q += 64) dst[q] = tr[q];
}

// RF of the event = partials summed in chunk order; block maxima of |RF|.
__global__ __launch_bounds__(256) void sim_reduce_kernel(SimArgs S, int n_chunks,
                                                         const float* __restrict__ partial,
                                                         float* __restrict__ rf, float* blockmax) {
  const size_t n_tot = static_cast<size_t>(S.n_elements) * S.n_samples;
  const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  float m = 0.0f;
  if (i < n_tot) {
    float s = 0.0f;
    for (int c = 0; c < n_chunks; ++c) s += partial[static_cast<size_t>(c) * n_tot + i];
    rf[i] = s;
    m = fabsf(s);
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o, 64));
  __shared__ float wm[4];
  if ((threadIdx.x & 63) == 0) wm[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) blockmax[blockIdx.x] = fmaxf(fmaxf(wm[0], wm[1]), fmaxf(wm[2], wm[3]));
}

// RF += z * (peak * noise_amp) when peak > 0 (simulator.cpp:405-416).
__global__ __launch_bounds__(256) void sim_noise_kernel(SimArgs S, int n_blocks,
                                                        const float* __restrict__ blockmax,
                                                        const float* __restrict__ z,
                                                        float* __restrict__ rf) {
  __shared__ float peak_s;
  if (threadIdx.x < 64) {
    float m = 0.0f;
    for (int b = threadIdx.x; b < n_blocks; b += 64) m = fmaxf(m, blockmax[b]);
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o, 64));
    if (threadIdx.x == 0) peak_s = m;
  }
  __syncthreads();
  const float peak = peak_s;
  if (!(peak > 0.0f)) return;
  const float sigma = peak * S.noise_amp;
  const size_t n_tot = static_cast<size_t>(S.n_elements) * S.n_samples;
  const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n_tot) rf[i] = rf[i] + (z[i] * sigma + 0.0f);
}

size_t sim_incident_smem(const SimArgs& S) { return 4 * sizeof(float) * static_cast<size_t>(S.inc_cap); }
size_t sim_echo_smem(const SimArgs& S) { return 4 * sizeof(float) * static_cast<size_t>(S.n_samples); }
int sim_chunks(int n_scat) { return (n_scat + kSimChunk - 1) / kSimChunk; }

void launch_sim_event(const SimArgs& S, int ev, const float* scat, int n_scat, const SimIncident& I,
                      float* partial, float* blockmax, const float* z, float* rf, hipStream_t st) {
  static const hipError_t a1 = hipFuncSetAttribute(
      reinterpret_cast<const void*>(sim_incident_kernel), hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
  static const hipError_t a2 = hipFuncSetAttribute(
      reinterpret_cast<const void*>(sim_echo_kernel), hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
  (void)a1;
  (void)a2;
  const size_t n_tot = static_cast<size_t>(S.n_elements) * S.n_samples;
  const int nb = static_cast<int>((n_tot + 255) / 256);
  const int nc = sim_chunks(n_scat);
  if (n_scat > 0) {
    hipLaunchKernelGGL(sim_incident_kernel, dim3((n_scat + 3) / 4), dim3(256), sim_incident_smem(S), st, S, ev,
                       scat, n_scat, I);
    hipLaunchKernelGGL(sim_echo_kernel, dim3(nc, (S.n_elements + 3) / 4), dim3(256), sim_echo_smem(S), st, S,
                       n_scat, I, partial);
  }
  hipLaunchKernelGGL(sim_reduce_kernel, dim3(nb), dim3(256), 0, st, S, n_scat > 0 ? nc : 0, partial, rf,
                     blockmax);
  if (S.noise_amp > 0.0f && z)
    hipLaunchKernelGGL(sim_noise_kernel, dim3(nb), dim3(256), 0, st, S, nb, blockmax, z, rf);
}

}  // namespace usxd
