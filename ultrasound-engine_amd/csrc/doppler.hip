// doppler.hip -- slow-time and frame-to-frame estimators on beamformed IQ,
// on gfx950: clutter (wall) filtering, colour flow (lag-one autocorrelation,
// Kasai / Loupas), power Doppler, axial displacement (coarse integer lag +
// phase refinement) and least-squares axial strain.
//
// Reference: core/src/doppler.cpp (apply_wall_filter 51-75, color_flow
// 77-176, power_doppler 178-186, estimate_axial_displacement 188-273,
// axial_strain 275-307).
//
// Layout. An ensemble is n_frames frames of n_pix = n0 * n1 complex pixels,
// frame-major ([frame][axis0][axis1], axis1 fastest), as the frames lie one
// after the other. One thread owns one pixel; consecutive lanes take
// consecutive depths, so every frame is read with coalesced 8-byte loads.
//
// Colour flow is split in two passes. The reference sums, per pixel, over the
// (2 kl + 1) x (2 ka + 1) window and over the ensemble in one running sum.
// Here the ensemble sums of each pixel (zero lag power, slow-time lag one,
// axial lag one) are formed once in the first pass, in the reference's
// ensemble order, and the second pass adds those over the window, depth
// first then line: the same terms, regrouped (float rounding only).
// The wall filter is applied in registers in the first pass, so the filtered
// ensemble is never written for colour flow or power Doppler.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "kernels.hpp"

namespace usxd {

namespace {

constexpr float kPiF = static_cast<float>(3.14159265358979323846);

__device__ __forceinline__ float2 cmul_conj(float2 a, float2 b) {
  // conj(a) * b with std::complex<float>'s products and sums
  return make_float2(a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ float cnorm(float2 z) { return z.x * z.x + z.y * z.y; }
// std::clamp(v, 0, 1): NaN passes through
__device__ __forceinline__ float clamp01(float v) { return v < 0.0f ? 0.0f : (1.0f < v ? 1.0f : v); }

// The NC-register copy of one pixel's ensemble (NC >= n), filtered by the
// row-major n x n projector M when M is not null: z'_i = sum_j M[i][j] z_j,
// j in order, each product (z_j re * m, z_j im * m) added to the running sum
// (doppler.cpp:66-73).
template <int NC>
__device__ __forceinline__ void load_filtered(const float2* __restrict__ in, size_t n_pix, size_t k,
                                              int n, const float* __restrict__ M, float2 (&z)[NC]) {
  float2 raw[NC];
#pragma unroll
  for (int j = 0; j < NC; ++j) raw[j] = j < n ? in[static_cast<size_t>(j) * n_pix + k] : make_float2(0.f, 0.f);
  if (!M) {
#pragma unroll
    for (int j = 0; j < NC; ++j) z[j] = raw[j];
    return;
  }
#pragma unroll
  for (int i = 0; i < NC; ++i) {
    float2 acc = make_float2(0.0f, 0.0f);
    if (i < n) {
      const float* row = M + i * n;
#pragma unroll
      for (int j = 0; j < NC; ++j) {
        if (j < n) {
          const float m = row[j];
          acc.x += raw[j].x * m;
          acc.y += raw[j].y * m;
        }
      }
    }
    z[i] = acc;
  }
}

// Ensemble lengths above the register copies: the same sums from memory.
__device__ __forceinline__ float2 filtered_at(const float2* __restrict__ in, size_t n_pix, size_t k,
                                              int n, const float* __restrict__ M, int i) {
  if (!M) return in[static_cast<size_t>(i) * n_pix + k];
  float2 acc = make_float2(0.0f, 0.0f);
  for (int j = 0; j < n; ++j) {
    const float2 v = in[static_cast<size_t>(j) * n_pix + k];
    const float m = M[i * n + j];
    acc.x += v.x * m;
    acc.y += v.y * m;
  }
  return acc;
}

}  // namespace

// ---------------------------------------------------------------------------
// Wall filter: out[i][k] = sum_j M[i][j] in[j][k].
// ---------------------------------------------------------------------------
template <int NC>
__global__ __launch_bounds__(256) void wall_filter_kernel(int n, size_t n_pix,
                                                          const float2* __restrict__ in,
                                                          const float* __restrict__ M,
                                                          float2* __restrict__ out) {
  const size_t k = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (k >= n_pix) return;
  if constexpr (NC > 0) {
    float2 z[NC];
    load_filtered<NC>(in, n_pix, k, n, M, z);
#pragma unroll
    for (int i = 0; i < NC; ++i)
      if (i < n) out[static_cast<size_t>(i) * n_pix + k] = z[i];
  } else {
    for (int i = 0; i < n; ++i) out[static_cast<size_t>(i) * n_pix + k] = filtered_at(in, n_pix, k, n, M, i);
  }
}

// ---------------------------------------------------------------------------
// Power Doppler: p[k] = sum over the filtered frames, in order, of |z'|^2
// (doppler.cpp:178-186).
// ---------------------------------------------------------------------------
template <int NC>
__global__ __launch_bounds__(256) void power_doppler_kernel(int n, size_t n_pix,
                                                            const float2* __restrict__ in,
                                                            const float* __restrict__ M,
                                                            float* __restrict__ p) {
  const size_t k = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (k >= n_pix) return;
  float s = 0.0f;
  if constexpr (NC > 0) {
    float2 z[NC];
    load_filtered<NC>(in, n_pix, k, n, M, z);
#pragma unroll
    for (int i = 0; i < NC; ++i)
      if (i < n) s += cnorm(z[i]);
  } else {
    for (int i = 0; i < n; ++i) s += cnorm(filtered_at(in, n_pix, k, n, M, i));
  }
  p[k] = s;
}

// ---------------------------------------------------------------------------
// Colour flow, pass 1: per pixel, over the filtered ensemble in order,
//   lag[0][k] = sum_e |z_e|^2
//   lag[1..2][k] = sum_{e < n-1} conj(z_e) z_{e+1}       (slow time)
//   lag[3..4][k] = sum_e conj(z_e[k]) z_e[k + 1]           (axial; 0 on the last depth)
// (doppler.cpp:115-133 with the window loop taken out). A wavefront loads and
// filters 64 consecutive pixels and writes the first 63: the axial neighbour's
// filtered ensemble comes from the next lane by a shuffle, so every pixel is
// filtered once. Thread 0 of the grid also clears the power maximum.
// ---------------------------------------------------------------------------
constexpr int kLagOut = 63;  // pixels written per wavefront

template <int NC>
__global__ __launch_bounds__(256) void cf_lag_kernel(int n, int n0, int n1,
                                                     const float2* __restrict__ in,
                                                     const float* __restrict__ M,
                                                     float* __restrict__ lag,
                                                     unsigned int* __restrict__ pmax_bits) {
  const size_t n_pix = static_cast<size_t>(n0) * n1;
  const int lane = threadIdx.x & 63;
  const size_t wave = static_cast<size_t>(blockIdx.x) * (blockDim.x >> 6) + (threadIdx.x >> 6);
  const size_t k = wave * kLagOut + lane;
  if (blockIdx.x == 0 && threadIdx.x == 0) *pmax_bits = 0u;
  if (wave * kLagOut >= n_pix) return;  // wavefront-uniform
  const bool live = k < n_pix;
  const size_t kc = live ? k : n_pix - 1;  // clamped: every lane takes part in the shuffles
  const int j = static_cast<int>(kc % static_cast<size_t>(n1));
  const bool axial = j + 1 < n1;
  float r00 = 0.0f;
  float2 r01 = make_float2(0.0f, 0.0f), r10 = make_float2(0.0f, 0.0f);
  if constexpr (NC > 0) {
    float2 z[NC];
    load_filtered<NC>(in, n_pix, kc, n, M, z);
#pragma unroll
    for (int e = 0; e < NC; ++e) {
      if (e < n) {
        const float2 zn = make_float2(__shfl_down(z[e].x, 1, 64), __shfl_down(z[e].y, 1, 64));
        r00 += cnorm(z[e]);
        if (e + 1 < n) {
          const float2 p = cmul_conj(z[e], z[e + 1 < NC ? e + 1 : e]);
          r01.x += p.x;
          r01.y += p.y;
        }
        if (axial) {
          const float2 q = cmul_conj(z[e], zn);
          r10.x += q.x;
          r10.y += q.y;
        }
      }
    }
  } else {
    float2 cur = filtered_at(in, n_pix, kc, n, M, 0);
    for (int e = 0; e < n; ++e) {
      const float2 zn = make_float2(__shfl_down(cur.x, 1, 64), __shfl_down(cur.y, 1, 64));
      r00 += cnorm(cur);
      float2 nxt = make_float2(0.0f, 0.0f);
      if (e + 1 < n) {
        nxt = filtered_at(in, n_pix, kc, n, M, e + 1);
        const float2 p = cmul_conj(cur, nxt);
        r01.x += p.x;
        r01.y += p.y;
      }
      if (axial) {
        const float2 q = cmul_conj(cur, zn);
        r10.x += q.x;
        r10.y += q.y;
      }
      cur = nxt;
    }
  }
  if (!live || lane == kLagOut) return;
  lag[k] = r00;
  lag[n_pix + k] = r01.x;
  lag[2 * n_pix + k] = r01.y;
  lag[3 * n_pix + k] = r10.x;
  lag[4 * n_pix + k] = r10.y;
}

// ---------------------------------------------------------------------------
// Colour flow, pass 2: window sums and the estimates (doppler.cpp:115-161).
// A workgroup owns kCfLines lines x 64 depths; it stages the five lag maps of
// the tile and its halo (kl lines, ka depths either side; zeros off the
// grid, which add nothing) in LDS, sums each staged row over the 2 ka + 1
// depths, then each pixel over the 2 kl + 1 rows: the reference's window sum
// regrouped as depth-then-line (float rounding only). valid is set before the
// power threshold; the largest finite power goes to *pmax_bits (non-negative
// floats order as their bit patterns) by one atomic per workgroup.
// ---------------------------------------------------------------------------
constexpr int kCfLines = 4;
constexpr int kCfMaxKa = 16, kCfMaxKl = 4;
constexpr int kCfRows = kCfLines + 2 * kCfMaxKl;
constexpr int kCfCols = 64 + 2 * kCfMaxKa;

__device__ __forceinline__ void cf_estimate_pixel(const ColorFlowArgs& A, float R00, float2 R01,
                                                  float2 R10, float* v_out, float* var_out,
                                                  uint8_t* ok_out) {
  float v = 0.0f, var = 0.0f;
  uint8_t ok = 0;
  if (R00 > 0.0f) {
    const float coherence = hypotf(R01.x, R01.y) / R00;
    var = clamp01(2.0f * (1.0f - coherence));
    if (!(coherence < A.coherence_threshold)) {
      const float phase_slow = atan2f(R01.y, R01.x);
      float f_centre = A.fdem;
      if (A.loupas) {
        const float phase_fast = atan2f(R10.y, R10.x);
        const float f_est = phase_fast * A.c / (4.0f * kPiF * A.dz);
        if (f_est > 0.25f * A.fdem && f_est < 2.0f * A.fdem) f_centre = f_est;
      }
      if (!(f_centre <= 0.0f)) {
        v = -phase_slow * A.c * A.prf / (4.0f * kPiF * f_centre);
        ok = 1;
      }
    }
  }
  *v_out = v;
  *var_out = var;
  *ok_out = ok;
}

__global__ __launch_bounds__(256) void cf_estimate_kernel(ColorFlowArgs A, const float* __restrict__ lag,
                                                          float* __restrict__ velocity,
                                                          float* __restrict__ power,
                                                          float* __restrict__ variance,
                                                          uint8_t* __restrict__ valid,
                                                          unsigned int* pmax_bits) {
  __shared__ float stage[5][kCfRows][kCfCols];
  __shared__ float rows[5][kCfRows][64];
  __shared__ unsigned int wave_max[4];
  const size_t n_pix = static_cast<size_t>(A.n0) * A.n1;
  const int tid = threadIdx.x;
  const int n_jt = (A.n1 + 63) / 64;
  const int i0 = (blockIdx.x / n_jt) * kCfLines, j0 = (blockIdx.x % n_jt) * 64;
  const int nr = kCfLines + 2 * A.kl, nc = 64 + 2 * A.ka;
  for (int q = tid; q < nr * nc; q += 256) {
    const int r = q / nc, c = q - r * nc;
    const int ii = i0 - A.kl + r, jj = j0 - A.ka + c;
    const bool in = ii >= 0 && ii < A.n0 && jj >= 0 && jj < A.n1;
    const size_t kk = in ? static_cast<size_t>(ii) * A.n1 + jj : 0;
#pragma unroll
    for (int m = 0; m < 5; ++m) stage[m][r][c] = in ? lag[m * n_pix + kk] : 0.0f;
  }
  __syncthreads();
  for (int q = tid; q < nr * 64; q += 256) {
    const int r = q >> 6, c = q & 63;
#pragma unroll
    for (int m = 0; m < 5; ++m) {
      float s = 0.0f;
      for (int d = 0; d <= 2 * A.ka; ++d) s += stage[m][r][c + d];
      rows[m][r][c] = s;
    }
  }
  __syncthreads();
  const int li = tid >> 6, c = tid & 63;
  const int i = i0 + li, j = j0 + c;
  const bool live = i < A.n0 && j < A.n1;
  float R[5] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
  for (int d = 0; d <= 2 * A.kl; ++d) {
#pragma unroll
    for (int m = 0; m < 5; ++m) R[m] += rows[m][li + d][c];
  }
  float pw = 0.0f;
  if (live) {
    const size_t k = static_cast<size_t>(i) * A.n1 + j;
    cf_estimate_pixel(A, R[0], make_float2(R[1], R[2]), make_float2(R[3], R[4]), velocity + k,
                      variance + k, valid + k);
    power[k] = R[0];
    pw = R[0] == R[0] ? R[0] : 0.0f;
  }
  unsigned int mx = __float_as_uint(pw);
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) mx = max(mx, static_cast<unsigned int>(__shfl_xor(static_cast<int>(mx), o, 64)));
  if ((tid & 63) == 0) wave_max[tid >> 6] = mx;
  __syncthreads();
  if (tid == 0) {
    const unsigned int b = max(max(wave_max[0], wave_max[1]), max(wave_max[2], wave_max[3]));
    if (b) atomicMax(pmax_bits, b);
  }
}

// Window sums straight from memory, for windows larger than the LDS tile.
__global__ __launch_bounds__(256) void cf_estimate_global_kernel(ColorFlowArgs A, const float* __restrict__ lag,
                                                                 float* __restrict__ velocity,
                                                                 float* __restrict__ power,
                                                                 float* __restrict__ variance,
                                                                 uint8_t* __restrict__ valid,
                                                                 unsigned int* pmax_bits) {
  const size_t n_pix = static_cast<size_t>(A.n0) * A.n1;
  const size_t k = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (k >= n_pix) return;
  const int i = static_cast<int>(k / static_cast<size_t>(A.n1));
  const int j = static_cast<int>(k - static_cast<size_t>(i) * A.n1);
  float R00 = 0.0f;
  float2 R01 = make_float2(0.0f, 0.0f), R10 = make_float2(0.0f, 0.0f);
  for (int di = -A.kl; di <= A.kl; ++di) {
    const int ii = i + di;
    if (ii < 0 || ii >= A.n0) continue;
    for (int dj = -A.ka; dj <= A.ka; ++dj) {
      const int jj = j + dj;
      if (jj < 0 || jj >= A.n1) continue;
      const size_t kk = static_cast<size_t>(ii) * A.n1 + jj;
      R00 += lag[kk];
      R01.x += lag[n_pix + kk];
      R01.y += lag[2 * n_pix + kk];
      R10.x += lag[3 * n_pix + kk];
      R10.y += lag[4 * n_pix + kk];
    }
  }
  cf_estimate_pixel(A, R00, R01, R10, velocity + k, variance + k, valid + k);
  power[k] = R00;
  if (R00 == R00) atomicMax(pmax_bits, __float_as_uint(R00));
}

// Power threshold (doppler.cpp:165-174): below pmax * factor is not flow.
__global__ __launch_bounds__(256) void cf_threshold_kernel(size_t n_pix, const float* __restrict__ power,
                                                           uint8_t* __restrict__ valid,
                                                           const unsigned int* __restrict__ pmax_bits,
                                                           float factor) {
  const size_t k = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (k >= n_pix) return;
  const float pmax = __uint_as_float(*pmax_bits);
  if (!(pmax > 0.0f)) return;
  const float thr = pmax * factor;
  if (power[k] < thr) valid[k] = 0;
}

// ---------------------------------------------------------------------------
// Axial displacement between two frames (doppler.cpp:188-273): the local
// centre frequency from the first frame's axial phase gradient, the integer
// lag with the highest normalised correlation, then the residual phase at
// that lag. Every sum runs in the reference's order.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void displacement_kernel(DisplacementArgs A,
                                                           const float2* __restrict__ before,
                                                           const float2* __restrict__ after,
                                                           float* __restrict__ axial,
                                                           float* __restrict__ correlation) {
  const size_t n_pix = static_cast<size_t>(A.n0) * A.n1;
  const size_t k = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (k >= n_pix) return;
  const int i = static_cast<int>(k / static_cast<size_t>(A.n1));
  const int j = static_cast<int>(k - static_cast<size_t>(i) * A.n1);
  float2 R10 = make_float2(0.0f, 0.0f);
  for (int di = -A.kl; di <= A.kl; ++di) {
    const int ii = i + di;
    if (ii < 0 || ii >= A.n0) continue;
    for (int dj = -A.ka; dj <= A.ka; ++dj) {
      const int jj = j + dj;
      if (jj < 0 || jj + 1 >= A.n1) continue;
      const size_t kk = static_cast<size_t>(ii) * A.n1 + jj;
      const float2 p = cmul_conj(before[kk], before[kk + 1]);
      R10.x += p.x;
      R10.y += p.y;
    }
  }
  float out_axial = 0.0f, out_corr = 0.0f;
  float f_centre = A.fdem;
  const float f_est = atan2f(R10.y, R10.x) * A.c / (4.0f * kPiF * A.dz);
  if (f_est > 0.25f * A.fdem && f_est < 2.0f * A.fdem) f_centre = f_est;
  if (!(f_centre <= 0.0f)) {
    float2 best_R = make_float2(0.0f, 0.0f);
    float best_corr = -1.0f;
    int best_lag = 0;
    for (int lag = -A.lag_max; lag <= A.lag_max; ++lag) {
      float2 R = make_float2(0.0f, 0.0f);
      float ea = 0.0f, eb = 0.0f;
      for (int di = -A.kl; di <= A.kl; ++di) {
        const int ii = i + di;
        if (ii < 0 || ii >= A.n0) continue;
        const size_t row = static_cast<size_t>(ii) * A.n1;
        for (int dj = -A.ka; dj <= A.ka; ++dj) {
          const int jj = j + dj;
          const int jl = jj + lag;
          if (jj < 0 || jj >= A.n1 || jl < 0 || jl >= A.n1) continue;
          const float2 a = before[row + jj];
          const float2 b = after[row + jl];
          const float2 p = cmul_conj(a, b);
          R.x += p.x;
          R.y += p.y;
          ea += cnorm(a);
          eb += cnorm(b);
        }
      }
      if (ea <= 0.0f || eb <= 0.0f) continue;
      const float corr = hypotf(R.x, R.y) / sqrtf(ea * eb);
      if (corr > best_corr) {
        best_corr = corr;
        best_R = R;
        best_lag = lag;
      }
    }
    if (!(best_corr < 0.0f)) {
      out_corr = clamp01(best_corr);
      const float residual = -atan2f(best_R.y, best_R.x) * A.c / (4.0f * kPiF * f_centre);
      out_axial = static_cast<float>(best_lag) * A.dz + residual;
    }
  }
  axial[k] = out_axial;
  correlation[k] = out_corr;
}

// Least-squares slope of the displacement against depth over 2 half + 1
// depths (doppler.cpp:275-307).
__global__ __launch_bounds__(256) void strain_kernel(int n0, int n1, float dz, int half,
                                                     const float* __restrict__ axial,
                                                     float* __restrict__ strain) {
  const size_t n_pix = static_cast<size_t>(n0) * n1;
  const size_t k = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (k >= n_pix) return;
  const int i = static_cast<int>(k / static_cast<size_t>(n1));
  const int j = static_cast<int>(k - static_cast<size_t>(i) * n1);
  float sx = 0.0f, sy = 0.0f, sxx = 0.0f, sxy = 0.0f;
  int n = 0;
  for (int d = -half; d <= half; ++d) {
    const int jj = j + d;
    if (jj < 0 || jj >= n1) continue;
    const float x = static_cast<float>(d) * dz;
    const float y = axial[static_cast<size_t>(i) * n1 + jj];
    sx += x;
    sy += y;
    sxx += x * x;
    sxy += x * y;
    ++n;
  }
  float s = 0.0f;
  if (n >= 2) {
    const float fn = static_cast<float>(n);
    const float denom = fn * sxx - sx * sx;
    if (!(fabsf(denom) < 1e-20f)) s = (fn * sxy - sx * sy) / denom;
  }
  strain[k] = s;
}

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------
namespace {
inline dim3 pix_grid(size_t n_pix) { return dim3(static_cast<unsigned>((n_pix + 255) / 256)); }
}  // namespace

#define USX_ENS_DISPATCH(KERNEL, n, ...)                                                     \
  do {                                                                                       \
    if ((n) <= 8) hipLaunchKernelGGL(KERNEL<8>, __VA_ARGS__);                                \
    else if ((n) <= 16) hipLaunchKernelGGL(KERNEL<16>, __VA_ARGS__);                         \
    else if ((n) <= 32) hipLaunchKernelGGL(KERNEL<32>, __VA_ARGS__);                         \
    else hipLaunchKernelGGL(KERNEL<0>, __VA_ARGS__);                                         \
  } while (0)

void launch_wall_filter(int n, size_t n_pix, const float2* in, const float* M, float2* out,
                        hipStream_t st) {
  USX_ENS_DISPATCH(wall_filter_kernel, n, pix_grid(n_pix), dim3(256), 0, st, n, n_pix, in, M, out);
}

void launch_power_doppler(int n, size_t n_pix, const float2* in, const float* M, float* p,
                          hipStream_t st) {
  USX_ENS_DISPATCH(power_doppler_kernel, n, pix_grid(n_pix), dim3(256), 0, st, n, n_pix, in, M, p);
}

void launch_color_flow(const ColorFlowArgs& A, const float2* in, const float* M, float* lag,
                       float* velocity, float* power, float* variance, uint8_t* valid,
                       unsigned int* pmax_bits, float factor, hipStream_t st) {
  const size_t n_pix = static_cast<size_t>(A.n0) * A.n1;
  const size_t waves = (n_pix + kLagOut - 1) / kLagOut;
  const dim3 lag_grid(static_cast<unsigned>((waves + 3) / 4));
  USX_ENS_DISPATCH(cf_lag_kernel, A.n, lag_grid, dim3(256), 0, st, A.n, A.n0, A.n1, in, M, lag,
                   pmax_bits);
  if (A.ka <= kCfMaxKa && A.kl <= kCfMaxKl) {
    const unsigned tiles = static_cast<unsigned>(((A.n0 + kCfLines - 1) / kCfLines) * ((A.n1 + 63) / 64));
    hipLaunchKernelGGL(cf_estimate_kernel, dim3(tiles), dim3(256), 0, st, A, lag, velocity, power,
                       variance, valid, pmax_bits);
  } else {
    hipLaunchKernelGGL(cf_estimate_global_kernel, pix_grid(n_pix), dim3(256), 0, st, A, lag, velocity,
                       power, variance, valid, pmax_bits);
  }
  hipLaunchKernelGGL(cf_threshold_kernel, pix_grid(n_pix), dim3(256), 0, st, n_pix, power, valid,
                     pmax_bits, factor);
}

#undef USX_ENS_DISPATCH

void launch_displacement(const DisplacementArgs& A, const float2* before, const float2* after,
                         float* axial, float* correlation, hipStream_t st) {
  const size_t n_pix = static_cast<size_t>(A.n0) * A.n1;
  hipLaunchKernelGGL(displacement_kernel, pix_grid(n_pix), dim3(256), 0, st, A, before, after, axial,
                     correlation);
}

void launch_strain(int n0, int n1, float dz, int half, const float* axial, float* strain,
                   hipStream_t st) {
  const size_t n_pix = static_cast<size_t>(n0) * n1;
  hipLaunchKernelGGL(strain_kernel, pix_grid(n_pix), dim3(256), 0, st, n0, n1, dz, half, axial, strain);
}

}  // namespace usxd
