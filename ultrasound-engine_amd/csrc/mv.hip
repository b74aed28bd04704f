// mv.hip -- minimum-variance (Capon) combine on gfx950.
//
// Reference: combine_mv_windowed and cholesky_solve,
// core/src/beamformer.cpp:47-79, 137-214; coherence factor :525-534.
//
// Two stages per frame:
//
// mv_classify (one thread per pixel): active count n from the channel mask;
//   n == 0 -> 0; n < 4 -> the DAS sum (beamformer.cpp:142-146), with the
//   coherence factor; otherwise L = clamp(subarray, 2, max(2, n/2)) and the
//   pixel is appended to the list of its size bucket P = 16 ceil(L/16).
//
// mv_solve_fast<P, NR> (P <= 64, FP32, no refinement, no weight output: the
// default path) -- one wavefront per pixel, persistent, the lower triangle
// packed in LDS so up to 7 pixels of the largest bucket are resident per CU:
//   1. staging as below, every depth's loads in flight together;
//   2. covariance by the subarray shift recursion: lane m forms R(m,0)
//      directly and walks the m-th sub-diagonal,
//        R(i+1,k+1) = R(i,k) + D(i,k),
//        D(i,k) = sum_r x_r[i+n_sub] conj(x_r[k+n_sub]) - x_r[i] conj(x_r[k]),
//      O(L n_snap + L^2 nrow) products instead of O(L^2 n_snap); all of D at
//      once on MFMA (a rank-2 nrow product), so a walk step is one add;
//   3. loading as below;
//   4. left-looking Cholesky in blocks of 8 columns on the lower triangle
//      packed by columns, the forward
//      substitutions for 1 and for the subarray average folded into the
//      column loop;
//   5. out = (y1^H y2) / |y1|^2 * n with y1 = L^-1 1, y2 = L^-1 avg, which is
//      (w^H avg) / sum w * n for w = R^-1 1 without the back substitution.
//
// mv_solve<P, PREC> (the general path: P up to 96, BF16/FP16 covariance
// inputs, refinement, effective-weight output) -- persistent, one wavefront
// per pixel, LDS sized for the bucket:
//   1. active list; centre and +-range_avg neighbour vectors staged in LDS
//      (neighbours clipped at the grid edge, beamformer.cpp:159-161);
//   2. covariance R = sum over (depth, subarray start) of x x^H: with FP32
//      inputs by the shift recursion above (increments on MFMA), lane m
//      walking sub-diagonals m + 64 q; with BF16/FP16 inputs as a dense
//      rank-k update on MFMA,
//      complex expanded to real:
//        Re R = [xr xi][xr xi]^T,  Im R = [xi -xr][xr xi]^T,
//      v_mfma_f32_16x16x4_f32, two snapshots per instruction (k = re/im of
//      snapshot t, re/im of snapshot t+1), L padded to P with zero rows,
//      upper-triangle 16x16 blocks only;
//      R is kept in LDS as the full matrix when refinement is on, else as
//      its lower triangle packed by rows (half the LDS, so twice the pixels
//      resident per CU);
//   3. R /= n_snap; diagonal loading max(delta tr(R)/L, 1e-20) with the trace
//      summed in FP64;
//   4. in-place FP32 Cholesky in LDS, left-looking in blocks of 8 columns:
//      the finished columns' terms on MFMA, the in-block terms by the row's
//      lane (the strictly upper triangle keeps R for the residual); a
//      non-positive pivot -> DAS sum (beamformer.cpp:199);
//   5. triangular solves for R w = 1 with the vector in registers (lane i
//      holds w_i) and the pivot broadcast by readlane: no barriers;
//   6. optional single refinement step: r = 1 - R w in FP64, R d = r, w += d;
//   7. |sum w| < 1e-30 -> DAS; out = (w^H avg) / sum w * n with avg the
//      subarray average of the centre vector (beamformer.cpp:202-213);
//   8. coherence factor, or |out|^2 accumulation for incoherent compounding.
// precision 1 / 2 rounds the MFMA inputs to bf16 / fp16 (FP32 accumulate):
// the reported precision variants; 0 (FP32 inputs) is the parity path.
#include <hip/hip_runtime.h>
#include <hip/hip_bf16.h>
#include <hip/hip_fp16.h>

#include <algorithm>
#include <type_traits>

#include "kernels.hpp"

namespace usxd {

namespace {

typedef float f4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ float wsum(float v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
__device__ __forceinline__ double wsum_d(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
__device__ __forceinline__ float lane_bcast(float v, int k) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), k));
}

template <int PREC>
__device__ __forceinline__ float quant(float x) {
  if (PREC == 1) return __bfloat162float(__float2bfloat16(x));
  if (PREC == 2) return __half2float(__float2half(x));
  return x;
}

constexpr int kMaxBuckets = 16;  // P up to 256

// Element q of a per-lane slot array with q uniform across the wavefront
// (a select, so the array stays in registers).
template <int NS>
__device__ __forceinline__ float pick(const float (&v)[NS], int q) {
  float r = v[0];
#pragma unroll
  for (int k = 1; k < NS; ++k) r = q == k ? v[k] : r;
  return r;
}

// Coherence factor |sum y|^2 / (n sum |y|^2) over the active centre vector.
__device__ __forceinline__ float coherence(const float2* yc, const int* act, int na, int lane) {
  float cr = 0.0f, ci = 0.0f, inc = 0.0f;
  for (int a = lane; a < na; a += 64) {
    const float2 v = yc[act[a]];
    cr += v.x;
    ci += v.y;
    const float m = hypotf(v.x, v.y);
    inc += m * m;
  }
  cr = wsum(cr);
  ci = wsum(ci);
  inc = wsum(inc);
  const float mc = hypotf(cr, ci);
  return (inc > 0.0f) ? (mc * mc) / (static_cast<float>(na) * inc) : 0.0f;
}

__device__ __forceinline__ void store_out(const MvArgs& M, int pix, float2 out, float2* frame,
                                          float* acc) {
  if (M.incoherent) {
    const float m = hypotf(out.x, out.y);
    acc[pix] += m * m;
  } else {
    frame[pix] = out;
  }
}

// Sum of the active centre vector (the DAS fallback).
__device__ __forceinline__ float2 das_sum(const float2* yc, const int* act, int na, int lane) {
  float sr = 0.0f, si = 0.0f;
  for (int a = lane; a < na; a += 64) {
    sr += yc[act[a]].x;
    si += yc[act[a]].y;
  }
  return make_float2(wsum(sr), wsum(si));
}

struct Staged {
  int na;    // active channels
  int nrow;  // depths staged (centre and in-grid neighbours)
  int crow;  // row index of the centre depth
};

// Steps 1-2 of a pixel: the active channel list act[0..na) and the centre and
// +-range_avg neighbour vectors over those channels, rows[r * na + a]
// (neighbours clipped at the grid edge, beamformer.cpp:159-161).
__device__ __forceinline__ Staged stage_pixel(const MvArgs& M, const float2* __restrict__ Y,
                                              const uint32_t* __restrict__ mask, int pix, int lane,
                                              int* act, float2* rows) {
  const int n_ch = M.n_ch;
  const int nw = (n_ch + 31) / 32;
  const int i0 = pix / M.n1;
  const int j = pix - i0 * M.n1;
  const uint32_t* mp = mask + static_cast<size_t>(pix) * nw;
  int na = 0;
  for (int base = 0; base < n_ch; base += 64) {
    const int ch = base + lane;
    const bool on = ch < n_ch && ((mp[ch >> 5] >> (ch & 31)) & 1u);
    const unsigned long long bal = __ballot(on);
    if (on) act[na + __popcll(bal & ((1ull << lane) - 1ull))] = ch;
    na += __popcll(bal);
  }
  __syncthreads();
  int nrow = 0, crow = 0;
  for (int d = -M.range_avg; d <= M.range_avg; ++d) {
    const int k = j + d;
    if (k < 0 || k >= M.n1) continue;
    if (d == 0) crow = nrow;
    const float2* src = Y + (static_cast<size_t>(i0) * M.n1 + k) * n_ch;
    // A channel inactive at the neighbouring depth contributes 0 there
    // (its Y was never accumulated, beamformer.cpp:469); the channel
    // kernel does not write such entries, so test the neighbour's mask.
    const uint32_t* mk = mask + (static_cast<size_t>(i0) * M.n1 + k) * nw;
    for (int a = lane; a < na; a += 64) {
      const int ch = act[a];
      const bool on = d == 0 || ((mk[ch >> 5] >> (ch & 31)) & 1u);
      rows[nrow * na + a] = on ? src[ch] : make_float2(0.0f, 0.0f);
    }
    ++nrow;
  }
  __syncthreads();
  return Staged{na, nrow, crow};
}

// Lanes of a wavefront split into G groups of LP = 64 / G lanes, one pixel
// per group. group_bcast: v of lane k of the caller's group -- one readlane
// for a single group; with several groups one ds_bpermute (the LDS crossbar,
// no memory access; measured faster than G readlanes and G - 1 selects even
// on the Cholesky's critical path); group_sum, group_sum_d: the sum over the
// caller's group.
template <int G>
__device__ __forceinline__ float group_bcast(float v, int k, int lane) {
  if constexpr (G == 1) {
    return lane_bcast(v, k);
  } else {
    constexpr int LP = 64 / G;
    return __shfl(v, (lane & ~(LP - 1)) + k, 64);
  }
}
template <int G>
__device__ __forceinline__ float group_sum(float v) {
#pragma unroll
  for (int o = 32 / G; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
template <int G>
__device__ __forceinline__ double group_sum_d(double v) {
#pragma unroll
  for (int o = 32 / G; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
// This workgroup's slot in each round of a persistent grid. Workgroups are
// dealt to the XCDs round robin (b runs on XCD b mod X), so the slots are
// handed out in runs of Q per XCD: neighbouring pixels of a line, which stage
// each other's depth rows, then mostly meet in one XCD's L2 instead of each
// fetching those rows from HBM. Runs are short because list order follows
// depth and the work per pixel grows with it: one long run per XCD and round
// would give one XCD the deepest pixels every time (measured 3 % slower).
__device__ __forceinline__ int xcd_slot(int X) {
  constexpr int Q = 8;
  const int g = gridDim.x, b = blockIdx.x;
  if (X <= 1 || g % (X * Q) != 0) return b;
  const int x = b % X, r = b / X;
  return (r / Q) * (Q * X) + x * Q + r % Q;
}

template <int G>
__device__ __forceinline__ double group_bcast_d(double v, int k, int lane) {
  const long long u = __double_as_longlong(v);
  const float lo = group_bcast<G>(__int_as_float(static_cast<int>(u)), k, lane);
  const float hi = group_bcast<G>(__int_as_float(static_cast<int>(u >> 32)), k, lane);
  return __longlong_as_double((static_cast<long long>(__float_as_int(hi)) << 32) |
                              static_cast<unsigned int>(__float_as_int(lo)));
}
__device__ __forceinline__ int wave_max(int v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, 64));
  return v;
}
__device__ __forceinline__ int wave_min(int v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o, 64));
  return v;
}

// stage_pixel for one pixel per lane group, with the depth count fixed at
// NR = 2 range_avg + 1 and the row stride at RS >= n_ch (compile-time, so
// the NR rows of one index are one address plus immediate offsets): every
// row's mask word and sample for a channel are loaded before any is stored,
// so a lane's NR global loads are in flight together. Rows past the grid
// edge are zero (rows nrow .. NR-1). act and rows are the group's own.
template <int NR, int RS, int G>
__device__ __forceinline__ Staged stage_group(const MvArgs& M, const float2* __restrict__ Y,
                                              const uint32_t* __restrict__ mask, int pix, int lane,
                                              int* act, float2* rows) {
  constexpr int RA = NR / 2;
  constexpr int LP = 64 / G;
  const int sub = lane % LP, g = lane / LP;
  const unsigned long long gbits =
      G == 1 ? ~0ull : (((1ull << LP) - 1ull) << (g * LP));
  const int n_ch = M.n_ch;
  const int nw = (n_ch + 31) / 32;
  const int i0 = pix / M.n1;
  const int j = pix - i0 * M.n1;
  const uint32_t* mp = mask + static_cast<size_t>(pix) * nw;
  int na = 0;
  for (int base = 0; base < n_ch; base += LP) {
    const int ch = base + sub;
    const bool on = ch < n_ch && ((mp[ch >> 5] >> (ch & 31)) & 1u);
    const unsigned long long bal = __ballot(on) & gbits;
    if (on) act[na + __popcll(bal & ((1ull << lane) - 1ull))] = ch;
    na += __popcll(bal);
  }
  __syncthreads();
  const int dlo = max(-RA, -j), dhi = min(RA, M.n1 - 1 - j);  // in-grid depths j + d
  const size_t line = static_cast<size_t>(i0) * M.n1 + j;
#pragma unroll 2
  for (int a = sub; a < na; a += LP) {
    const int ch = act[a];
    float2 v[NR];
    uint32_t mw[NR];
#pragma unroll
    for (int r = 0; r < NR; ++r) {
      const size_t p = line + min(dlo + r, dhi);
      v[r] = Y[p * n_ch + ch];
      mw[r] = mask[p * nw + (ch >> 5)];
    }
#pragma unroll
    for (int r = 0; r < NR; ++r) {
      const int d = dlo + r;
      // a channel inactive at the neighbouring depth contributes 0 there
      // (beamformer.cpp:469); the centre's channels are active by definition
      const bool on = d <= dhi && (d == 0 || ((mw[r] >> (ch & 31)) & 1u));
      rows[r * RS + a] = on ? v[r] : make_float2(0.0f, 0.0f);
    }
  }
  __syncthreads();
  return Staged{na, dhi - dlo + 1, -dlo};
}

// stage_group for pixels whose active channels form one run from channel lo
// (every group of the wavefront): a single round of loads -- the pixel's mask
// words (for n_a) and each row's entries over the RS channels from lo (the
// run's longest, clamped to the channel count; those past the run are
// dropped), issued together -- so no active list and no load that waits on
// another. Rows as stage_group's.
template <int NR, int RS, int G>
__device__ __forceinline__ Staged stage_group_run(const MvArgs& M, const float2* __restrict__ Y,
                                                  const uint32_t* __restrict__ mask, int pix, int lo,
                                                  int lane, float2* rows) {
  constexpr int RA = NR / 2;
  constexpr int LP = 64 / G;
  constexpr int NS = (RS + LP - 1) / LP;  // channel slots per lane
  const int sub = lane % LP;
  const int n_ch = M.n_ch;
  const int nw = (n_ch + 31) / 32;  // <= 8
  const int i0 = pix / M.n1;
  const int j = pix - i0 * M.n1;
  const uint32_t* mp = mask + static_cast<size_t>(pix) * nw;
  const int dlo = max(-RA, -j), dhi = min(RA, M.n1 - 1 - j);  // in-grid depths j + d
  const size_t line = static_cast<size_t>(i0) * M.n1 + j;
  const uint32_t cw = sub < nw ? mp[sub] : 0u;  // lane w of the group: mask word w
  float2 v[NS][NR];
  uint32_t mw[NS][NR];
#pragma unroll
  for (int s = 0; s < NS; ++s) {
    const int ch = min(lo + sub + s * LP, n_ch - 1);
#pragma unroll
    for (int r = 0; r < NR; ++r) {
      const size_t p = line + min(dlo + r, dhi);
      v[s][r] = Y[p * n_ch + ch];
      mw[s][r] = mask[p * nw + (ch >> 5)];
    }
  }
  const int na = static_cast<int>(group_sum<G>(static_cast<float>(__popc(cw))));  // exact: <= 256
#pragma unroll
  for (int s = 0; s < NS; ++s) {
    const int a = sub + s * LP, ch = lo + a;
    if (a < na) {
#pragma unroll
      for (int r = 0; r < NR; ++r) {
        const int d = dlo + r;
        // a channel inactive at the neighbouring depth contributes 0 there
        // (beamformer.cpp:469); the centre's channels are active by definition
        const bool on = d <= dhi && (d == 0 || ((mw[s][r] >> (ch & 31)) & 1u));
        rows[r * RS + a] = on ? v[s][r] : make_float2(0.0f, 0.0f);
      }
    }
  }
  __syncthreads();
  return Staged{na, dhi - dlo + 1, -dlo};
}

}  // namespace

// Bytes of dynamic LDS for one wavefront of bucket P: the full matrix with
// refinement (its upper triangle keeps R for the residual), else the lower
// triangle packed by rows.
size_t mv_smem_bytes(int n_ch, int range_avg, int subarray, int P, bool refine) {
  const int na_max = subarray <= P ? n_ch : std::min(n_ch, 2 * P + 1);
  const size_t ints = (static_cast<size_t>(n_ch) * 4 + 15) / 16 * 16;
  const size_t matrix = refine ? static_cast<size_t>(P) * (P + 1) : static_cast<size_t>(P) * (P + 1) / 2;
  return ints + sizeof(float2) * (static_cast<size_t>(2 * range_avg + 1) * na_max + matrix +
                                  2 * static_cast<size_t>(P));
}

// Row stride of mv_solve_fast's staged vectors (float2) for bucket P, or 0
// when the channel count is beyond its instantiations. When the subarray
// exceeds P, every pixel of the bucket has L = n / 2 <= P, so at most
// 2P + 1 active channels: the rows take round8(2P + 1) instead of n_ch.
constexpr int mv_tight_stride(int P) { return (2 * P + 1 + 7) / 8 * 8; }
static int mv_fast_row_stride(int n_ch, int subarray, int P) {
  if (subarray > P && mv_tight_stride(P) < 128) return mv_tight_stride(P);
  return n_ch <= 128 ? 128 : (n_ch <= 256 ? 256 : 0);
}

// Pixels per wavefront of mv_solve_fast for bucket P: groups of 16, 32 or
// 64 lanes, one matrix row per lane.
static int mv_fast_groups(int P) { return P <= 16 ? 4 : (P <= 32 ? 2 : 1); }

// Bytes of mv_solve_fast's active list per pixel group: n_ch ints, and at
// least lp float2 (one sink entry per lane for the covariance walks, which
// run after the list is spent), a multiple of 16.
__host__ __device__ inline size_t mv_fast_ints_bytes(int n_ch, int lp) {
  const size_t a = (static_cast<size_t>(n_ch) * 4 + 15) / 16 * 16, b = static_cast<size_t>(lp) * 8;
  return a > b ? a : b;
}

// Bytes of dynamic LDS for mv_solve_fast: per pixel group, the packed lower
// triangle, the active list and NR rows at the fixed stride.
size_t mv_fast_smem_bytes(int n_ch, int range_avg, int subarray, int P) {
  const int groups = mv_fast_groups(P);
  return static_cast<size_t>(groups) *
         (mv_fast_ints_bytes(n_ch, 64 / groups) +
          sizeof(float2) * (static_cast<size_t>(2 * range_avg + 1) * mv_fast_row_stride(n_ch, subarray, P) +
                            static_cast<size_t>(P) * (P + 1) / 2));
}

__global__ void mv_zero_counts(int* counts) {
  if (threadIdx.x < kMaxBuckets) counts[threadIdx.x] = 0;
}

// Slot in bucket b's list: the workgroup's lanes count per bucket in LDS
// (their local slots), then one global atomic per (workgroup, bucket) claims
// a base; a workgroup's pixels of one bucket take consecutive slots. Every
// thread of the workgroup calls this (barriers).
__device__ __forceinline__ int bucket_slot(int* counts, int b, bool take) {
  __shared__ int hist[kMaxBuckets], base[kMaxBuckets];
  const int t = threadIdx.x;
  if (t < kMaxBuckets) hist[t] = 0;
  __syncthreads();
  const int local = take ? atomicAdd(&hist[b], 1) : 0;
  __syncthreads();
  if (t < kMaxBuckets) base[t] = hist[t] ? atomicAdd(&counts[t], hist[t]) : 0;
  __syncthreads();
  return take ? base[b] + local : 0;
}

__global__ __launch_bounds__(256) void mv_classify(MvArgs M, const float2* __restrict__ Y,
                                                   const uint32_t* __restrict__ mask,
                                                   float2* __restrict__ frame, float* acc,
                                                   int* counts, int* lists) {
  const int pix = blockIdx.x * blockDim.x + threadIdx.x;
  const int n_pix = M.n0 * M.n1;
  const bool valid = pix < n_pix;
  const int nw = (M.n_ch + 31) / 32;
  const uint32_t* mp = mask + static_cast<size_t>(valid ? pix : 0) * nw;
  int na = 0, lo = -1, hi = -1;  // active channels, the first and the last
  if (valid)
    for (int w = 0; w < nw; ++w) {
      const uint32_t bits = mp[w];
      na += __popc(bits);
      if (bits && lo < 0) lo = w * 32 + __ffs(bits) - 1;
      if (bits) hi = w * 32 + 31 - __clz(bits);
    }
  if (valid && na == 0 && !M.incoherent) frame[pix] = make_float2(0.0f, 0.0f);
  if (valid && M.status) M.status[pix] = na == 0 ? 3 : (na < 4 ? 1 : 0);
  if (valid && na > 0 && na < 4 && M.weights) {
    float2* gw = M.weights + static_cast<size_t>(pix) * M.n_ch;
    for (int w = 0; w < nw; ++w)
      for (uint32_t bits = mp[w]; bits; bits &= bits - 1)
        gw[w * 32 + __ffs(bits) - 1] = make_float2(1.0f, 0.0f);
  }
  if (valid && na > 0 && na < 4) {
    const float2* yc = Y + static_cast<size_t>(pix) * M.n_ch;
    float sr = 0.0f, si = 0.0f, inc = 0.0f;
    for (int w = 0; w < nw; ++w) {
      uint32_t bits = mp[w];
      while (bits) {
        const int b = __ffs(bits) - 1;
        bits &= bits - 1;
        const float2 v = yc[w * 32 + b];
        sr += v.x;
        si += v.y;
        const float m = hypotf(v.x, v.y);
        inc += m * m;
      }
    }
    float2 out = make_float2(sr, si);
    if (M.cf) {
      const float mc = hypotf(sr, si);
      const float f = (inc > 0.0f) ? (mc * mc) / (static_cast<float>(na) * inc) : 0.0f;
      out.x *= f;
      out.y *= f;
    }
    store_out(M, pix, out, frame, acc);
  }
  const bool take = valid && na >= 4;
  const int L = min(max(M.subarray, 2), max(2, na / 2));
  const int b = (L + 15) / 16 - 1;  // size bucket, P = 16 (b + 1)
  // the fast kernel's pixels of L = P - 1, P: their own count, from the back
  // of the bucket's array
  const bool full = mv_fast_config(M) && b < 4 && L >= 16 * (b + 1) - 1;
  const int cb = full ? kFullList + b : b;
  const int slot = bucket_slot(counts, take ? cb : -1, take);
  // the fast kernel's entries carry the start of a contiguous run
  const bool run = mv_fast_config(M) && b < 4 && n_pix <= kPixMask + 1 && hi - lo + 1 == na;
  if (take)
    lists[static_cast<size_t>(b) * n_pix + (full ? n_pix - 1 - slot : slot)] =
        run ? pix | ((lo + 1) << kRunShift) : pix;
}

// Trailing update of the general path's Cholesky block [j0, j0 + 8) on MFMA,
// in the full matrix (rows of stride P + 1 complex, factor in the lower
// triangle): S(r, j0 + c) = A(r, j0 + c) - sum_{k < j0} L(r, k) conj(L(j0 + c, k))
// for rows j0 + c <= r < L, written back in place. Complex products expanded
// to real: with a = L(r, k), b = L(j0 + c, k),
//   Re S -= [a_r a_i] . [b_r  b_i],   Im S -= [a_r a_i] . [-b_i  b_r],
// a (rows x 2 j0) . (2 j0 x 16) real product on v_mfma_f32_16x16x4_f32: 16 rows
// per tile, 2 complex columns per step, accumulated onto the matrix entries
// with B negated. Lane l supplies A[row l % 16][k-part l / 16] and
// B[k-part l / 16][column l % 16] and holds D[4 (l / 16) + i][l % 16]. Rows of
// a tile at or past L give values never stored. packed: the lower triangle
// by rows, (i, k) at i (i + 1) / 2 + k; entries read outside it (k > i, from
// rows never stored) still lie inside the P (P + 1) / 2 allocation for i, k < P.
template <int P>
__device__ __forceinline__ void block_update_full(float* A, bool packed, int lane, int j0, int L) {
  auto base = [packed](int r) { return 2 * (packed ? (r * (r + 1)) >> 1 : r * (P + 1)); };
  const int lr = lane & 15, lq = lane >> 4;
  const int c = lr >> 1, part = lr & 1;
  const int kp = lq & 1;
  // -B[2k + kp][2c + part]: one component of b times a lane constant sign
  const int comp = part ^ kp;
  const float sgn = (part == 1 && kp == 0) ? 1.0f : -1.0f;
  const int col = j0 + c;  // < P: j0 < L <= P, both multiples of 8
  const int t1 = (L + 15) >> 4;
  const float* bq = A + base(col) + comp;
  for (int mt = j0 >> 4; mt < t1; ++mt) {
    f4 acc;
#pragma unroll
    for (int i = 0; i < 4; ++i) acc[i] = A[base(16 * mt + 4 * lq + i) + 2 * col + part];
    const float* arow = A + base(16 * mt + lr) + kp;
#pragma unroll 2
    for (int st = 0; st < (j0 >> 1); ++st) {
      const int k = 2 * st + (lq >> 1);
      acc = __builtin_amdgcn_mfma_f32_16x16x4f32(arow[2 * k], sgn * bq[2 * k], acc, 0, 0, 0);
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int r = 16 * mt + 4 * lq + i;
      if (r >= col && r < L && col < L) A[base(r) + 2 * col + part] = acc[i];
    }
  }
}

// Walk increments of the general path on MFMA (as walk_increments_mfma of the
// fast path, for one pixel per wavefront): D(i, k) = sum_r x_r[i + n_sub]
// conj(x_r[k + n_sub]) - x_r[i] conj(x_r[k]) for k <= i < L into the lower
// triangle (packed by rows, or rows of stride P + 1), from the nrow staged
// rows of stride na; indices clamped to L - 2 (reads stay in the staged
// entries 0 .. na - 1; row L - 1 gets values the walk overwrites unread).
template <int P>
__device__ __forceinline__ void walk_increments_full(float* A, bool packed, const float* rows,
                                                     int na, int nrow, int ns, int L, int lane) {
  auto base = [packed](int r) { return 2 * (packed ? (r * (r + 1)) >> 1 : r * (P + 1)); };
  const int lr = lane & 15, lq = lane >> 4;
  const int c = lr >> 1, part = lr & 1;
  const int hi = lq >> 1, comp = lq & 1;
  const int bcomp = part ^ comp;
  const float bsgn = ((part == 1 && comp == 0) ? -1.0f : 1.0f) * (hi ? 1.0f : -1.0f);
  const int off = hi ? ns : 0;
  const int t1 = (L + 15) >> 4, c1 = (L + 7) >> 3;
  for (int mt = 0; mt < t1; ++mt) {
    const float* ra = rows + 2 * (min(16 * mt + lr, L - 2) + off) + comp;
    const int cbn = min(2 * mt + 2, c1);
    for (int cb = 0; cb < cbn; ++cb) {
      const float* rb = rows + 2 * (min(8 * cb + c, L - 2) + off) + bcomp;
      f4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
      for (int r = 0; r < nrow; ++r)
        acc = __builtin_amdgcn_mfma_f32_16x16x4f32(ra[2 * r * na], bsgn * rb[2 * r * na], acc, 0, 0, 0);
      const int col = 8 * cb + c;
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int row = 16 * mt + 4 * lq + e;
        if (row >= col && row < L) A[base(row) + 2 * col + part] = acc[e];
      }
    }
  }
}

template <int P, int PREC>
__global__ __launch_bounds__(64) void mv_solve(MvArgs M, const float2* __restrict__ Y,
                                               const uint32_t* __restrict__ mask,
                                               float2* __restrict__ frame, float* acc,
                                               const int* __restrict__ counts,
                                               const int* __restrict__ lists, int na_max) {
  constexpr int NB = P / 16;
  constexpr int NS = (P + 63) / 64;  // matrix rows (and vector elements) per lane
  constexpr int NPAIRS = NB * (NB + 1) / 2;
  constexpr int BUCKET = P / 16 - 1;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int lane = threadIdx.x;
  const int n_ch = M.n_ch;
  const int nw = (n_ch + 31) / 32;
  const int n_pix = M.n0 * M.n1;
  int* act = reinterpret_cast<int*>(smem_raw);
  float2* rows = reinterpret_cast<float2*>(smem_raw + (static_cast<size_t>(n_ch) * 4 + 15) / 16 * 16);
  // A with refinement: P rows of stride P + 1 (conflict-free column access),
  // strictly upper triangle R (kept for the residual), lower triangle +
  // diagonal the Cholesky factor. Without refinement the upper triangle is
  // never read: the lower triangle alone, packed by rows, halves the LDS per
  // pixel. dg[a] = (loaded diagonal of R, 1 / L_aa).
  constexpr int S = P + 1;
  const bool packed = !M.refine;
  auto ix = [packed](int i, int k) { return packed ? ((i * (i + 1)) >> 1) + k : i * S + k; };
  float2* A = rows + static_cast<size_t>(2 * M.range_avg + 1) * na_max;
  float2* dg = A + (packed ? P * (P + 1) / 2 : P * S);
  float2* rv = dg + P;  // scratch vector

  const int count = counts[BUCKET];
  for (int item = xcd_slot(M.n_xcd); item < count; item += gridDim.x) {
    const int pix = lists[static_cast<size_t>(BUCKET) * n_pix + item];
    __syncthreads();  // LDS of the previous pixel is free

    // ---- 1-2. active list, centre and neighbour rows ------------------------
    const Staged sg = stage_pixel(M, Y, mask, pix, lane, act, rows);
    const int na = sg.na, nrow = sg.nrow, crow = sg.crow;
    const float2* yc = Y + static_cast<size_t>(pix) * n_ch;
    const int L = min(max(M.subarray, 2), max(2, na / 2));
    const int ns = na - L + 1;
    const int n_snap = nrow * ns;

    // ---- 3. covariance on MFMA ----------------------------------------------
    const int li = lane & 15, kk = lane >> 4;
    const int tsub = kk >> 1, comp = kk & 1;
    // FP16 inputs: scale by a power of two so the largest component sits
    // near 2^8 (half overflows at 65504); the scale leaves relative rounding
    // unchanged and is divided back out of R exactly.
    float sc = 1.0f, unsc2 = 1.0f;
    if (PREC == 2) {
      float mx = 0.0f;
      for (int q = lane; q < nrow * na; q += 64)
        mx = fmaxf(mx, fmaxf(fabsf(rows[q].x), fabsf(rows[q].y)));
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) mx = fmaxf(mx, __shfl_xor(mx, o, 64));
      if (mx > 0.0f) {
        const int e = ilogbf(mx) - 8;
        sc = ldexpf(1.0f, -e);
        unsc2 = ldexpf(1.0f, 2 * e);
      }
    }
    const float inv = unsc2 / static_cast<float>(n_snap);
    if constexpr (PREC == 0) {
      // FP32 inputs: the subarray shift recursion along the sub-diagonals,
      // lane owns m = lane + 64 q,
      //   R(m,0) = sum_r sum_{a<ns} x_r[a+m] conj(x_r[a]),
      //   R(m+t+1,t+1) = R(m+t,t) + D(m+t,t),
      //   D(i,k) = sum_r x_r[i+ns] conj(x_r[k+ns]) - x_r[i] conj(x_r[k]),
      // O(L n_snap + L^2 nrow) products instead of the rank-k update's
      // O(L^2 n_snap). Lanes past their diagonal's end read clamped indices
      // and store nothing.
      float cr[NS], ci[NS];
#pragma unroll
      for (int q = 0; q < NS; ++q) {
        cr[q] = 0.0f;
        ci[q] = 0.0f;
      }
      for (int r = 0; r < nrow; ++r) {
        const float2* x = rows + r * na;
#pragma unroll 2
        for (int a = 0; a < ns; ++a) {
          const float2 b = x[a];
#pragma unroll
          for (int q = 0; q < NS; ++q) {
            const float2 v = x[a + min(lane + 64 * q, L - 1)];
            cr[q] = fmaf(v.x, b.x, fmaf(v.y, b.y, cr[q]));
            ci[q] = fmaf(v.y, b.x, fmaf(-v.x, b.y, ci[q]));
          }
        }
      }
      // Increments D on MFMA into the lower triangle, then lane m walks its
      // sub-diagonals in place: at (m+t, t) it reads D, stores R and adds;
      // the reads of 4 steps go out before their stores (a lane touches only
      // its own diagonals).
      walk_increments_full<P>(reinterpret_cast<float*>(A), packed,
                              reinterpret_cast<const float*>(rows), na, nrow, ns, L, lane);
      __syncthreads();
      for (int t0 = 0; t0 < L; t0 += 4) {
        float2 d[NS][4];
#pragma unroll
        for (int q = 0; q < NS; ++q)
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            const int i = min(lane + 64 * q + t0 + u, L - 1);
            d[q][u] = A[ix(i, min(t0 + u, i))];  // past the diagonal's end: read, never stored
          }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const int t = t0 + u;
#pragma unroll
          for (int q = 0; q < NS; ++q) {
            const int m = lane + 64 * q, i = m + t;
            if (i < L) {
              const float2 v = make_float2(cr[q] * inv, ci[q] * inv);
              if (m == 0) {
                dg[t] = make_float2(v.x, 0.0f);
              } else {
                A[ix(i, t)] = v;                                   // factor input, lower
                if (!packed) A[t * S + i] = make_float2(v.x, -v.y);  // R, upper
              }
            }
            cr[q] += d[q][u].x;
            ci[q] += d[q][u].y;
          }
        }
      }
    } else {
      // All upper-triangle block pairs in one pass: per snapshot pair, one
      // operand load per 16-row block, then 2 MFMAs per block pair.
      f4 accR[NPAIRS], accI[NPAIRS];
#pragma unroll
      for (int q = 0; q < NPAIRS; ++q) {
        accR[q] = f4{0.0f, 0.0f, 0.0f, 0.0f};
        accI[q] = accR[q];
      }
      bool rowok[NB];
#pragma unroll
      for (int bb = 0; bb < NB; ++bb) rowok[bb] = bb * 16 + li < L;
      int ri = 0, s0 = tsub;  // snapshot t + tsub as (row, subarray start)
      if (s0 >= ns) {
        s0 -= ns;
        ++ri;
      }
      for (int t = 0; t < n_snap; t += 2) {
        const bool tv = t + tsub < n_snap;
        const float2* xrow = rows + ri * na + s0 + li;
        float ar[NB], aim[NB];
#pragma unroll
        for (int bb = 0; bb < NB; ++bb) {
          float2 x = make_float2(0.0f, 0.0f);
          if (tv && rowok[bb]) x = xrow[bb * 16];
          ar[bb] = quant<PREC>(sc * (comp ? x.y : x.x));
          aim[bb] = quant<PREC>(sc * (comp ? -x.x : x.y));
        }
        int q = 0;
#pragma unroll
        for (int bi = 0; bi < NB; ++bi) {
#pragma unroll
          for (int bj = bi; bj < NB; ++bj, ++q) {
            accR[q] = __builtin_amdgcn_mfma_f32_16x16x4f32(ar[bi], ar[bj], accR[q], 0, 0, 0);
            accI[q] = __builtin_amdgcn_mfma_f32_16x16x4f32(aim[bi], ar[bj], accI[q], 0, 0, 0);
          }
        }
        s0 += 2;
        if (s0 >= ns) {
          s0 -= ns;
          ++ri;
          if (s0 >= ns) {  // ns == 1: two rows per step
            s0 -= ns;
            ++ri;
          }
        }
      }
      // D[4*(lane/16) + e][lane%16] -> R[bi*16 + row][bj*16 + col]
      int q = 0;
#pragma unroll
      for (int bi = 0; bi < NB; ++bi) {
#pragma unroll
        for (int bj = bi; bj < NB; ++bj, ++q) {
#pragma unroll
          for (int e = 0; e < 4; ++e) {
            const int row = bi * 16 + kk * 4 + e;
            const int col = bj * 16 + li;
            if (row < L && col < L && row <= col) {
              const float2 v = make_float2(accR[q][e] * inv, accI[q][e] * inv);
              if (row == col) {
                dg[row] = make_float2(v.x, 0.0f);
              } else {
                if (!packed) A[row * S + col] = v;            // R, upper
                A[ix(col, row)] = make_float2(v.x, -v.y);  // factor input, lower
              }
            }
          }
        }
      }
    }
    __syncthreads();

    // ---- 4. diagonal loading, trace in FP64 -----------------------------------
    double tr = 0.0;
    for (int a = lane; a < L; a += 64) tr += static_cast<double>(dg[a].x);
    tr = wsum_d(tr);
    const float eps = fmaxf(static_cast<float>(static_cast<double>(M.loading) * tr / L), 1e-20f);
    for (int a = lane; a < L; a += 64) {
      const float d = dg[a].x + eps;
      dg[a] = make_float2(d, 0.0f);
      A[ix(a, a)] = make_float2(d, 0.0f);
    }
    __syncthreads();

    // ---- 5. Cholesky, left-looking in blocks of CB columns: lane i owns rows
    // i, i + 64 --------------------------------------------------------------
    //   L_ij = (A_ij - sum_{k<j} L_ik conj(L_jk)) / L_jj, L_jj = sqrt(A_jj - sum |L_jk|^2),
    // the reference's sums (beamformer.cpp:48-63). For the block [j0, j0 + CB)
    // the finished columns' terms k < j0 are subtracted on MFMA in place
    // (block_update_full); lane i then loads its rows of the block and finishes
    // them column by column with the in-block terms k = j0 .. j - 1, taken by
    // readlane from the lane that owns row j, k increasing. Two barriers per
    // block.
    constexpr int CB = 8;
    bool ok = true;
    for (int j0 = 0; j0 < L && ok; j0 += CB) {
      if (j0 > 0) {
        block_update_full<P>(reinterpret_cast<float*>(A), packed, lane, j0, L);
        __syncthreads();
      }
      float sr[NS][CB], si[NS][CB], rjs[CB];
#pragma unroll
      for (int q = 0; q < NS; ++q) {
        const int row = lane + 64 * q;
#pragma unroll
        for (int c = 0; c < CB; ++c) {
          const int j = j0 + c;
          float2 v = make_float2(0.0f, 0.0f);
          if (j < L && row >= j && row < L) v = A[ix(row, j)];
          sr[q][c] = v.x;
          si[q][c] = v.y;
        }
      }
#pragma unroll
      for (int c = 0; c < CB; ++c) {
        const int j = j0 + c;
        rjs[c] = 0.0f;
        if (j >= L || !ok) continue;
#pragma unroll
        for (int u = 0; u < c; ++u) {  // in-block terms, k = j0 + u: -= L_ik conj(L_jk)
          float tr_[NS], ti_[NS];
#pragma unroll
          for (int q = 0; q < NS; ++q) {
            tr_[q] = sr[q][u];
            ti_[q] = si[q][u];
          }
          const float br = lane_bcast(pick<NS>(tr_, j >> 6), j & 63);
          const float bi = lane_bcast(pick<NS>(ti_, j >> 6), j & 63);
#pragma unroll
          for (int q = 0; q < NS; ++q) {
            const float ar = sr[q][u], ai = si[q][u];
            sr[q][c] = fmaf(-ar, br, fmaf(-ai, bi, sr[q][c]));
            si[q][c] = fmaf(-ai, br, fmaf(ar, bi, si[q][c]));
          }
        }
        float dq[NS];
#pragma unroll
        for (int q = 0; q < NS; ++q) dq[q] = sr[q][c];
        const float d = lane_bcast(pick<NS>(dq, j >> 6), j & 63);
        if (!(d > 0.0f)) {
          ok = false;
          continue;
        }
        const float ljj = sqrtf(d);
        const float rj = 1.0f / ljj;
        rjs[c] = rj;
#pragma unroll
        for (int q = 0; q < NS; ++q) {
          const int row = lane + 64 * q;
          const bool piv = row == j;
          sr[q][c] = piv ? ljj : sr[q][c] * rj;  // rows above j are not stored
          si[q][c] = piv ? 0.0f : si[q][c] * rj;
        }
      }
      if (!ok) break;
#pragma unroll
      for (int q = 0; q < NS; ++q) {
        const int row = lane + 64 * q;
#pragma unroll
        for (int c = 0; c < CB; ++c) {
          const int j = j0 + c;
          if (j < L && row >= j && row < L) {
            A[ix(row, j)] = make_float2(sr[q][c], si[q][c]);
            if (row == j) dg[j].y = rjs[c];
          }
        }
      }
      __syncthreads();
    }
    float2 out;
    bool das = !ok;
    if (!ok) {
      float sr = 0.0f, si = 0.0f;
      for (int a = lane; a < na; a += 64) {
        sr += yc[act[a]].x;
        si += yc[act[a]].y;
      }
      out = make_float2(wsum(sr), wsum(si));
    } else {
      // ---- 6. solves with the vector in registers (element i + 64 q in slot q
      // of lane i) and the pivot element broadcast by readlane: no barriers ----
      float wr[NS], wi[NS];
      auto solve = [&](float (&br)[NS], float (&bi)[NS]) {
        for (int k = 0; k < L; ++k) {  // L y = b
          const float rk = dg[k].y;
          const float yr = lane_bcast(pick<NS>(br, k >> 6), k & 63) * rk;
          const float yi = lane_bcast(pick<NS>(bi, k >> 6), k & 63) * rk;
#pragma unroll
          for (int q = 0; q < NS; ++q) {
            const int row = lane + 64 * q;
            if (row == k) {
              br[q] = yr;
              bi[q] = yi;
            } else if (row > k && row < L) {
              const float2 l = A[ix(row, k)];
              br[q] = fmaf(-l.x, yr, fmaf(l.y, yi, br[q]));
              bi[q] = fmaf(-l.x, yi, fmaf(-l.y, yr, bi[q]));
            }
          }
        }
        for (int k = L - 1; k >= 0; --k) {  // L^H x = y
          const float rk = dg[k].y;
          const float xr = lane_bcast(pick<NS>(br, k >> 6), k & 63) * rk;
          const float xi = lane_bcast(pick<NS>(bi, k >> 6), k & 63) * rk;
#pragma unroll
          for (int q = 0; q < NS; ++q) {
            const int row = lane + 64 * q;
            if (row == k) {
              br[q] = xr;
              bi[q] = xi;
            } else if (row < k) {
              const float2 l = A[ix(k, row)];  // -= conj(L[k][row]) * x_k
              br[q] = fmaf(-l.x, xr, fmaf(-l.y, xi, br[q]));
              bi[q] = fmaf(-l.x, xi, fmaf(l.y, xr, bi[q]));
            }
          }
        }
      };
#pragma unroll
      for (int q = 0; q < NS; ++q) {
        wr[q] = lane + 64 * q < L ? 1.0f : 0.0f;
        wi[q] = 0.0f;
      }
      solve(wr, wi);

      if (M.refine) {  // ---- 7. one refinement step, FP64 residual ----------
#pragma unroll
        for (int q = 0; q < NS; ++q) {
          const int row = lane + 64 * q;
          if (row < P) rv[row] = row < L ? make_float2(wr[q], wi[q]) : make_float2(0.0f, 0.0f);
        }
        __syncthreads();
        float rr_[NS], ri_[NS];
#pragma unroll
        for (int q = 0; q < NS; ++q) {
          const int a = lane + 64 * q;
          rr_[q] = 0.0f;
          ri_[q] = 0.0f;
          if (a < L) {
            double rr = 1.0 - static_cast<double>(dg[a].x) * wr[q];
            double ri = -static_cast<double>(dg[a].x) * wi[q];
            for (int b = 0; b < L; ++b) {
              if (b == a) continue;
              float2 rab = (a < b) ? A[a * S + b] : A[b * S + a];
              if (a > b) rab.y = -rab.y;
              const float2 wb = rv[b];
              rr = fma(-static_cast<double>(rab.x), wb.x, fma(static_cast<double>(rab.y), wb.y, rr));
              ri = fma(-static_cast<double>(rab.x), wb.y, fma(-static_cast<double>(rab.y), wb.x, ri));
            }
            rr_[q] = static_cast<float>(rr);
            ri_[q] = static_cast<float>(ri);
          }
        }
        solve(rr_, ri_);
#pragma unroll
        for (int q = 0; q < NS; ++q) {
          if (lane + 64 * q < L) {
            wr[q] += rr_[q];
            wi[q] += ri_[q];
          }
        }
      }

      // ---- 8. weights on the subarray-averaged centre snapshot --------------
      float dr = 0.0f, di = 0.0f, orr = 0.0f, oi = 0.0f;
      const float2* cen = rows + crow * na;
#pragma unroll
      for (int q = 0; q < NS; ++q) {
        const int a = lane + 64 * q;
        if (a < L) {
          float sr = 0.0f, si = 0.0f;
          for (int s0 = 0; s0 < ns; ++s0) {
            sr += cen[s0 + a].x;
            si += cen[s0 + a].y;
          }
          sr /= static_cast<float>(ns);
          si /= static_cast<float>(ns);
          dr += wr[q];
          di += wi[q];
          orr += wr[q] * sr + wi[q] * si;  // conj(w) * avg
          oi += wr[q] * si - wi[q] * sr;
        }
      }
      dr = wsum(dr);
      di = wsum(di);
      orr = wsum(orr);
      oi = wsum(oi);
      if (hypotf(dr, di) < 1e-30f) {
        das = true;
        float sr = 0.0f, si = 0.0f;
        for (int a = lane; a < na; a += 64) {
          sr += yc[act[a]].x;
          si += yc[act[a]].y;
        }
        out = make_float2(wsum(sr), wsum(si));
      } else {
        const double dd = static_cast<double>(dr) * dr + static_cast<double>(di) * di;
        const float nf = static_cast<float>(na);
        out = make_float2(
            static_cast<float>((static_cast<double>(orr) * dr + static_cast<double>(oi) * di) / dd) * nf,
            static_cast<float>((static_cast<double>(oi) * dr - static_cast<double>(orr) * di) / dd) * nf);
        if (M.weights) {
          // g_c = n / (n_sub sum w) * sum_{a = max(0, c - n_sub + 1)}^{min(L - 1, c)} conj(w_a):
          // inclusive prefix sums of conj(w) over the L entries (a wave scan per
          // slot, slot q offset by the total of the slots before it).
          float carry_r = 0.0f, carry_i = 0.0f;
#pragma unroll
          for (int q = 0; q < NS; ++q) {
            const int a = lane + 64 * q;
            float pr = a < L ? wr[q] : 0.0f, pi = a < L ? -wi[q] : 0.0f;
#pragma unroll
            for (int o = 1; o < 64; o <<= 1) {
              const float qr = __shfl_up(pr, o, 64), qi = __shfl_up(pi, o, 64);
              if (lane >= o) {
                pr += qr;
                pi += qi;
              }
            }
            pr += carry_r;
            pi += carry_i;
            if (a < P) rv[a] = make_float2(pr, pi);
            carry_r = lane_bcast(pr, 63);
            carry_i = lane_bcast(pi, 63);
          }
          __syncthreads();
          const double sc = static_cast<double>(na) / (static_cast<double>(ns) * dd);
          float2* gw = M.weights + static_cast<size_t>(pix) * n_ch;
          for (int c = lane; c < na; c += 64) {
            float2 acc = rv[min(L - 1, c)];
            if (c - ns >= 0) {
              const float2 lo = rv[c - ns];
              acc.x -= lo.x;
              acc.y -= lo.y;
            }
            // acc / (dr + j di) * n / n_sub
            gw[act[c]] = make_float2(
                static_cast<float>((static_cast<double>(acc.x) * dr + static_cast<double>(acc.y) * di) * sc),
                static_cast<float>((static_cast<double>(acc.y) * dr - static_cast<double>(acc.x) * di) * sc));
          }
        }
      }
    }
    if (M.status && lane == 0 && das) M.status[pix] = 2;
    if (M.weights && das) {
      float2* gw = M.weights + static_cast<size_t>(pix) * n_ch;
      for (int a = lane; a < na; a += 64) gw[act[a]] = make_float2(1.0f, 0.0f);
    }
    if (M.cf) {
      const float f = coherence(yc, act, na, lane);
      out.x *= f;
      out.y *= f;
    }
    if (lane == 0) store_out(M, pix, out, frame, acc);
  }
}

// The FP32 path without refinement or weight output, for P <= 64 (one matrix
// row per lane). Same staging, loading and fallbacks as mv_solve; three
// changes of formulation, each exact in real arithmetic:
//
// * Covariance by the subarray shift recursion. With
//     R(i,k) = sum_r sum_{a<n_sub} x_r[a+i] conj(x_r[a+k]),
//   R(i+1,k+1) = R(i,k) + sum_r (x_r[i+n_sub] conj(x_r[k+n_sub]) - x_r[i] conj(x_r[k])),
//   so lane m forms the first-column entry R(m,0) directly (n_snap products)
//   and walks the m-th sub-diagonal, the increments of every step formed at
//   once on MFMA (walk_increments_mfma), one add per step:
//   O(L n_snap + L^2 nrow) per pixel instead of O(L^2 n_snap).
// * The forward substitutions run inside the Cholesky column loop
//   (column-oriented: once column j of the factor is known, y_j = b_j / L_jj
//   and b_i -= L_ij y_j below it), for b = 1 and b = avg.
// * No back substitution: with w = R^-1 1, sum w = 1^H R^-1 1 = |y1|^2 and
//   w^H avg = y1^H y2 (y1 = L^-1 1, y2 = L^-1 avg), so
//   out = (y1^H y2) / |y1|^2 * n.
// The fast path's lower triangle, packed by rows: (i >= k, k) at
// i (i + 1) / 2 + k. A row's entries are contiguous, so one step along the
// columns is one element for every row, and the operands of the MFMA trailing
// update sit at fixed distances from a per-lane base (immediate offsets).
__device__ __forceinline__ int tri_at(int i, int k) { return ((i * (i + 1)) >> 1) + k; }

// Trailing update of the Cholesky block [j0, j0 + 8) on MFMA, for every
// pixel group gg of the wavefront: S(r, j0 + c) = A(r, j0 + c) -
// sum_{k < j0} L(r, k) conj(L(j0 + c, k)), written back into the packed lower
// triangle in place (rows r >= j0 + c of the columns j0 + c < L of the group).
// Complex products are expanded to real: with a = L(r, k), b = L(j0 + c, k),
//   Re S -= [a_r a_i] . [b_r  b_i],   Im S -= [a_r a_i] . [-b_i  b_r],
// a (rows x 2 j0) . (2 j0 x 16) real product on v_mfma_f32_16x16x4_f32: 16
// rows per tile, 4 real k (2 complex columns) per step, accumulated onto the
// triangle's entries with B negated. Lane l supplies A[row l % 16][k-part
// l / 16] and B[k-part l / 16][column l % 16] and holds D[4 (l / 16) + i][l % 16].
// The row tiles from the block's down to the last are accumulated together
// (NT of them), so a step's B operand is read and signed once for all; with
// the triangle packed by rows every operand of step st lies 4 st floats past
// its lane's base. Tiles and steps have wavefront-uniform bounds; rows and
// columns a group does not own produce values that are never stored.
template <int NT>
__device__ __forceinline__ void block_update_tiles(float* tri, int lane, int j0, int t0, int Lg) {
  const int lr = lane & 15, lq = lane >> 4;
  const int c = lr >> 1, part = lr & 1;
  const int kp = lq & 1, kh = lq >> 1;  // this lane's complex column k = 2 st + kh
  // -B[2k + kp][2c + part] = -b_r, -b_i, +b_i, -b_r for (part, kp) = (0,0), (0,1),
  // (1,0), (1,1): one component of b times a lane constant sign
  const int comp = part ^ kp;
  const float sgn = (part == 1 && kp == 0) ? 1.0f : -1.0f;
  const int col = j0 + c;
  int ib = 2 * tri_at(col, kh) + comp;  // float indices into tri (kept LDS-typed)
  int ia[NT];
  f4 acc[NT];
#pragma unroll
  for (int m = 0; m < NT; ++m) {
    ia[m] = 2 * tri_at(16 * (t0 + m) + lr, kh) + kp;
#pragma unroll
    for (int i = 0; i < 4; ++i) acc[m][i] = tri[2 * tri_at(16 * (t0 + m) + 4 * lq + i, col) + part];
  }
  for (int s4 = 0; s4 < (j0 >> 1); s4 += 4) {  // j0 / 2 is a multiple of 4
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const float b = sgn * tri[ib + 4 * u];
#pragma unroll
      for (int m = 0; m < NT; ++m)
        acc[m] = __builtin_amdgcn_mfma_f32_16x16x4f32(tri[ia[m] + 4 * u], b, acc[m], 0, 0, 0);
    }
    ib += 16;
#pragma unroll
    for (int m = 0; m < NT; ++m) ia[m] += 16;
  }
#pragma unroll
  for (int m = 0; m < NT; ++m) {
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int r = 16 * (t0 + m) + 4 * lq + i;
      if (r >= col && col < Lg) tri[2 * tri_at(r, col) + part] = acc[m][i];
    }
  }
}

// tri_off: byte offset of the triangle in a group's LDS, group_bytes: LDS per
// group; Lw: the wavefront's largest row count.
template <int P, int G>
__device__ __forceinline__ void block_update_mfma(unsigned char* smem, size_t group_bytes,
                                                  size_t tri_off, int lane, int j0, int Lw,
                                                  int L) {
  constexpr int LP = 64 / G;
  constexpr int T = P / 16;
  const int t0 = j0 >> 4, nt = min((Lw + 15) >> 4, T) - t0;  // >= 1: j0 < Lw
#pragma unroll
  for (int gg = 0; gg < G; ++gg) {
    float* tri = reinterpret_cast<float*>(smem + gg * group_bytes + tri_off);
    const int Lg = G == 1 ? L : __shfl(L, gg * LP, 64);
    if (T >= 4 && nt >= 4) block_update_tiles<T >= 4 ? 4 : 1>(tri, lane, j0, t0, Lg);
    else if (T >= 3 && nt == 3) block_update_tiles<T >= 3 ? 3 : 1>(tri, lane, j0, t0, Lg);
    else if (T >= 2 && nt == 2) block_update_tiles<T >= 2 ? 2 : 1>(tri, lane, j0, t0, Lg);
    else block_update_tiles<1>(tri, lane, j0, t0, Lg);
  }
}

// Increments of the sub-diagonal walks on MFMA, for every pixel group gg of
// the wavefront:
//   D(i, k) = sum_r x_r[i + n_sub] conj(x_r[k + n_sub]) - x_r[i] conj(x_r[k]),  k <= i < L,
// a rank-2 NR product H H^H - Lo Lo^H of the NR staged rows' high (offset
// n_sub) and low windows, written into the packed lower triangle at (i, k);
// R(i+1, k+1) = R(i, k) + D(i, k) then leaves one add per entry to the walks.
// Complex expanded to real as in block_update_mfma: 16 rows x 8 complex
// columns per tile, and per MFMA the 4 real terms of one row r (high re, high
// im, low re, low im -- the low ones with B negated). Indices are clamped to
// L - 2, so every read lies in the staged entries 0 .. n - 1; row L - 1 gets
// values the walks overwrite unread. The increments come out multiplied by
// the group's scale (B carries it). rows_off / tri_off: byte offsets of the
// rows and of the triangle in a group's LDS, group_bytes: LDS per group.
template <int P, int NR, int RS, int G>
__device__ __forceinline__ void walk_increments_mfma(unsigned char* smem, size_t group_bytes,
                                                     size_t rows_off, size_t tri_off, int lane,
                                                     int L, int ns, float scale) {
  constexpr int LP = 64 / G;
  const int lr = lane & 15, lq = lane >> 4;
  const int c = lr >> 1, part = lr & 1;
  const int hi = lq >> 1, comp = lq & 1;  // this lane's k-part: window and component
  // B[k-part][2c + part] = Re: [w_r w_i], Im: [-w_i w_r], the low window's
  // negated: component part ^ comp of w times a lane constant sign
  const int bcomp = part ^ comp;
  const float bsgn = ((part == 1 && comp == 0) ? -1.0f : 1.0f) * (hi ? 1.0f : -1.0f);
#pragma unroll
  for (int gg = 0; gg < G; ++gg) {
    const float* rows = reinterpret_cast<const float*>(smem + gg * group_bytes + rows_off);
    float* tri = reinterpret_cast<float*>(smem + gg * group_bytes + tri_off);
    const int Lg = G == 1 ? L : __shfl(L, gg * LP, 64);
    const int nsg = G == 1 ? ns : __shfl(ns, gg * LP, 64);  // every lane takes part
    const float bs = bsgn * (G == 1 ? scale : __shfl(scale, gg * LP, 64));
    const int off = hi ? nsg : 0;
    const int t1 = (Lg + 15) >> 4, c1 = (Lg + 7) >> 3;
    for (int mt = 0; mt < t1; ++mt) {
      const int i = min(16 * mt + lr, Lg - 2) + off;
      float a[NR];
#pragma unroll
      for (int r = 0; r < NR; ++r) a[r] = rows[2 * (r * RS + i) + comp];
      const int cbn = min(2 * mt + 2, c1);  // column blocks reaching the tile's rows
      for (int cb = 0; cb < cbn; ++cb) {
        const int j = min(8 * cb + c, Lg - 2) + off;
        f4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int r = 0; r < NR; ++r)
          acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a[r], bs * rows[2 * (r * RS + j) + bcomp],
                                                     acc, 0, 0, 0);
        const int col = 8 * cb + c;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const int row = 16 * mt + 4 * lq + e;
          if (row >= col && row < Lg) tri[2 * tri_at(row, col) + part] = acc[e];
        }
      }
    }
  }
}

template <int P, int NR, int RS, int G, bool REF>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(3))) void mv_solve_fast(MvArgs M, const float2* __restrict__ Y,
                                                    const uint32_t* __restrict__ mask,
                                                    float2* __restrict__ frame, float* acc,
                                                    const int* __restrict__ counts,
                                                    const int* __restrict__ lists) {
  constexpr int LP = 64 / G;  // lanes per pixel
  static_assert(P <= LP, "one matrix row per lane of the pixel's group");
  constexpr int BUCKET = P / 16 - 1;
  constexpr int TRI = P * (P + 1) / 2;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int lane = threadIdx.x;
  const int g = lane / LP, sub = lane % LP;
  const int n_ch = M.n_ch;
  const int n_pix = M.n0 * M.n1;
  // per group: the lower triangle packed by rows (tri_at), act (n_ch ints,
  // at least LP float2: once staged, the walks' sinks A[TRI + m]), NR rows at
  // stride RS
  const size_t ints = mv_fast_ints_bytes(n_ch, LP);
  const size_t group_bytes = sizeof(float2) * (TRI + NR * RS) + ints;
  const size_t tri_off = 0, rows_off = sizeof(float2) * TRI + ints;
  unsigned char* mine = smem_raw + g * group_bytes;
  float2* A = reinterpret_cast<float2*>(mine);
  int* act = reinterpret_cast<int*>(mine + sizeof(float2) * TRI);
  float2* rows = reinterpret_cast<float2*>(mine + rows_off);
  auto at = [](int i, int k) { return tri_at(i, k); };
  const int rb = tri_at(sub, 0);  // this lane's row

  // Two lists per bucket (mv_classify): the pixels of L <= P - 2 from the
  // front of the bucket's array, solved with the right-hand sides as two
  // rows bordering the matrix (BORDER), then those of L = P - 1 and P from
  // its back, whose bordering rows would need lanes P and P + 1, with the
  // forward substitutions carried through the column loop instead.
  auto solve = [&](auto border_tag) {
  constexpr bool BORDER = decltype(border_tag)::value;
  const int count = counts[BORDER ? BUCKET : kFullList + BUCKET];
  const int* list = lists + static_cast<size_t>(BUCKET) * n_pix;
  // a group past the end of the list repeats the wavefront's first pixel and
  // stores nothing; each round reads the next round's entry ahead
  auto entry = [&](int it) {
    const int q = it + g < count ? it + g : it;
    return it < count ? list[BORDER ? q : n_pix - 1 - q] : 0;
  };
  const int stride = gridDim.x * G;
  int next = entry(xcd_slot(M.n_xcd) * G);
  for (int it = xcd_slot(M.n_xcd) * G; it < count; it += stride) {
    const bool valid = it + g < count;
    const int e = G == 1 ? __builtin_amdgcn_readfirstlane(next) : next;  // one pixel: uniform
    next = entry(it + stride);
    const int pix = e & kPixMask, lo1 = e >> kRunShift;
    __syncthreads();  // LDS of the previous pixels is free
    Staged sg;
    if constexpr (G == 1 && RS <= 2 * LP) {  // a run's single round of loads, registers allowing
      if (__all(lo1 > 0)) sg = stage_group_run<NR, RS, G>(M, Y, mask, pix, lo1 - 1, lane, rows);
      else sg = stage_group<NR, RS, G>(M, Y, mask, pix, lane, act, rows);
    } else {
      sg = stage_group<NR, RS, G>(M, Y, mask, pix, lane, act, rows);
    }
    const int na = sg.na, nrow = sg.nrow;
    const int L = min(max(M.subarray, 2), max(2, na / 2));
    const int ns = na - L + 1;
    const float inv = 1.0f / static_cast<float>(nrow * ns);
    // uniform loop bounds, kept scalar
    const int Lw = __builtin_amdgcn_readfirstlane(wave_max(L));
    const int nsw = __builtin_amdgcn_readfirstlane(wave_max(ns));

    // ---- covariance: first column, then the sub-diagonal walks ---------------
    // Loops have uniform bounds and fixed row counts (NR rows, the missing
    // ones zero) so the loads of a step are all issued before its products;
    // lanes past their diagonal's end (or their pixel's size) read clamped
    // indices and store nothing.
    float cr, ci;
    {
      float ar[2] = {0.0f, 0.0f}, ai[2] = {0.0f, 0.0f};
      const int m = min(sub, L - 1);
      if constexpr (G == 1) {
        // one pixel: n_sub is uniform, so pairs need no clamping; odd tail after
        int a = 0;
        for (; a + 1 < ns; a += 2) {  // R(m,0) += x[a+m] conj(x[a])
          float2 p[NR][2], q[NR][2];
#pragma unroll
          for (int r = 0; r < NR; ++r) {
            const float2* x = rows + r * RS;
#pragma unroll
            for (int u = 0; u < 2; ++u) {
              p[r][u] = x[a + u + m];
              q[r][u] = x[a + u];
            }
          }
#pragma unroll
          for (int r = 0; r < NR; ++r) {
#pragma unroll
            for (int u = 0; u < 2; ++u) {
              const float2 pp = p[r][u], qq = q[r][u];
              ar[u] = fmaf(pp.x, qq.x, fmaf(pp.y, qq.y, ar[u]));
              ai[u] = fmaf(pp.y, qq.x, fmaf(-pp.x, qq.y, ai[u]));
            }
          }
        }
        if (a < ns) {
#pragma unroll
          for (int r = 0; r < NR; ++r) {
            const float2* x = rows + r * RS;
            const float2 pp = x[a + m], qq = x[a];
            ar[0] = fmaf(pp.x, qq.x, fmaf(pp.y, qq.y, ar[0]));
            ai[0] = fmaf(pp.y, qq.x, fmaf(-pp.x, qq.y, ai[0]));
          }
        }
      } else {  // groups: pairs below every group's n_sub unclamped, then clamped
        const int nsmin = __builtin_amdgcn_readfirstlane(wave_min(ns));
        int a = 0;
        for (; a + 1 < nsmin; a += 2) {
          float2 p[NR][2], q[NR][2];
#pragma unroll
          for (int r = 0; r < NR; ++r) {
            const float2* x = rows + r * RS;
#pragma unroll
            for (int u = 0; u < 2; ++u) {
              p[r][u] = x[a + u + m];
              q[r][u] = x[a + u];
            }
          }
#pragma unroll
          for (int r = 0; r < NR; ++r) {
#pragma unroll
            for (int u = 0; u < 2; ++u) {
              const float2 pp = p[r][u], qq = q[r][u];
              ar[u] = fmaf(pp.x, qq.x, fmaf(pp.y, qq.y, ar[u]));
              ai[u] = fmaf(pp.y, qq.x, fmaf(-pp.x, qq.y, ai[u]));
            }
          }
        }
        for (; a < nsw; a += 2) {  // R(m,0) += x[a+m] conj(x[a])
          float2 p[NR][2], q[NR][2];
#pragma unroll
          for (int r = 0; r < NR; ++r) {
            const float2* x = rows + r * RS;
#pragma unroll
            for (int u = 0; u < 2; ++u) {
              const int au = min(a + u, ns - 1);
              p[r][u] = x[au + m];
              q[r][u] = a + u < ns ? x[au] : make_float2(0.0f, 0.0f);
            }
          }
#pragma unroll
          for (int r = 0; r < NR; ++r) {
#pragma unroll
            for (int u = 0; u < 2; ++u) {
              const float2 pp = p[r][u], qq = q[r][u];
              ar[u & 1] = fmaf(pp.x, qq.x, fmaf(pp.y, qq.y, ar[u & 1]));
              ai[u & 1] = fmaf(pp.y, qq.x, fmaf(-pp.x, qq.y, ai[u & 1]));
            }
          }
        }
      }
      cr = (ar[0] + ar[1]) * inv;
      ci = (ai[0] + ai[1]) * inv;
    }
    // Walk increments D (times 1 / N) on MFMA into the triangle, then lane m
    // walks its sub-diagonal in place: at (m+t, t) it reads D(m+t, t), stores
    // R(m+t, t) and adds, R(m+t+1, t+1) = R(m+t, t) + D(m+t, t). A lane touches
    // only its own diagonal, so the reads of 8 steps go out before their
    // stores. Past its pixel's rows a lane walks on through rows the solve
    // never reads (overwritten first where they border the matrix), and past
    // the triangle it reads and writes its sink entry A[TRI + m] in the spent
    // active list (or an entry between TRI and there, garbage for garbage):
    // the stores need no condition, share the reads' addresses and seldom
    // meet in one bank.
    const int sink = TRI + sub;
    walk_increments_mfma<P, NR, RS, G>(smem_raw, group_bytes, rows_off, tri_off, lane, L, ns, inv);
    __syncthreads();
    int pos = rb;  // at(sub + t, t); at(i + 1, k + 1) = at(i, k) + i + 2
    for (int t0 = 0; t0 < Lw; t0 += 8) {
      const int q = sub + 2 + t0;
      float2 d[8];
      int pc[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        pc[u] = min(pos, sink);
        pos += q + u;
        d[u] = A[pc[u]];
      }
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        // (the main diagonal's imaginary part, rounding noise, is never read:
        // the Cholesky takes the pivot's real part only)
        A[pc[u]] = make_float2(cr, ci);
        cr += d[u].x;
        ci += d[u].y;
      }
    }
    __syncthreads();  // the main diagonal was written by lane 0 of each group
    // ---- diagonal loading max(delta tr / L, 1e-20), the trace summed in FP64
    // (exactly, for FP32 entries of one sign within 2^29 of each other) ------
    float2* diag = A + at(min(sub, P - 1), min(sub, P - 1));
    const double tr = group_sum_d<G>(sub < L ? static_cast<double>(diag->x) : 0.0);
    const float eps = fmaxf(static_cast<float>(static_cast<double>(M.loading) * tr / L), 1e-20f);
    // subarray average of the centre vector: the second right-hand side
    float b2r = 0.0f, b2i = 0.0f;
    if (sub < L) {
      diag->x += eps;
      const float2* cen = rows + sg.crow * RS + sub;
      float pr[4] = {0.0f, 0.0f, 0.0f, 0.0f}, pi[4] = {0.0f, 0.0f, 0.0f, 0.0f};
      int s0 = 0;
      for (; s0 + 8 <= ns; s0 += 8) {  // whole groups of 8 snapshots, then the rest
        float2 v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) v[u] = cen[s0 + u];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          pr[u & 3] += v[u].x;
          pi[u & 3] += v[u].y;
        }
      }
      for (; s0 < ns; ++s0) {
        const float2 x = cen[s0];
        pr[0] += x.x;
        pi[0] += x.y;
      }
      b2r = ((pr[0] + pr[1]) + (pr[2] + pr[3])) / static_cast<float>(ns);
      b2i = ((pi[0] + pi[1]) + (pi[2] + pi[3])) / static_cast<float>(ns);
      // The right-hand sides 1 and avg enter as two rows below the matrix:
      // the Cholesky factor of [R B; B^H *], B = [1 avg], has rows L and
      // L + 1 equal to conj(L^-1 1) and conj(L^-1 avg), so the column loop's
      // row-per-lane work does both forward substitutions (rows L + 1 < P:
      // the buckets hold L <= P - 2).
      if (BORDER) {
        A[at(L, sub)] = make_float2(1.0f, 0.0f);
        A[at(L + 1, sub)] = make_float2(b2r, -b2i);
      }
    }
    const float avr = b2r, avi = b2i;  // the subarray average, kept for the refined weights
    __syncthreads();

    // ---- Cholesky (of the bordered matrix when BORDER), left-looking in
    // blocks of CB columns over the L columns of R and its L (L + 2) rows. For
    // the block [j0, j0 + CB) the terms of the finished columns,
    // s_c = A(i, j0+c) - sum_{k<j0} L(i,k) conj(L(j0+c,k)), are formed on MFMA
    // in the triangle (block_update_mfma); lane i then finishes its row of the
    // block column by column with the in-block terms k = j0 .. j0+c-1 (taken
    // by readlane from the registers of lane j0+c), in increasing k as the
    // reference (beamformer.cpp:48-63).
    constexpr int CB = 8;
    // Bordered pixels: the rows above a column's pivot (and past the bordered
    // matrix, and a group's columns past its size) may hold any value -- their
    // sums and factor entries are never stored, broadcast or read (a broadcast
    // reads the pivot row, the trailing update the stored entries, and the
    // bordering rows lie below every pivot) -- so they go unmasked.
    constexpr bool LOOSE = BORDER;
    float rjj = 0.0f;  // 1 / L_jj of this lane's row (the refinement's solves)
    // !BORDER: the forward substitutions' vectors, lane i holding entry i
    float b1r = (!BORDER && sub < L) ? 1.0f : 0.0f, b1i = 0.0f;
    bool ok = true;  // this group's pivots so far are positive
    const int L2 = BORDER ? L + 2 : L;  // rows (of the bordered matrix)
    for (int j0 = 0; j0 < Lw; j0 += CB) {
      // Lanes outside [j0, L2) keep zero partial sums, so their factor
      // entries are 0 (and !BORDER: their b stays as it is).
      if (j0 > 0) {
        block_update_mfma<P, G>(smem_raw, group_bytes, tri_off, lane, j0, BORDER ? Lw + 2 : Lw, L);
        __syncthreads();
      }
      float sr[CB], si[CB], lr[CB], li[CB];
#pragma unroll
      for (int c = 0; c < CB; ++c) {
        sr[c] = 0.0f;
        si[c] = 0.0f;
      }
      // (the block's first column as a scalar, so every column offset and
      // every readlane's lane select is scalar arithmetic, with no per-column
      // readfirstlane and hazard wait)
      const int j0s = __builtin_amdgcn_readfirstlane(j0);
      {
        // branch-free: every lane reads (its own entry, or the triangle's
        // last: its row's base plus a scalar column, clamped) and keeps the
        // value where its row reaches the column
        const bool rowon = sub >= j0s && sub < L2;
#pragma unroll
        for (int c = 0; c < CB; ++c) {
          const int jc = j0s + c;
          const bool on = LOOSE || (rowon && jc <= sub);
          const float2 v = A[min(rb + jc, TRI - 1)];
          sr[c] = on ? v.x : 0.0f;
          si[c] = on ? v.y : 0.0f;
        }
      }
      // Column by column, branch-free (selects): rows below the pivot take
      // L_ij (rows L and L + 1 included: the forward substitutions). A group
      // whose pixel has no column j (j >= L) changes nothing. Each column is
      // stored as soon as it is finished, and the pivot row's in-block
      // entries k <= j - 2 come back from LDS as one broadcast per group --
      // only k = j - 1 by group_bcast (readlane, ds_bpermute).
      float2 rowv[CB][CB];  // rowv[c][q] = L(j0 + c, j0 + q)
#pragma unroll
      for (int c = 0; c < CB; ++c) {
        const int j = j0s + c;
        if (j >= Lw) break;
        const bool col = G == 1 || j < L;  // one pixel: j < Lw = L
        {  // row j, columns j0 .. j - 2 (stored by now)
          const int nb = tri_at(j, j0s);
#pragma unroll
          for (int q = 0; q + 1 < c; ++q) rowv[c][q] = A[nb + q];
        }
#pragma unroll
        for (int q = 0; q < c; ++q) {  // in-block terms, k = j0 + q
          float br, bi;
          if (q + 1 < c) {
            br = rowv[c][q].x;
            bi = rowv[c][q].y;
          } else {
            br = group_bcast<G>(lr[q], j, lane);
            bi = group_bcast<G>(li[q], j, lane);
          }
          sr[c] = fmaf(-lr[q], br, fmaf(-li[q], bi, sr[c]));
          si[c] = fmaf(-li[q], br, fmaf(lr[q], bi, si[c]));
        }
        const float d = group_bcast<G>(sr[c], j, lane);
        ok = ok && (!col || d > 0.0f);
        // 1 / L_jj in one instruction (v_rsq_f32, ~1 ulp); L_jj itself is
        // never needed (the pivot row stores 0 or it in its place, unread).
        // A group whose pixel has no column j takes 0.
        const float rj = col ? __builtin_amdgcn_rsqf(d) : 0.0f;
        const bool piv = col && sub == j, below = col && sub > j;
        if (REF) rjj = piv ? rj : rjj;
        // rows not below the pivot take l = 0: their entries of this column are
        // never read (rows above j are not stored, and the factor's diagonal is
        // not read back), so b -= l y leaves them as they are without a select
        // (LOOSE: no b, and any value will do)
        const float ar = (LOOSE || below) ? sr[c] * rj : 0.0f;
        const float ai = (LOOSE || below) ? si[c] * rj : 0.0f;
        lr[c] = ar;
        li[c] = ai;
        if constexpr (!BORDER) {
          // y_j = b_j / L_jj at the pivot row, b_i -= L_ij y_j below it
          const float y1r = group_bcast<G>(b1r, j, lane) * rj, y1i = group_bcast<G>(b1i, j, lane) * rj;
          const float y2r = group_bcast<G>(b2r, j, lane) * rj, y2i = group_bcast<G>(b2i, j, lane) * rj;
          const float n1r = fmaf(-ar, y1r, fmaf(ai, y1i, b1r));
          const float n1i = fmaf(-ar, y1i, fmaf(-ai, y1r, b1i));
          const float n2r = fmaf(-ar, y2r, fmaf(ai, y2i, b2r));
          const float n2i = fmaf(-ar, y2i, fmaf(-ai, y2r, b2i));
          b1r = piv ? y1r : n1r;
          b1i = piv ? y1i : n1i;
          b2r = piv ? y2r : n2r;
          b2i = piv ? y2i : n2i;
        }
        if (sub >= j && sub < L2) A[rb + j] = make_float2(ar, ai);  // (j >= L: rows never read)
      }
      __syncthreads();
      if (!__any(ok)) break;  // every group's factorisation failed
    }
    const bool on = sub < L;
    const float nf = static_cast<float>(na);
    // e1 = conj(y1), e2 = conj(y2) with y1 = L^-1 1 and y2 = L^-1 avg: rows L
    // and L + 1 of the bordered factor, or the substitutions' vectors
    float2 e1, e2;
    if constexpr (BORDER) {
      e1 = on ? A[at(L, sub)] : make_float2(0.0f, 0.0f);
      e2 = on ? A[at(L + 1, sub)] : make_float2(0.0f, 0.0f);
    } else {
      e1 = on ? make_float2(b1r, -b1i) : make_float2(0.0f, 0.0f);
      e2 = on ? make_float2(b2r, -b2i) : make_float2(0.0f, 0.0f);
    }
    bool das;
    float2 out;
    if constexpr (!REF) {
      const float dn = group_sum<G>(fmaf(e1.x, e1.x, e1.y * e1.y));     // |y1|^2 = sum w
      const float nr = group_sum<G>(fmaf(e1.x, e2.x, e1.y * e2.y));     // y1^H y2 = w^H avg
      const float ni = group_sum<G>(fmaf(e1.x, -e2.y, e1.y * e2.x));
      das = !ok || !(dn >= 1e-30f);
      out = make_float2(nr / dn * nf, ni / dn * nf);
    } else {
      // ---- one step of iterative refinement ---------------------------------
      // w = L^-H y1 (back substitution), the residual r = 1 - R w in FP64, the
      // correction c = R^-1 r with the same FP32 factor, w += c; then the
      // reference's (w^H avg) / (sum w) * n (beamformer.cpp:198-213). R is not
      // kept: R w = (1/N) X (X^H w) + eps w over the N = nrow * n_sub staged
      // snapshots x_{r,s}[a] = rows_r[s + a], two FP64 correlations.
      // Column-oriented substitutions, lane i holding z_i: the pivot lane's
      // z_j / L_jj is broadcast and the rows it reaches are updated.
      auto forward = [&](float& zr, float& zi) {  // z <- L^-1 z
        for (int j = 0; j < Lw; ++j) {
          const bool col = j < L;
          const float yr = group_bcast<G>(zr * rjj, j, lane), yi = group_bcast<G>(zi * rjj, j, lane);
          const bool below = col && sub > j && sub < L;
          const float2 l = A[below ? at(sub, j) : 0];  // L(i, j)
          const float nr_ = fmaf(-l.x, yr, fmaf(l.y, yi, zr)), ni_ = fmaf(-l.x, yi, fmaf(-l.y, yr, zi));
          if (col && sub == j) {
            zr = yr;
            zi = yi;
          } else if (below) {
            zr = nr_;
            zi = ni_;
          }
        }
      };
      auto backward = [&](float& zr, float& zi) {  // z <- L^-H z
        for (int j = Lw - 1; j >= 0; --j) {
          const bool col = j < L;
          const float xr = group_bcast<G>(zr * rjj, j, lane), xi = group_bcast<G>(zi * rjj, j, lane);
          const bool above = col && sub < j;
          const float2 l = A[above ? at(j, sub) : 0];  // L(j, i): z_i -= conj(L(j, i)) x_j
          const float nr_ = fmaf(-l.x, xr, fmaf(-l.y, xi, zr)), ni_ = fmaf(-l.x, xi, fmaf(l.y, xr, zi));
          if (col && sub == j) {
            zr = xr;
            zi = xi;
          } else if (above) {
            zr = nr_;
            zi = ni_;
          }
        }
      };
      float wr = e1.x, wi = -e1.y;
      backward(wr, wi);
      double xr = 0.0, xi = 0.0;  // (X X^H w)_a, a = sub
      for (int r = 0; r < NR; ++r) {
        const float2* x = rows + r * RS;
        for (int s0 = 0; s0 < nsw; s0 += LP) {
          const int s = s0 + sub;
          double tr = 0.0, ti = 0.0;  // t_s = x_s^H w
          for (int k = 0; k < Lw; ++k) {
            const float wkr = group_bcast<G>(wr, k, lane), wki = group_bcast<G>(wi, k, lane);
            if (k < L && s < ns) {
              const float2 v = x[s + k];
              tr += static_cast<double>(v.x) * wkr + static_cast<double>(v.y) * wki;
              ti += static_cast<double>(v.x) * wki - static_cast<double>(v.y) * wkr;
            }
          }
          const int qn = min(LP, nsw - s0);  // snapshots of this chunk (wavefront-uniform)
          for (int q = 0; q < qn; ++q) {
            const double tqr = group_bcast_d<G>(tr, q, lane), tqi = group_bcast_d<G>(ti, q, lane);
            if (s0 + q < ns && on) {
              const float2 v = x[s0 + q + sub];
              xr += static_cast<double>(v.x) * tqr - static_cast<double>(v.y) * tqi;
              xi += static_cast<double>(v.x) * tqi + static_cast<double>(v.y) * tqr;
            }
          }
        }
      }
      const double invn = 1.0 / static_cast<double>(nrow * ns);
      float cr = on ? static_cast<float>(1.0 - (invn * xr + static_cast<double>(eps) * wr)) : 0.0f;
      float ci = on ? static_cast<float>(-(invn * xi + static_cast<double>(eps) * wi)) : 0.0f;
      forward(cr, ci);
      backward(cr, ci);
      wr += cr;
      wi += ci;
      const float sr_ = group_sum<G>(on ? wr : 0.0f), si_ = group_sum<G>(on ? wi : 0.0f);     // sum w
      const float nr = group_sum<G>(on ? fmaf(wr, avr, wi * avi) : 0.0f);                      // w^H avg
      const float ni = group_sum<G>(on ? fmaf(wr, avi, -(wi * avr)) : 0.0f);
      const float d2 = fmaf(sr_, sr_, si_ * si_);
      das = !ok || !(sqrtf(d2) >= 1e-30f);
      // (w^H avg) / (sum w), complex, times n
      out = make_float2((nr * sr_ + ni * si_) / d2 * nf, (ni * sr_ - nr * si_) / d2 * nf);
    }
    // DAS fallback and coherence factor over the group's active centre vector
    // (the staged centre row: yc[act[a]] for a < na)
    const float2* cen = rows + sg.crow * RS;
    float sr = 0.0f, si = 0.0f, inc = 0.0f;
    for (int a = sub; a < na; a += LP) {
      const float2 v = cen[a];
      sr += v.x;
      si += v.y;
      if (M.cf) {
        const float m = hypotf(v.x, v.y);
        inc += m * m;
      }
    }
    sr = group_sum<G>(sr);
    si = group_sum<G>(si);
    if (das) out = make_float2(sr, si);
    if (M.cf) {
      inc = group_sum<G>(inc);
      const float mc = hypotf(sr, si);
      const float f = (inc > 0.0f) ? (mc * mc) / (nf * inc) : 0.0f;
      out.x *= f;
      out.y *= f;
    }
    if (valid && sub == 0) {
      if (M.status && das) M.status[pix] = 2;
      store_out(M, pix, out, frame, acc);
    }
  }
  };
  solve(std::true_type{});
  solve(std::false_type{});
}

template <int P>
static void launch_bucket(const MvArgs& M, const float2* Y, const uint32_t* mask, float2* frame,
                          float* acc, const int* counts, const int* lists, hipStream_t st) {
  const bool fast = P <= 64 && mv_fast_config(M);
  const size_t smem = fast ? mv_fast_smem_bytes(M.n_ch, M.range_avg, M.subarray, P)
                           : mv_smem_bytes(M.n_ch, M.range_avg, M.subarray, P, M.refine != 0);
  const int na_max = M.subarray <= P ? M.n_ch : std::min(M.n_ch, 2 * P + 1);
  const int n_pix = M.n0 * M.n1;
  // Persistent grid: as many single-wavefront workgroups per CU as can be
  // resident at once -- what the runtime reports for this kernel and LDS
  // size, bounded by LDS (160 KB over smem rounded up to the 1280-byte
  // allocation granule) and by VGPRs (4 SIMDs x min(8, 512 / VGPRs per wave))
  // -- and by the tuned occupancy when one is given. The pixels are dealt to
  // the workgroups in equal shares, so the grid is sized to fill every CU
  // exactly: a CU with room to spare may be handed more workgroups than its
  // neighbours (measured 704 -> 804 us at P = 48 when 12 would fit but 11 per
  // CU were launched), and one workgroup more than fits waits for another to
  // end (measured 572 -> 797 us at P = 48 when 16 bytes more per workgroup,
  // 14096 instead of 11 granules' 14080, left room for 10 but 11 were
  // launched on the runtime's report).
  const int cap = M.occupancy > 0 ? M.occupancy : 32;
  const int lds_cap = static_cast<int>(160 * 1024 / ((smem + 1279) / 1280 * 1280));
#define USX_LAUNCH(KERNEL, ...)                                                                   \
  do {                                                                                           \
    static const int vgpr_cap = [] {                                                             \
      (void)hipFuncSetAttribute(reinterpret_cast<const void*>(KERNEL),                           \
                                hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);         \
      hipFuncAttributes fa{};                                                                    \
      if (hipFuncGetAttributes(&fa, reinterpret_cast<const void*>(KERNEL)) != hipSuccess ||      \
          fa.numRegs <= 0)                                                                       \
        return 16;                                                                               \
      return 4 * std::max(1, std::min(8, 512 / ((fa.numRegs + 7) / 8 * 8)));                    \
    }();                                                                                         \
    int resident = 0;                                                                            \
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&resident, reinterpret_cast<const void*>(KERNEL), \
                                                     64, smem) != hipSuccess || resident <= 0)   \
      resident = 32;                                                                             \
    const int per_cu = std::max(1, std::min(std::min(cap, resident), std::min(lds_cap, vgpr_cap))); \
    const int grid = std::min(n_pix, M.n_cu * per_cu);                                           \
    hipLaunchKernelGGL(KERNEL, dim3(grid), dim3(64), smem, st, M, Y, mask, frame, acc, counts,   \
                       lists __VA_OPT__(, ) __VA_ARGS__);                                        \
  } while (0)
  if constexpr (P <= 64) {
    if (fast) {
      constexpr int G = P <= 16 ? 4 : (P <= 32 ? 2 : 1);  // mv_fast_groups(P)
      constexpr int RT = mv_tight_stride(P) < 128 ? mv_tight_stride(P) : 128;
      const int rs = mv_fast_row_stride(M.n_ch, M.subarray, P);
#define USX_FAST_NR(RSV, REFV)                                                        \
  do {                                                                                \
    if (M.range_avg == 0) USX_LAUNCH((mv_solve_fast<P, 1, RSV, G, REFV>));            \
    else if (M.range_avg == 1) USX_LAUNCH((mv_solve_fast<P, 3, RSV, G, REFV>));       \
    else USX_LAUNCH((mv_solve_fast<P, 5, RSV, G, REFV>));                             \
  } while (0)
#define USX_FAST(REFV)                                                                \
  do {                                                                                \
    if (rs == RT && RT < 128) USX_FAST_NR(RT, REFV);                                  \
    else if (rs == 128) USX_FAST_NR(128, REFV);                                       \
    else USX_FAST_NR(256, REFV);                                                      \
  } while (0)
      if (M.refine) USX_FAST(true);
      else USX_FAST(false);
#undef USX_FAST
#undef USX_FAST_NR
      return;
    }
  }
  if (M.precision == 1) USX_LAUNCH((mv_solve<P, 1>), na_max);
  else if (M.precision == 2) USX_LAUNCH((mv_solve<P, 2>), na_max);
  else USX_LAUNCH((mv_solve<P, 0>), na_max);
#undef USX_LAUNCH
}

int mv_max_bucket_p(int n_ch, int subarray) {
  const int lmax = std::min(std::max(subarray, 2), std::max(2, n_ch / 2));
  return (lmax + 15) / 16 * 16;
}

void launch_mv(const MvArgs& M, const float2* Y, const uint32_t* mask, float2* frame, float* acc,
               int* counts, int* lists, hipStream_t st) {
  const int n_pix = M.n0 * M.n1;
  hipLaunchKernelGGL(mv_zero_counts, dim3(1), dim3(64), 0, st, counts);
  hipLaunchKernelGGL(mv_classify, dim3((n_pix + 255) / 256), dim3(256), 0, st, M, Y, mask, frame,
                     acc, counts, lists);
  const int pmax = mv_max_bucket_p(M.n_ch, M.subarray);
  if (pmax >= 16) launch_bucket<16>(M, Y, mask, frame, acc, counts, lists, st);
  if (pmax >= 32) launch_bucket<32>(M, Y, mask, frame, acc, counts, lists, st);
  if (pmax >= 48) launch_bucket<48>(M, Y, mask, frame, acc, counts, lists, st);
  if (pmax >= 64) launch_bucket<64>(M, Y, mask, frame, acc, counts, lists, st);
  if (pmax >= 80) launch_bucket<80>(M, Y, mask, frame, acc, counts, lists, st);
  if (pmax >= 96) launch_bucket<96>(M, Y, mask, frame, acc, counts, lists, st);
  if (pmax >= 112) launch_bucket<112>(M, Y, mask, frame, acc, counts, lists, st);
  if (pmax >= 128) launch_bucket<128>(M, Y, mask, frame, acc, counts, lists, st);
}

}  // namespace usxd
