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
// mv_solve<P> (persistent, one wavefront per pixel, LDS sized for the
// bucket so small systems run at high occupancy):
//   1. active list; centre and +-range_avg neighbour vectors staged in LDS
//      (neighbours clipped at the grid edge, beamformer.cpp:159-161);
//   2. covariance R = sum over (depth, subarray start) of x x^H as a dense
//      rank-k update on MFMA, complex expanded to real:
//        Re R = [xr xi][xr xi]^T,  Im R = [xi -xr][xr xi]^T,
//      v_mfma_f32_16x16x4_f32, two snapshots per instruction (k = re/im of
//      snapshot t, re/im of snapshot t+1), L padded to P with zero rows,
//      upper-triangle 16x16 blocks only;
//   3. R /= n_snap; diagonal loading max(delta tr(R)/L, 1e-20) with the trace
//      summed in FP64;
//   4. in-place FP32 Cholesky in LDS, element-parallel right-looking updates
//      (the strictly upper triangle keeps R for the residual); a non-positive
//      pivot -> DAS sum (beamformer.cpp:199);
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

}  // namespace

// Bytes of dynamic LDS for one wavefront of bucket P.
size_t mv_smem_bytes(int n_ch, int range_avg, int subarray, int P) {
  const int na_max = subarray <= P ? n_ch : std::min(n_ch, 2 * P + 1);
  const size_t ints = (static_cast<size_t>(n_ch) * 4 + 15) / 16 * 16;
  return ints + sizeof(float2) * (static_cast<size_t>(2 * range_avg + 1) * na_max +
                                  static_cast<size_t>(P) * (P + 1) + 2 * static_cast<size_t>(P));
}

__global__ void mv_zero_counts(int* counts) {
  if (threadIdx.x < kMaxBuckets) counts[threadIdx.x] = 0;
}

// Slot in bucket b's list, one atomic per (wavefront, bucket): lanes of a
// wavefront that join the same bucket take consecutive slots.
__device__ __forceinline__ int bucket_slot(int* counts, int b, bool take) {
  int slot = 0;
  for (int q = 0; q < kMaxBuckets; ++q) {
    const unsigned long long m = __ballot(take && b == q);
    if (m == 0) continue;
    const int leader = __ffsll(static_cast<long long>(m)) - 1;
    int base = 0;
    if (static_cast<int>(threadIdx.x & 63) == leader) base = atomicAdd(&counts[q], __popcll(m));
    base = __shfl(base, leader, 64);
    if (take && b == q)
      slot = base + __popcll(m & ((1ull << (threadIdx.x & 63)) - 1ull));
  }
  return slot;
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
  int na = 0;
  if (valid)
    for (int w = 0; w < nw; ++w) na += __popc(mp[w]);
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
  const int b = take ? (L + 15) / 16 - 1 : -1;
  const int slot = bucket_slot(counts, b, take);
  if (take) lists[static_cast<size_t>(b) * n_pix + slot] = pix;
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
  // A: P rows of stride P + 1 (conflict-free column access). Strictly upper
  // triangle: R (kept for the residual); lower triangle + diagonal: the
  // Cholesky factor. dg[a] = (loaded diagonal of R, 1 / L_aa).
  constexpr int S = P + 1;
  float2* A = rows + static_cast<size_t>(2 * M.range_avg + 1) * na_max;
  float2* dg = A + P * S;
  float2* rv = dg + P;  // scratch vector

  const int count = counts[BUCKET];
  for (int item = blockIdx.x; item < count; item += gridDim.x) {
    const int pix = lists[static_cast<size_t>(BUCKET) * n_pix + item];
    const int i0 = pix / M.n1;
    const int j = pix - i0 * M.n1;
    __syncthreads();  // LDS of the previous pixel is free

    // ---- 1. active list ------------------------------------------------------
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
    const float2* yc = Y + static_cast<size_t>(pix) * n_ch;
    const int L = min(max(M.subarray, 2), max(2, na / 2));
    const int ns = na - L + 1;

    // ---- 2. stage centre and neighbour rows ---------------------------------
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
    {
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
                A[row * S + col] = v;                       // R, upper
                A[col * S + row] = make_float2(v.x, -v.y);  // factor input, lower
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
      A[a * S + a] = make_float2(d, 0.0f);
    }
    __syncthreads();

    // ---- 5. Cholesky, left-looking by columns: lane i owns rows i, i + 64 ----
    //   L_ij = (A_ij - sum_{k<j} L_ik conj(L_jk)) / L_jj, L_jj = sqrt(A_jj - sum |L_jk|^2),
    // the reference's sums (beamformer.cpp:48-63); one barrier per column.
    bool ok = true;
    for (int j = 0; j < L; ++j) {
      float sr[NS], si[NS];
#pragma unroll
      for (int q = 0; q < NS; ++q) {
        const int row = lane + 64 * q;
        sr[q] = 0.0f;
        si[q] = 0.0f;
        if (row >= j && row < L) {
          const float2* ri_ = A + row * S;
          const float2* rj = A + j * S;
          float ar = ri_[j].x, ai = ri_[j].y;
#pragma unroll 4
          for (int k = 0; k < j; ++k) {
            const float2 a = ri_[k], b = rj[k];  // -= a * conj(b)
            ar = fmaf(-a.x, b.x, fmaf(-a.y, b.y, ar));
            ai = fmaf(-a.y, b.x, fmaf(a.x, b.y, ai));
          }
          sr[q] = ar;
          si[q] = ai;
        }
      }
      const float d = lane_bcast(pick<NS>(sr, j >> 6), j & 63);
      if (!(d > 0.0f)) {
        ok = false;
        break;
      }
      const float ljj = sqrtf(d);
      const float rj = 1.0f / ljj;
#pragma unroll
      for (int q = 0; q < NS; ++q) {
        const int row = lane + 64 * q;
        if (row == j) {
          A[j * S + j] = make_float2(ljj, 0.0f);
          dg[j].y = rj;
        } else if (row > j && row < L) {
          A[row * S + j] = make_float2(sr[q] * rj, si[q] * rj);
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
              const float2 l = A[row * S + k];
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
              const float2 l = A[k * S + row];  // -= conj(L[k][row]) * x_k
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

template <int P>
static void launch_bucket(const MvArgs& M, const float2* Y, const uint32_t* mask, float2* frame,
                          float* acc, const int* counts, const int* lists, hipStream_t st) {
  const size_t smem = mv_smem_bytes(M.n_ch, M.range_avg, M.subarray, P);
  const int na_max = M.subarray <= P ? M.n_ch : std::min(M.n_ch, 2 * P + 1);
  const int n_pix = M.n0 * M.n1;
  // persistent grid: as many single-wavefront workgroups per CU as LDS allows
  // (capped at 32, or at the tuned occupancy)
  const int cap = M.occupancy > 0 ? M.occupancy : 32;
  const int per_cu = std::max(1, std::min(cap, static_cast<int>(160 * 1024 / smem)));
  const int grid = std::min(n_pix, 256 * per_cu);
#define USX_MV(PR)                                                                            \
  do {                                                                                        \
    static const hipError_t attr = hipFuncSetAttribute(                                       \
        reinterpret_cast<const void*>(mv_solve<P, PR>),                                       \
        hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);                              \
    (void)attr;                                                                               \
    hipLaunchKernelGGL((mv_solve<P, PR>), dim3(grid), dim3(64), smem, st, M, Y, mask, frame,  \
                       acc, counts, lists, na_max);                                           \
  } while (0)
  if (M.precision == 1) USX_MV(1);
  else if (M.precision == 2) USX_MV(2);
  else USX_MV(0);
#undef USX_MV
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
}

}  // namespace usxd
