// mv.hip -- minimum-variance (Capon) combine on gfx950.
//
// Reference: combine_mv_windowed and cholesky_solve,
// core/src/beamformer.cpp:47-79, 137-214; coherence factor :525-534.
//
// One wavefront per pixel:
//   1. active list of the centre pixel (bit mask from the channel kernel),
//      L = clamp(subarray, 2, max(2, n/2)), n_sub = n - L + 1;
//      fewer than 4 active channels -> DAS sum (beamformer.cpp:142-146);
//   2. the channel vectors of the centre depth and its +-range_avg neighbours
//      (clipped at the grid edge) are staged in LDS;
//   3. covariance R = sum over (depth, subarray start) of x x^H as a dense
//      rank-k update on MFMA, complex expanded to real:
//        Re R = [xr xi][xr xi]^T,  Im R = [xi -xr][xr xi]^T,
//      v_mfma_f32_16x16x4_f32, two snapshots per instruction (k = re/im of
//      snapshot t, re/im of snapshot t+1), L padded to a multiple of 16 with
//      zero rows; only upper-triangle 16x16 blocks are formed;
//   4. R /= n_snap, diagonal loading max(delta * tr(R) / L, 1e-20) with the
//      trace summed in FP64;
//   5. in-place FP32 Cholesky in LDS (the strictly upper triangle keeps R
//      for the refinement residual), forward/back substitution for R w = 1;
//      failure (non-positive pivot) -> DAS sum (beamformer.cpp:199);
//   6. optional refinement: r = 1 - R w in FP64, R d = r, w += d;
//   7. |sum w| < 1e-30 -> DAS sum; out = (w^H avg) / sum w * n_active with
//      avg the subarray average of the centre vector (beamformer.cpp:205-213);
//   8. coherence factor over the centre vector, or |out|^2 accumulation for
//      incoherent compounding.
// precision 1 / 2 rounds the MFMA inputs to bf16 / fp16 (FP32 accumulate),
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

template <int PREC>
__device__ __forceinline__ float quant(float x) {
  if (PREC == 1) return __bfloat162float(__float2bfloat16(x));
  if (PREC == 2) return __half2float(__float2half(x));
  return x;
}

struct MvSmem {
  int* act;        // n_ch
  float2* rows;    // nrow_max * n_ch
  float2* A;       // P * P : strictly upper = R, lower + diag = Cholesky factor
  float2* dg;      // P : loaded diagonal of R (real)
  float2* w;       // P
  float2* r;       // P
};

constexpr int kPairsPerGroup = 4;

}  // namespace

size_t mv_smem_bytes(int n_ch, int range_avg, int pmax) {
  const size_t ints = (static_cast<size_t>(n_ch) * 4 + 15) / 16 * 16;
  return ints + sizeof(float2) * (static_cast<size_t>(2 * range_avg + 1) * n_ch +
                                  static_cast<size_t>(pmax) * pmax + 3 * static_cast<size_t>(pmax));
}

template <int PREC>
__global__ __launch_bounds__(64) void mv_kernel(MvArgs M, const float2* __restrict__ Y,
                                                const uint32_t* __restrict__ mask,
                                                float2* __restrict__ frame, float* acc_out,
                                                int pmax) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int pix = blockIdx.x;
  const int i0 = pix / M.n1;
  const int j = pix - i0 * M.n1;
  const int lane = threadIdx.x;
  const int n_ch = M.n_ch;
  const int nw = (n_ch + 31) / 32;
  const int nrow_max = 2 * M.range_avg + 1;

  MvSmem S;
  S.act = reinterpret_cast<int*>(smem_raw);
  S.rows = reinterpret_cast<float2*>(smem_raw + (static_cast<size_t>(n_ch) * 4 + 15) / 16 * 16);
  S.A = S.rows + static_cast<size_t>(nrow_max) * n_ch;
  S.dg = S.A + static_cast<size_t>(pmax) * pmax;
  S.w = S.dg + pmax;
  S.r = S.w + pmax;

  // ---- 1. active list of the centre pixel --------------------------------
  const uint32_t* mp = mask + static_cast<size_t>(pix) * nw;
  int na = 0;
  for (int base = 0; base < n_ch; base += 64) {
    const int ch = base + lane;
    const bool on = ch < n_ch && ((mp[ch >> 5] >> (ch & 31)) & 1u);
    const unsigned long long bal = __ballot(on);
    const int pos = na + __popcll(bal & ((1ull << lane) - 1ull));
    if (on) S.act[pos] = ch;
    na += __popcll(bal);
  }
  __syncthreads();

  const float2* yc = Y + static_cast<size_t>(pix) * n_ch;
  auto finish = [&](float2 out) {
    if (M.cf) {
      float cr = 0.0f, ci = 0.0f, inc = 0.0f;
      for (int a = lane; a < na; a += 64) {
        const float2 v = yc[S.act[a]];
        cr += v.x;
        ci += v.y;
        const float m = hypotf(v.x, v.y);
        inc += m * m;
      }
      cr = wsum(cr);
      ci = wsum(ci);
      inc = wsum(inc);
      const float mc = hypotf(cr, ci);
      const float f = (inc > 0.0f) ? (mc * mc) / (static_cast<float>(na) * inc) : 0.0f;
      out.x *= f;
      out.y *= f;
    }
    if (lane == 0) {
      if (M.incoherent) {
        const float m = hypotf(out.x, out.y);
        acc_out[pix] += m * m;
      } else {
        frame[pix] = out;
      }
    }
  };
  auto das = [&]() {
    float sr = 0.0f, si = 0.0f;
    for (int a = lane; a < na; a += 64) {
      const float2 v = yc[S.act[a]];
      sr += v.x;
      si += v.y;
    }
    return make_float2(wsum(sr), wsum(si));
  };

  if (na == 0) {
    if (lane == 0 && !M.incoherent) frame[pix] = make_float2(0.0f, 0.0f);
    return;
  }
  if (na < 4) {
    finish(das());
    return;
  }
  const int L = min(max(M.subarray, 2), max(2, na / 2));
  const int ns = na - L + 1;
  const int P = (L + 15) & ~15;
  const int nb = P / 16;

  // ---- 2. stage centre and neighbour rows --------------------------------
  int nrow = 0, crow = 0;
  for (int d = -M.range_avg; d <= M.range_avg; ++d) {
    const int k = j + d;
    if (k < 0 || k >= M.n1) continue;
    if (d == 0) crow = nrow;
    const float2* src = Y + (static_cast<size_t>(i0) * M.n1 + k) * n_ch;
    for (int a = lane; a < na; a += 64) S.rows[nrow * na + a] = src[S.act[a]];
    ++nrow;
  }
  __syncthreads();
  const int n_snap = nrow * ns;

  // ---- 3. covariance on MFMA ---------------------------------------------
  const int li = lane & 15, kk = lane >> 4;
  const int tsub = kk >> 1, comp = kk & 1;
  const int n_pairs = nb * (nb + 1) / 2;
  const float inv = 1.0f / static_cast<float>(n_snap);
  for (int p0 = 0; p0 < n_pairs; p0 += kPairsPerGroup) {
    f4 accR[kPairsPerGroup], accI[kPairsPerGroup];
    int pbi[kPairsPerGroup], pbj[kPairsPerGroup];
#pragma unroll
    for (int q = 0; q < kPairsPerGroup; ++q) {
      accR[q] = f4{0.0f, 0.0f, 0.0f, 0.0f};
      accI[q] = accR[q];
      // decode pair index p0+q -> (bi <= bj), row-major over the upper triangle
      int p = p0 + q, bi = 0;
      while (p >= nb - bi && bi < nb) {
        p -= nb - bi;
        ++bi;
      }
      pbi[q] = bi;
      pbj[q] = bi + p;
    }
    for (int t = 0; t < n_snap; t += 2) {
      const int ts = t + tsub;
      const bool tv = ts < n_snap;
      const int ri = tv ? ts / ns : 0;
      const int s0 = tv ? ts - ri * ns : 0;
      const float2* xrow = S.rows + ri * na + s0;
#pragma unroll
      for (int q = 0; q < kPairsPerGroup; ++q) {
        if (p0 + q < n_pairs) {
          const int ai = pbi[q] * 16 + li, aj = pbj[q] * 16 + li;
          float2 xi2 = make_float2(0.0f, 0.0f), xj2 = xi2;
          if (tv && ai < L) xi2 = xrow[ai];
          if (tv && aj < L) xj2 = xrow[aj];
          const float ar = quant<PREC>(comp ? xi2.y : xi2.x);
          const float aim = quant<PREC>(comp ? -xi2.x : xi2.y);
          const float b = quant<PREC>(comp ? xj2.y : xj2.x);
          accR[q] = __builtin_amdgcn_mfma_f32_16x16x4f32(ar, b, accR[q], 0, 0, 0);
          accI[q] = __builtin_amdgcn_mfma_f32_16x16x4f32(aim, b, accI[q], 0, 0, 0);
        }
      }
    }
    // D[4*(lane/16) + e][lane%16] -> R[bi*16 + row][bj*16 + col]
#pragma unroll
    for (int q = 0; q < kPairsPerGroup; ++q) {
      if (p0 + q < n_pairs) {
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const int row = pbi[q] * 16 + kk * 4 + e;
          const int col = pbj[q] * 16 + li;
          if (row < L && col < L && row <= col) {
            const float2 v = make_float2(accR[q][e] * inv, accI[q][e] * inv);
            if (row == col) {
              S.dg[row] = make_float2(v.x, 0.0f);
            } else {
              S.A[row * P + col] = v;                           // R (upper)
              S.A[col * P + row] = make_float2(v.x, -v.y);      // factor input (lower)
            }
          }
        }
      }
    }
  }
  __syncthreads();

  // ---- 4. diagonal loading (trace in FP64) -------------------------------
  double tr = 0.0;
  for (int a = lane; a < L; a += 64) tr += static_cast<double>(S.dg[a].x);
  tr = wsum_d(tr);
  const float eps = fmaxf(static_cast<float>(static_cast<double>(M.loading) * tr / L), 1e-20f);
  for (int a = lane; a < L; a += 64) {
    const float d = S.dg[a].x + eps;
    S.dg[a] = make_float2(d, 0.0f);
    S.A[a * P + a] = make_float2(d, 0.0f);
    S.w[a] = make_float2(1.0f, 0.0f);
  }
  __syncthreads();

  // ---- 5. Cholesky (lower, in place) and solve ---------------------------
  bool ok = true;
  for (int k = 0; k < L; ++k) {
    const float d = S.A[k * P + k].x;
    if (!(d > 0.0f)) {
      ok = false;
      break;
    }
    const float lkk = sqrtf(d);
    for (int i = k + 1 + lane; i < L; i += 64) {
      const float2 v = S.A[i * P + k];
      S.A[i * P + k] = make_float2(v.x / lkk, v.y / lkk);
    }
    if (lane == 0) S.A[k * P + k] = make_float2(lkk, 0.0f);
    __syncthreads();
    const int m = L - k - 1;
    for (int q = lane; q < m * m; q += 64) {
      const int r = q / m, c = q - r * m;
      if (c <= r) {
        const int ii = k + 1 + r, jj = k + 1 + c;
        const float2 lik = S.A[ii * P + k], ljk = S.A[jj * P + k];
        float2 a = S.A[ii * P + jj];
        // a -= lik * conj(ljk)
        a.x -= lik.x * ljk.x + lik.y * ljk.y;
        a.y -= lik.y * ljk.x - lik.x * ljk.y;
        S.A[ii * P + jj] = a;
      }
    }
    __syncthreads();
  }
  if (!ok) {
    finish(das());
    return;
  }

  auto solve = [&](float2* b) {
    // forward: L y = b
    for (int k = 0; k < L; ++k) {
      const float lkk = S.A[k * P + k].x;
      const float2 bk = b[k];
      const float2 yk = make_float2(bk.x / lkk, bk.y / lkk);
      __syncthreads();
      if (lane == 0) b[k] = yk;
      for (int i = k + 1 + lane; i < L; i += 64) {
        const float2 l = S.A[i * P + k];
        float2 v = b[i];
        v.x -= l.x * yk.x - l.y * yk.y;
        v.y -= l.x * yk.y + l.y * yk.x;
        b[i] = v;
      }
      __syncthreads();
    }
    // back: L^H x = y
    for (int k = L - 1; k >= 0; --k) {
      const float lkk = S.A[k * P + k].x;
      const float2 bk = b[k];
      const float2 xk = make_float2(bk.x / lkk, bk.y / lkk);
      __syncthreads();
      if (lane == 0) b[k] = xk;
      for (int i = lane; i < k; i += 64) {
        const float2 l = S.A[k * P + i];  // conj(L[k][i]) * xk
        float2 v = b[i];
        v.x -= l.x * xk.x + l.y * xk.y;
        v.y -= l.x * xk.y - l.y * xk.x;
        b[i] = v;
      }
      __syncthreads();
    }
  };
  solve(S.w);

  // ---- 6. one refinement step, FP64 residual ------------------------------
  if (M.refine) {
    for (int a = lane; a < L; a += 64) {
      double rr = 1.0, ri = 0.0;
      const float2 wa = S.w[a];
      rr -= static_cast<double>(S.dg[a].x) * wa.x;
      ri -= static_cast<double>(S.dg[a].x) * wa.y;
      for (int b = 0; b < L; ++b) {
        if (b == a) continue;
        // R[a][b]: upper stored at [min][max]; lower = conj
        float2 rab = (a < b) ? S.A[a * P + b] : S.A[b * P + a];
        if (a > b) rab.y = -rab.y;
        const float2 wb = S.w[b];
        rr -= static_cast<double>(rab.x) * wb.x - static_cast<double>(rab.y) * wb.y;
        ri -= static_cast<double>(rab.x) * wb.y + static_cast<double>(rab.y) * wb.x;
      }
      S.r[a] = make_float2(static_cast<float>(rr), static_cast<float>(ri));
    }
    __syncthreads();
    solve(S.r);
    for (int a = lane; a < L; a += 64) {
      float2 wa = S.w[a];
      wa.x += S.r[a].x;
      wa.y += S.r[a].y;
      S.w[a] = wa;
    }
    __syncthreads();
  }

  // ---- 7. weights applied to the subarray-averaged centre snapshot -------
  float dr = 0.0f, di = 0.0f, orr = 0.0f, oi = 0.0f;
  const float2* cen = S.rows + crow * na;
  for (int a = lane; a < L; a += 64) {
    const float2 wa = S.w[a];
    dr += wa.x;
    di += wa.y;
    float sr = 0.0f, si = 0.0f;
    for (int s0 = 0; s0 < ns; ++s0) {
      sr += cen[s0 + a].x;
      si += cen[s0 + a].y;
    }
    sr /= static_cast<float>(ns);
    si /= static_cast<float>(ns);
    // conj(w) * avg
    orr += wa.x * sr + wa.y * si;
    oi += wa.x * si - wa.y * sr;
  }
  dr = wsum(dr);
  di = wsum(di);
  orr = wsum(orr);
  oi = wsum(oi);
  if (hypotf(dr, di) < 1e-30f) {
    finish(das());
    return;
  }
  // (out / den) * n_active
  const double dd = static_cast<double>(dr) * dr + static_cast<double>(di) * di;
  const float qr = static_cast<float>((static_cast<double>(orr) * dr + static_cast<double>(oi) * di) / dd);
  const float qi = static_cast<float>((static_cast<double>(oi) * dr - static_cast<double>(orr) * di) / dd);
  const float nf = static_cast<float>(na);
  finish(make_float2(qr * nf, qi * nf));
}

void launch_mv(const MvArgs& M, const float2* Y, const uint32_t* mask, float2* frame, float* acc,
               hipStream_t st) {
  const int lmax = std::min(std::max(M.subarray, 2), std::max(2, M.n_ch / 2));
  const int pmax = (lmax + 15) & ~15;
  const size_t smem = mv_smem_bytes(M.n_ch, M.range_avg, pmax);
  const dim3 grid(static_cast<unsigned>(M.n0) * M.n1), block(64);
  if (M.precision == 1)
    hipLaunchKernelGGL(mv_kernel<1>, grid, block, smem, st, M, Y, mask, frame, acc, pmax);
  else if (M.precision == 2)
    hipLaunchKernelGGL(mv_kernel<2>, grid, block, smem, st, M, Y, mask, frame, acc, pmax);
  else
    hipLaunchKernelGGL(mv_kernel<0>, grid, block, smem, st, M, Y, mask, frame, acc, pmax);
}

}  // namespace usxd
