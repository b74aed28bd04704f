// apod.hip -- learned apodization on gfx950: a 256 -> 256 -> 256 -> 256 -> 256
// MLP (tanh-form GELU after the first three layers) that maps a pixel's delayed
// channel vector (128 complex samples, rectangular aperture) to 128 complex
// channel weights. BF16 operands on MFMA (v_mfma_f32_16x16x32_bf16), FP32
// accumulation; biases, pre-activations and the output in FP32.
//
// This has no counterpart in the reference. The network is trained on the
// effective channel weights of the minimum-variance beamformer (mv.hip),
// and the frame it produces is out = sum_c (1 + o_c) Y_c over the active
// channels: a zero network is DAS with the rectangular aperture.
//
// Layout. Features are interleaved (re, im) per channel: f = 2 ch + {0, 1}.
// Every GEMM runs transposed, out^T = W^T in^T, so samples are the MFMA N
// dimension (lane % 16) and features the M and K dimensions. A 16x16 output
// tile leaves lane (g = lane / 16, r = lane % 16) holding features
// 16 m + 4 g + i (i = 0..3) of sample r; two neighbouring tiles (m = 2 kb,
// 2 kb + 1) are then exactly the B operand of K block kb of the next layer
// with the K order
//   f(kb, g, j) = 32 kb + 4 g + j        (j < 4)
//               = 32 kb + 16 + 4 g + j-4 (j >= 4),
// so activations never leave registers between layers. The weights are
// stored pre-swizzled in that K order: fragment (m, kb) is 64 lanes x 8
// bf16 = 1 KB contiguous, lane (g, r) element j = W[f(kb, g, j)][16 m + r]
// (forward) or W[16 m + r][f(kb, g, j)] (backward, dIn = W dOut). The sum
// over K does not depend on which K index the hardware pairs with which
// (g, j), only on A and B using the same map, which they do by construction.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

#include "kernels.hpp"

namespace usxd {

namespace {

typedef float f4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
typedef unsigned int u4 __attribute__((ext_vector_type(4)));

constexpr int kF = kApodFeatures;    // 256
constexpr int kL = kApodLayers;      // 4
constexpr int kMT = kF / 16;         // 16 output tiles
constexpr int kKB = kF / 32;         // 8 K blocks
constexpr int kLayerBf16 = kF * kF;  // swizzled bf16 per layer
constexpr int kNT = 1;               // 16-sample tiles per wavefront (training)
constexpr int kInferWaves = 8;       // wavefronts (and 16-pixel tiles) per inference workgroup
static_assert(2 * kInferWaves == kMT, "inference: each wavefront owns two output tiles");

// FP32 -> BF16, round to nearest even (NaN stays NaN).
__device__ __forceinline__ uint32_t f2bf(float x) {
  uint32_t u = __float_as_uint(x);
  if ((u & 0x7fffffffu) > 0x7f800000u) return 0x7fc0u;
  u += 0x7fffu + ((u >> 16) & 1u);
  return u >> 16;
}
__device__ __forceinline__ float bf2f(uint32_t h) { return __uint_as_float(h << 16); }
// Two floats -> packed BF16 pair (low = a), round to nearest even: one
// v_cvt_pk_bf16_f32.
__device__ __forceinline__ uint32_t pack2(float a, float b) {
  typedef __bf16 bf2 __attribute__((ext_vector_type(2)));
  typedef float fl2 __attribute__((ext_vector_type(2)));
  return __builtin_bit_cast(uint32_t, __builtin_convertvector(fl2{a, b}, bf2));
}

__device__ __forceinline__ bf8 as_bf8(u4 v) { return __builtin_bit_cast(bf8, v); }

// GELU in its tanh form, 0.5 z (1 + tanh(k (z + 0.044715 z^3))), k = sqrt(2 / pi),
// and its derivative; tanh(u) = 1 - 2 / (1 + e^{2u}) (saturates correctly at
// both ends: e^{2u} = inf gives 1, e^{2u} = 0 gives -1). The forward value is
// evaluated as the same function in sigmoid form, z / (1 + e^{-2a}) with
// a = k (z + c z^3), i.e. z / (1 + 2^(z (p + q z^2))), p = -2 k log2(e),
// q = p c: five VALU operations and two transcendentals (-inf..inf maps to
// -0..inf as the tanh form does).
constexpr float kGeluK = 0.7978845608028654f;
constexpr float kGeluC = 0.044715f;
constexpr float kGeluP = static_cast<float>(-2.0 * 0.7978845608028654 * 1.4426950408889634);
constexpr float kGeluQ = static_cast<float>(-2.0 * 0.7978845608028654 * 1.4426950408889634 * 0.044715);
__device__ __forceinline__ float tanh_fast(float u) {
  const float e = __expf(2.0f * u);
  return 1.0f - 2.0f * __builtin_amdgcn_rcpf(1.0f + e);
}
__device__ __forceinline__ float gelu(float z) {
  const float u = z * __builtin_fmaf(kGeluQ, z * z, kGeluP);
  return z * __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(u));
}
__device__ __forceinline__ float gelu_grad(float z) {
  const float t = tanh_fast(kGeluK * (z + kGeluC * z * z * z));
  return 0.5f * (1.0f + t) + 0.5f * z * (1.0f - t * t) * kGeluK * (1.0f + 3.0f * kGeluC * z * z);
}

// Swizzled position of W[in][out] in the forward (A = W^T) and backward
// (A = W) fragment layouts.
__host__ __device__ __forceinline__ int frag_pos(int row, int k) {
  const int m = row >> 4, r = row & 15, kb = k >> 5, q = k & 31;
  const int g = (q & 15) >> 2, j = (q & 3) + ((q >> 4) << 2);
  return ((m * kKB + kb) * 64 + g * 16 + r) * 8 + j;
}

// One layer on a wavefront's NT sample tiles: acc[m][t] = bias + A_m . act[.][t]
// with A fragments from `w` (LDS or global, 16 bytes per lane per fragment).
// The fragments of tile m + 1 are loaded while tile m's MFMAs run; the
// scheduling barrier keeps the compiler from hoisting every load of the
// layer (128 fragments) to the top, which would need 512 VGPRs.
template <bool kGlobal, int NT = kNT, int M0 = 0, int M1 = kMT>
__device__ __forceinline__ void layer(const u4* w, const float* bias, const u4 (&act)[kKB][NT],
                                      f4 (&acc)[kMT][NT], int lane) {
  const int g = lane >> 4;
  u4 cur[kKB], nxt[kKB];
  // w holds the fragments of tiles M0 .. M1-1
  auto load = [&](int m, u4 (&dst)[kKB]) {
    const u4* wm = w + (m - M0) * kKB * 64;
    // Global fragments: an opaque per-tile uniform base, so the compiler
    // addresses each load as SGPR base + lane offset instead of hoisting
    // 128 per-lane 64-bit addresses out of the layer loop.
    if (kGlobal) asm volatile("" : "+s"(wm));
#pragma unroll
    for (int kb = 0; kb < kKB; ++kb) dst[kb] = wm[kb * 64 + lane];
  };
  load(M0, cur);
#pragma unroll
  for (int m = M0; m < M1; ++m) {
    if (m + 1 < M1) load(m + 1, nxt);
    f4 b = f4{0.0f, 0.0f, 0.0f, 0.0f};
    if (bias) {
      const float4 bb = *reinterpret_cast<const float4*>(bias + 16 * m + 4 * g);
      b = f4{bb.x, bb.y, bb.z, bb.w};
    }
#pragma unroll
    for (int t = 0; t < NT; ++t) acc[m][t] = b;
#pragma unroll
    for (int kb = 0; kb < kKB; ++kb)
#pragma unroll
      for (int t = 0; t < NT; ++t)
        acc[m][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(as_bf8(cur[kb]), as_bf8(act[kb][t]),
                                                            acc[m][t], 0, 0, 0);
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int kb = 0; kb < kKB; ++kb) cur[kb] = nxt[kb];
  }
}

// Network input of one pixel from its channel vector: the active channels
// (mask bit set) rotated by conj(sum y) / |sum y| and scaled to unit mean
// power, in BF16. Lane (g, r) of tile t holds pixel r's features for K
// blocks kb = 0..7: channels 16 kb + 2 g + {0, 1} and 16 kb + 8 + 2 g + {0, 1}.
// Inactive channels and invalid pixels give 0. Returns the active count.
struct PixelSums {
  float sr, si, e;  // sum y (re, im) and sum |y|^2 over the active channels
};
// The loads of load_features, issued ahead of their use: lane (g, r)'s
// channels of the pixel and the pixel's mask words.
struct RawPixel {
  float4 v[kKB][2];
  uint32_t mw[4];
};
__device__ __forceinline__ void fetch_pixel(const float2* Y, const uint32_t* mask, int n_ch,
                                            long pix, bool valid, int lane, RawPixel& R) {
  const int g = lane >> 4;
  const int nw = (n_ch + 31) / 32;
#pragma unroll
  for (int q = 0; q < 4; ++q) R.mw[q] = valid ? mask[pix * nw + q] : 0u;
#pragma unroll
  for (int kb = 0; kb < kKB; ++kb) {
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int ch = 16 * kb + 8 * h + 2 * g;
      const uint32_t bits = R.mw[ch >> 5] >> (ch & 31);
      R.v[kb][h] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
      if (bits & 3u) R.v[kb][h] = *reinterpret_cast<const float4*>(Y + pix * n_ch + ch);
    }
  }
}
// fetch_pixel for the inference kernel through buffer resources over Y
// (n_pix x 128 float2) and the mask words (n_pix x 4): one 16-byte mask load,
// then lane (g, r)'s 16 channel pairs at one lane offset plus immediates. A
// pair with no active channel, and a pixel past the frame, reads zeros
// without touching memory (its offset fails the range check).
__device__ __forceinline__ u4 mask_row(__amdgpu_buffer_rsrc_t mrs, uint32_t pix) {
  return __builtin_amdgcn_raw_buffer_load_b128(mrs, pix * 16u, 0, 0);
}
__device__ __forceinline__ void fetch_pixel_rsrc(__amdgpu_buffer_rsrc_t yrs, const u4& m,
                                                 uint32_t pix, int lane, RawPixel& R) {
  const int g = lane >> 4;
  R.mw[0] = m[0];
  R.mw[1] = m[1];
  R.mw[2] = m[2];
  R.mw[3] = m[3];
  const uint32_t vo = (pix * kApodChannels + 2u * g) * 8u;
#pragma unroll
  for (int kb = 0; kb < kKB; ++kb) {
    const uint32_t w = R.mw[kb >> 1] >> (2 * g);
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const uint32_t on = w & (3u << (16 * (kb & 1) + 8 * h));
      const uint32_t o = on ? vo : 0xFFFFF000u;
      R.v[kb][h] = __builtin_bit_cast(
          float4, __builtin_amdgcn_raw_buffer_load_b128(yrs, o + (16 * kb + 8 * h) * 8, 0, 0));
    }
  }
}
__device__ __forceinline__ int features_of(const RawPixel& R, u4 (&x)[kKB], PixelSums* sums) {
  // A fetched pair holds an active channel, so its 32-channel block was
  // written by the channel-vector kernel, which stores exactly 0 for the
  // block's inactive channels: the entries need no masking here.
  const float4 (&v)[kKB][2] = R.v;
  float sr = 0.0f, si = 0.0f, e = 0.0f;
  const int na = __popc(R.mw[0]) + __popc(R.mw[1]) + __popc(R.mw[2]) + __popc(R.mw[3]);
#pragma unroll
  for (int kb = 0; kb < kKB; ++kb) {
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const float4 y = v[kb][h];
      sr += y.x + y.z;
      si += y.y + y.w;
      e = __builtin_fmaf(y.x, y.x, e);
      e = __builtin_fmaf(y.y, y.y, e);
      e = __builtin_fmaf(y.z, y.z, e);
      e = __builtin_fmaf(y.w, y.w, e);
    }
  }
#pragma unroll
  for (int o = 16; o < 64; o <<= 1) {
    sr += __shfl_xor(sr, o, 64);
    si += __shfl_xor(si, o, 64);
    e += __shfl_xor(e, o, 64);
  }
  if (sums) *sums = PixelSums{sr, si, e};
  // conj(sum y) / |sum y| and 1 / rms through the hardware reciprocal square
  // root (an ulp or two; the features are rounded to BF16 next)
  float cr = 1.0f, ci = 0.0f, s = 0.0f;
  if (e > 0.0f && na > 0) {
    const float m2 = __builtin_fmaf(sr, sr, si * si);
    if (m2 > 0.0f) {
      const float im = __builtin_amdgcn_rsqf(m2);
      cr = sr * im;
      ci = -si * im;
    }
    s = __builtin_amdgcn_rsqf(e * __builtin_amdgcn_rcpf(static_cast<float>(na)));
  }
  cr *= s;
  ci *= s;
#pragma unroll
  for (int kb = 0; kb < kKB; ++kb) {
    const float4 a = v[kb][0], b = v[kb][1];
    x[kb] = u4{pack2(__builtin_fmaf(a.x, cr, -a.y * ci), __builtin_fmaf(a.x, ci, a.y * cr)),
               pack2(__builtin_fmaf(a.z, cr, -a.w * ci), __builtin_fmaf(a.z, ci, a.w * cr)),
               pack2(__builtin_fmaf(b.x, cr, -b.y * ci), __builtin_fmaf(b.x, ci, b.y * cr)),
               pack2(__builtin_fmaf(b.z, cr, -b.w * ci), __builtin_fmaf(b.z, ci, b.w * cr))};
  }
  return na;
}
__device__ __forceinline__ int load_features(const float2* Y, const uint32_t* mask, int n_ch,
                                             long pix, bool valid, int lane, u4 (&x)[kKB],
                                             PixelSums* sums = nullptr) {
  RawPixel R;
  fetch_pixel(Y, mask, n_ch, pix, valid, lane, R);
  return features_of(R, x, sums);
}

// The next layer's B operand from output tiles 2 kb and 2 kb + 1, through GELU.
template <int NT = kNT>
__device__ __forceinline__ void activate(const f4 (&acc)[kMT][NT], u4 (&act)[kKB][NT]) {
#pragma unroll
  for (int kb = 0; kb < kKB; ++kb)
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      const f4 a = acc[2 * kb][t], b = acc[2 * kb + 1][t];
      act[kb][t] = u4{pack2(gelu(a[0]), gelu(a[1])), pack2(gelu(a[2]), gelu(a[3])),
                      pack2(gelu(b[0]), gelu(b[1])), pack2(gelu(b[2]), gelu(b[3]))};
      // bound the interleaving of the erff evaluations (register pressure)
      __builtin_amdgcn_sched_barrier(0);
    }
}

}  // namespace

// ---------------------------------------------------------------------------
// Inference, weight-stationary: one workgroup = kInferWaves wavefronts and
// kInferWaves 16-pixel sample tiles (128 pixels, consecutive depths of a
// line). Wavefront w owns output tiles 2w and 2w + 1 of every layer -- the
// features of K block w of the next layer -- and holds their 16 weight
// fragments in registers (64 VGPRs; the next layer's are loaded from L2 while
// this layer runs). The activations of the 128 pixels sit in LDS in the B
// operand layout ([tile][K block][lane], 16 bytes per lane, 64 KB a layer,
// two buffers): each 16-byte fragment read feeds two MFMAs (both of the
// wavefront's output tiles), and the weights never pass through LDS. After
// GELU a wavefront's two output tiles are exactly K block w of the next
// layer's input, one 16-byte store per sample tile. One barrier per layer.
// Each output tile accumulates K blocks 0..7 in order onto its bias (no
// K block is skipped: a conditional MFMA costs more VALU in accumulator
// copies than the products it saves, and zero features add exactly
// nothing). The last layer's outputs o_c stay in the owning wavefront, which
// forms its channels' share of sum (1 + o_c) y_c per pixel; the shares meet
// in LDS. The kernel is VALU-bound (GELU: two quarter-rate transcendentals
// per activation), not bound by the MFMAs or by LDS.
// ---------------------------------------------------------------------------
constexpr int kActU4 = kInferWaves * kKB * 64;  // one activation buffer (u4), 64 KB

size_t apod_infer_smem_bytes() {
  return sizeof(u4) * 2 * kActU4 + sizeof(float2) * kInferWaves * kInferWaves * 16 +
         sizeof(float) * kL * kF + 8;  // + the biases and the tile's flag bytes
}

// Fragments (l, 2 wv + q, kb) through a buffer resource: one lane offset
// (16 lane) and a scalar offset per fragment, so no per-lane 64-bit
// addresses are kept across the tile loop.
__device__ __forceinline__ void load_weights(__amdgpu_buffer_rsrc_t wrs, int l, int wv, int lane,
                                             u4 (&w)[2][kKB]) {
  const int base = (l * (kLayerBf16 / 8) + (2 * wv) * kKB * 64) * 16;
#pragma unroll
  for (int q = 0; q < 2; ++q)
#pragma unroll
    for (int kb = 0; kb < kKB; ++kb)
      w[q][kb] = __builtin_amdgcn_raw_buffer_load_b128(wrs, 16 * lane, base + (q * kKB + kb) * 1024, 0);
}

// A layer runs over the workgroup's kInferWaves sample tiles two at a time
// (four accumulator chains), software-pipelined: the MFMAs of a pair of
// sample tiles (ws_products, here) are issued before the VALU work on the
// previous pair's outputs (GELU, two quarter-rate transcendentals per value,
// or the last layer's weighted sums), so the matrix pipe and the VALU work
// together.
__device__ __forceinline__ void ws_products(const u4 (&w)[2][kKB], const float4& b0,
                                            const float4& b1, const u4* in, int t, int lane,
                                            f4 (&acc)[2][2]) {
#pragma unroll
  for (int u = 0; u < 2; ++u) {
    acc[0][u] = f4{b0.x, b0.y, b0.z, b0.w};
    acc[1][u] = f4{b1.x, b1.y, b1.z, b1.w};
  }
  u4 cur[2], nxt[2];
  cur[0] = in[(t * kKB) * 64 + lane];
  cur[1] = in[((t + 1) * kKB) * 64 + lane];
#pragma unroll
  for (int kb = 0; kb < kKB; ++kb) {
    if (kb + 1 < kKB) {
      nxt[0] = in[(t * kKB + kb + 1) * 64 + lane];
      nxt[1] = in[((t + 1) * kKB + kb + 1) * 64 + lane];
    }
    acc[0][0] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(as_bf8(w[0][kb]), as_bf8(cur[0]), acc[0][0], 0, 0, 0);
    acc[1][0] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(as_bf8(w[1][kb]), as_bf8(cur[0]), acc[1][0], 0, 0, 0);
    acc[0][1] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(as_bf8(w[0][kb]), as_bf8(cur[1]), acc[0][1], 0, 0, 0);
    acc[1][1] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(as_bf8(w[1][kb]), as_bf8(cur[1]), acc[1][1], 0, 0, 0);
    cur[0] = nxt[0];
    cur[1] = nxt[1];
  }
}


// K block wv of the next layer for sample tiles t, t + 1: GELU of the
// wavefront's two output tiles.
__device__ __forceinline__ void ws_activate(const f4 (&acc)[2][2], int t, u4* out, int wv,
                                            int lane) {
#pragma unroll
  for (int u = 0; u < 2; ++u) {
    const f4 a = acc[0][u], b = acc[1][u];
    out[((t + u) * kKB + wv) * 64 + lane] =
        u4{pack2(gelu(a[0]), gelu(a[1])), pack2(gelu(a[2]), gelu(a[3])),
           pack2(gelu(b[0]), gelu(b[1])), pack2(gelu(b[2]), gelu(b[3]))};
  }
}

// One hidden layer: the products of a pair of sample tiles are issued
// before the GELU of the previous pair.
__device__ __forceinline__ void ws_layer(const u4 (&w)[2][kKB], const float* bias, const u4* in,
                                         u4* out, int wv, int lane) {
  const int g = lane >> 4;
  const float4 b0 = *reinterpret_cast<const float4*>(bias + 16 * (2 * wv) + 4 * g);
  const float4 b1 = *reinterpret_cast<const float4*>(bias + 16 * (2 * wv + 1) + 4 * g);
  static_assert(kInferWaves % 4 == 0, "sample tiles in pairs of pairs");
  f4 acc0[2][2], acc1[2][2];
  ws_products(w, b0, b1, in, 0, lane, acc0);
#pragma unroll
  for (int t = 2; t < kInferWaves; t += 4) {
    ws_products(w, b0, b1, in, t, lane, acc1);
    ws_activate(acc0, t - 2, out, wv, lane);
    if (t + 2 < kInferWaves) ws_products(w, b0, b1, in, t + 2, lane, acc0);
    ws_activate(acc1, t, out, wv, lane);
  }
}

// The last layer's view of the tile's channel vectors: buffer resources over
// Y and the mask words, lane (g, r)'s byte offsets for pixel r of sample
// tile 0 (channel pair 16 wv + 2 g, mask word wv / 2; sample tile t adds 16 t
// pixels as a scalar offset) and the position of the pair's bits in that
// word (2 g + 16 (wv % 2); the second pair's 8 above).
struct LastIn {
  __amdgpu_buffer_rsrc_t yrs, mrs;
  uint32_t yo, mo, shift;
};
// What the last layer reads for sample tiles t, t + 1: lane (g, r)'s mask word
// and its two channel pairs of pixel r (a pixel past the frame reads zeros:
// range check).
struct LastLoads {
  uint32_t mw[2];
  float4 y[2][2];
};
__device__ __forceinline__ void last_loads(const LastIn& in, int t, LastLoads& L) {
#pragma unroll
  for (int u = 0; u < 2; ++u) {
    const uint32_t st = static_cast<uint32_t>(16 * (t + u));
    L.mw[u] = __builtin_amdgcn_raw_buffer_load_b32(in.mrs, in.mo, st * 16u, 0);
#pragma unroll
    for (int h = 0; h < 2; ++h)
      L.y[u][h] = __builtin_bit_cast(float4, __builtin_amdgcn_raw_buffer_load_b128(
                                                 in.yrs, in.yo + 64 * h, st * (kApodChannels * 8u), 0));
  }
}
// o of channels 16 wv + 2 g + {0, 1} (tile 2 wv) and 16 wv + 8 + 2 g + {0, 1}
// (tile 2 wv + 1) of pixel r of sample tiles t, t + 1: the wavefront's
// shares of sum (1 + o_c) y_c go to `part`. Inactive channels (whose entries
// may hold anything) are selected away, so their (1 + o) multiplies a zero.
// The selects are per channel: a select of a whole pair lets the compiler
// turn it into a load behind the mask word's.
__device__ __forceinline__ void last_finish(const f4 (&acc)[2][2], int t, int wv, int lane,
                                            const LastLoads& L, uint32_t shift, float2* part) {
  const int g = lane >> 4, r = lane & 15;
#pragma unroll
  for (int u = 0; u < 2; ++u) {
    const uint32_t mw = L.mw[u] >> shift;
    float lr = 0.0f, li = 0.0f;
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const bool on0 = mw & (1u << (8 * h)), on1 = mw & (2u << (8 * h));
      const float4 yv = L.y[u][h];
      const float4 y = {on0 ? yv.x : 0.0f, on0 ? yv.y : 0.0f, on1 ? yv.z : 0.0f, on1 ? yv.w : 0.0f};
      const f4 o = acc[h][u];
      const float g0r = 1.0f + o[0], g1r = 1.0f + o[2];
      lr = __builtin_fmaf(g0r, y.x, lr);
      lr = __builtin_fmaf(-o[1], y.y, lr);
      lr = __builtin_fmaf(g1r, y.z, lr);
      lr = __builtin_fmaf(-o[3], y.w, lr);
      li = __builtin_fmaf(g0r, y.y, li);
      li = __builtin_fmaf(o[1], y.x, li);
      li = __builtin_fmaf(g1r, y.w, li);
      li = __builtin_fmaf(o[3], y.z, li);
    }
#pragma unroll
    for (int o = 16; o < 64; o <<= 1) {
      lr += __shfl_xor(lr, o, 64);
      li += __shfl_xor(li, o, 64);
    }
    if (g == 0) part[(wv * kInferWaves + t + u) * 16 + r] = make_float2(lr, li);
  }
}

// A pair of sample tiles whose 32 pixels have no active channel among the
// wavefront's 16 (bit wv of neither tile's flag byte in fl) contributes
// nothing: its loads, products and sums are skipped and its shares are 0.
__device__ __forceinline__ void last_zero(int t, int wv, int lane, float2* part) {
  const int g = lane >> 4, r = lane & 15;
  if (g == 0) {
    part[(wv * kInferWaves + t) * 16 + r] = make_float2(0.0f, 0.0f);
    part[(wv * kInferWaves + t + 1) * 16 + r] = make_float2(0.0f, 0.0f);
  }
}

// The last layer: as ws_layer, with each pair's channel vectors loaded one
// pair ahead of their use, so their memory latency hides behind a pair's
// products and the previous pair's sums. Once the last products are issued
// the weights' registers are free, and the next tile's channel vectors (at
// pixel npix of the wavefront's next sample tile, mask row nm) are fetched
// into R behind the last two pairs' sums. fl: the tile's flag bytes (byte t:
// the 16-channel blocks holding an active channel of sample tile t).
__device__ __forceinline__ void ws_last_layer(const u4 (&w)[2][kKB], const float* bias,
                                              const u4* in, int wv, int lane, const LastIn& li,
                                              float2* part, const u4& nm, uint32_t npix,
                                              RawPixel& R, uint64_t fl) {
  const int g = lane >> 4;
  const float4 b0 = *reinterpret_cast<const float4*>(bias + 16 * (2 * wv) + 4 * g);
  const float4 b1 = *reinterpret_cast<const float4*>(bias + 16 * (2 * wv + 1) + 4 * g);
  auto live = [&](int t) { return (((fl >> (8 * t)) | (fl >> (8 * t + 8))) >> wv) & 1u; };
  f4 acc0[2][2], acc1[2][2];
  LastLoads L0, L1;
  const bool on0 = live(0);
  if (on0) {
    last_loads(li, 0, L0);
    ws_products(w, b0, b1, in, 0, lane, acc0);
  }
  bool onA = on0;
#pragma unroll
  for (int t = 2; t < kInferWaves; t += 4) {
    const bool onB = live(t);
    if (onB) {
      last_loads(li, t, L1);
      ws_products(w, b0, b1, in, t, lane, acc1);
    }
    if (onA) last_finish(acc0, t - 2, wv, lane, L0, li.shift, part);
    else last_zero(t - 2, wv, lane, part);
    if (t + 2 < kInferWaves) {
      onA = live(t + 2);
      if (onA) {
        last_loads(li, t + 2, L0);
        ws_products(w, b0, b1, in, t + 2, lane, acc0);
      }
    } else {
      fetch_pixel_rsrc(li.yrs, nm, npix, lane, R);
    }
    if (onB) last_finish(acc1, t, wv, lane, L1, li.shift, part);
    else last_zero(t, wv, lane, part);
  }
}

// Stage this wavefront's 16 pixels (sample tile wv) from their fetched
// channel vectors into the activation buffer. Inactive channels are zero
// features, so the first layer's products over them add exactly nothing.
// Also this sample tile's flag byte: bit b set when one of its 16 pixels has
// an active channel among 16 b .. 16 b + 15 (the mask words are the same in
// the four lanes of a pixel).
__device__ __forceinline__ int stage_input(const RawPixel& R, int wv, int lane, u4* act,
                                           PixelSums* ps, uint8_t* flags) {
  u4 x[kKB];
  const int na = features_of(R, x, ps);
#pragma unroll
  for (int kb = 0; kb < kKB; ++kb) act[(wv * kKB + kb) * 64 + lane] = x[kb];
  uint32_t fb = 0;
#pragma unroll
  for (int b = 0; b < kKB; ++b)
    fb |= ((R.mw[b >> 1] >> (16 * (b & 1))) & 0xFFFFu) ? (1u << b) : 0u;
#pragma unroll
  for (int o = 1; o < 16; o <<= 1) fb |= __shfl_xor(fb, o, 64);
  if (lane == 0) flags[wv] = static_cast<uint8_t>(fb);
  return na;
}

// Persistent: workgroup b takes the 128-pixel tiles b, b + gridDim.x, ...
// (one workgroup per CU: 140 KB of LDS). The next tile's mask rows are loaded
// before the last layer and its channel vectors once the last layer's
// products are issued (into the registers its weights leave), so most of
// their memory latency hides behind the last sums; the first two layers'
// weights follow the staging.
__global__ __launch_bounds__(64 * kInferWaves, 1) void apod_infer_kernel(ApodNet N, int n_pix, int n_ch,
                                                         const float2* __restrict__ Y,
                                                         const uint32_t* __restrict__ mask,
                                                         float2* __restrict__ frame, float* acc_out,
                                                         int cf, int incoherent) {
  extern __shared__ u4 smem[];
  u4* act0 = smem;
  u4* act1 = smem + kActU4;
  float2* part = reinterpret_cast<float2*>(smem + 2 * kActU4);
  // the biases in LDS: a layer's bias reads then wait on no vector-memory
  // load (in-order counters would hold them behind the weights in flight)
  float* sbias = reinterpret_cast<float*>(part + kInferWaves * kInferWaves * 16);
  uint8_t* flags = reinterpret_cast<uint8_t*>(sbias + kL * kF);
  static_assert(kInferWaves == 8, "one flag byte per sample tile, 8 to a 64-bit word");
  for (int q = threadIdx.x; q < kL * kF; q += 64 * kInferWaves) sbias[q] = N.bias[q];
  const int tid = threadIdx.x, lane = tid & 63;
  const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int g = lane >> 4, r = lane & 15;
  const int n_tiles = (n_pix + 16 * kInferWaves - 1) / (16 * kInferWaves);
  (void)n_ch;  // kApodChannels: the host admits no other channel count
  // the forward fragments, 4 layers x 128 KB, as a raw buffer
  const __amdgpu_buffer_rsrc_t wsrc =
      __builtin_amdgcn_make_buffer_rsrc(N.wf, 0, kL * kLayerBf16 * 2, 0x00020000);
  // Y and the mask words as raw buffers (the host keeps n_pix x 1 KB of
  // channel vectors under 4 GiB): 32-bit offsets, zeros past the frame
  LastIn lin;
  lin.yrs = __builtin_amdgcn_make_buffer_rsrc(const_cast<float2*>(Y), 0,
                                              static_cast<int>(static_cast<uint32_t>(n_pix) * (kApodChannels * 8u)),
                                              0x00020000);
  lin.mrs = __builtin_amdgcn_make_buffer_rsrc(const_cast<uint32_t*>(mask), 0,
                                              static_cast<int>(static_cast<uint32_t>(n_pix) * 16u),
                                              0x00020000);
  lin.shift = 2 * g + 16 * (wv & 1);
  int tile = blockIdx.x;
  u4 wa[2][kKB], wb[2][kKB];
  load_weights(wsrc, 0, wv, lane, wa);
  int na;
  PixelSums ps;
  {
    const uint32_t pix = static_cast<uint32_t>(tile) * (16 * kInferWaves) + 16 * wv + r;
    RawPixel R;
    fetch_pixel_rsrc(lin.yrs, mask_row(lin.mrs, pix), pix, lane, R);
    na = stage_input(R, wv, lane, act0, &ps, flags);
  }
  load_weights(wsrc, 1, wv, lane, wb);
  __syncthreads();
  while (tile < n_tiles) {
    const uint32_t base = static_cast<uint32_t>(tile) * (16 * kInferWaves);
    lin.yo = ((base + r) * kApodChannels + 16 * wv + 2 * g) * 8u;
    lin.mo = ((base + r) * 4u + (wv >> 1)) * 4u;
    // the tile's flag bytes, read before the first layer's barrier (the next
    // staging, which rewrites them, follows three more barriers)
    const uint64_t fl64 = *reinterpret_cast<const uint64_t*>(flags);
    const uint64_t fl = (static_cast<uint64_t>(__builtin_amdgcn_readfirstlane(
                             static_cast<uint32_t>(fl64 >> 32))) << 32) |
                        __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(fl64));
    ws_layer(wa, sbias + 0 * kF, act0, act1, wv, lane);
    load_weights(wsrc, 2, wv, lane, wa);
    __syncthreads();
    ws_layer(wb, sbias + 1 * kF, act1, act0, wv, lane);
    load_weights(wsrc, 3, wv, lane, wb);
    __syncthreads();
    ws_layer(wa, sbias + 2 * kF, act0, act1, wv, lane);
    __syncthreads();
    // the next tile's mask rows now, its channel vectors during the last
    // layer's final sums; then act0 (free) takes them
    const int next = tile + gridDim.x;
    const uint32_t npix = static_cast<uint32_t>(next) * (16 * kInferWaves) + 16 * wv + r;
    const u4 nm = mask_row(lin.mrs, npix);
    RawPixel R;
    ws_last_layer(wb, sbias + 3 * kF, act1, wv, lane, lin, part, nm, npix, R, fl);
    PixelSums nps;
    const int nna = stage_input(R, wv, lane, act0, &nps, flags);
    load_weights(wsrc, 0, wv, lane, wa);
    load_weights(wsrc, 1, wv, lane, wb);
    __syncthreads();
    // out = sum over active channels of (1 + o_c) y_c (the wavefronts' shares
    // in order); fewer than 4 active channels -> the DAS sum, as minimum
    // variance does.
    const long pix = static_cast<long>(base) + 16 * wv + r;
    if (pix < n_pix && g == 0) {
      float lr = 0.0f, li = 0.0f;
#pragma unroll
      for (int w = 0; w < kInferWaves; ++w) {
        const float2 sp = part[(w * kInferWaves + wv) * 16 + r];
        lr += sp.x;
        li += sp.y;
      }
      float2 out = na >= 4 ? make_float2(lr, li) : make_float2(ps.sr, ps.si);
      if (cf) {
        const float mc2 = ps.sr * ps.sr + ps.si * ps.si;
        const float f = (ps.e > 0.0f && na > 0) ? mc2 / (static_cast<float>(na) * ps.e) : 0.0f;
        out.x *= f;
        out.y *= f;
      }
      if (incoherent) {
        acc_out[pix] += out.x * out.x + out.y * out.y;
      } else {
        frame[pix] = out;
      }
    }
    na = nna;
    ps = nps;
    tile = next;
  }
}

// ---------------------------------------------------------------------------
// Training set rows from a frame's channel vectors and minimum-variance
// weights: X (bf16 features, the inference kernel's input), G (weights as
// interleaved floats), mask words and the sample weight 1 / n_active (0 when
// minimum variance fell back to DAS or the pixel has < 4 active channels).
// One wavefront = 32 listed pixels, the inference kernel's lane mapping.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(64) void apod_gather_kernel(int n_list, const int* __restrict__ pixels,
                                                         int n_ch, const float2* __restrict__ Y,
                                                         const uint32_t* __restrict__ mask,
                                                         const float2* __restrict__ weights,
                                                         const int* __restrict__ status,
                                                         ApodData D, long dst) {
  const int lane = threadIdx.x, g = lane >> 4, r = lane & 15;
  const int nw = (n_ch + 31) / 32;
#pragma unroll
  for (int t = 0; t < kNT; ++t) {
    const int i = blockIdx.x * (16 * kNT) + 16 * t + r;
    const bool valid = i < n_list;
    const long pix = valid ? pixels[i] : 0;
    u4 x[kKB];
    const int na = load_features(Y, mask, n_ch, pix, valid, lane, x);
    if (!valid) continue;
    const long row = dst + i;
    uint2* xr = reinterpret_cast<uint2*>(D.x + row * kF);
#pragma unroll
    for (int kb = 0; kb < kKB; ++kb) {
      xr[(32 * kb + 4 * g) / 4] = make_uint2(x[kb].x, x[kb].y);
      xr[(32 * kb + 16 + 4 * g) / 4] = make_uint2(x[kb].z, x[kb].w);
    }
    // labels, mask and sample weight: 16 lanes per pixel row
    const float4* src = reinterpret_cast<const float4*>(weights + pix * n_ch);
    float4* gd = reinterpret_cast<float4*>(D.g + row * kF);
    for (int q = g; q < kF / 4; q += 4) gd[q] = src[q];
    if (g == 0) {
#pragma unroll
      for (int q = 0; q < 4; ++q) D.mask[row * 4 + q] = mask[pix * nw + q];
      D.sw[row] = (status[pix] == 0 && na >= 4) ? 1.0f / static_cast<float>(na) : 0.0f;
    }
  }
}

// ---------------------------------------------------------------------------
// Training step, forward and backward, one wavefront per 16 * kNT samples
// with weight fragments read from global memory (L2-resident: 1 MB for both
// layouts). Writes, for the weight-gradient kernel, the layer inputs H_l and
// the output gradients dZ_l ([sample][feature], bf16), and the FP32
// pre-activations for the GELU derivative; every store is a lane's row base
// plus an immediate offset. Loss partial per wavefront.
//
// Loss: (1/B) sum_b sw_b sum_{c active} |1 + o_c - g_c|^2 (o, g complex).
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t mix32(uint64_t x) {
  x ^= x >> 33;
  x *= 0xff51afd7ed558ccdull;
  x ^= x >> 33;
  x *= 0xc4ceb9fe1a85ec53ull;
  x ^= x >> 33;
  return static_cast<uint32_t>(x);
}

__device__ __forceinline__ long sample_index(const ApodStep& S, int b) {
  if (S.random) {
    const uint64_t h = mix32(S.seed * 0x9e3779b97f4a7c15ull + static_cast<uint64_t>(S.step) * 0x100000001b3ull + b);
    return static_cast<long>(h % static_cast<uint64_t>(S.n_data));
  }
  return b < S.n_data ? b : -1;
}

// A lane's B-operand fragments (8 bf16 per K block) into row `row` of a
// [sample][256] bf16 buffer: features 32 kb + 4 g + 0..3 and 32 kb + 16 + 4 g + 0..3.
__device__ __forceinline__ void store_frags(uint16_t* row, const u4 (&v)[kKB], int g) {
  uint2* p = reinterpret_cast<uint2*>(row) + g;
#pragma unroll
  for (int kb = 0; kb < kKB; ++kb) {
    p[8 * kb] = make_uint2(v[kb].x, v[kb].y);
    p[8 * kb + 4] = make_uint2(v[kb].z, v[kb].w);
  }
}

__global__ __launch_bounds__(64) void apod_train_kernel(ApodNet N, ApodData D, ApodStep S,
                                                        ApodWork W) {
  const int lane = threadIdx.x, g = lane >> 4, r = lane & 15;
  const int B = S.batch;
  const int b0 = blockIdx.x * (16 * kNT);
  const u4* wf = reinterpret_cast<const u4*>(N.wf);
  const u4* wb = reinterpret_cast<const u4*>(N.wb);
  long src[kNT];
  u4 act[kKB][kNT];
#pragma unroll
  for (int t = 0; t < kNT; ++t) {
    src[t] = sample_index(S, b0 + 16 * t + r);
    const uint2* xr = reinterpret_cast<const uint2*>(D.x + (src[t] < 0 ? 0 : src[t]) * kF) + g;
#pragma unroll
    for (int kb = 0; kb < kKB; ++kb) {
      uint2 lo = make_uint2(0u, 0u), hi = lo;
      if (src[t] >= 0) {
        lo = xr[8 * kb];
        hi = xr[8 * kb + 4];
      }
      act[kb][t] = u4{lo.x, lo.y, hi.x, hi.y};
    }
  }
  f4 acc[kMT][kNT];
#pragma nounroll
  for (int l = 0; l < kL; ++l) {
#pragma unroll
    for (int t = 0; t < kNT; ++t) {
      u4 v[kKB];
#pragma unroll
      for (int kb = 0; kb < kKB; ++kb) v[kb] = act[kb][t];
      store_frags(W.h + (static_cast<size_t>(l) * B + b0 + 16 * t + r) * kF, v, g);
    }
    layer<true>(wf + l * (kLayerBf16 / 8), N.bias + l * kF, act, acc, lane);
    if (l + 1 < kL) {
#pragma unroll
      for (int t = 0; t < kNT; ++t) {
        float4* zr = reinterpret_cast<float4*>(W.z + (static_cast<size_t>(l) * B + b0 + 16 * t + r) * kF) + g;
#pragma unroll
        for (int m = 0; m < kMT; ++m) {
          const f4 a = acc[m][t];
          zr[4 * m] = make_float4(a[0], a[1], a[2], a[3]);
        }
      }
      activate(acc, act);
    }
  }
  // loss and dO (FP32); dZ is then rounded to bf16, the operand of the
  // backward GEMMs and of the weight gradients.
  float loss = 0.0f;
#pragma unroll
  for (int t = 0; t < kNT; ++t) {
    const long s = src[t];
    float sw = 0.0f;
    uint32_t mw[4] = {0u, 0u, 0u, 0u};
    if (s >= 0) {
      sw = D.sw[s];
#pragma unroll
      for (int q = 0; q < 4; ++q) mw[q] = D.mask[s * 4 + q];
    }
    const float gs = S.grad_scale * sw;
    const float4* lr = reinterpret_cast<const float4*>(D.g + (s < 0 ? 0 : s) * kF) + g;
#pragma unroll
    for (int m = 0; m < kMT; ++m) {
      float4 lab = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
      if (sw != 0.0f) lab = lr[4 * m];
      const float lv[4] = {lab.x, lab.y, lab.z, lab.w};
      const int ch = 8 * m + 2 * g;
      const uint32_t bits = mw[ch >> 5] >> (ch & 31);
      f4 d;
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const bool on = (bits >> (i >> 1)) & 1u;
        const float pred = acc[m][t][i] + ((i & 1) ? 0.0f : 1.0f);
        const float e = on ? pred - lv[i] : 0.0f;
        loss += sw * e * e;
        d[i] = gs * e;
      }
      acc[m][t] = d;
    }
  }
  // backward: dZ_l (acc) -> bf16 fragments -> dH_l = W_l dZ_l -> dZ_{l-1} = dH_l * gelu'(Z_{l-1})
#pragma nounroll
  for (int l = kL - 1; l >= 0; --l) {
#pragma unroll
    for (int t = 0; t < kNT; ++t) {
      u4 v[kKB];
#pragma unroll
      for (int kb = 0; kb < kKB; ++kb) {
        const f4 a = acc[2 * kb][t], c = acc[2 * kb + 1][t];
        v[kb] = u4{pack2(a[0], a[1]), pack2(a[2], a[3]), pack2(c[0], c[1]), pack2(c[2], c[3])};
        act[kb][t] = v[kb];
      }
      store_frags(W.dz + (static_cast<size_t>(l) * B + b0 + 16 * t + r) * kF, v, g);
    }
    if (l == 0) break;
    layer<true>(wb + l * (kLayerBf16 / 8), nullptr, act, acc, lane);
#pragma unroll
    for (int t = 0; t < kNT; ++t) {
      const float4* zr = reinterpret_cast<const float4*>(W.z + (static_cast<size_t>(l - 1) * B + b0 + 16 * t + r) * kF) + g;
#pragma unroll
      for (int m = 0; m < kMT; ++m) {
        const float4 z = zr[4 * m];
        acc[m][t][0] *= gelu_grad(z.x);
        acc[m][t][1] *= gelu_grad(z.y);
        acc[m][t][2] *= gelu_grad(z.z);
        acc[m][t][3] *= gelu_grad(z.w);
        __builtin_amdgcn_sched_barrier(0);
      }
    }
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) loss += __shfl_xor(loss, o, 64);
  if (lane == 0) W.loss_part[blockIdx.x] = loss * S.loss_scale;
}

// ---------------------------------------------------------------------------
// Weight gradients dW_l[in][out] = sum_b H_l[b][in] dZ_l[b][out] over one K
// split of the batch, bf16 MFMA with the samples as K. Workgroup = 4
// wavefronts on a 64 x 64 block (2 x 2 tiles each); per 32-sample step the
// two 32 x 64 row blocks are staged transposed in LDS ([feature][sample]) so
// each operand fragment is one 16-byte read. Grid (16 blocks, splits, 4).
// ---------------------------------------------------------------------------
constexpr int kDwPad = 40;  // bf16 per transposed LDS row (32 samples + 8)

__global__ __launch_bounds__(256) void apod_dw_kernel(ApodWork W, int B, int ks) {
  __shared__ __align__(16) uint16_t ht[64 * kDwPad], dt[64 * kDwPad];
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int g = lane >> 4, r = lane & 15;
  const int blk = blockIdx.x, split = blockIdx.y, l = blockIdx.z;
  const int in0 = (blk >> 2) * 64, out0 = (blk & 3) * 64;
  const int wi = (wv >> 1) * 32, wo = (wv & 1) * 32;
  const uint16_t* h = W.h + static_cast<size_t>(l) * B * kF;
  const uint16_t* dz = W.dz + static_cast<size_t>(l) * B * kF;
  f4 acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int c = 0; c < 2; ++c) acc[a][c] = f4{0.0f, 0.0f, 0.0f, 0.0f};
  // staging: thread -> (sample sr, 8 features from fc)
  const int sr = tid >> 3, fc = (tid & 7) * 8;
  const int k_lo = split * ks, k_hi = k_lo + ks;
  for (int k0 = k_lo; k0 < k_hi; k0 += 32) {
    const u4 hv = *reinterpret_cast<const u4*>(h + static_cast<size_t>(k0 + sr) * kF + in0 + fc);
    const u4 dv = *reinterpret_cast<const u4*>(dz + static_cast<size_t>(k0 + sr) * kF + out0 + fc);
    __syncthreads();
    const uint32_t he[4] = {hv.x, hv.y, hv.z, hv.w}, de[4] = {dv.x, dv.y, dv.z, dv.w};
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      ht[(fc + e) * kDwPad + sr] = static_cast<uint16_t>(he[e >> 1] >> (16 * (e & 1)));
      dt[(fc + e) * kDwPad + sr] = static_cast<uint16_t>(de[e >> 1] >> (16 * (e & 1)));
    }
    __syncthreads();
    u4 av[2], bv[2];
#pragma unroll
    for (int a = 0; a < 2; ++a) av[a] = *reinterpret_cast<const u4*>(ht + (wi + 16 * a + r) * kDwPad + 8 * g);
#pragma unroll
    for (int c = 0; c < 2; ++c) bv[c] = *reinterpret_cast<const u4*>(dt + (wo + 16 * c + r) * kDwPad + 8 * g);
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int c = 0; c < 2; ++c)
        acc[a][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(as_bf8(av[a]), as_bf8(bv[c]), acc[a][c], 0, 0, 0);
  }
  float* p = W.partial + (static_cast<size_t>(l) * gridDim.y + split) * kLayerBf16;
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int c = 0; c < 2; ++c)
#pragma unroll
      for (int i = 0; i < 4; ++i)
        p[(in0 + wi + 16 * a + 4 * g + i) * kF + out0 + wo + 16 * c + r] = acc[a][c][i];
}

// Bias gradient partials db_l[chunk][out] = sum over the chunk's samples of
// dZ_l[b][out]; grid (kDbChunks, 4), one thread per output feature.
constexpr int kDbChunks = 32;

__global__ __launch_bounds__(256) void apod_db_kernel(ApodWork W, int B) {
  const int out = threadIdx.x, chunk = blockIdx.x, l = blockIdx.y;
  const int per = (B + kDbChunks - 1) / kDbChunks;
  const int lo = chunk * per, hi = min(B, lo + per);
  const uint16_t* dz = W.dz + static_cast<size_t>(l) * B * kF + out;
  float s = 0.0f;
  for (int b = lo; b < hi; ++b) s += bf2f(dz[static_cast<size_t>(b) * kF]);
  W.db[(l * kDbChunks + chunk) * kF + out] = s;
}

// Loss of the step: the wavefront partials summed in order.
__global__ void apod_loss_kernel(const float* part, int n, float* out) {
  if (threadIdx.x != 0) return;
  double s = 0.0;
  for (int i = 0; i < n; ++i) s += part[i];
  *out = static_cast<float>(s);
}

// Gradient reduction over the K splits, then either the plain gradient
// (grads != nullptr) or an Adam step on the FP32 master parameters; the
// bf16 swizzled copies are rewritten from the master either way.
__global__ __launch_bounds__(256) void apod_update_kernel(ApodNet N, ApodWork W, int splits,
                                                          ApodAdam opt, float* grads) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= kApodParams) return;
  const int nw = kL * kLayerBf16;
  float gr;
  if (p < nw) {
    const int l = p / kLayerBf16, e = p - l * kLayerBf16;
    float s = 0.0f;
    for (int k = 0; k < splits; ++k) s += W.partial[(static_cast<size_t>(l) * splits + k) * kLayerBf16 + e];
    gr = s;
  } else {
    const int l = (p - nw) / kF, out = (p - nw) - l * kF;
    float s = 0.0f;
    for (int k = 0; k < kDbChunks; ++k) s += W.db[(l * kDbChunks + k) * kF + out];
    gr = s;
  }
  if (grads) {
    grads[p] = gr;
    return;
  }
  float w = N.master[p];
  const float m = opt.beta1 * N.m[p] + (1.0f - opt.beta1) * gr;
  const float v = opt.beta2 * N.v[p] + (1.0f - opt.beta2) * gr * gr;
  N.m[p] = m;
  N.v[p] = v;
  const float mh = m / opt.bc1, vh = v / opt.bc2;
  const float decay = p < nw ? opt.weight_decay : 0.0f;
  w -= opt.lr * (mh / (sqrtf(vh) + opt.eps) + decay * w);
  N.master[p] = w;
  if (p < nw) {
    const int l = p / kLayerBf16, e = p - l * kLayerBf16;
    const int in = e / kF, out = e - in * kF;
    const uint16_t h = static_cast<uint16_t>(f2bf(w));
    N.wf[l * kLayerBf16 + frag_pos(out, in)] = h;
    N.wb[l * kLayerBf16 + frag_pos(in, out)] = h;
  } else {
    N.bias[p - nw] = w;
  }
}

// The bf16 swizzled copies and the FP32 bias from the master parameters.
__global__ __launch_bounds__(256) void apod_pack_kernel(ApodNet N) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= kApodParams) return;
  const int nw = kL * kLayerBf16;
  const float w = N.master[p];
  if (p < nw) {
    const int l = p / kLayerBf16, e = p - l * kLayerBf16;
    const int in = e / kF, out = e - in * kF;
    const uint16_t h = static_cast<uint16_t>(f2bf(w));
    N.wf[l * kLayerBf16 + frag_pos(out, in)] = h;
    N.wb[l * kLayerBf16 + frag_pos(in, out)] = h;
  } else {
    N.bias[p - nw] = w;
  }
}

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------
size_t apod_infer_smem() { return apod_infer_smem_bytes(); }

void launch_apod_pack(const ApodNet& N, hipStream_t st) {
  hipLaunchKernelGGL(apod_pack_kernel, dim3((kApodParams + 255) / 256), dim3(256), 0, st, N);
}

void launch_apod_infer(const ApodNet& N, int n_pix, int n_ch, const float2* Y,
                       const uint32_t* mask, float2* frame, float* acc, int cf, int incoherent,
                       int n_cu, hipStream_t st) {
  const int per_block = kInferWaves * 16;
  static const hipError_t attr = hipFuncSetAttribute(
      reinterpret_cast<const void*>(apod_infer_kernel), hipFuncAttributeMaxDynamicSharedMemorySize,
      static_cast<int>(apod_infer_smem()));
  (void)attr;
  const int n_tiles = (n_pix + per_block - 1) / per_block;  // one persistent workgroup per CU
  hipLaunchKernelGGL(apod_infer_kernel, dim3(std::max(1, std::min(n_tiles, n_cu))), dim3(64 * kInferWaves),
                     apod_infer_smem(), st, N, n_pix, n_ch, Y, mask, frame, acc, cf, incoherent);
}

void launch_apod_gather(int n_list, const int* pixels, int n_ch, const float2* Y,
                        const uint32_t* mask, const float2* weights, const int* status,
                        const ApodData& D, long dst, hipStream_t st) {
  hipLaunchKernelGGL(apod_gather_kernel, dim3((n_list + 16 * kNT - 1) / (16 * kNT)), dim3(64), 0,
                     st, n_list, pixels, n_ch, Y, mask, weights, status, D, dst);
}

int apod_splits(int batch) { return batch >= 256 ? batch / 256 : 1; }

void launch_apod_step(const ApodNet& N, const ApodData& D, const ApodStep& S, const ApodWork& W,
                      const ApodAdam& opt, float* grads, float* loss_out, hipStream_t st) {
  const int B = S.batch;
  const int nwaves = B / (16 * kNT);
  hipLaunchKernelGGL(apod_train_kernel, dim3(nwaves), dim3(64), 0, st, N, D, S, W);
  const int splits = apod_splits(B);
  hipLaunchKernelGGL(apod_dw_kernel, dim3(16, splits, kL), dim3(256), 0, st, W, B, B / splits);
  hipLaunchKernelGGL(apod_db_kernel, dim3(kDbChunks, kL), dim3(kF), 0, st, W, B);
  hipLaunchKernelGGL(apod_loss_kernel, dim3(1), dim3(64), 0, st, W.loss_part, nwaves, loss_out);
  hipLaunchKernelGGL(apod_update_kernel, dim3((kApodParams + 255) / 256), dim3(256), 0, st, N, W,
                     splits, opt, grads);
}

}  // namespace usxd
