// beamform.hip -- delay, interpolate, rotate and combine on gfx950.
//
// Reference: Beamformer::beamform, core/src/beamformer.cpp:297-549.
//
// Thread mapping. A workgroup owns a tile of 64 consecutive axis1 (depth)
// pixels of one lateral line, one pixel per lane; its 4 wavefronts split the
// channel loop of those pixels (wave w takes channels lo + w, lo + w + 4, ...)
// so four times as many wavefronts hide the latency of the sample gathers.
// Tile indices are remapped so that each XCD (8 per MI355X, each with its own
// L2) receives a contiguous range of tiles (neighbouring lines reuse IQ).
//
// Per pixel, the transmit terms of each event slot (delay, illumination
// weight times transmit phasor) are computed once (split across the waves,
// shared through LDS) and held in registers; then each receive channel's
// terms (aperture weight, delay, phasor) are computed once per (pixel,
// channel) and the event slots accumulated:
//
//   Y_ch = (a_ch e^{j phi_rx}) * sum_e v_e(idx) * (tw_e e^{j phi_tx,e})
//
// which is the reference's Y[ch] += v * (e^{j phi_tx} e^{j phi_rx}) * (a*tw)
// (beamformer.cpp:470-478) with the receive factor taken out of the event
// sum. The interpolation index, whose rounding decides which samples are
// read, is evaluated in the reference's exact operation order. DAS, DMAS and
// coherence-factor sums are linear over channels, so each wave keeps partial
// sums and the workgroup adds the four in a fixed order through LDS.
#include <hip/hip_runtime.h>

#include <climits>

#include "device_math.hpp"
#include <type_traits>

#include "kernels.hpp"

namespace usxd {

namespace {

__device__ __forceinline__ int wave_min(int v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o, 64));
  return v;
}
__device__ __forceinline__ int wave_max(int v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, 64));
  return v;
}

// A tile is 4 neighbouring lateral lines x 16 consecutive depths, one pixel
// per lane (lane = 16 * line offset + depth offset). Against 64 depths of one
// line it narrows the union of the 64 pixels' receive apertures, which the
// wavefront's channel loop covers: the aperture's half width grows with depth
// (z / 2F#), and 16 depths spread it a quarter as much, while four lines
// shift it by under two elements.
constexpr int kTileLines = 4, kTileDepths = 16;

__device__ __forceinline__ int tiles_per_line_block(const BfArgs& A) {
  return (A.g.n1 + kTileDepths - 1) / kTileDepths;
}

// Pixel (line, depth) of lane r of tile t.
__device__ __forceinline__ void tile_pixel(const BfArgs& A, int t, int r, int* i0, int* j) {
  const int nj = tiles_per_line_block(A);
  const int ti = t / nj, tj = t - ti * nj;
  *i0 = ti * kTileLines + (r >> 4);
  *j = tj * kTileDepths + (r & 15);
}

// This workgroup's tile, with XCD-aware remapping of the workgroup index.
// Workgroups are dealt to the XCDs round robin (workgroup b runs on XCD
// b mod n_xcd), so tile (b mod X) * per + b / X gives each XCD, and its own
// L2, a contiguous range of line blocks.
__device__ __forceinline__ int tile_index(const BfArgs& A) {
  const int nb = gridDim.x;
  const int b = blockIdx.x;
  const int X = A.n_xcd;
  const int per = (nb + X - 1) / X;
  // Blocks beyond an uneven last XCD slice fall back to the identity map.
  return (nb % X != 0) ? b : (b % X) * per + b / X;
}

__device__ __forceinline__ void tile_of(const BfArgs& A, int* i0, int* j) {
  tile_pixel(A, tile_index(A), static_cast<int>(threadIdx.x & 63), i0, j);
}

struct TxSlot {
  float tau, tr, ti;
  int ev;
};

// Transmit terms of one event for pixel p (beamformer.cpp:425-451).
__device__ __forceinline__ TxSlot tx_slot(const BfArgs& A, int ev, float w, const Pt& p) {
  TxSlot s;
  s.ev = ev;
  s.tau = tx_time(A.tx[ev], p, A.c, A.inv_c);
  if (A.fdem > 0.0f) {
    float sn, cs;
    rot_of(A.w0 * s.tau, &cs, &sn);
    s.tr = w * cs;
    s.ti = w * sn;
  } else {
    s.tr = w;
    s.ti = 0.0f;
  }
  return s;
}

// Illumination weight of event ev for pixel p (coherent and incoherent
// compounding: the weight only scales the event's contribution).
__device__ __forceinline__ float tx_w(const BfArgs& A, int ev, const Pt& p) {
  return A.transmit_weighting ? tx_weight<false>(A.tx[ev], p, A.lambda) : 1.0f;
}

// Best-illuminating event (Select compounding, beamformer.cpp:394-408):
// first strict maximum of the illumination weight, -1 if none is lit.
__device__ __forceinline__ int select_event(const BfArgs& A, const Pt& p, float* wbest) {
  int best = -1;
  float bw = 0.0f;
  for (int e = A.ev_begin; e < A.ev_end; ++e) {
    const float w = tx_weight<true>(A.tx[e], p, A.lambda);
    if (w > bw) {
      bw = w;
      best = e;
    }
  }
  *wbest = bw;
  return best;
}

// Linear interpolation of a trace at a fractional index (beamformer.cpp:27-39);
// zero when floor(idx) < 0 or floor(idx) + 1 >= n (and for NaN). The load is
// unconditional, at an index clamped into the trace, and the result is
// selected afterwards: no branch around the load, so the loads of all event
// slots can be in flight together (the value is the branchy version's, bit
// for bit).
template <bool IQ>
__device__ __forceinline__ float2 sample(const void* data, size_t base, int n, float idx) {
  const bool ok = idx >= 0.0f && idx < static_cast<float>(n - 1);
  const int k = ok ? static_cast<int>(floorf(idx)) : 0;
  const int k1 = min(k + 1, n - 1);
  const float f = idx - static_cast<float>(k);
  const float g = 1.0f - f;
  float2 r;
  if (IQ) {
    const float2* t = static_cast<const float2*>(data) + base;
    const float2 a = t[k], b = t[k1];
    r = make_float2(fmaf(b.x, f, a.x * g), fmaf(b.y, f, a.y * g));
  } else {
    const float* t = static_cast<const float*>(data) + base;
    r = make_float2(fmaf(t[k1], f, t[k] * g), 0.0f);
  }
  return ok ? r : make_float2(0.0f, 0.0f);
}

// The same sample from a trace addressed as a buffer resource whose extent is
// exactly the trace: a load outside [0, n) returns 0 instead of touching
// memory, so an index outside the trace needs no clamping, both samples come
// in one load (16 bytes of IQ, 8 of RF), and the address is a 32-bit lane
// offset from a scalar base. The interpolation is a + f (b - a), the
// reference's a (1 - f) + b f (beamformer.cpp:27-39) to an ulp or so.
template <bool IQ>
__device__ __forceinline__ __amdgpu_buffer_rsrc_t trace_rsrc(const void* data, size_t base, int n) {
  constexpr unsigned kBytes = IQ ? 8 : 4;
  char* p = const_cast<char*>(static_cast<const char*>(data)) + base * kBytes;
  // dword 3 of a raw buffer descriptor on gfx9-family parts (32-bit data, no swizzle)
  return __builtin_amdgcn_make_buffer_rsrc(p, 0, static_cast<int>(static_cast<unsigned>(n) * kBytes),
                                           0x00020000);
}
template <bool IQ, bool CHK = true>
__device__ __forceinline__ float2 sample_rsrc(__amdgpu_buffer_rsrc_t t, float nm1, float idx,
                                              uint32_t soff = 0) {
  typedef unsigned int u2 __attribute__((ext_vector_type(2)));
  typedef unsigned int u4 __attribute__((ext_vector_type(4)));
  // Outside [0, n - 1) (and for NaN) the index moves to 2^24: its byte
  // offset lies past every trace, so both samples read 0, and its fraction
  // is 0, so the result is the reference's 0 without a select on the value.
  // !CHK: the caller has shown that idx lies inside (gathers_inside).
  const bool ok = !CHK || (idx >= 0.0f && idx < nm1);
  const float id = ok ? idx : 16777216.0f;
  const int k = static_cast<int>(id);             // == floor(id): id >= 0
  const float f = __builtin_amdgcn_fractf(id);    // id - floor(id), exact
  if (IQ) {
    const u4 ab = __builtin_amdgcn_raw_buffer_load_b128(t, k * 8, soff, 0);
    const float ax = __uint_as_float(ab.x), ay = __uint_as_float(ab.y);
    return make_float2(fmaf(f, __uint_as_float(ab.z) - ax, ax), fmaf(f, __uint_as_float(ab.w) - ay, ay));
  } else {
    const u2 ab = __builtin_amdgcn_raw_buffer_load_b64(t, k * 4, soff, 0);
    const float a = __uint_as_float(ab.x);
    return make_float2(fmaf(f, __uint_as_float(ab.y) - a, a), 0.0f);
  }
}

// A pixel's receive aperture: d = max(half aperture, 1e-9) and r = 1/d, once
// per pixel (dynamic_aperture_weight, geometry.cpp:123-130); none for
// pixels at or above the face.
struct RxAperture {
  float d, r;
  bool any;
};
__device__ __forceinline__ RxAperture rx_aperture(const BfArgs& A, const Pt& p, bool valid) {
  RxAperture ap;
  const float depth = p.z - 0.0f;
  ap.any = valid && depth > 0.0f;
  const float half = 0.5f * depth / smax(A.f_number, 1e-3f);
  ap.d = smax(half, 1e-9f);
  ap.r = 1.0f / ap.d;
  return ap;
}
// Receive weight of channel ch (*on: the reference's weight is non-zero).
template <bool HANN>
__device__ __forceinline__ float rx_apod(const BfArgs& A, const RxAperture& ap, float ex, const Pt& p,
                                        bool* on) {
  if (!ap.any) {
    *on = false;
    return 0.0f;
  }
  return rx_weight<HANN>(A.rx_window, A.tukey_alpha, ex - p.x, ap.d, ap.r, A.cos1_lo, A.cos1_hi, on);
}

// Conservative channel range of the receive aperture of p: |x_e - x_p| <= z/(2 F#).
__device__ __forceinline__ void aperture_range(const BfArgs& A, const Pt& p, bool valid, int* lo,
                                               int* hi) {
  int l = INT_MAX, h = -1;
  if (valid && p.z > 0.0f) {
    const float half = 0.5f * p.z / smax(A.f_number, 1e-3f);
    const float c0 = 0.5f * static_cast<float>(A.n_ch - 1);
    const float fl = (p.x - half) / A.pitch + c0;
    const float fh = (p.x + half) / A.pitch + c0;
    l = fl < -1.0f ? 0 : static_cast<int>(floorf(fl)) - 1;
    h = fh > static_cast<float>(A.n_ch) ? A.n_ch - 1 : static_cast<int>(ceilf(fh)) + 1;
    l = max(l, 0);
    h = min(h, A.n_ch - 1);
    if (!(half == half)) {  // NaN geometry: scan everything, exact tests decide
      l = 0;
      h = A.n_ch - 1;
    }
  }
  // wavefront-uniform (the reductions leave every lane the same value): keep
  // them in scalar registers so the channel loop and its addresses are scalar
  *lo = __builtin_amdgcn_readfirstlane(wave_min(l));
  *hi = __builtin_amdgcn_readfirstlane(wave_max(h));
}

// Element x of the channels a wavefront visits, ch = first + 4 i (i < 128,
// i.e. up to 512 channels), one per lane in two registers: entry i (i
// wavefront-uniform) is a readlane, with no memory access and no wait in
// the channel loop. (For the channel-vector kernel, whose stores of Y keep
// the compiler from scalarising a load of A.ex: a vector load and a
// vmcnt(0) wait per channel.)
struct WaveEx {
  float r0, r1;
  // lane i: channel first + 32 (i / 8) + w + 4 (i % 8) (and i + 64): the
  // wavefront's channels (every fourth, from w) of each 32-channel block from
  // first
  __device__ __forceinline__ void load(const BfArgs& A, int first, int w, int lane) {
    const int c0 = first + 32 * (lane >> 3) + w + 4 * (lane & 7), c1 = c0 + 32 * 8;
    r0 = c0 < A.n_ch ? A.ex[c0] : 0.0f;
    r1 = c1 < A.n_ch ? A.ex[c1] : 0.0f;
    // wait here (vmcnt(0)), once, so that the channel loop's readlanes of
    // r0 / r1 carry no wait for these loads (which would also wait for the
    // loop's own gathers and stores)
    __builtin_amdgcn_s_waitcnt(0x0F70);
  }
  // (both readlanes, then a scalar select: a VALU select of r0 / r1 would be
  // formed only in the active lanes, and lane i may be inactive here)
  __device__ __forceinline__ float at(int i) const {
    const int v0 = __builtin_amdgcn_readlane(__float_as_int(r0), i & 63);
    const int v1 = __builtin_amdgcn_readlane(__float_as_int(r1), i & 63);
    return __int_as_float(i < 64 ? v0 : v1);
  }
};

}  // namespace

// Transmit slots of the tile's pixels, computed by the four waves together
// and returned in registers (slot s of this lane's pixel in sl[s]).
template <int NE>
__device__ __forceinline__ int tile_slots(const BfArgs& A, const Pt& p, bool valid, int ev_first,
                                          int n_slots, int select, TxSlot (&sl)[NE],
                                          TxSlot (*shared)[64]) {
  // the wavefront index, uniform: kept scalar so channel indices and trace
  // addresses derived from it are scalar too
  const int lane = threadIdx.x & 63, w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  int nsl = n_slots;
  if (select) {
    nsl = 1;
    if (w == 0) {
      float wb;
      const int e = valid ? select_event(A, p, &wb) : -1;
      shared[0][lane] = tx_slot(A, e < 0 ? A.ev_begin : e, (e >= 0 && wb > 0.0f) ? 1.0f : 0.0f, p);
    }
  } else {
    for (int s = w; s < n_slots; s += 4)
      shared[s][lane] = tx_slot(A, ev_first + s, valid ? tx_w(A, ev_first + s, p) : 0.0f, p);
  }
  __syncthreads();
#pragma unroll
  for (int s = 0; s < NE; ++s)
    if (s < nsl) sl[s] = shared[s][lane];
  return nsl;
}

// Whether every gather of this wavefront -- its pixels, their live slots and
// the channel range lo..hi -- reads strictly inside the traces, by a margin of
// two samples: then its channel loop needs no bounds test per gather. The
// receive time of a pixel lies between depth / c (an element on the face is
// at least that far) and the time to the farther end of lo..hi (the elements
// lie along x, evenly spaced, so distance grows towards the ends). Every
// pixel must also be deeper than 1e-12 m, which keeps the squared distances
// normal for rx_time_normal. Wavefront-uniform.
template <int NE>
__device__ __forceinline__ bool gathers_inside(const BfArgs& A, const Pt& p, bool valid,
                                               const TxSlot (&sl)[NE], int nsl, int lo, int hi) {
  bool ok = true;
  if (valid && p.z > 0.0f && lo <= hi) {
    float tmin = sl[0].tau, tmax = sl[0].tau;
#pragma unroll
    for (int s = 1; s < NE; ++s)
      if (s < nsl) {
        tmin = fminf(tmin, sl[s].tau);
        tmax = fmaxf(tmax, sl[s].tau);
      }
    const float rmin = p.z * A.inv_c;
    const float rmax = fmaxf(rx_time(A.ex[lo], p, A.c, A.inv_c), rx_time(A.ex[hi], p, A.c, A.inv_c));
    const float ilo = (tmin + rmin - A.t0) * A.fs;
    const float ihi = (tmax + rmax - A.t0) * A.fs;
    ok = ilo >= 2.0f && ihi <= static_cast<float>(A.n_s - 1) - 2.0f && p.z >= 1e-12f;
  }
  return __builtin_amdgcn_readfirstlane(wave_min(ok ? 1 : 0)) != 0;
}

// Sum over the event slots of one channel, then the receive rotation and
// apodization: this lane's Y_ch (see the header). `ch` is wavefront-uniform,
// and so is the event of slot s unless SEL (Select compounding: the slot's
// own event): each slot's trace then starts at a scalar address and only the
// sample index is per lane.
template <int NE, bool IQ, bool SEL, bool CHK = true, bool T0Z = false>
__device__ __forceinline__ float2 channel_value(const BfArgs& A, const void* data, const Pt& p,
                                                const TxSlot (&sl)[NE], int nsl, int ch, float ex,
                                                float a, int ev_first) {
  // (nsl <= NE slots are live; !CHK: gathers_inside holds for this wavefront;
  // T0Z: t0 == 0, where (tau + trx) - t0 is tau + trx exactly)
  const float trx = CHK ? rx_time(ex, p, A.c, A.inv_c) : rx_time_normal(ex, p, A.c, A.inv_c);
  const float nm1 = static_cast<float>(A.n_s - 1);
  float2 acc = make_float2(0.0f, 0.0f);
  // NE = 1 and NE = 11 are launched only with exactly that many slots
  // (launch_fused / launch_channels), so every slot is live at compile time
  if (NE == 1 || NE == 11) nsl = NE;
  // slots in groups of G: one uniform branch per group, and inside it all
  // the group's loads issued together (a slot past nsl reads slot 0's trace
  // and is not added; slots past NE are not formed at all)
  constexpr int G = (NE == 11 || NE == 12) ? 6 : (NE < 4 ? NE : 4);
  // Consecutive events whose gathers all lie inside the traces (!SEL, !CHK):
  // one buffer resource for the channel, based at its trace in event ev_first
  // and spanning the launch's events, with slot s's trace reached by the
  // scalar offset s (n_ch n_s) samples (the host keeps 16 events' traces
  // within 4 GiB). With bounds tests each trace keeps its own resource, whose
  // range test returns 0 past the trace.
  const uint32_t ev_bytes = static_cast<uint32_t>(A.n_ch) * A.n_s * (IQ ? 8u : 4u);
  const __amdgpu_buffer_rsrc_t chr = trace_rsrc<IQ>(
      data, (static_cast<size_t>(ev_first) * A.n_ch + ch) * A.n_s, (NE - 1) * A.n_ch * A.n_s + A.n_s);
#pragma unroll
  for (int s0 = 0; s0 < NE; s0 += G) {
    if (s0 < nsl) {
      float2 v[G];
#pragma unroll
      for (int q = 0; q < G; ++q) {
        const int s = s0 + q;
        if (s >= NE) break;
        const bool on = s < nsl;
        const float idx = T0Z ? (sl[on ? s : 0].tau + trx) * A.fs : (sl[on ? s : 0].tau + trx - A.t0) * A.fs;
        const int ev = SEL ? sl[on ? s : 0].ev : ev_first + (on ? s : 0);
        const size_t base = (static_cast<size_t>(ev) * A.n_ch + ch) * A.n_s;
        if (SEL) {
          v[q] = sample<IQ>(data, base, A.n_s, idx);  // per-lane event: per-lane trace
        } else if (!CHK) {
          v[q] = sample_rsrc<IQ, CHK>(chr, nm1, idx, static_cast<uint32_t>(on ? s : 0) * ev_bytes);
        } else {
          v[q] = sample_rsrc<IQ, CHK>(trace_rsrc<IQ>(data, base, A.n_s), nm1, idx);
        }
      }
#pragma unroll
      for (int q = 0; q < G; ++q) {
        const int s = s0 + q;
        if (s < NE && s < nsl) {
          acc.x = fmaf(v[q].x, sl[s].tr, fmaf(-v[q].y, sl[s].ti, acc.x));
          acc.y = fmaf(v[q].x, sl[s].ti, fmaf(v[q].y, sl[s].tr, acc.y));
        }
      }
    }
  }
  float rr = a, ri = 0.0f;
  if (A.fdem > 0.0f) {
    float sn, cs;
    rot_of(A.w0 * trx, &cs, &sn);
    rr = a * cs;
    ri = a * sn;
  }
  return make_float2(acc.x * rr - acc.y * ri, acc.x * ri + acc.y * rr);
}

// ---------------------------------------------------------------------------
// Fused kernel: DAS / DMAS (+ coherence factor) straight to the frame.
// NE = event slots per launch (compile-time, <= 16).
//   select    : one slot per pixel, the pixel's best event, weight 1
//   accumulate: add to the existing frame (DAS over event chunks; linear)
// Launch bounds as bf_channels_kernel's (measured 0.181 -> 0.157 ms at
// 256x640, 128 channels, 11 transmits).
// ---------------------------------------------------------------------------
// HANN: the receive window is Hann (rx_weight<true>).
template <int NE, int COMB, bool CF, bool IQ, bool HANN>
__global__ __launch_bounds__(256, NE <= 12 ? 4 : 1) void bf_fused_kernel(BfArgs A, const void* data, float2* frame,
                                                       int ev_first, int n_slots, int select,
                                                       int accumulate) {
  constexpr int NPART = 10;  // das(2) s(2) q(2) coh(2) inc na
  __shared__ TxSlot shared_slots[NE][64];
  __shared__ float part[4][NPART][64];
  int i0, j;
  tile_of(A, &i0, &j);
  // the wavefront index, uniform: kept scalar so channel indices and trace
  // addresses derived from it are scalar too
  const int lane = threadIdx.x & 63, w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const bool valid = i0 < A.g.n0 && j < A.g.n1;
  const Pt p = valid ? grid_point(A.g, i0, j) : Pt{0.0f, 0.0f, 1.0f};
  TxSlot sl[NE];
  // Select compounding launches one slot per pixel: only NE = 1 kernels carry it
  const bool sel = NE == 1 && select;
  const int nsl = tile_slots<NE>(A, p, valid, ev_first, n_slots, sel, sl, shared_slots);

  int lo, hi;
  aperture_range(A, p, valid, &lo, &hi);

  float2 das = make_float2(0.0f, 0.0f);
  float2 ds = das, dq = das, coh = das;
  float inc = 0.0f;
  int na = 0;
  const RxAperture ap = rx_aperture(A, p, valid);
  // the channel loop twice: without bounds tests when every gather of the
  // wavefront is inside the traces, with them otherwise
  auto channels = [&](auto checked, auto t0_zero) {
    constexpr bool CHK = decltype(checked)::value;
    constexpr bool T0Z = decltype(t0_zero)::value;
  for (int ch = lo + w; ch <= hi; ch += 4) {
    bool on;
    // (a scalar load: this kernel stores nothing before its channel loop)
    const float ex = A.ex[ch];
    const float a = rx_apod<HANN>(A, ap, ex, p, &on);
    if (!on) continue;
    ++na;
    const float2 y = sel ? channel_value<NE, IQ, true, true>(A, data, p, sl, nsl, ch, ex, a, ev_first)
                         : channel_value<NE, IQ, false, CHK, T0Z>(A, data, p, sl, nsl, ch, ex, a, ev_first);
    if (COMB == 0) {
      das.x += y.x;
      das.y += y.y;
    } else {
      const float m = hypotf(y.x, y.y);
      if (m > 0.0f) {
        const float sq = sqrtf(m);
        const float vr = y.x / sq, vi = y.y / sq;
        ds.x += vr;
        ds.y += vi;
        dq.x = fmaf(vr, vr, fmaf(-vi, vi, dq.x));
        dq.y = fmaf(vr, vi, fmaf(vi, vr, dq.y));
      }
    }
    if (CF) {
      coh.x += y.x;
      coh.y += y.y;
      const float m = hypotf(y.x, y.y);
      inc = fmaf(m, m, inc);
    }
  }
  };
  if (!sel && gathers_inside<NE>(A, p, valid, sl, nsl, lo, hi)) {
    if (A.t0 == 0.0f)
      channels(std::false_type{}, std::true_type{});
    else
      channels(std::false_type{}, std::false_type{});
  } else {
    channels(std::true_type{}, std::false_type{});
  }
  const float mine[NPART] = {das.x, das.y, ds.x, ds.y, dq.x, dq.y, coh.x, coh.y, inc,
                             static_cast<float>(na)};
#pragma unroll
  for (int q = 0; q < NPART; ++q) part[w][q][lane] = mine[q];
  __syncthreads();
  if (w != 0 || !valid) return;
  float tot[NPART];
#pragma unroll
  for (int q = 0; q < NPART; ++q)
    tot[q] = ((part[0][q][lane] + part[1][q][lane]) + part[2][q][lane]) + part[3][q][lane];
  const int n_act = static_cast<int>(tot[9]);
  float2 out = make_float2(0.0f, 0.0f);
  if (n_act > 0) {
    if (COMB == 0) {
      out = make_float2(tot[0], tot[1]);
    } else {
      // ((s*s - q) * 0.5) / (0.5 * max(1, n - 1))  (beamformer.cpp:125-129)
      const float sr = tot[2], si = tot[3];
      const float ar = (sr * sr - si * si - tot[4]) * 0.5f;
      const float ai = (sr * si + si * sr - tot[5]) * 0.5f;
      const float nrm = 0.5f * smax(1.0f, static_cast<float>(n_act - 1));
      out = make_float2(ar / nrm, ai / nrm);
    }
    if (CF) {
      const float mc = hypotf(tot[6], tot[7]);
      const float f = (tot[8] > 0.0f) ? (mc * mc) / (static_cast<float>(n_act) * tot[8]) : 0.0f;
      out.x *= f;
      out.y *= f;
    }
  }
  float2* o = frame + static_cast<size_t>(i0) * A.g.n1 + j;
  if (accumulate) {
    const float2 prev = *o;
    out.x += prev.x;
    out.y += prev.y;
  }
  *o = out;
}

// ---------------------------------------------------------------------------
// Channel-vector kernel: Y[pixel][ch] over event chunks, plus the active
// mask (one bit per channel, a > 0). rect: MV's rectangular aperture
// (beamformer.cpp:385). Only channels inside the tile's aperture union are
// written, whole 32-channel blocks of it, an inactive channel of a written
// block as exactly 0 (the mask marks the active ones; consumers read no
// entry of an unwritten block, and the learned-apodization kernels sum the
// channel pairs of written blocks unmasked). Y is accumulated across chunk
// launches; the first writes.
// Values are staged per 32-channel block in LDS so each pixel's row segment
// is written as one contiguous 256-byte run. Up to 12 slots, launch bounds
// ask for 4 workgroups (16 wavefronts) per CU, i.e. at most 128 VGPRs: the
// gathers are latency-bound, and the extra resident wavefront per SIMD is
// worth the few registers spilled (measured: 0.228 -> 0.204 ms at 256x640,
// 128 channels, 11 transmits); 16 slots would spill too much.
// ---------------------------------------------------------------------------
constexpr int kMaskWordsMax = 16;  // up to 512 channels

template <int NE, bool IQ, bool HANN>
__global__ __launch_bounds__(256, NE <= 12 ? 4 : 1) void bf_channels_kernel(BfArgs A, const void* data, float2* Y,
                                                          uint32_t* mask, int ev_first,
                                                          int n_slots, int select, int first,
                                                          int rect) {
  __shared__ TxSlot shared_slots[NE][64];
  __shared__ float2 blk[64][33];  // one 32-channel block of the 64 pixels (+1: banks)
  __shared__ uint32_t bits[64][kMaskWordsMax];
  int i0, j;
  tile_of(A, &i0, &j);
  const int tid = threadIdx.x, lane = tid & 63, w = __builtin_amdgcn_readfirstlane(tid >> 6);
  const bool valid = i0 < A.g.n0 && j < A.g.n1;
  const Pt p = valid ? grid_point(A.g, i0, j) : Pt{0.0f, 0.0f, 1.0f};
  const int nw = (A.n_ch + 31) / 32;
  for (int q = tid; q < 64 * kMaskWordsMax; q += 256) bits[q / kMaskWordsMax][q % kMaskWordsMax] = 0;
  TxSlot sl[NE];
  // Select compounding launches one slot per pixel: only NE = 1 kernels carry it
  const bool sel = NE == 1 && select;
  const int nsl = tile_slots<NE>(A, p, valid, ev_first, n_slots, sel, sl, shared_slots);

  int lo, hi;
  aperture_range(A, p, valid, &lo, &hi);
  const RxAperture ap = rx_aperture(A, p, valid);
  const int tix = tile_index(A);
  const size_t pix = static_cast<size_t>(i0) * A.g.n1 + j;  // this lane's pixel (when valid)
  // the channel loop twice: without bounds tests when every gather of the
  // wavefront is inside the traces, with them otherwise (gathers_inside is
  // wavefront-uniform and the four wavefronts share pixels and channel range,
  // so all take the same copy, and all meet at the same barriers). Each
  // 32-channel block: wavefront w forms channels c0 + w + 4 k (k < 8) of the
  // 64 pixels -- every fourth channel, so a block at the aperture's edge
  // leaves the wavefronts equal shares of its active channels -- into the
  // block's LDS rows; after a barrier each wavefront writes 16 pixels' rows
  // out (2 pixels x 256 contiguous bytes per store).
  WaveEx exs{0.0f, 0.0f};
  const int cb = lo & ~31;
  if (lo <= hi) exs.load(A, cb, w, lane);
  auto channels = [&](auto checked, auto t0_zero) {
    constexpr bool CHK = decltype(checked)::value;
    constexpr bool T0Z = decltype(t0_zero)::value;
    for (int c0 = cb; c0 <= hi; c0 += 32) {
      uint32_t mybits = 0u;
      const int ib = (c0 - cb) >> 2;  // the block's first index in exs
#pragma unroll 1
      for (int k = 0; k < 8; ++k) {
        const int ch = c0 + w + 4 * k;
        float2 y = make_float2(0.0f, 0.0f);
        if (valid && ch >= lo && ch <= hi) {
          const float ex = exs.at(ib + k);
          bool on;
          float a = rx_apod<HANN>(A, ap, ex, p, &on);
          if (rect) {  // beamformer.cpp:385: a > 0 ? 1 : 0, so a NaN weight is inactive here
            on = on && a == a;
            a = 1.0f;
          }
          if (on) {
            mybits |= 1u << (w + 4 * k);
            y = sel ? channel_value<NE, IQ, true, true>(A, data, p, sl, nsl, ch, ex, a, ev_first)
                    : channel_value<NE, IQ, false, CHK, T0Z>(A, data, p, sl, nsl, ch, ex, a, ev_first);
          }
        }
        blk[lane][w + 4 * k] = y;
      }
      if (mybits) atomicOr(&bits[lane][c0 >> 5], mybits);
      __syncthreads();
      // write the block: wavefront w's pixels 16 w .. 16 w + 15, two per store
      // (lanes 0-31 and 32-63 the 32 channels of one pixel each), read from
      // LDS before the barrier that lets the next block's writes in
      const int kk = lane & 31;
      float2 v[8];
#pragma unroll
      for (int q = 0; q < 8; ++q) v[q] = blk[16 * w + 2 * q + (lane >> 5)][kk];
      __syncthreads();
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        const int r = 16 * w + 2 * q + (lane >> 5);
        int ir, jr;
        tile_pixel(A, tix, r, &ir, &jr);
        if (c0 + kk < A.n_ch && jr < A.g.n1 && ir < A.g.n0) {
          float2* dst = Y + (static_cast<size_t>(ir) * A.g.n1 + jr) * A.n_ch + c0 + kk;
          float2 o = v[q];
          if (!first) {
            const float2 prev = *dst;
            o.x += prev.x;
            o.y += prev.y;
          }
          *dst = o;
        }
      }
    }
  };
  if (lo <= hi) {
    if (!sel && gathers_inside<NE>(A, p, valid, sl, nsl, lo, hi)) {
      if (A.t0 == 0.0f)
        channels(std::false_type{}, std::true_type{});
      else
        channels(std::false_type{}, std::false_type{});
    } else {
      channels(std::true_type{}, std::false_type{});
    }
  }
  __syncthreads();
  if (valid && w == 0) {
    uint32_t* mp = mask + pix * nw;
    for (int q = 0; q < nw; ++q) mp[q] = bits[lane][q];
  }
}

// ---------------------------------------------------------------------------
// Combine kernel from channel vectors: DAS / DMAS (+ CF). incoherent: add
// |out|^2 into acc instead of writing the frame (beamformer.cpp:536).
// ---------------------------------------------------------------------------
template <int COMB, bool CF>
__global__ __launch_bounds__(256) void bf_combine_kernel(int n_pix, int n_ch, const float2* Y,
                                                         const uint32_t* mask, float2* frame,
                                                         float* acc_out, int incoherent) {
  const int pix = blockIdx.x * blockDim.x + threadIdx.x;
  if (pix >= n_pix) return;
  const int nw = (n_ch + 31) / 32;
  const float2* yp = Y + static_cast<size_t>(pix) * n_ch;
  const uint32_t* mp = mask + static_cast<size_t>(pix) * nw;
  float2 das = make_float2(0.0f, 0.0f), ds = das, dq = das, coh = das;
  float inc = 0.0f;
  int na = 0;
  for (int w = 0; w < nw; ++w) {
    uint32_t bits = mp[w];
    while (bits) {
      const int b = __ffs(bits) - 1;
      bits &= bits - 1;
      const float2 y = yp[w * 32 + b];
      ++na;
      if (COMB == 0) {
        das.x += y.x;
        das.y += y.y;
      } else {
        const float m = hypotf(y.x, y.y);
        if (m > 0.0f) {
          const float sq = sqrtf(m);
          const float vr = y.x / sq, vi = y.y / sq;
          ds.x += vr;
          ds.y += vi;
          dq.x += vr * vr - vi * vi;
          dq.y += vr * vi + vi * vr;
        }
      }
      if (CF) {
        coh.x += y.x;
        coh.y += y.y;
        const float m = hypotf(y.x, y.y);
        inc += m * m;
      }
    }
  }
  if (na == 0) {
    if (!incoherent) frame[pix] = make_float2(0.0f, 0.0f);
    return;
  }
  float2 out;
  if (COMB == 0) {
    out = das;
  } else {
    const float ar = (ds.x * ds.x - ds.y * ds.y - dq.x) * 0.5f;
    const float ai = (ds.x * ds.y + ds.y * ds.x - dq.y) * 0.5f;
    const float nrm = 0.5f * smax(1.0f, static_cast<float>(na - 1));
    out = make_float2(ar / nrm, ai / nrm);
  }
  if (CF) {
    const float mc = hypotf(coh.x, coh.y);
    const float f = (inc > 0.0f) ? (mc * mc) / (static_cast<float>(na) * inc) : 0.0f;
    out.x *= f;
    out.y *= f;
  }
  if (incoherent) {
    const float m = hypotf(out.x, out.y);
    acc_out[pix] += m * m;
  } else {
    frame[pix] = out;
  }
}

// Incoherent compounding: frame = (sqrt(acc), 0) (beamformer.cpp:541-543).
__global__ void incoherent_finish_kernel(int n_pix, const float* acc, float2* frame) {
  const int pix = blockIdx.x * blockDim.x + threadIdx.x;
  if (pix < n_pix) frame[pix] = make_float2(sqrtf(acc[pix]), 0.0f);
}

__global__ void fill_zero_kernel(float* p, size_t n) {
  const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0.0f;
}

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------

static int tile_blocks(const BfArgs& A) {
  return ((A.g.n0 + kTileLines - 1) / kTileLines) * ((A.g.n1 + kTileDepths - 1) / kTileDepths);
}

template <int NE, bool IQ>
static void launch_fused_ne(const BfArgs& A, const void* data, float2* frame, int ev_first,
                            int n_slots, int select, int accumulate, int comb, bool cf,
                            hipStream_t st) {
  const dim3 grid(tile_blocks(A)), block(256);
#define USX_FUSED(C_, CF_)                                                                        \
  do {                                                                                            \
    if (A.rx_window == 1)                                                                         \
      hipLaunchKernelGGL((bf_fused_kernel<NE, C_, CF_, IQ, true>), grid, block, 0, st, A, data,   \
                         frame, ev_first, n_slots, select, accumulate);                           \
    else                                                                                          \
      hipLaunchKernelGGL((bf_fused_kernel<NE, C_, CF_, IQ, false>), grid, block, 0, st, A, data,  \
                         frame, ev_first, n_slots, select, accumulate);                           \
  } while (0)
  if (comb == 0 && !cf) USX_FUSED(0, false);
  else if (comb == 0 && cf) USX_FUSED(0, true);
  else if (!cf) USX_FUSED(1, false);
  else USX_FUSED(1, true);
#undef USX_FUSED
}

template <bool IQ>
static void launch_fused_iq(const BfArgs& A, const void* data, float2* frame, int ev_first,
                            int n_slots, int select, int accumulate, int comb, bool cf,
                            hipStream_t st) {
  if (n_slots <= 1)
    launch_fused_ne<1, IQ>(A, data, frame, ev_first, n_slots, select, accumulate, comb, cf, st);
  else if (n_slots <= 4)
    launch_fused_ne<4, IQ>(A, data, frame, ev_first, n_slots, select, accumulate, comb, cf, st);
  else if (n_slots <= 8)
    launch_fused_ne<8, IQ>(A, data, frame, ev_first, n_slots, select, accumulate, comb, cf, st);
  else if (n_slots == 11)  // 11 transmits: no idle 12th slot in the second load group
    launch_fused_ne<11, IQ>(A, data, frame, ev_first, n_slots, select, accumulate, comb, cf, st);
  else if (n_slots <= 12)
    launch_fused_ne<12, IQ>(A, data, frame, ev_first, n_slots, select, accumulate, comb, cf, st);
  else
    launch_fused_ne<16, IQ>(A, data, frame, ev_first, n_slots, select, accumulate, comb, cf, st);
}

void launch_fused(const BfArgs& A, const void* data, float2* frame, int ev_first, int n_slots,
                  int select, int accumulate, int comb, bool cf, bool iq, hipStream_t st) {
  if (iq)
    launch_fused_iq<true>(A, data, frame, ev_first, n_slots, select, accumulate, comb, cf, st);
  else
    launch_fused_iq<false>(A, data, frame, ev_first, n_slots, select, accumulate, comb, cf, st);
}

template <int NE, bool IQ>
static void launch_channels_ne(const BfArgs& A, const void* data, float2* Y, uint32_t* mask,
                               int ev_first, int n_slots, int select, int first, int rect,
                               hipStream_t st) {
  if (A.rx_window == 1)
    hipLaunchKernelGGL((bf_channels_kernel<NE, IQ, true>), dim3(tile_blocks(A)), dim3(256), 0, st, A,
                       data, Y, mask, ev_first, n_slots, select, first, rect);
  else
    hipLaunchKernelGGL((bf_channels_kernel<NE, IQ, false>), dim3(tile_blocks(A)), dim3(256), 0, st, A,
                       data, Y, mask, ev_first, n_slots, select, first, rect);
}

template <bool IQ>
static void launch_channels_iq(const BfArgs& A, const void* data, float2* Y, uint32_t* mask,
                               int ev_first, int n_slots, int select, int first, int rect,
                               hipStream_t st) {
  if (n_slots <= 1)
    launch_channels_ne<1, IQ>(A, data, Y, mask, ev_first, n_slots, select, first, rect, st);
  else if (n_slots <= 4)
    launch_channels_ne<4, IQ>(A, data, Y, mask, ev_first, n_slots, select, first, rect, st);
  else if (n_slots <= 8)
    launch_channels_ne<8, IQ>(A, data, Y, mask, ev_first, n_slots, select, first, rect, st);
  else if (n_slots == 11)
    launch_channels_ne<11, IQ>(A, data, Y, mask, ev_first, n_slots, select, first, rect, st);
  else if (n_slots <= 12)
    launch_channels_ne<12, IQ>(A, data, Y, mask, ev_first, n_slots, select, first, rect, st);
  else
    launch_channels_ne<16, IQ>(A, data, Y, mask, ev_first, n_slots, select, first, rect, st);
}

void launch_channels(const BfArgs& A, const void* data, float2* Y, uint32_t* mask, int ev_first,
                     int n_slots, int select, int first, int rect, bool iq, hipStream_t st) {
  if (iq)
    launch_channels_iq<true>(A, data, Y, mask, ev_first, n_slots, select, first, rect, st);
  else
    launch_channels_iq<false>(A, data, Y, mask, ev_first, n_slots, select, first, rect, st);
}

void launch_combine(int n_pix, int n_ch, const float2* Y, const uint32_t* mask, float2* frame,
                    float* acc, int incoherent, int comb, bool cf, hipStream_t st) {
  const dim3 grid((n_pix + 255) / 256), block(256);
#define USX_COMB(C_, CF_) \
  hipLaunchKernelGGL((bf_combine_kernel<C_, CF_>), grid, block, 0, st, n_pix, n_ch, Y, mask, \
                     frame, acc, incoherent)
  if (comb == 0 && !cf) USX_COMB(0, false);
  else if (comb == 0 && cf) USX_COMB(0, true);
  else if (!cf) USX_COMB(1, false);
  else USX_COMB(1, true);
#undef USX_COMB
}

void launch_incoherent_finish(int n_pix, const float* acc, float2* frame, hipStream_t st) {
  hipLaunchKernelGGL(incoherent_finish_kernel, dim3((n_pix + 255) / 256), dim3(256), 0, st,
                     n_pix, acc, frame);
}

void launch_fill_zero(float* p, size_t n, hipStream_t st) {
  if (n == 0) return;
  hipLaunchKernelGGL(fill_zero_kernel, dim3(static_cast<unsigned>((n + 255) / 256)), dim3(256),
                     0, st, p, n);
}

}  // namespace usxd
