// context.hip -- the C ABI (include/usx_gpu.h): contexts, pooled device
// buffers, descriptor validation, launch plans and stage timing.
//
// Launch plan of usx_beamform (reference: beamformer.cpp:297-549):
//   DAS, coherent            fused kernel, events in chunks of <= 16 slots,
//                            later chunks accumulate (DAS is linear)
//   DAS/DMAS, Select         fused kernel, one slot (each pixel's best event)
//   DMAS, coherent <= 16 ev  fused kernel
//   DMAS, coherent > 16 ev   channel vectors accumulated over chunks, then
//                            the combine kernel
//   any, Incoherent          per event: channel vectors, then combine or MV
//                            with |out|^2 accumulation; sqrt at the end
//   MV                       channel vectors with the rectangular aperture,
//                            then the MV kernel (mv.hip)
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

#include "kernels.hpp"
#include "usx_gpu.h"

using usxd::BfArgs;

namespace {

struct DevBuf {
  void* p = nullptr;
  size_t n = 0;
};

// Host-side copy of the last uploaded descriptor table, so unchanged tables
// (the steady state of a pipeline) are not re-uploaded every frame.
struct Cached {
  DevBuf dev;
  std::vector<unsigned char> host;
};

}  // namespace

struct usx_ctx {
  int device = -1;
  int n_cu = 256;   // compute units (hipDeviceAttributeMultiprocessorCount)
  int n_xcd = 8;    // accelerator complex dies, each with its own L2 (hipDeviceAttributeNumberOfXccs)
  void* scratch = nullptr;        // this device's allocation-free scratch (see bind_scratch)
  hipError_t scratch_err = hipSuccess;
  hipStream_t stream = nullptr;
  hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
  hipEvent_t user_ev[8] = {};
  std::string err;
  DevBuf in, out, iq, Y, mask, acc, mv_counts, mv_lists;
  Cached taps, rot, ex, tx, ax0, ax1, s0, c0;
  // profiling: pending (name, start, stop) event triples and the totals
  bool profiling = false;
  std::vector<hipEvent_t> ev_pool;
  std::vector<std::pair<std::string, std::pair<hipEvent_t, hipEvent_t>>> pending;
  std::map<std::string, std::pair<long, double>> totals;
  // learned apodization: parameters, optimizer state, training set, scratch
  bool apod_loaded = false;
  int apod_t = 0;
  DevBuf ap_master, ap_m, ap_v, ap_wf, ap_wb;
  DevBuf ds_x, ds_g, ds_mask, ds_sw;
  size_t ds_n = 0, ds_cap = 0;
  DevBuf tr_h, tr_z, tr_dz, tr_partial, tr_db, tr_loss_part, tr_losses, tr_grads;
  DevBuf mv_w, mv_status, ap_pixels;
  // simulator
  Cached s_ex, s_delays, s_firing, s_fw, s_nf, s_pulse, s_noise;
  DevBuf s_scat, s_live, s_t0, s_n, s_pos, s_wave, s_partial, s_bmax, s_rf;
  // display tail
  DevBuf disp_env, disp_bmax, disp_img, disp_frame;
  // Doppler and speckle tracking
  Cached dop_reject;
  DevBuf dop_in, dop_in2, dop_out, dop_lag, dop_f0, dop_f1, dop_f2, dop_valid, dop_pmax;
};

// ---- scratch of the allocation-free entry point -----------------------------
// One fixed array per GPU (each device that loads this code object holds its
// own copy until the process ends). In host code `scratch` is only a handle
// for hipGetSymbolAddress, never a device address: every use goes through the
// per-device pointer bind_scratch returns.
static __device__ __align__(256) unsigned char scratch[USX_NOALLOC_SCRATCH_BYTES];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// The device address of `scratch` on `device`, which must be current:
// hipGetSymbolAddress once per device, the result kept for every later context.
hipError_t bind_scratch(int device, void** out) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  *out = nullptr;
  if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
  if (!g_scratch_ptr[device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch_ptr[device] = p;
  }
  *out = g_scratch_ptr[device];
  return hipSuccess;
}
}  // namespace

namespace {
// Set once the process starts exiting; later destroys skip HIP calls, which
// may run after the runtime has been torn down.
bool g_exiting = false;
void mark_exiting() { g_exiting = true; }
}  // namespace

namespace {

struct Fail {
  usx_status st;
};

[[noreturn]] void fail(usx_ctx* c, usx_status st, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  c->err = buf;
  throw Fail{st};
}

void hip_check(usx_ctx* c, hipError_t e, const char* what) {
  if (e != hipSuccess) fail(c, USX_ERR_RUNTIME, "%s: %s", what, hipGetErrorString(e));
}

void* ensure(usx_ctx* c, DevBuf& b, size_t bytes) {
  if (bytes == 0) bytes = 16;
  if (b.n < bytes) {
    if (b.p) hip_check(c, hipFree(b.p), "hipFree");
    b.p = nullptr;
    b.n = 0;
    hip_check(c, hipMalloc(&b.p, bytes), "hipMalloc");
    b.n = bytes;
  }
  return b.p;
}

void* upload(usx_ctx* c, Cached& t, const void* src, size_t bytes) {
  if (t.dev.n < (bytes == 0 ? 16 : bytes)) t.host.clear();  // reallocation loses the contents
  void* d = ensure(c, t.dev, bytes);
  if (bytes == 0) return d;
  if (t.host.size() != bytes || std::memcmp(t.host.data(), src, bytes) != 0) {
    t.host.assign(static_cast<const unsigned char*>(src),
                  static_cast<const unsigned char*>(src) + bytes);
    hip_check(c, hipMemcpyAsync(d, t.host.data(), bytes, hipMemcpyHostToDevice, c->stream),
              "hipMemcpyAsync(descriptor)");
  }
  return d;
}

template <typename F>
usx_status guarded(usx_ctx* c, F&& f) {
  if (!c) return USX_ERR_INVALID;
  try {
    hip_check(c, hipSetDevice(c->device), "hipSetDevice");
    f();
    return USX_OK;
  } catch (const Fail& e) {
    return e.st;
  } catch (const std::exception& e) {
    c->err = e.what();
    return USX_ERR_RUNTIME;
  }
}

void check_demod(usx_ctx* c, const usx_demod_desc* d) {
  if (!d) fail(c, USX_ERR_INVALID, "usx_demodulate: null descriptor");
  if (d->n_traces <= 0 || d->n_in <= 0) fail(c, USX_ERR_INVALID, "demodulate: input has no RF samples");
  if (d->decimation < 1) fail(c, USX_ERR_INVALID, "demodulate: decimation must be >= 1");
  if (d->n_out != (d->n_in + d->decimation - 1) / d->decimation)
    fail(c, USX_ERR_INVALID, "demodulate: n_out must be ceil(n_in / decimation)");
  if (d->n_taps < 1 || !d->taps || !d->rotator)
    fail(c, USX_ERR_INVALID, "demodulate: taps and rotator tables are required");
}

// Subarray length as minimum variance uses it: clamp(L, 2, ...) lifts
// anything below 2 (beamformer.cpp:147).
int mv_sub(const usx_bf_desc* b) { return std::max(2, b->mv_subarray); }

void check_bf(usx_ctx* c, const usx_bf_desc* b) {
  if (!b) fail(c, USX_ERR_INVALID, "usx_beamform: null descriptor");
  if (b->n_events <= 0) fail(c, USX_ERR_INVALID, "ChannelData has no transmit events");
  if (b->n_channels <= 0) fail(c, USX_ERR_INVALID, "Probe.n_elements must be > 0");
  if (b->n_channels > 512) fail(c, USX_ERR_INVALID, "beamform: at most 512 channels are supported");
  if (b->n_samples <= 0) fail(c, USX_ERR_INVALID, "Acquisition.n_samples must be > 0");
  // the beamforming kernels reach the traces of up to 16 consecutive events
  // through one buffer resource and 32-bit offsets
  if (16ull * static_cast<unsigned long long>(b->n_channels) * b->n_samples * (b->is_iq ? 8u : 4u) >
      0xFFFFFFFFull)
    fail(c, USX_ERR_INVALID, "beamform: at most 256 MiB of channel data per transmit event");
  if (b->event_begin < 0 || b->event_end < b->event_begin || b->event_end > b->n_events)
    fail(c, USX_ERR_INVALID, "beamform: event range [%d, %d) outside [0, %d]", b->event_begin,
         b->event_end, b->n_events);
  if (b->n_axis0 <= 0 || b->n_axis1 <= 0) fail(c, USX_ERR_INVALID, "ScanGrid dimensions must be > 0");
  if (!b->element_x || !b->tx || !b->axis0 || !b->axis1)
    fail(c, USX_ERR_INVALID, "beamform: element, transmit and grid tables are required");
  if (b->grid_type == 1 && (!b->axis0_sin || !b->axis0_cos))
    fail(c, USX_ERR_INVALID, "beamform: sector grids need the sin/cos tables");
  if (b->grid_type != 0 && b->grid_type != 1) fail(c, USX_ERR_INVALID, "beamform: unknown grid type");
  // BeamformerConfig is validated when a Beamformer is constructed
  // (beamformer.hpp:110), not per call: a config assigned afterwards
  // (pymodule.cpp:430-431, 617-618) runs with the reference's clamps instead,
  // f_number through max(f, 1e-3) (geometry.cpp:127), the subarray through
  // clamp(L, 2, ...) (beamformer.cpp:147) and the loading through
  // max(delta tr / L, 1e-20) (beamformer.cpp:187). A negative range average
  // has no clamp there (its scratch vector of 2 r + 1 entries cannot be
  // built, beamformer.cpp:357), so it stays an error for minimum variance.
  if (b->combiner == 2 && b->mv_range_average < 0)
    fail(c, USX_ERR_INVALID, "BeamformerConfig.mv_range_average must be >= 0");
  if (b->rx_window < 0 || b->rx_window > 4) fail(c, USX_ERR_INVALID, "unknown window");
  if (b->combiner < 0 || b->combiner > 3) fail(c, USX_ERR_INVALID, "unknown combiner");
  if (b->combiner == 3 && b->n_channels != USX_APOD_CHANNELS)
    fail(c, USX_ERR_INVALID, "learned apodization: the network takes %d channels, got %d",
         USX_APOD_CHANNELS, b->n_channels);
  // the inference kernel addresses the channel vectors (1 KiB a pixel) with
  // 32-bit offsets, its last 128-pixel tile included
  if (b->combiner == 3 && static_cast<long long>(b->n_axis0) * b->n_axis1 > usxd::kApodMaxPixels)
    fail(c, USX_ERR_INVALID, "learned apodization: at most %lld pixels per frame, got %lld",
         usxd::kApodMaxPixels, static_cast<long long>(b->n_axis0) * b->n_axis1);
  if (b->compounding < 0 || b->compounding > 2) fail(c, USX_ERR_INVALID, "unknown compounding");
  if (b->is_iq && !(b->demodulation_frequency > 0))
    fail(c, USX_ERR_INVALID, "IQ ChannelData needs a positive demodulation_frequency");
  if (b->mv_precision < 0 || b->mv_precision > 2) fail(c, USX_ERR_INVALID, "unknown mv_precision");
  if (b->tile_slots < 0 || b->tile_slots > 16 || b->tile_mv_occupancy < 0 || b->tile_mv_occupancy > 32)
    fail(c, USX_ERR_INVALID, "beamform: tile_slots must be in 0..16 and tile_mv_occupancy in 0..32");
  if (b->combiner == 2) {
    const int pmax = usxd::mv_max_bucket_p(b->n_channels, mv_sub(b));
    if (pmax > usxd::kMvMaxP)
      fail(c, USX_ERR_INVALID,
           "MV: subarray length up to %d with %d channels exceeds the supported %d",
           std::min(mv_sub(b), b->n_channels / 2), b->n_channels, usxd::kMvMaxP);
    const size_t lds =
        usxd::mv_smem_bytes(b->n_channels, b->mv_range_average, mv_sub(b), pmax, b->mv_refine != 0);
    if (lds > 160 * 1024)
      fail(c, USX_ERR_INVALID,
           "MV: %d channels with range average %d need %zu B of LDS per pixel (limit 163840)",
           b->n_channels, b->mv_range_average, lds);
  }
}

hipEvent_t pool_event(usx_ctx* c) {
  if (!c->ev_pool.empty()) {
    hipEvent_t e = c->ev_pool.back();
    c->ev_pool.pop_back();
    return e;
  }
  hipEvent_t e;
  hip_check(c, hipEventCreate(&e), "hipEventCreate");
  return e;
}

// Enqueue `fn` and, when profiling, bracket it with events named `name`.
template <typename F>
void timed(usx_ctx* c, const char* name, F&& fn) {
  if (!c->profiling) {
    fn();
    return;
  }
  hipEvent_t a = pool_event(c), b = pool_event(c);
  hip_check(c, hipEventRecord(a, c->stream), "hipEventRecord");
  fn();
  hip_check(c, hipEventRecord(b, c->stream), "hipEventRecord");
  c->pending.push_back({name, {a, b}});
}

void drain_profile(usx_ctx* c) {
  for (auto& p : c->pending) {
    float ms = 0;
    hip_check(c, hipEventSynchronize(p.second.second), "hipEventSynchronize");
    hip_check(c, hipEventElapsedTime(&ms, p.second.first, p.second.second), "hipEventElapsedTime");
    auto& t = c->totals[p.first];
    t.first += 1;
    t.second += ms;
    c->ev_pool.push_back(p.second.first);
    c->ev_pool.push_back(p.second.second);
  }
  c->pending.clear();
}

// Device copies of a beamform descriptor's host tables.
struct BfTables {
  const float* ex = nullptr;
  const usx_tx_desc* tx = nullptr;
  const float* ax0 = nullptr;
  const float* ax1 = nullptr;
  const float* s0 = nullptr;
  const float* c0 = nullptr;
};

// Where the C library's cosf returns exactly 1.0f on [0, 2 pi]: x <= lo and
// x >= hi. The reference's Hann weight 0.5 (1 - cos(2 pi t)) (types.cpp:133)
// is exactly 0 there, so these bounds decide its receive aperture. Found by
// walking the floats from 2^-12 (where a correctly rounded cosine leaves 1.0f)
// and down from (2 pi)f, once per process, with the same libm the host path
// uses; volatile keeps the compiler from evaluating cosf itself.
struct CosOne {
  float lo, hi;
};
CosOne cos_one_bounds() {
  static const CosOne bounds = [] {
    auto cosf_lib = [](float x) {
      volatile float v = x;
      return std::cos(static_cast<float>(v));
    };
    CosOne b;
    float x = std::ldexp(1.0f, -12);
    while (cosf_lib(std::nextafter(x, 1.0f)) == 1.0f) x = std::nextafter(x, 1.0f);
    while (x > 0.0f && cosf_lib(x) != 1.0f) x = std::nextafter(x, 0.0f);
    b.lo = x;
    const float two_pi = 2.0f * usxd::kPi;
    float y = two_pi;
    if (cosf_lib(y) != 1.0f) {
      b.hi = INFINITY;  // no float near 2 pi gives 1.0f
    } else {
      while (cosf_lib(std::nextafter(y, 0.0f)) == 1.0f) y = std::nextafter(y, 0.0f);
      b.hi = y;
    }
    return b;
  }();
  return bounds;
}

BfArgs args_of(const usx_ctx* c, const usx_bf_desc* b, const BfTables& t) {
  BfArgs A{};
  A.n_ch = b->n_channels;
  A.n_s = b->n_samples;
  A.n_ev = b->n_events;
  A.ev_begin = b->event_begin;
  A.ev_end = b->event_end;
  A.fs = b->sampling_frequency;
  A.t0 = b->t0;
  A.c = b->speed_of_sound;
  A.fdem = b->is_iq ? b->demodulation_frequency : 0.0f;
  // (2.0f * kPi) * fdem, the reference's left-to-right product (beamformer.cpp:390, 446).
  A.w0 = (2.0f * usxd::kPi) * A.fdem;
  A.lambda = b->wavelength;
  A.pitch = b->pitch;
  A.f_number = b->f_number;
  A.rx_window = b->rx_window;
  A.tukey_alpha = b->tukey_alpha;
  A.transmit_weighting = b->transmit_weighting;
  A.n_xcd = c->n_xcd;
  A.inv_c = 1.0f / b->speed_of_sound;
  const CosOne one = cos_one_bounds();
  A.cos1_lo = one.lo;
  A.cos1_hi = one.hi;
  A.ex = t.ex;
  A.tx = t.tx;
  A.g.type = b->grid_type;
  A.g.n0 = b->n_axis0;
  A.g.n1 = b->n_axis1;
  A.g.axis0 = t.ax0;
  A.g.axis1 = t.ax1;
  A.g.s0 = b->grid_type == 1 ? t.s0 : t.ax0;
  A.g.c0 = b->grid_type == 1 ? t.c0 : t.ax0;
  A.g.apex_x = b->apex_x;
  A.g.apex_z = b->apex_z;
  return A;
}

BfArgs make_args(usx_ctx* c, const usx_bf_desc* b) {
  BfTables t;
  t.ex = static_cast<const float*>(upload(c, c->ex, b->element_x, sizeof(float) * b->n_channels));
  t.tx = static_cast<const usx_tx_desc*>(upload(c, c->tx, b->tx, sizeof(usx_tx_desc) * b->n_events));
  t.ax0 = static_cast<const float*>(upload(c, c->ax0, b->axis0, sizeof(float) * b->n_axis0));
  t.ax1 = static_cast<const float*>(upload(c, c->ax1, b->axis1, sizeof(float) * b->n_axis1));
  if (b->grid_type == 1) {
    t.s0 = static_cast<const float*>(upload(c, c->s0, b->axis0_sin, sizeof(float) * b->n_axis0));
    t.c0 = static_cast<const float*>(upload(c, c->c0, b->axis0_cos, sizeof(float) * b->n_axis0));
  }
  return args_of(c, b, t);
}

// Launch-shape tunables of one beamforming call (what the autotuner picks):
// event slots per beamforming launch (1..16) and persistent MV workgroups per
// CU (0: as many as LDS allows). From the descriptor's tile fields; 0 = default.
struct Tile {
  int slots = 16;
  int mv_occupancy = 0;
};

Tile tile_of(const usx_bf_desc* b) {
  Tile t;
  if (b->tile_slots > 0) t.slots = std::min(16, b->tile_slots);
  t.mv_occupancy = std::max(0, b->tile_mv_occupancy);
  return t;
}

usxd::ApodNet apod_net(usx_ctx* c) {
  usxd::ApodNet N{};
  N.master = static_cast<float*>(c->ap_master.p);
  N.m = static_cast<float*>(c->ap_m.p);
  N.v = static_cast<float*>(c->ap_v.p);
  N.wf = static_cast<uint16_t*>(c->ap_wf.p);
  N.wb = static_cast<uint16_t*>(c->ap_wb.p);
  N.bias = N.master ? N.master + usxd::kApodLayers * usxd::kApodFeatures * usxd::kApodFeatures : nullptr;
  return N;
}

// Device scratch of one beamforming launch sequence.
struct Scratch {
  float2* Y = nullptr;
  uint32_t* mask = nullptr;
  float* acc = nullptr;
  int* counts = nullptr;
  int* lists = nullptr;
};

// Bytes of each scratch buffer the launch sequence of b needs (0: unused).
struct ScratchSize {
  size_t y = 0, mask = 0, acc = 0, counts = 0, lists = 0;
};

ScratchSize scratch_size(const usx_bf_desc* b, const Tile& T) {
  ScratchSize z;
  const size_t n_pix = static_cast<size_t>(b->n_axis0) * b->n_axis1;
  const int n_ev = b->event_end - b->event_begin;
  if (n_ev == 0) return z;
  const int comb = b->combiner;
  const bool select = b->compounding == 2, incoh = b->compounding == 1, cf = b->coherence_factor != 0;
  const bool need_y = comb >= 2 || incoh || (!select && !(comb == 0 && !cf) && n_ev > T.slots);
  if (need_y) {
    z.y = n_pix * b->n_channels * sizeof(float2);
    z.mask = n_pix * ((b->n_channels + 31) / 32) * sizeof(uint32_t);
  }
  if (incoh) z.acc = n_pix * sizeof(float);
  if (comb == 2) {
    z.counts = 16 * sizeof(int);
    z.lists = static_cast<size_t>(usxd::mv_max_bucket_p(b->n_channels, mv_sub(b)) / 16) * n_pix * sizeof(int);
  }
  return z;
}

Scratch context_scratch(usx_ctx* c, const ScratchSize& z) {
  Scratch S;
  if (z.y) S.Y = static_cast<float2*>(ensure(c, c->Y, z.y));
  if (z.mask) S.mask = static_cast<uint32_t*>(ensure(c, c->mask, z.mask));
  if (z.acc) S.acc = static_cast<float*>(ensure(c, c->acc, z.acc));
  if (z.counts) S.counts = static_cast<int*>(ensure(c, c->mv_counts, z.counts));
  if (z.lists) S.lists = static_cast<int*>(ensure(c, c->mv_lists, z.lists));
  return S;
}

// Enqueue the beamforming of device-resident data into a device frame:
// kernel launches only (no allocation, copy or memset), so the sequence can
// be captured into a HIP graph. d_weights / d_status (MV only, may be null):
// the effective channel weights and per-pixel status of the MV solve.
void enqueue_beamform(usx_ctx* c, const usx_bf_desc* b, const BfArgs& A, const Scratch& S,
                      const Tile& T, const void* d_data, float2* d_frame, float2* d_weights,
                      int* d_status) {
  const bool iq = b->is_iq != 0;
  const int n_pix = b->n_axis0 * b->n_axis1;
  const int n_ev = b->event_end - b->event_begin;
  const bool cf = b->coherence_factor != 0;
  const int comb = b->combiner;
  const bool select = b->compounding == 2;
  const bool incoh = b->compounding == 1;
  const int slots = T.slots;
  hipStream_t st = c->stream;

  if (n_ev == 0) {  // nothing to beamform: the reference leaves the frame zero
    timed(c, "fill_zero", [&] { usxd::launch_fill_zero(reinterpret_cast<float*>(d_frame), 2 * static_cast<size_t>(n_pix), st); });
    return;
  }
  const size_t ny = static_cast<size_t>(n_pix) * b->n_channels;
  float2* Y = S.Y;
  uint32_t* mk = S.mask;

  usxd::MvArgs M{};
  M.n0 = b->n_axis0;
  M.n1 = b->n_axis1;
  M.n_ch = b->n_channels;
  M.subarray = mv_sub(b);
  M.loading = b->mv_diagonal_loading;
  M.range_avg = b->mv_range_average;
  M.cf = cf;
  M.precision = b->mv_precision;
  M.refine = b->mv_refine;
  M.occupancy = T.mv_occupancy;
  M.n_cu = c->n_cu;
  M.n_xcd = c->n_xcd;
  if (comb == 3 && !c->apod_loaded)
    fail(c, USX_ERR_INVALID, "learned apodization: no network loaded (usx_apod_set_params)");
  auto mv = [&](float* acc) {
    if (comb == 3) {
      timed(c, "apod_infer", [&] {
        usxd::launch_apod_infer(apod_net(c), n_pix, b->n_channels, Y, mk, d_frame, acc, cf,
                                acc != nullptr, c->n_cu, st);
      });
      return;
    }
    M.weights = d_weights;
    M.status = d_status;
    if (d_weights)
      timed(c, "fill_zero", [&] { usxd::launch_fill_zero(reinterpret_cast<float*>(d_weights), 2 * ny, st); });
    timed(c, "mv", [&] { usxd::launch_mv(M, Y, mk, d_frame, acc, S.counts, S.lists, st); });
  };

  if (incoh) {
    timed(c, "fill_zero", [&] { usxd::launch_fill_zero(S.acc, n_pix, st); });
    for (int e = b->event_begin; e < b->event_end; ++e) {
      timed(c, "bf_channels", [&] { usxd::launch_channels(A, d_data, Y, mk, e, 1, 0, 1, comb >= 2, iq, st); });
      if (comb >= 2) {
        M.incoherent = 1;
        mv(S.acc);
      } else {
        timed(c, "bf_combine", [&] { usxd::launch_combine(n_pix, b->n_channels, Y, mk, d_frame, S.acc, 1, comb, cf, st); });
      }
    }
    timed(c, "incoherent_finish", [&] { usxd::launch_incoherent_finish(n_pix, S.acc, d_frame, st); });
    return;
  }

  if (comb >= 2) {
    if (select) {
      timed(c, "bf_channels", [&] { usxd::launch_channels(A, d_data, Y, mk, b->event_begin, 1, 1, 1, 1, iq, st); });
    } else {
      for (int e0 = b->event_begin; e0 < b->event_end; e0 += slots) {
        const int n = std::min(slots, b->event_end - e0);
        timed(c, "bf_channels", [&] { usxd::launch_channels(A, d_data, Y, mk, e0, n, 0, e0 == b->event_begin, 1, iq, st); });
      }
    }
    M.incoherent = 0;
    mv(nullptr);
    return;
  }

  if (select) {
    timed(c, "bf_fused", [&] { usxd::launch_fused(A, d_data, d_frame, b->event_begin, 1, 1, 0, comb, cf, iq, st); });
    return;
  }
  if (comb == 0 && !cf) {
    for (int e0 = b->event_begin; e0 < b->event_end; e0 += slots) {
      const int n = std::min(slots, b->event_end - e0);
      timed(c, "bf_fused", [&] { usxd::launch_fused(A, d_data, d_frame, e0, n, 0, e0 != b->event_begin, 0, false, iq, st); });
    }
    return;
  }
  if (n_ev <= slots) {
    timed(c, "bf_fused", [&] { usxd::launch_fused(A, d_data, d_frame, b->event_begin, n_ev, 0, 0, comb, cf, iq, st); });
    return;
  }
  for (int e0 = b->event_begin; e0 < b->event_end; e0 += slots) {
    const int n = std::min(slots, b->event_end - e0);
    timed(c, "bf_channels", [&] { usxd::launch_channels(A, d_data, Y, mk, e0, n, 0, e0 == b->event_begin, 0, iq, st); });
  }
  timed(c, "bf_combine", [&] { usxd::launch_combine(n_pix, b->n_channels, Y, mk, d_frame, nullptr, 0, comb, cf, st); });
}

// Upload, size the context's scratch, enqueue.
void run_beamform(usx_ctx* c, const usx_bf_desc* b, const void* d_data, float2* d_frame,
                  float2* d_weights = nullptr, int* d_status = nullptr) {
  const BfArgs A = make_args(c, b);
  const Tile T = tile_of(b);
  const Scratch S = context_scratch(c, scratch_size(b, T));
  enqueue_beamform(c, b, A, S, T, d_data, d_frame, d_weights, d_status);
}

void enqueue_demod(usx_ctx* c, const usx_demod_desc* d, const float* taps, const float2* rot,
                   const float* d_rf, float2* d_iq) {
  timed(c, "demod", [&] {
    usxd::launch_demod(d->n_traces, d->n_in, d->n_out, d->decimation, d->n_taps, taps, rot, d_rf,
                       d_iq, c->stream);
  });
}

void run_demod(usx_ctx* c, const usx_demod_desc* d, const float* d_rf, float2* d_iq) {
  const float* taps = static_cast<const float*>(upload(c, c->taps, d->taps, sizeof(float) * d->n_taps));
  const float2* rot =
      static_cast<const float2*>(upload(c, c->rot, d->rotator, sizeof(usx_c32) * d->n_in));
  enqueue_demod(c, d, taps, rot, d_rf, d_iq);
}

void check_launch(usx_ctx* c, const char* what) { hip_check(c, hipGetLastError(), what); }



// ---- learned apodization helpers -------------------------------------------

void apod_alloc(usx_ctx* c) {
  const size_t np = usxd::kApodParams;
  const size_t nw = static_cast<size_t>(usxd::kApodLayers) * usxd::kApodFeatures * usxd::kApodFeatures;
  ensure(c, c->ap_master, np * sizeof(float));
  ensure(c, c->ap_m, np * sizeof(float));
  ensure(c, c->ap_v, np * sizeof(float));
  ensure(c, c->ap_wf, nw * sizeof(uint16_t));
  ensure(c, c->ap_wb, nw * sizeof(uint16_t));
}

// Training scratch for batches of up to `batch` rows.
usxd::ApodWork apod_work(usx_ctx* c, int batch) {
  const size_t B = static_cast<size_t>(batch), F = usxd::kApodFeatures;
  const int splits = usxd::apod_splits(batch);
  usxd::ApodWork W{};
  W.h = static_cast<uint16_t*>(ensure(c, c->tr_h, 4 * B * F * sizeof(uint16_t)));
  W.z = static_cast<float*>(ensure(c, c->tr_z, 3 * B * F * sizeof(float)));
  W.dz = static_cast<uint16_t*>(ensure(c, c->tr_dz, 4 * B * F * sizeof(uint16_t)));
  W.partial = static_cast<float*>(ensure(c, c->tr_partial, 4 * static_cast<size_t>(splits) * F * F * sizeof(float)));
  W.db = static_cast<float*>(ensure(c, c->tr_db, 4 * 32 * F * sizeof(float)));
  W.loss_part = static_cast<float*>(ensure(c, c->tr_loss_part, (B / 16 + 1) * sizeof(float)));
  return W;
}

usxd::ApodData apod_data(usx_ctx* c) {
  usxd::ApodData D{};
  D.x = static_cast<uint16_t*>(c->ds_x.p);
  D.g = static_cast<float*>(c->ds_g.p);
  D.mask = static_cast<uint32_t*>(c->ds_mask.p);
  D.sw = static_cast<float*>(c->ds_sw.p);
  return D;
}

void apod_reserve(usx_ctx* c, size_t cap) {
  const size_t F = usxd::kApodFeatures;
  ensure(c, c->ds_x, cap * F * sizeof(uint16_t));
  ensure(c, c->ds_g, cap * F * sizeof(float));
  ensure(c, c->ds_mask, cap * 4 * sizeof(uint32_t));
  ensure(c, c->ds_sw, cap * sizeof(float));
  c->ds_cap = cap;
  c->ds_n = 0;
}

size_t data_bytes(const usx_bf_desc* b) {
  return static_cast<size_t>(b->n_events) * b->n_channels * b->n_samples *
         (b->is_iq ? sizeof(usx_c32) : sizeof(float));
}

}  // namespace

extern "C" {

usx_status usx_ctx_create(int device, usx_ctx** out) {
  if (!out) return USX_ERR_INVALID;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return USX_ERR_NO_DEVICE;
  if (device < 0) device = 0;
  if (device >= n) return USX_ERR_NO_DEVICE;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) != hipSuccess) return USX_ERR_NO_DEVICE;
  if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) return USX_ERR_NO_DEVICE;
  static bool registered = (std::atexit(mark_exiting), true);
  (void)registered;
  usx_ctx* c = new usx_ctx();
  c->device = device;
  int v = 0;
  if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess && v > 0)
    c->n_cu = v;
  if (hipDeviceGetAttribute(&v, hipDeviceAttributeNumberOfXccs, device) == hipSuccess && v > 0)
    c->n_xcd = v;
  if (hipSetDevice(device) != hipSuccess || hipStreamCreate(&c->stream) != hipSuccess) {
    delete c;
    return USX_ERR_RUNTIME;
  }
  c->scratch_err = bind_scratch(device, &c->scratch);
  for (auto& e : c->ev) {
    if (hipEventCreate(&e) != hipSuccess) {
      usx_ctx_destroy(c);
      return USX_ERR_RUNTIME;
    }
  }
  *out = c;
  return USX_OK;
}

usx_status usx_ctx_destroy(usx_ctx* c) {
  if (!c) return USX_OK;
  if (g_exiting) {  // the runtime may already be gone; the OS reclaims everything
    delete c;
    return USX_OK;
  }
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  for (DevBuf* b : {&c->in, &c->out, &c->iq, &c->Y, &c->mask, &c->acc, &c->mv_counts, &c->mv_lists,
                    &c->ap_master, &c->ap_m, &c->ap_v, &c->ap_wf, &c->ap_wb, &c->ds_x, &c->ds_g,
                    &c->ds_mask, &c->ds_sw, &c->tr_h, &c->tr_z, &c->tr_dz, &c->tr_partial,
                    &c->tr_db, &c->tr_loss_part, &c->tr_losses, &c->tr_grads, &c->mv_w,
                    &c->mv_status, &c->ap_pixels, &c->s_scat, &c->s_live, &c->s_t0, &c->s_n,
                    &c->s_pos, &c->s_wave, &c->s_partial, &c->s_bmax, &c->s_rf, &c->disp_env,
                    &c->disp_bmax, &c->disp_img, &c->disp_frame, &c->dop_in, &c->dop_in2,
                    &c->dop_out, &c->dop_lag, &c->dop_f0, &c->dop_f1, &c->dop_f2, &c->dop_valid,
                    &c->dop_pmax})
    if (b->p) (void)hipFree(b->p);
  for (Cached* t : {&c->taps, &c->rot, &c->ex, &c->tx, &c->ax0, &c->ax1, &c->s0, &c->c0, &c->s_ex,
                    &c->s_delays, &c->s_firing, &c->s_fw, &c->s_nf, &c->s_pulse, &c->s_noise,
                    &c->dop_reject})
    if (t->dev.p) (void)hipFree(t->dev.p);
  for (auto& e : c->ev)
    if (e) (void)hipEventDestroy(e);
  for (auto& e : c->user_ev)
    if (e) (void)hipEventDestroy(e);
  for (auto& p : c->pending) {
    (void)hipEventDestroy(p.second.first);
    (void)hipEventDestroy(p.second.second);
  }
  for (auto& e : c->ev_pool) (void)hipEventDestroy(e);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return USX_OK;
}

const char* usx_ctx_last_error(const usx_ctx* c) { return c ? c->err.c_str() : "null context"; }

void* usx_ctx_stream(usx_ctx* c) { return c ? static_cast<void*>(c->stream) : nullptr; }

int usx_ctx_device(const usx_ctx* c) { return c ? c->device : -1; }

usx_status usx_ctx_synchronize(usx_ctx* c) {
  return guarded(c, [&] { hip_check(c, hipStreamSynchronize(c->stream), "hipStreamSynchronize"); });
}

usx_status usx_ctx_event_record(usx_ctx* c, int slot) {
  return guarded(c, [&] {
    if (slot < 0 || slot >= 8) fail(c, USX_ERR_INVALID, "event slot %d outside 0..7", slot);
    if (!c->user_ev[slot]) hip_check(c, hipEventCreate(&c->user_ev[slot]), "hipEventCreate");
    hip_check(c, hipEventRecord(c->user_ev[slot], c->stream), "hipEventRecord");
  });
}

usx_status usx_ctx_event_elapsed(usx_ctx* c, int a, int b, float* ms) {
  return guarded(c, [&] {
    if (a < 0 || a >= 8 || b < 0 || b >= 8 || !c->user_ev[a] || !c->user_ev[b] || !ms)
      fail(c, USX_ERR_INVALID, "event slots %d/%d not recorded", a, b);
    hip_check(c, hipEventSynchronize(c->user_ev[b]), "hipEventSynchronize");
    hip_check(c, hipEventElapsedTime(ms, c->user_ev[a], c->user_ev[b]), "hipEventElapsedTime");
  });
}

usx_status usx_ctx_profile(usx_ctx* c, int enable) {
  return guarded(c, [&] {
    drain_profile(c);
    c->totals.clear();
    c->profiling = enable != 0;
  });
}

usx_status usx_ctx_profile_report(usx_ctx* c, char* buf, size_t len) {
  return guarded(c, [&] {
    if (!buf || len == 0) fail(c, USX_ERR_INVALID, "profile report needs a buffer");
    drain_profile(c);
    std::string s = "{";
    bool first = true;
    for (const auto& t : c->totals) {
      char item[256];
      snprintf(item, sizeof(item), "%s\"%s\": [%ld, %.6f]", first ? "" : ", ", t.first.c_str(),
               t.second.first, t.second.second);
      s += item;
      first = false;
    }
    s += "}";
    if (s.size() + 1 > len) fail(c, USX_ERR_INVALID, "profile report buffer too small");
    std::memcpy(buf, s.c_str(), s.size() + 1);
  });
}

const char* usx_build_info(void) {
  return "usx_hip gfx950 -O3 -ffp-contract=off correctly-rounded fp32 div/sqrt";
}

usx_status usx_demodulate(usx_ctx* c, const usx_demod_desc* d, const float* rf, usx_c32* iq,
                          int memory) {
  return guarded(c, [&] {
    check_demod(c, d);
    if (!rf || !iq) fail(c, USX_ERR_INVALID, "usx_demodulate: null data pointer");
    const size_t nin = static_cast<size_t>(d->n_traces) * d->n_in;
    const size_t nout = static_cast<size_t>(d->n_traces) * d->n_out;
    if (memory == USX_DEVICE) {
      run_demod(c, d, rf, reinterpret_cast<float2*>(iq));
      check_launch(c, "demodulate launch");
      return;
    }
    float* d_rf = static_cast<float*>(ensure(c, c->in, nin * sizeof(float)));
    float2* d_iq = static_cast<float2*>(ensure(c, c->iq, nout * sizeof(float2)));
    hip_check(c, hipMemcpyAsync(d_rf, rf, nin * sizeof(float), hipMemcpyHostToDevice, c->stream),
              "hipMemcpyAsync(rf)");
    run_demod(c, d, d_rf, d_iq);
    check_launch(c, "demodulate launch");
    hip_check(c, hipMemcpyAsync(iq, d_iq, nout * sizeof(float2), hipMemcpyDeviceToHost, c->stream),
              "hipMemcpyAsync(iq)");
    hip_check(c, hipStreamSynchronize(c->stream), "hipStreamSynchronize");
  });
}

usx_status usx_beamform(usx_ctx* c, const usx_bf_desc* b, const void* data, usx_c32* frame,
                        int memory) {
  return guarded(c, [&] {
    check_bf(c, b);
    if (!data || !frame) fail(c, USX_ERR_INVALID, "usx_beamform: null data pointer");
    const size_t npix = static_cast<size_t>(b->n_axis0) * b->n_axis1;
    if (memory == USX_DEVICE) {
      run_beamform(c, b, data, reinterpret_cast<float2*>(frame));
      check_launch(c, "beamform launch");
      return;
    }
    void* d_in = ensure(c, c->in, data_bytes(b));
    float2* d_out = static_cast<float2*>(ensure(c, c->out, npix * sizeof(float2)));
    hip_check(c, hipMemcpyAsync(d_in, data, data_bytes(b), hipMemcpyHostToDevice, c->stream),
              "hipMemcpyAsync(data)");
    run_beamform(c, b, d_in, d_out);
    check_launch(c, "beamform launch");
    hip_check(c, hipMemcpyAsync(frame, d_out, npix * sizeof(float2), hipMemcpyDeviceToHost,
                                c->stream),
              "hipMemcpyAsync(frame)");
    hip_check(c, hipStreamSynchronize(c->stream), "hipStreamSynchronize");
  });
}

usx_status usx_process_frame(usx_ctx* c, const usx_demod_desc* d, const usx_bf_desc* b,
                             const float* rf, usx_c32* frame, int memory, usx_stage_times* t) {
  return guarded(c, [&] {
    check_demod(c, d);
    check_bf(c, b);
    if (!rf || !frame) fail(c, USX_ERR_INVALID, "usx_process_frame: null data pointer");
    if (!b->is_iq) fail(c, USX_ERR_INVALID, "usx_process_frame: beamform descriptor must describe IQ");
    if (b->n_events * b->n_channels != d->n_traces || b->n_samples != d->n_out)
      fail(c, USX_ERR_INVALID, "usx_process_frame: demodulate and beamform shapes disagree");
    const size_t nin = static_cast<size_t>(d->n_traces) * d->n_in;
    const size_t nout = static_cast<size_t>(d->n_traces) * d->n_out;
    const size_t npix = static_cast<size_t>(b->n_axis0) * b->n_axis1;
    const auto w0 = std::chrono::steady_clock::now();
    // Stage events only when the caller asks for stage times: each record is
    // a marker packet between the kernels, a few microseconds of idle GPU.
    if (t) hip_check(c, hipEventRecord(c->ev[0], c->stream), "hipEventRecord");
    const float* d_rf = rf;
    float2* d_out = reinterpret_cast<float2*>(frame);
    if (memory != USX_DEVICE) {
      float* p = static_cast<float*>(ensure(c, c->in, nin * sizeof(float)));
      hip_check(c, hipMemcpyAsync(p, rf, nin * sizeof(float), hipMemcpyHostToDevice, c->stream),
                "hipMemcpyAsync(rf)");
      d_rf = p;
      d_out = static_cast<float2*>(ensure(c, c->out, npix * sizeof(float2)));
    }
    float2* d_iq = static_cast<float2*>(ensure(c, c->iq, nout * sizeof(float2)));
    run_demod(c, d, d_rf, d_iq);
    check_launch(c, "demodulate launch");
    if (t) hip_check(c, hipEventRecord(c->ev[1], c->stream), "hipEventRecord");
    run_beamform(c, b, d_iq, d_out);
    check_launch(c, "beamform launch");
    if (memory != USX_DEVICE)
      hip_check(c, hipMemcpyAsync(frame, d_out, npix * sizeof(float2), hipMemcpyDeviceToHost,
                                  c->stream),
                "hipMemcpyAsync(frame)");
    if (t) hip_check(c, hipEventRecord(c->ev[2], c->stream), "hipEventRecord");
    if (memory != USX_DEVICE || t) {
      hip_check(c, hipStreamSynchronize(c->stream), "hipStreamSynchronize");
    }
    if (t) {
      float a = 0, bb = 0;
      hip_check(c, hipEventElapsedTime(&a, c->ev[0], c->ev[1]), "hipEventElapsedTime");
      hip_check(c, hipEventElapsedTime(&bb, c->ev[1], c->ev[2]), "hipEventElapsedTime");
      t->demodulate_ms = a;
      t->beamform_ms = bb;
      t->total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - w0).count();
    }
  });
}

usx_status usx_process_frame_noalloc(usx_ctx* c, const usx_demod_desc* d, const usx_bf_desc* b,
                                     const float* rf, usx_c32* frame) {
  return guarded(c, [&] {
    check_demod(c, d);
    check_bf(c, b);
    if (!rf || !frame) fail(c, USX_ERR_INVALID, "usx_process_frame_noalloc: null data pointer");
    if (!b->is_iq) fail(c, USX_ERR_INVALID, "usx_process_frame_noalloc: beamform descriptor must describe IQ");
    if (b->n_events * b->n_channels != d->n_traces || b->n_samples != d->n_out)
      fail(c, USX_ERR_INVALID, "usx_process_frame_noalloc: demodulate and beamform shapes disagree");
    if (!c->scratch)
      fail(c, USX_ERR_RUNTIME, "usx_process_frame_noalloc: scratch array unavailable (hipGetSymbolAddress: %s)",
           hipGetErrorString(c->scratch_err));
    // Lay out every buffer of the call in the scratch array, 256-byte aligned:
    // descriptor tables, IQ, then what the beamforming launch sequence needs.
    const Tile T = tile_of(b);
    const ScratchSize z = scratch_size(b, T);
    const bool sector = b->grid_type == 1;
    const size_t sizes[] = {sizeof(float) * d->n_taps, sizeof(usx_c32) * d->n_in,
                            sizeof(float) * b->n_channels, sizeof(usx_tx_desc) * b->n_events,
                            sizeof(float) * b->n_axis0, sizeof(float) * b->n_axis1,
                            sector ? sizeof(float) * b->n_axis0 : 0, sector ? sizeof(float) * b->n_axis0 : 0,
                            sizeof(usx_c32) * static_cast<size_t>(d->n_traces) * d->n_out,
                            z.y, z.mask, z.acc, z.counts, z.lists};
    constexpr int kParts = sizeof(sizes) / sizeof(sizes[0]);
    size_t off[kParts], used = 0;
    for (int i = 0; i < kParts; ++i) {
      off[i] = used;
      used += (sizes[i] + 255) / 256 * 256;
    }
    if (used > USX_NOALLOC_SCRATCH_BYTES)
      fail(c, USX_ERR_INVALID,
           "usx_process_frame_noalloc: this frame needs %zu B of scratch; the allocation-free entry "
           "point holds at most %zu B (use usx_process_frame)",
           used, static_cast<size_t>(USX_NOALLOC_SCRATCH_BYTES));
    unsigned char* base = static_cast<unsigned char*>(c->scratch);
    auto at = [&](int i) -> void* { return sizes[i] ? base + off[i] : nullptr; };
    const void* host[] = {d->taps, d->rotator, b->element_x, b->tx, b->axis0, b->axis1,
                          b->axis0_sin, b->axis0_cos};
    for (int i = 0; i < 8; ++i)
      if (sizes[i])
        hip_check(c, hipMemcpyAsync(at(i), host[i], sizes[i], hipMemcpyHostToDevice, c->stream),
                  "hipMemcpyAsync(descriptor)");
    BfTables t;
    t.ex = static_cast<const float*>(at(2));
    t.tx = static_cast<const usx_tx_desc*>(at(3));
    t.ax0 = static_cast<const float*>(at(4));
    t.ax1 = static_cast<const float*>(at(5));
    t.s0 = static_cast<const float*>(at(6));
    t.c0 = static_cast<const float*>(at(7));
    const BfArgs A = args_of(c, b, t);
    float2* d_iq = static_cast<float2*>(at(8));
    Scratch S;
    S.Y = static_cast<float2*>(at(9));
    S.mask = static_cast<uint32_t*>(at(10));
    S.acc = static_cast<float*>(at(11));
    S.counts = static_cast<int*>(at(12));
    S.lists = static_cast<int*>(at(13));
    enqueue_demod(c, d, static_cast<const float*>(at(0)), static_cast<const float2*>(at(1)), rf, d_iq);
    check_launch(c, "demodulate launch");
    enqueue_beamform(c, b, A, S, T, d_iq, reinterpret_cast<float2*>(frame), nullptr, nullptr);
    check_launch(c, "beamform launch");
  });
}

usx_status usx_mv_weights(usx_ctx* c, const usx_bf_desc* b, const void* data, usx_c32* frame,
                          usx_c32* weights, int* status, int memory) {
  return guarded(c, [&] {
    check_bf(c, b);
    if (b->combiner != 2) fail(c, USX_ERR_INVALID, "usx_mv_weights: combiner must be MV");
    if (b->compounding == 1) fail(c, USX_ERR_INVALID, "usx_mv_weights: incoherent compounding has no weights");
    if (!data || !frame || !weights) fail(c, USX_ERR_INVALID, "usx_mv_weights: null data pointer");
    const size_t npix = static_cast<size_t>(b->n_axis0) * b->n_axis1;
    const size_t nw = npix * b->n_channels;
    if (memory == USX_DEVICE) {
      run_beamform(c, b, data, reinterpret_cast<float2*>(frame), reinterpret_cast<float2*>(weights), status);
      check_launch(c, "mv_weights launch");
      return;
    }
    void* d_in = ensure(c, c->in, data_bytes(b));
    float2* d_out = static_cast<float2*>(ensure(c, c->out, npix * sizeof(float2)));
    float2* d_w = static_cast<float2*>(ensure(c, c->mv_w, nw * sizeof(float2)));
    int* d_st = static_cast<int*>(ensure(c, c->mv_status, npix * sizeof(int)));
    hip_check(c, hipMemcpyAsync(d_in, data, data_bytes(b), hipMemcpyHostToDevice, c->stream), "hipMemcpyAsync(data)");
    run_beamform(c, b, d_in, d_out, d_w, d_st);
    check_launch(c, "mv_weights launch");
    hip_check(c, hipMemcpyAsync(frame, d_out, npix * sizeof(float2), hipMemcpyDeviceToHost, c->stream), "hipMemcpyAsync(frame)");
    hip_check(c, hipMemcpyAsync(weights, d_w, nw * sizeof(float2), hipMemcpyDeviceToHost, c->stream), "hipMemcpyAsync(weights)");
    if (status)
      hip_check(c, hipMemcpyAsync(status, d_st, npix * sizeof(int), hipMemcpyDeviceToHost, c->stream), "hipMemcpyAsync(status)");
    hip_check(c, hipStreamSynchronize(c->stream), "hipStreamSynchronize");
  });
}

usx_status usx_apod_set_params(usx_ctx* c, const float* params) {
  return guarded(c, [&] {
    if (!params) fail(c, USX_ERR_INVALID, "usx_apod_set_params: null parameters");
    apod_alloc(c);
    const size_t bytes = sizeof(float) * usxd::kApodParams;
    hip_check(c, hipMemcpyAsync(c->ap_master.p, params, bytes, hipMemcpyHostToDevice, c->stream), "hipMemcpyAsync(params)");
    usxd::launch_fill_zero(static_cast<float*>(c->ap_m.p), usxd::kApodParams, c->stream);
    usxd::launch_fill_zero(static_cast<float*>(c->ap_v.p), usxd::kApodParams, c->stream);
    usxd::launch_apod_pack(apod_net(c), c->stream);
    check_launch(c, "apod pack launch");
    hip_check(c, hipStreamSynchronize(c->stream), "hipStreamSynchronize");
    c->apod_loaded = true;
    c->apod_t = 0;
  });
}

usx_status usx_apod_get_params(usx_ctx* c, float* params) {
  return guarded(c, [&] {
    if (!params) fail(c, USX_ERR_INVALID, "usx_apod_get_params: null buffer");
    if (!c->apod_loaded) fail(c, USX_ERR_INVALID, "learned apodization: no network loaded");
    hip_check(c, hipMemcpyAsync(params, c->ap_master.p, sizeof(float) * usxd::kApodParams,
                                hipMemcpyDeviceToHost, c->stream), "hipMemcpyAsync(params)");
    hip_check(c, hipStreamSynchronize(c->stream), "hipStreamSynchronize");
  });
}

usx_status usx_apod_dataset_reserve(usx_ctx* c, size_t capacity) {
  return guarded(c, [&] {
    if (capacity == 0) fail(c, USX_ERR_INVALID, "usx_apod_dataset_reserve: capacity must be > 0");
    apod_reserve(c, capacity);
  });
}

usx_status usx_apod_dataset_size(usx_ctx* c, size_t* n) {
  return guarded(c, [&] {
    if (!n) fail(c, USX_ERR_INVALID, "usx_apod_dataset_size: null output");
    *n = c->ds_n;
  });
}

usx_status usx_apod_dataset_append(usx_ctx* c, const usx_demod_desc* d, const usx_bf_desc* b,
                                   const void* data, int memory, const int* pixels, int n_pixels,
                                   usx_c32* frame) {
  return guarded(c, [&] {
    check_bf(c, b);
    if (d) check_demod(c, d);
    if (b->combiner != 2) fail(c, USX_ERR_INVALID, "usx_apod_dataset_append: labels come from MV (combiner 2)");
    if (b->compounding == 1) fail(c, USX_ERR_INVALID, "usx_apod_dataset_append: incoherent compounding has no weights");
    if (b->n_channels != USX_APOD_CHANNELS)
      fail(c, USX_ERR_INVALID, "learned apodization: the network takes %d channels, got %d",
           USX_APOD_CHANNELS, b->n_channels);
    if (!data || !pixels || n_pixels < 0) fail(c, USX_ERR_INVALID, "usx_apod_dataset_append: null data or pixel list");
    if (c->ds_n + static_cast<size_t>(n_pixels) > c->ds_cap)
      fail(c, USX_ERR_INVALID, "usx_apod_dataset_append: %zu + %d rows exceed the reserved %zu",
           c->ds_n, n_pixels, c->ds_cap);
    const int npix = b->n_axis0 * b->n_axis1;
    for (int i = 0; i < n_pixels; ++i)
      if (pixels[i] < 0 || pixels[i] >= npix)
        fail(c, USX_ERR_INVALID, "usx_apod_dataset_append: pixel %d outside the %d-pixel grid", pixels[i], npix);
    if (d && (!b->is_iq || b->n_events * b->n_channels != d->n_traces || b->n_samples != d->n_out))
      fail(c, USX_ERR_INVALID, "usx_apod_dataset_append: demodulate and beamform shapes disagree");
    const void* d_data = data;
    if (memory != USX_DEVICE) {
      const size_t bytes = d ? sizeof(float) * static_cast<size_t>(d->n_traces) * d->n_in : data_bytes(b);
      void* p = ensure(c, c->in, bytes);
      hip_check(c, hipMemcpyAsync(p, data, bytes, hipMemcpyHostToDevice, c->stream), "hipMemcpyAsync(data)");
      d_data = p;
    }
    if (d) {
      const size_t nout = static_cast<size_t>(d->n_traces) * d->n_out;
      float2* d_iq = static_cast<float2*>(ensure(c, c->iq, nout * sizeof(float2)));
      run_demod(c, d, static_cast<const float*>(d_data), d_iq);
      d_data = d_iq;
    }
    float2* d_out = static_cast<float2*>(ensure(c, c->out, npix * sizeof(float2)));
    float2* d_w = static_cast<float2*>(ensure(c, c->mv_w, static_cast<size_t>(npix) * b->n_channels * sizeof(float2)));
    int* d_st = static_cast<int*>(ensure(c, c->mv_status, npix * sizeof(int)));
    run_beamform(c, b, d_data, d_out, d_w, d_st);
    int* d_pix = static_cast<int*>(ensure(c, c->ap_pixels, sizeof(int) * (n_pixels + 1)));
    hip_check(c, hipMemcpyAsync(d_pix, pixels, sizeof(int) * n_pixels, hipMemcpyHostToDevice, c->stream), "hipMemcpyAsync(pixels)");
    if (n_pixels > 0)
      timed(c, "apod_gather", [&] {
        usxd::launch_apod_gather(n_pixels, d_pix, b->n_channels, static_cast<const float2*>(c->Y.p),
                                 static_cast<const uint32_t*>(c->mask.p), d_w, d_st, apod_data(c),
                                 static_cast<long>(c->ds_n), c->stream);
      });
    check_launch(c, "apod dataset launch");
    if (frame)
      hip_check(c, hipMemcpyAsync(frame, d_out, npix * sizeof(float2), hipMemcpyDeviceToHost, c->stream), "hipMemcpyAsync(frame)");
    hip_check(c, hipStreamSynchronize(c->stream), "hipStreamSynchronize");
    c->ds_n += n_pixels;
  });
}

usx_status usx_apod_train(usx_ctx* c, const usx_apod_train_desc* t, float* losses) {
  return guarded(c, [&] {
    if (!t) fail(c, USX_ERR_INVALID, "usx_apod_train: null descriptor");
    if (!c->apod_loaded) fail(c, USX_ERR_INVALID, "learned apodization: no network loaded");
    if (c->ds_n == 0) fail(c, USX_ERR_INVALID, "usx_apod_train: the training set is empty");
    if (t->batch < 256 || t->batch % 256 != 0 || t->batch > (1 << 20))
      fail(c, USX_ERR_INVALID, "usx_apod_train: batch must be a multiple of 256 up to 2^20");
    if (t->n_steps < 0) fail(c, USX_ERR_INVALID, "usx_apod_train: n_steps must be >= 0");
    const usxd::ApodWork W = apod_work(c, t->batch);
    float* d_losses = static_cast<float*>(ensure(c, c->tr_losses, sizeof(float) * (t->n_steps + 1)));
    const usxd::ApodNet N = apod_net(c);
    const usxd::ApodData D = apod_data(c);
    for (int k = 0; k < t->n_steps; ++k) {
      const int step = ++c->apod_t;
      usxd::ApodStep S{};
      S.batch = t->batch;
      S.n_data = static_cast<long>(c->ds_n);
      S.random = 1;
      S.seed = t->seed;
      S.step = step;
      S.grad_scale = 2.0f / static_cast<float>(t->batch);
      S.loss_scale = 1.0f / static_cast<float>(t->batch);
      usxd::ApodAdam opt{};
      opt.lr = t->lr;
      opt.beta1 = t->beta1;
      opt.beta2 = t->beta2;
      opt.eps = t->eps;
      opt.weight_decay = t->weight_decay;
      opt.bc1 = static_cast<float>(1.0 - std::pow(static_cast<double>(t->beta1), step));
      opt.bc2 = static_cast<float>(1.0 - std::pow(static_cast<double>(t->beta2), step));
      timed(c, "apod_step", [&] { usxd::launch_apod_step(N, D, S, W, opt, nullptr, d_losses + k, c->stream); });
    }
    check_launch(c, "apod train launch");
    if (losses && t->n_steps > 0)
      hip_check(c, hipMemcpyAsync(losses, d_losses, sizeof(float) * t->n_steps, hipMemcpyDeviceToHost, c->stream), "hipMemcpyAsync(losses)");
    hip_check(c, hipStreamSynchronize(c->stream), "hipStreamSynchronize");
  });
}

usx_status usx_apod_gradient(usx_ctx* c, int n, const uint16_t* x, const float* g,
                             const uint32_t* mask, const float* sw, float* loss, float* grads) {
  return guarded(c, [&] {
    if (!c->apod_loaded) fail(c, USX_ERR_INVALID, "learned apodization: no network loaded");
    if (n <= 0 || n > 4096) fail(c, USX_ERR_INVALID, "usx_apod_gradient: n must be in 1..4096");
    if (!x || !g || !mask || !sw || !loss || !grads) fail(c, USX_ERR_INVALID, "usx_apod_gradient: null pointer");
    // the batch becomes a temporary training set (the stored one is kept)
    const size_t F = usxd::kApodFeatures;
    const int B = n <= 256 ? (n + 31) / 32 * 32 : (n + 255) / 256 * 256;
    DevBuf tx{}, tg{}, tm{}, ts{};
    auto cleanup = [&] {
      for (DevBuf* p : {&tx, &tg, &tm, &ts})
        if (p->p) (void)hipFree(p->p);
    };
    try {
      usxd::ApodData D{};
      D.x = static_cast<uint16_t*>(ensure(c, tx, n * F * sizeof(uint16_t)));
      D.g = static_cast<float*>(ensure(c, tg, n * F * sizeof(float)));
      D.mask = static_cast<uint32_t*>(ensure(c, tm, n * 4 * sizeof(uint32_t)));
      D.sw = static_cast<float*>(ensure(c, ts, n * sizeof(float)));
      hip_check(c, hipMemcpyAsync(D.x, x, n * F * sizeof(uint16_t), hipMemcpyHostToDevice, c->stream), "hipMemcpyAsync(x)");
      hip_check(c, hipMemcpyAsync(D.g, g, n * F * sizeof(float), hipMemcpyHostToDevice, c->stream), "hipMemcpyAsync(g)");
      hip_check(c, hipMemcpyAsync(D.mask, mask, n * 4 * sizeof(uint32_t), hipMemcpyHostToDevice, c->stream), "hipMemcpyAsync(mask)");
      hip_check(c, hipMemcpyAsync(D.sw, sw, n * sizeof(float), hipMemcpyHostToDevice, c->stream), "hipMemcpyAsync(sw)");
      const usxd::ApodWork W = apod_work(c, B);
      float* d_grads = static_cast<float*>(ensure(c, c->tr_grads, sizeof(float) * (usxd::kApodParams + 1)));
      float* d_loss = d_grads + usxd::kApodParams;
      usxd::ApodStep S{};
      S.batch = B;
      S.n_data = n;
      S.random = 0;
      S.grad_scale = 2.0f / static_cast<float>(n);
      S.loss_scale = 1.0f / static_cast<float>(n);
      usxd::ApodAdam opt{};
      usxd::launch_apod_step(apod_net(c), D, S, W, opt, d_grads, d_loss, c->stream);
      check_launch(c, "apod gradient launch");
      hip_check(c, hipMemcpyAsync(grads, d_grads, sizeof(float) * usxd::kApodParams, hipMemcpyDeviceToHost, c->stream), "hipMemcpyAsync(grads)");
      hip_check(c, hipMemcpyAsync(loss, d_loss, sizeof(float), hipMemcpyDeviceToHost, c->stream), "hipMemcpyAsync(loss)");
      hip_check(c, hipStreamSynchronize(c->stream), "hipStreamSynchronize");
    } catch (...) {
      cleanup();
      throw;
    }
    cleanup();
  });
}

usx_status usx_simulate(usx_ctx* c, const usx_sim_desc* d, int n_scat, const float* scatterers,
                        float* rf, int memory) {
  return guarded(c, [&] {
    if (!d) fail(c, USX_ERR_INVALID, "usx_simulate: null descriptor");
    if (d->n_elements <= 0 || d->n_samples <= 0 || d->n_events < 0)
      fail(c, USX_ERR_INVALID, "usx_simulate: empty probe or acquisition");
    if (n_scat < 0 || (n_scat > 0 && !scatterers) || !rf)
      fail(c, USX_ERR_INVALID, "usx_simulate: null scatterers or output");
    if (!d->element_x || !d->delays || !d->pulse || d->pulse_n < 1 || d->oversampling < 1)
      fail(c, USX_ERR_INVALID, "usx_simulate: element, delay and pulse tables are required");
    if (d->noise_amp > 0.0f && !d->noise)
      fail(c, USX_ERR_INVALID, "usx_simulate: noise_amp > 0 needs the standard-normal table");
    const int ne = d->n_elements, nev = d->n_events;
    usxd::SimArgs S{};
    S.n_elements = ne;
    S.n_samples = d->n_samples;
    S.fs = d->sampling_frequency;
    S.t0 = d->t0;
    S.prf = d->prf;
    S.c = d->speed_of_sound;
    S.fs_os = d->sampling_frequency * static_cast<float>(d->oversampling);
    S.wavelength = d->wavelength;
    S.element_width = d->element_width;
    S.pulse_n = d->pulse_n;
    S.pulse_half = d->pulse_half;
    S.atten = d->attenuation;
    S.max_window = d->max_window;
    S.trim_amp = d->trim_amp;
    S.noise_amp = d->noise_amp;
    S.spreading = d->spreading;
    S.directivity = d->directivity;
    S.motion = d->motion;
    S.full_diffraction = d->full_diffraction;
    // incident buffer: ceil((max_window + 2 pulse_half) fs_os) + 2, plus 2 spare
    S.inc_cap = static_cast<int>(std::ceil((d->max_window + 2.0f * d->pulse_half) * S.fs_os)) + 4;
    if (S.inc_cap < 8) S.inc_cap = 8;
    if (usxd::sim_incident_smem(S) > 160 * 1024 || usxd::sim_echo_smem(S) > 160 * 1024)
      fail(c, USX_ERR_INVALID, "usx_simulate: incident window (%d samples) or record length (%d) too long",
           S.inc_cap, S.n_samples);
    // firing lists (simulator.cpp:229-233)
    std::vector<int> firing(static_cast<size_t>(std::max(nev, 1)) * ne), nfire(std::max(nev, 1));
    std::vector<float> fw(firing.size());
    for (int e = 0; e < nev; ++e) {
      int n = 0;
      const size_t row = static_cast<size_t>(e) * ne;
      const bool has = d->has_apodization && d->has_apodization[e] && d->apodization;
      if (has)
        for (int i = 0; i < ne; ++i) {
          const float a = d->apodization[row + i];
          if (a != 0.0f) {
            firing[row + n] = i;
            fw[row + n] = a;
            ++n;
          }
        }
      if (n == 0)
        for (int i = 0; i < ne; ++i) {
          // every element fires: weight 1 for an empty apodization, else its
          // own (zero) weight, as path_amp reads it
          firing[row + n] = i;
          fw[row + n] = has ? d->apodization[row + i] : 1.0f;
          ++n;
        }
      nfire[e] = n;
    }
    S.ex = static_cast<const float*>(upload(c, c->s_ex, d->element_x, sizeof(float) * ne));
    S.delays = static_cast<const float*>(upload(c, c->s_delays, d->delays, sizeof(float) * nev * ne));
    S.firing = static_cast<const int*>(upload(c, c->s_firing, firing.data(), sizeof(int) * firing.size()));
    S.fire_w = static_cast<const float*>(upload(c, c->s_fw, fw.data(), sizeof(float) * fw.size()));
    S.n_firing = static_cast<const int*>(upload(c, c->s_nf, nfire.data(), sizeof(int) * nfire.size()));
    S.pulse = static_cast<const float*>(upload(c, c->s_pulse, d->pulse, sizeof(float) * d->pulse_n));
    const size_t n_tot = static_cast<size_t>(ne) * d->n_samples;
    const float* z = nullptr;
    if (d->noise_amp > 0.0f)
      z = static_cast<const float*>(upload(c, c->s_noise, d->noise, sizeof(float) * n_tot * nev));
    const size_t ns = static_cast<size_t>(std::max(n_scat, 1));
    float* d_scat = static_cast<float*>(ensure(c, c->s_scat, sizeof(float) * 7 * ns));
    if (n_scat > 0)
      hip_check(c, hipMemcpyAsync(d_scat, scatterers, sizeof(float) * 7 * n_scat, hipMemcpyHostToDevice, c->stream),
                "hipMemcpyAsync(scatterers)");
    usxd::SimIncident I{};
    I.live = static_cast<int*>(ensure(c, c->s_live, sizeof(int) * ns));
    I.t0 = static_cast<float*>(ensure(c, c->s_t0, sizeof(float) * ns));
    I.n = static_cast<int*>(ensure(c, c->s_n, sizeof(int) * ns));
    I.pos = static_cast<float4*>(ensure(c, c->s_pos, sizeof(float4) * ns));
    I.wave = static_cast<float*>(ensure(c, c->s_wave, sizeof(float) * ns * S.inc_cap));
    const int nc = std::max(1, usxd::sim_chunks(n_scat));
    float* partial = static_cast<float*>(ensure(c, c->s_partial, sizeof(float) * n_tot * nc));
    float* bmax = static_cast<float*>(ensure(c, c->s_bmax, sizeof(float) * ((n_tot + 255) / 256 + 1)));
    float* d_rf = memory == USX_DEVICE
                      ? rf
                      : static_cast<float*>(ensure(c, c->s_rf, sizeof(float) * n_tot * std::max(nev, 1)));
    for (int e = 0; e < nev; ++e)
      timed(c, "simulate", [&] {
        usxd::launch_sim_event(S, e, d_scat, n_scat, I, partial, bmax, z ? z + n_tot * e : nullptr,
                               d_rf + n_tot * e, c->stream);
      });
    check_launch(c, "simulate launch");
    if (memory != USX_DEVICE) {
      hip_check(c, hipMemcpyAsync(rf, d_rf, sizeof(float) * n_tot * nev, hipMemcpyDeviceToHost, c->stream),
                "hipMemcpyAsync(rf)");
      hip_check(c, hipStreamSynchronize(c->stream), "hipStreamSynchronize");
    }
  });
}

}  // extern "C"

// ---- frame-sequence plans and the autotuner --------------------------------

struct usx_plan {
  int n = 0;
  std::vector<usx_demod_desc> d;
  std::vector<usx_bf_desc> b;
  std::vector<DevBuf> tables;  // per frame: taps, rotator, element x, transmits, axes
  std::vector<const float*> taps;
  std::vector<const float2*> rot;
  std::vector<BfArgs> args;
  std::vector<const float*> rf;
  std::vector<float2*> frames;
  DevBuf iq, Y, mask, acc, counts, lists;
  Scratch scratch;
  float2* d_iq = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

namespace {

size_t align256(size_t n) { return (n + 255) / 256 * 256; }

void plan_release(usx_plan* p) {
  if (!p) return;
  if (!g_exiting) {
    if (p->exec) (void)hipGraphExecDestroy(p->exec);
    if (p->graph) (void)hipGraphDestroy(p->graph);
    for (DevBuf& t : p->tables)
      if (t.p) (void)hipFree(t.p);
    for (DevBuf* t : {&p->iq, &p->Y, &p->mask, &p->acc, &p->counts, &p->lists})
      if (t->p) (void)hipFree(t->p);
  }
  delete p;
}

// The whole launch sequence of a plan on the context stream (kernels only).
void plan_enqueue(usx_ctx* c, usx_plan* p) {
  for (int f = 0; f < p->n; ++f) {
    enqueue_demod(c, &p->d[f], p->taps[f], p->rot[f], p->rf[f], p->d_iq);
    enqueue_beamform(c, &p->b[f], p->args[f], p->scratch, tile_of(&p->b[f]), p->d_iq,
                     p->frames[f], nullptr, nullptr);
  }
}

// Copy the host tables of frame f into one device blob.
void plan_tables(usx_ctx* c, usx_plan* p, int f) {
  const usx_demod_desc& d = p->d[f];
  const usx_bf_desc& b = p->b[f];
  const size_t sz[8] = {sizeof(float) * d.n_taps, sizeof(usx_c32) * d.n_in,
                        sizeof(float) * b.n_channels, sizeof(usx_tx_desc) * b.n_events,
                        sizeof(float) * b.n_axis0, sizeof(float) * b.n_axis1,
                        b.grid_type == 1 ? sizeof(float) * b.n_axis0 : 0,
                        b.grid_type == 1 ? sizeof(float) * b.n_axis0 : 0};
  const void* src[8] = {d.taps, d.rotator, b.element_x, b.tx, b.axis0, b.axis1, b.axis0_sin,
                        b.axis0_cos};
  size_t off[8], total = 0;
  for (int k = 0; k < 8; ++k) {
    off[k] = total;
    total += align256(sz[k]);
  }
  std::vector<unsigned char> host(total, 0);
  for (int k = 0; k < 8; ++k)
    if (sz[k]) std::memcpy(host.data() + off[k], src[k], sz[k]);
  DevBuf& t = p->tables[f];
  ensure(c, t, total);
  hip_check(c, hipMemcpy(t.p, host.data(), total, hipMemcpyHostToDevice), "hipMemcpy(plan tables)");
  auto at = [&](int k) { return static_cast<const void*>(static_cast<unsigned char*>(t.p) + off[k]); };
  p->taps[f] = static_cast<const float*>(at(0));
  p->rot[f] = static_cast<const float2*>(at(1));
  BfTables bt;
  bt.ex = static_cast<const float*>(at(2));
  bt.tx = static_cast<const usx_tx_desc*>(at(3));
  bt.ax0 = static_cast<const float*>(at(4));
  bt.ax1 = static_cast<const float*>(at(5));
  bt.s0 = static_cast<const float*>(at(6));
  bt.c0 = static_cast<const float*>(at(7));
  p->args[f] = args_of(c, &b, bt);
}

}  // namespace

extern "C" {

usx_status usx_buffer_alloc(usx_ctx* c, size_t bytes, void** dev) {
  return guarded(c, [&] {
    if (!dev) fail(c, USX_ERR_INVALID, "usx_buffer_alloc: null output");
    *dev = nullptr;
    hip_check(c, hipMalloc(dev, bytes == 0 ? 16 : bytes), "hipMalloc");
  });
}

usx_status usx_buffer_free(usx_ctx* c, void* dev) {
  if (!dev || g_exiting) return USX_OK;
  return guarded(c, [&] { hip_check(c, hipFree(dev), "hipFree"); });
}

usx_status usx_buffer_upload(usx_ctx* c, void* dev, const void* host, size_t bytes) {
  return guarded(c, [&] {
    if (!dev || (!host && bytes)) fail(c, USX_ERR_INVALID, "usx_buffer_upload: null pointer");
    hip_check(c, hipMemcpyAsync(dev, host, bytes, hipMemcpyHostToDevice, c->stream), "hipMemcpyAsync(upload)");
    hip_check(c, hipStreamSynchronize(c->stream), "hipStreamSynchronize");
  });
}

usx_status usx_buffer_download(usx_ctx* c, void* host, const void* dev, size_t bytes) {
  return guarded(c, [&] {
    if (!dev || (!host && bytes)) fail(c, USX_ERR_INVALID, "usx_buffer_download: null pointer");
    hip_check(c, hipMemcpyAsync(host, dev, bytes, hipMemcpyDeviceToHost, c->stream), "hipMemcpyAsync(download)");
    hip_check(c, hipStreamSynchronize(c->stream), "hipStreamSynchronize");
  });
}

usx_status usx_buffer_zero(usx_ctx* c, void* dev, size_t bytes) {
  return guarded(c, [&] {
    if (!dev || bytes % 4) fail(c, USX_ERR_INVALID, "usx_buffer_zero: null pointer or size not a multiple of 4");
    usxd::launch_fill_zero(static_cast<float*>(dev), bytes / 4, c->stream);
    check_launch(c, "fill launch");
  });
}

usx_status usx_plan_create(usx_ctx* c, int n, const usx_demod_desc* d, const usx_bf_desc* b,
                           const float* const* rf, usx_c32* const* frames, usx_plan** out) {
  return guarded(c, [&] {
    if (!out) fail(c, USX_ERR_INVALID, "usx_plan_create: null output");
    *out = nullptr;
    if (n <= 0 || !d || !b || !rf || !frames) fail(c, USX_ERR_INVALID, "usx_plan_create: empty frame list");
    for (int f = 0; f < n; ++f) {
      check_demod(c, &d[f]);
      check_bf(c, &b[f]);
      if (!b[f].is_iq || b[f].n_events * b[f].n_channels != d[f].n_traces || b[f].n_samples != d[f].n_out)
        fail(c, USX_ERR_INVALID, "usx_plan_create: frame %d: demodulate and beamform shapes disagree", f);
      if (!rf[f] || !frames[f]) fail(c, USX_ERR_INVALID, "usx_plan_create: frame %d: null device pointer", f);
      if (b[f].combiner == 3 && !c->apod_loaded)
        fail(c, USX_ERR_INVALID, "learned apodization: no network loaded (usx_apod_set_params)");
    }
    usx_plan* p = new usx_plan();
    try {
      p->n = n;
      p->d.assign(d, d + n);
      p->b.assign(b, b + n);
      p->tables.resize(n);
      p->taps.resize(n);
      p->rot.resize(n);
      p->args.resize(n);
      p->rf.assign(rf, rf + n);
      p->frames.resize(n);
      for (int f = 0; f < n; ++f) p->frames[f] = reinterpret_cast<float2*>(frames[f]);
      ScratchSize z;
      size_t niq = 0;
      for (int f = 0; f < n; ++f) {
        plan_tables(c, p, f);
        const ScratchSize s = scratch_size(&b[f], tile_of(&b[f]));
        z.y = std::max(z.y, s.y);
        z.mask = std::max(z.mask, s.mask);
        z.acc = std::max(z.acc, s.acc);
        z.counts = std::max(z.counts, s.counts);
        z.lists = std::max(z.lists, s.lists);
        niq = std::max(niq, static_cast<size_t>(d[f].n_traces) * d[f].n_out);
      }
      p->d_iq = static_cast<float2*>(ensure(c, p->iq, niq * sizeof(float2)));
      if (z.y) p->scratch.Y = static_cast<float2*>(ensure(c, p->Y, z.y));
      if (z.mask) p->scratch.mask = static_cast<uint32_t*>(ensure(c, p->mask, z.mask));
      if (z.acc) p->scratch.acc = static_cast<float*>(ensure(c, p->acc, z.acc));
      if (z.counts) p->scratch.counts = static_cast<int*>(ensure(c, p->counts, z.counts));
      if (z.lists) p->scratch.lists = static_cast<int*>(ensure(c, p->lists, z.lists));
      // one eager pass (first-use initialisation of the kernels happens outside
      // the capture), then the capture of the whole sequence
      const bool prof = c->profiling;
      c->profiling = false;
      plan_enqueue(c, p);
      check_launch(c, "plan launch");
      hip_check(c, hipStreamSynchronize(c->stream), "hipStreamSynchronize");
      hip_check(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
      try {
        plan_enqueue(c, p);
      } catch (...) {
        hipGraph_t g = nullptr;
        (void)hipStreamEndCapture(c->stream, &g);
        if (g) (void)hipGraphDestroy(g);
        c->profiling = prof;
        throw;
      }
      hip_check(c, hipStreamEndCapture(c->stream, &p->graph), "hipStreamEndCapture");
      c->profiling = prof;
      hip_check(c, hipGraphInstantiate(&p->exec, p->graph, nullptr, nullptr, 0), "hipGraphInstantiate");
    } catch (...) {
      plan_release(p);
      throw;
    }
    *out = p;
  });
}

usx_status usx_plan_launch(usx_ctx* c, usx_plan* p, int use_graph) {
  return guarded(c, [&] {
    if (!p) fail(c, USX_ERR_INVALID, "usx_plan_launch: null plan");
    if (use_graph) {
      hip_check(c, hipGraphLaunch(p->exec, c->stream), "hipGraphLaunch");
    } else {
      plan_enqueue(c, p);
      check_launch(c, "plan launch");
    }
  });
}

usx_status usx_plan_destroy(usx_ctx* c, usx_plan* p) {
  if (!p) return USX_OK;
  if (c && !g_exiting) {
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
  }
  plan_release(p);
  return USX_OK;
}

usx_status usx_autotune(usx_ctx* c, const usx_demod_desc* d, usx_bf_desc* b, const float* rf,
                        usx_c32* frame, int reps, float* best_ms) {
  return guarded(c, [&] {
    check_demod(c, d);
    check_bf(c, b);
    if (!rf || !frame) fail(c, USX_ERR_INVALID, "usx_autotune: null device pointer");
    if (!b->is_iq || b->n_events * b->n_channels != d->n_traces || b->n_samples != d->n_out)
      fail(c, USX_ERR_INVALID, "usx_autotune: demodulate and beamform shapes disagree");
    reps = std::max(1, reps);
    const int n_ev = b->event_end - b->event_begin;
    std::vector<std::pair<int, int>> cand;  // (slots, mv occupancy)
    const int slot_opts[3] = {4, 8, 16};
    const int occ_opts[4] = {0, 4, 8, 16};
    for (int s : slot_opts) {
      if (s > 4 && s / 2 >= n_ev) continue;  // no more launches to merge
      if (b->combiner == 2) {
        for (int o : occ_opts) cand.push_back({s, o});
      } else {
        cand.push_back({s, 0});
      }
    }
    const size_t niq = static_cast<size_t>(d->n_traces) * d->n_out;
    float2* d_iq = static_cast<float2*>(ensure(c, c->iq, niq * sizeof(float2)));
    const bool prof = c->profiling;
    c->profiling = false;
    hipEvent_t e0 = c->ev[0], e1 = c->ev[1];
    float best = 1e30f;
    std::pair<int, int> pick = {16, 0};
    for (const auto& cd : cand) {
      usx_bf_desc t = *b;
      t.tile_slots = cd.first;
      t.tile_mv_occupancy = cd.second;
      run_demod(c, d, rf, d_iq);
      run_beamform(c, &t, d_iq, reinterpret_cast<float2*>(frame));  // warm-up
      hip_check(c, hipEventRecord(e0, c->stream), "hipEventRecord");
      for (int r = 0; r < reps; ++r) {
        run_demod(c, d, rf, d_iq);
        run_beamform(c, &t, d_iq, reinterpret_cast<float2*>(frame));
      }
      hip_check(c, hipEventRecord(e1, c->stream), "hipEventRecord");
      check_launch(c, "autotune launch");
      hip_check(c, hipEventSynchronize(e1), "hipEventSynchronize");
      float ms = 0.0f;
      hip_check(c, hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
      ms /= static_cast<float>(reps);
      if (ms < best) {
        best = ms;
        pick = cd;
      }
    }
    c->profiling = prof;
    b->tile_slots = pick.first;
    b->tile_mv_occupancy = pick.second;
    if (best_ms) *best_ms = best;
  });
}

}  // extern "C"

// ---- display tail -----------------------------------------------------------

namespace {

usxd::DisplayArgs display_args(usx_ctx* c, const usx_bf_desc* b, const usx_display_desc* p) {
  const BfArgs A = make_args(c, b);
  usxd::DisplayArgs D{};
  D.g = A.g;
  D.apply_tgc = p->apply_tgc;
  D.attenuation = p->attenuation_db_cm_mhz;
  D.f_mhz = p->center_frequency_hz * 1e-6f;
  D.dynamic_range_db = p->dynamic_range_db;
  D.gain_db = p->gain_db;
  return D;
}

// Envelope + TGC of a device frame, the peak to the host, the smoothed
// reference updated (pipeline.cpp:58-66), then the log compression.
void run_display(usx_ctx* c, const usx_bf_desc* b, const usx_display_desc* p, const float2* d_frame,
                 bool rf, float* d_image, float* reference) {
  if (rf && b->n_axis1 > 8192)
    fail(c, USX_ERR_INVALID, "display: RF frames are limited to 8192 samples along axis1");
  const usxd::DisplayArgs D = display_args(c, b, p);
  const size_t n_pix = static_cast<size_t>(b->n_axis0) * b->n_axis1;
  const int nb = usxd::display_blocks(D, rf);
  float* env = static_cast<float*>(ensure(c, c->disp_env, n_pix * sizeof(float)));
  float* bmax = static_cast<float*>(ensure(c, c->disp_bmax, (nb + 2) * sizeof(float)));
  float* ref_dev = bmax + nb;
  timed(c, "display_envelope", [&] { usxd::launch_display_envelope(D, d_frame, rf, env, bmax, c->stream); });
  // the frame's peak from the block maxima, on the host, for the smoothing
  std::vector<float> host(nb);
  hip_check(c, hipMemcpyAsync(host.data(), bmax, nb * sizeof(float), hipMemcpyDeviceToHost, c->stream),
            "hipMemcpyAsync(peaks)");
  hip_check(c, hipStreamSynchronize(c->stream), "hipStreamSynchronize");
  float peak = 0.0f;
  for (float v : host) peak = std::max(peak, v);
  float ref = *reference;
  if (peak > 0.0f) {
    const float a = std::clamp(p->reference_smoothing, 0.0f, 0.999f);
    ref = (ref > 0.0f) ? (a * ref + (1.0f - a) * peak) : peak;
  }
  *reference = ref;
  const float use = ref > 0.0f ? ref : -1.0f;
  hip_check(c, hipMemcpyAsync(ref_dev, &use, sizeof(float), hipMemcpyHostToDevice, c->stream),
            "hipMemcpyAsync(reference)");
  timed(c, "display_compress", [&] { usxd::launch_display_compress(D, env, bmax, nb, ref_dev, d_image, c->stream); });
}

// The grid fields of a beamform descriptor (the display tail needs no more).
void check_grid(usx_ctx* c, const usx_bf_desc* b) {
  if (!b) fail(c, USX_ERR_INVALID, "display: null grid descriptor");
  if (b->n_axis0 <= 0 || b->n_axis1 <= 0) fail(c, USX_ERR_INVALID, "ScanGrid dimensions must be > 0");
  if (!b->axis0 || !b->axis1) fail(c, USX_ERR_INVALID, "display: grid axes are required");
  if (b->grid_type != 0 && b->grid_type != 1) fail(c, USX_ERR_INVALID, "beamform: unknown grid type");
  if (b->grid_type == 1 && (!b->axis0_sin || !b->axis0_cos))
    fail(c, USX_ERR_INVALID, "beamform: sector grids need the sin/cos tables");
  if (b->n_events < 0 || b->n_channels < 0) fail(c, USX_ERR_INVALID, "display: negative sizes");
}

void check_display(usx_ctx* c, const usx_display_desc* p, const float* reference) {
  if (!p || !reference) fail(c, USX_ERR_INVALID, "display: null descriptor or reference");
}

}  // namespace

extern "C" {

usx_status usx_display(usx_ctx* c, const usx_bf_desc* b, const usx_display_desc* p,
                       const usx_c32* frame, int rf, float* image, float* reference, int memory) {
  return guarded(c, [&] {
    check_grid(c, b);
    check_display(c, p, reference);
    if (!frame || !image) fail(c, USX_ERR_INVALID, "usx_display: null data pointer");
    const size_t n_pix = static_cast<size_t>(b->n_axis0) * b->n_axis1;
    if (memory == USX_DEVICE) {
      run_display(c, b, p, reinterpret_cast<const float2*>(frame), rf != 0, image, reference);
      check_launch(c, "display launch");
      return;
    }
    float2* d_frame = static_cast<float2*>(ensure(c, c->out, n_pix * sizeof(float2)));
    float* d_image = static_cast<float*>(ensure(c, c->disp_img, n_pix * sizeof(float)));
    hip_check(c, hipMemcpyAsync(d_frame, frame, n_pix * sizeof(float2), hipMemcpyHostToDevice, c->stream),
              "hipMemcpyAsync(frame)");
    run_display(c, b, p, d_frame, rf != 0, d_image, reference);
    check_launch(c, "display launch");
    hip_check(c, hipMemcpyAsync(image, d_image, n_pix * sizeof(float), hipMemcpyDeviceToHost, c->stream),
              "hipMemcpyAsync(image)");
    hip_check(c, hipStreamSynchronize(c->stream), "hipStreamSynchronize");
  });
}

usx_status usx_process_display(usx_ctx* c, const usx_demod_desc* d, const usx_bf_desc* b,
                               const usx_display_desc* p, const float* rf, float* image,
                               float* reference, int memory, usx_stage_times* t) {
  return guarded(c, [&] {
    check_demod(c, d);
    check_bf(c, b);
    check_display(c, p, reference);
    if (!rf || !image) fail(c, USX_ERR_INVALID, "usx_process_display: null data pointer");
    if (!b->is_iq || b->n_events * b->n_channels != d->n_traces || b->n_samples != d->n_out)
      fail(c, USX_ERR_INVALID, "usx_process_display: demodulate and beamform shapes disagree");
    const size_t nin = static_cast<size_t>(d->n_traces) * d->n_in;
    const size_t nout = static_cast<size_t>(d->n_traces) * d->n_out;
    const size_t n_pix = static_cast<size_t>(b->n_axis0) * b->n_axis1;
    const auto w0 = std::chrono::steady_clock::now();
    hip_check(c, hipEventRecord(c->ev[0], c->stream), "hipEventRecord");
    const float* d_rf = rf;
    float* d_image = image;
    if (memory != USX_DEVICE) {
      float* q = static_cast<float*>(ensure(c, c->in, nin * sizeof(float)));
      hip_check(c, hipMemcpyAsync(q, rf, nin * sizeof(float), hipMemcpyHostToDevice, c->stream),
                "hipMemcpyAsync(rf)");
      d_rf = q;
      d_image = static_cast<float*>(ensure(c, c->disp_img, n_pix * sizeof(float)));
    }
    float2* d_iq = static_cast<float2*>(ensure(c, c->iq, nout * sizeof(float2)));
    float2* d_frame = static_cast<float2*>(ensure(c, c->disp_frame, n_pix * sizeof(float2)));
    run_demod(c, d, d_rf, d_iq);
    hip_check(c, hipEventRecord(c->ev[1], c->stream), "hipEventRecord");
    run_beamform(c, b, d_iq, d_frame);
    check_launch(c, "beamform launch");
    hip_check(c, hipEventRecord(c->ev[2], c->stream), "hipEventRecord");
    run_display(c, b, p, d_frame, false, d_image, reference);
    check_launch(c, "display launch");
    hip_check(c, hipEventRecord(c->ev[3], c->stream), "hipEventRecord");
    if (memory != USX_DEVICE)
      hip_check(c, hipMemcpyAsync(image, d_image, n_pix * sizeof(float), hipMemcpyDeviceToHost, c->stream),
                "hipMemcpyAsync(image)");
    hip_check(c, hipStreamSynchronize(c->stream), "hipStreamSynchronize");
    if (t) {
      float a = 0, bb = 0, cc = 0;
      hip_check(c, hipEventElapsedTime(&a, c->ev[0], c->ev[1]), "hipEventElapsedTime");
      hip_check(c, hipEventElapsedTime(&bb, c->ev[1], c->ev[2]), "hipEventElapsedTime");
      hip_check(c, hipEventElapsedTime(&cc, c->ev[2], c->ev[3]), "hipEventElapsedTime");
      t->demodulate_ms = a;
      t->beamform_ms = bb;
      t->postprocess_ms = cc;
      t->total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - w0).count();
    }
  });
}

}  // extern "C"

// ---- Doppler and speckle tracking (doppler.cpp) ---------------------------

namespace {

size_t ensemble_pixels(usx_ctx* c, const usx_ensemble_desc* e, int min_frames, const char* who) {
  if (!e) fail(c, USX_ERR_INVALID, "%s: null descriptor", who);
  if (e->n_frames < min_frames)
    fail(c, USX_ERR_INVALID, "%s: need at least %d frames in the ensemble", who, min_frames);
  if (e->n_axis0 <= 0 || e->n_axis1 <= 0) fail(c, USX_ERR_INVALID, "ScanGrid dimensions must be > 0");
  return static_cast<size_t>(e->n_axis0) * e->n_axis1;
}

// The projector on the device, or nullptr for no filter.
const float* reject_on_device(usx_ctx* c, const usx_ensemble_desc* e) {
  if (!e->reject) return nullptr;
  const size_t bytes = sizeof(float) * static_cast<size_t>(e->n_frames) * e->n_frames;
  return static_cast<const float*>(upload(c, c->dop_reject, e->reject, bytes));
}

// The ensemble on the device: the caller's pointer, or an upload of it.
const float2* ensemble_on_device(usx_ctx* c, const usx_ensemble_desc* e, const usx_c32* ens,
                                 int memory, DevBuf& buf) {
  if (memory == USX_DEVICE) return reinterpret_cast<const float2*>(ens);
  const size_t bytes = sizeof(float2) * e->n_frames * ensemble_pixels(c, e, 1, "ensemble");
  void* d = ensure(c, buf, bytes);
  hip_check(c, hipMemcpyAsync(d, ens, bytes, hipMemcpyHostToDevice, c->stream), "hipMemcpyAsync(ensemble)");
  return static_cast<const float2*>(d);
}

template <typename T>
T* out_on_device(usx_ctx* c, T* host_or_dev, size_t n, int memory, DevBuf& buf) {
  if (memory == USX_DEVICE) return host_or_dev;
  return static_cast<T*>(ensure(c, buf, n * sizeof(T)));
}

template <typename T>
void download(usx_ctx* c, T* host, const T* dev, size_t n, int memory, const char* what) {
  if (memory == USX_DEVICE) return;
  hip_check(c, hipMemcpyAsync(host, dev, n * sizeof(T), hipMemcpyDeviceToHost, c->stream), what);
}

void finish(usx_ctx* c, int memory, const char* what) {
  check_launch(c, what);
  if (memory != USX_DEVICE) hip_check(c, hipStreamSynchronize(c->stream), "hipStreamSynchronize");
}

}  // namespace

extern "C" {

usx_status usx_wall_filter(usx_ctx* c, const usx_ensemble_desc* e, const usx_c32* ensemble,
                           usx_c32* out, int memory) {
  return guarded(c, [&] {
    const size_t n_pix = ensemble_pixels(c, e, 1, "apply_wall_filter");
    if (!ensemble || !out) fail(c, USX_ERR_INVALID, "apply_wall_filter: null data pointer");
    if (static_cast<const void*>(ensemble) == static_cast<const void*>(out))
      fail(c, USX_ERR_INVALID, "apply_wall_filter: the output must not overwrite the ensemble");
    const size_t n = n_pix * e->n_frames;
    const float* M = reject_on_device(c, e);
    const float2* d_in = ensemble_on_device(c, e, ensemble, memory, c->dop_in);
    float2* d_out = out_on_device(c, reinterpret_cast<float2*>(out), n, memory, c->dop_out);
    timed(c, "wall_filter", [&] { usxd::launch_wall_filter(e->n_frames, n_pix, d_in, M, d_out, c->stream); });
    download(c, reinterpret_cast<float2*>(out), d_out, n, memory, "hipMemcpyAsync(filtered)");
    finish(c, memory, "wall filter launch");
  });
}

usx_status usx_power_doppler(usx_ctx* c, const usx_ensemble_desc* e, const usx_c32* ensemble,
                             float* power, int memory) {
  return guarded(c, [&] {
    const size_t n_pix = ensemble_pixels(c, e, 1, "power_doppler");
    if (!ensemble || !power) fail(c, USX_ERR_INVALID, "power_doppler: null data pointer");
    const float* M = reject_on_device(c, e);
    const float2* d_in = ensemble_on_device(c, e, ensemble, memory, c->dop_in);
    float* d_p = out_on_device(c, power, n_pix, memory, c->dop_f0);
    timed(c, "power_doppler", [&] { usxd::launch_power_doppler(e->n_frames, n_pix, d_in, M, d_p, c->stream); });
    download(c, power, d_p, n_pix, memory, "hipMemcpyAsync(power)");
    finish(c, memory, "power doppler launch");
  });
}

usx_status usx_color_flow(usx_ctx* c, const usx_ensemble_desc* e, const usx_color_flow_desc* f,
                          const usx_c32* ensemble, float* velocity, float* power, float* variance,
                          uint8_t* valid, int memory) {
  return guarded(c, [&] {
    const size_t n_pix = ensemble_pixels(c, e, 2, "color_flow");
    if (!f) fail(c, USX_ERR_INVALID, "color_flow: null descriptor");
    if (!ensemble || !velocity || !power || !variance || !valid)
      fail(c, USX_ERR_INVALID, "color_flow: null data pointer");
    usxd::ColorFlowArgs A;
    A.n = e->n_frames;
    A.n0 = e->n_axis0;
    A.n1 = e->n_axis1;
    A.ka = std::max(0, f->kernel_axial);
    A.kl = std::max(0, f->kernel_lateral);
    A.loupas = f->loupas;
    A.fdem = f->demodulation_frequency;
    A.c = f->speed_of_sound;
    A.prf = f->prf;
    A.dz = f->dz;
    A.coherence_threshold = f->coherence_threshold;
    const float* M = reject_on_device(c, e);
    const float2* d_in = ensemble_on_device(c, e, ensemble, memory, c->dop_in);
    float* lag = static_cast<float*>(ensure(c, c->dop_lag, 5 * n_pix * sizeof(float)));
    unsigned int* pmax = static_cast<unsigned int*>(ensure(c, c->dop_pmax, 16));
    float* d_v = out_on_device(c, velocity, n_pix, memory, c->dop_f0);
    float* d_p = out_on_device(c, power, n_pix, memory, c->dop_f1);
    float* d_w = out_on_device(c, variance, n_pix, memory, c->dop_f2);
    uint8_t* d_ok = out_on_device(c, valid, n_pix, memory, c->dop_valid);
    timed(c, "color_flow", [&] {
      usxd::launch_color_flow(A, d_in, M, lag, d_v, d_p, d_w, d_ok, pmax, f->power_factor, c->stream);
    });
    download(c, velocity, d_v, n_pix, memory, "hipMemcpyAsync(velocity)");
    download(c, power, d_p, n_pix, memory, "hipMemcpyAsync(power)");
    download(c, variance, d_w, n_pix, memory, "hipMemcpyAsync(variance)");
    download(c, valid, d_ok, n_pix, memory, "hipMemcpyAsync(valid)");
    finish(c, memory, "color flow launch");
  });
}

usx_status usx_axial_displacement(usx_ctx* c, const usx_displacement_desc* d,
                                  const usx_c32* before, const usx_c32* after, float* axial,
                                  float* correlation, int memory) {
  return guarded(c, [&] {
    if (!d) fail(c, USX_ERR_INVALID, "estimate_axial_displacement: null descriptor");
    if (d->n_axis0 <= 0 || d->n_axis1 <= 0) fail(c, USX_ERR_INVALID, "ScanGrid dimensions must be > 0");
    if (!before || !after || !axial || !correlation)
      fail(c, USX_ERR_INVALID, "estimate_axial_displacement: null data pointer");
    const size_t n_pix = static_cast<size_t>(d->n_axis0) * d->n_axis1;
    usxd::DisplacementArgs A;
    A.n0 = d->n_axis0;
    A.n1 = d->n_axis1;
    A.ka = std::max(1, d->kernel_axial);
    A.kl = std::max(0, d->kernel_lateral);
    A.lag_max = std::max(0, d->max_lag);
    A.fdem = d->demodulation_frequency;
    A.c = d->speed_of_sound;
    A.dz = d->dz;
    const usx_ensemble_desc one{1, d->n_axis0, d->n_axis1, nullptr};
    const float2* d_a = ensemble_on_device(c, &one, before, memory, c->dop_in);
    const float2* d_b = ensemble_on_device(c, &one, after, memory, c->dop_in2);
    float* d_ax = out_on_device(c, axial, n_pix, memory, c->dop_f0);
    float* d_co = out_on_device(c, correlation, n_pix, memory, c->dop_f1);
    timed(c, "displacement", [&] { usxd::launch_displacement(A, d_a, d_b, d_ax, d_co, c->stream); });
    download(c, axial, d_ax, n_pix, memory, "hipMemcpyAsync(axial)");
    download(c, correlation, d_co, n_pix, memory, "hipMemcpyAsync(correlation)");
    finish(c, memory, "displacement launch");
  });
}

usx_status usx_axial_strain(usx_ctx* c, int n_axis0, int n_axis1, float dz, int fit_length,
                            const float* axial, float* strain, int memory) {
  return guarded(c, [&] {
    if (n_axis0 <= 0 || n_axis1 <= 0) fail(c, USX_ERR_INVALID, "ScanGrid dimensions must be > 0");
    if (!axial || !strain) fail(c, USX_ERR_INVALID, "axial_strain: null data pointer");
    const size_t n_pix = static_cast<size_t>(n_axis0) * n_axis1;
    const float* d_ax = axial;
    if (memory != USX_DEVICE) {
      float* q = static_cast<float*>(ensure(c, c->dop_f1, n_pix * sizeof(float)));
      hip_check(c, hipMemcpyAsync(q, axial, n_pix * sizeof(float), hipMemcpyHostToDevice, c->stream),
                "hipMemcpyAsync(axial)");
      d_ax = q;
    }
    float* d_s = out_on_device(c, strain, n_pix, memory, c->dop_f0);
    const int half = std::max(1, fit_length / 2);
    timed(c, "strain", [&] { usxd::launch_strain(n_axis0, n_axis1, dz, half, d_ax, d_s, c->stream); });
    download(c, strain, d_s, n_pix, memory, "hipMemcpyAsync(strain)");
    finish(c, memory, "strain launch");
  });
}

}  // extern "C"
