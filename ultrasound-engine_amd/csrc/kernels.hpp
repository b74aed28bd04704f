// kernels.hpp -- launch interface between the C-ABI layer (context.hip) and
// the gfx950 kernels (demod.hip, beamform.hip, mv.hip). Internal to the
// library; nothing here crosses the C ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "device_math.hpp"
#include "usx_gpu.h"

namespace usxd {

// Everything the per-pixel kernels read, with device pointers.
struct BfArgs {
  GridView g;
  int n_ch, n_s, n_ev;
  int ev_begin, ev_end;
  float fs, t0, c, fdem;
  float w0;      // (2 pi) * fdem, rounded as the reference forms it
  float lambda, pitch;
  const float* ex;           // element x, n_ch
  const usx_tx_desc* tx;     // n_ev
  float f_number;
  int rx_window;
  float tukey_alpha;
  int transmit_weighting;
  int n_xcd;                 // XCDs of the device: the tile remap gives each a contiguous range
  float inv_c;               // 1 / speed of sound, correctly rounded (divisions by c: div_r)
  float cos1_lo, cos1_hi;    // the C library's cosf(x) is exactly 1.0f for x <= lo and x >= hi
                             // on [0, 2 pi]: where the Hann receive weight is exactly 0
};

// demod.hip
void launch_demod(int n_traces, int n_in, int n_out, int dec, int n_taps, const float* taps,
                  const float2* rotator, const float* rf, float2* iq, hipStream_t st);

// beamform.hip
void launch_fused(const BfArgs& A, const void* data, float2* frame, int ev_first, int n_slots,
                  int select, int accumulate, int comb, bool cf, bool iq, hipStream_t st);
void launch_channels(const BfArgs& A, const void* data, float2* Y, uint32_t* mask, int ev_first,
                     int n_slots, int select, int first, int rect, bool iq, hipStream_t st);
void launch_combine(int n_pix, int n_ch, const float2* Y, const uint32_t* mask, float2* frame,
                    float* acc, int incoherent, int comb, bool cf, hipStream_t st);
void launch_incoherent_finish(int n_pix, const float* acc, float2* frame, hipStream_t st);
void launch_fill_zero(float* p, size_t n, hipStream_t st);

// mv.hip
struct MvArgs {
  int n0, n1, n_ch;
  int subarray;
  float loading;
  int range_avg;
  int cf;
  int incoherent;
  int precision;   // 0 FP32, 1 BF16, 2 FP16 covariance inputs
  int refine;      // one step of iterative refinement (FP64 residual)
  // Optional outputs (coherent compounding only; nullptr when unused):
  // weights[pix][ch] the effective per-channel weights g with
  // out = sum_ch g_ch Y_ch before the coherence factor (the caller zeroes
  // the buffer: only active channels are written); status[pix] 0 solved,
  // 1 fewer than 4 active channels (DAS), 2 solve failed (DAS), 3 no
  // active channel.
  float2* weights;
  int* status;
  int occupancy;   // persistent workgroups per CU per bucket (0: as many as LDS allows, <= 32)
  int n_cu;        // compute units of the device (sizes the persistent grids)
  int n_xcd;       // XCDs of the device (the persistent grids give each a contiguous run of pixels)
};
// Whether the fast MV kernel (FP32, one matrix row per lane) serves the size
// buckets P <= 64 of this configuration; the general kernel serves the rest.
__host__ __device__ inline bool mv_fast_config(const MvArgs& M) {
  return M.precision == 0 && M.weights == nullptr && M.range_avg >= 0 && M.range_avg <= 2 &&
         M.n_ch <= 256;
}
// Pixel counts: counts[b] for size bucket b (matrices padded to
// P = 16 (b + 1)), counts[kFullList + b] for the fast kernel's pixels of
// L = P - 1 and P in bucket b (b < 4), which sit at the back of the bucket's
// list. counts: 16 ints; lists: (mv_max_bucket_p / 16) * n_pix ints (device scratch).
constexpr int kFullList = 8;
// Entries of the fast kernel's lists (grids of at most 2^21 pixels): the
// pixel in the low kRunShift bits and, when its active channels form one
// contiguous run, the run's first channel + 1 above them (0: not a run).
constexpr int kRunShift = 21;
constexpr int kPixMask = (1 << kRunShift) - 1;
void launch_mv(const MvArgs& M, const float2* Y, const uint32_t* mask, float2* frame, float* acc,
               int* counts, int* lists, hipStream_t st);
// Largest covariance size bucket (L rounded up to 16) a configuration can produce.
int mv_max_bucket_p(int n_ch, int subarray);
// Dynamic LDS per wavefront of the general MV kernel for bucket P (refine:
// the full matrix, else its packed lower triangle).
size_t mv_smem_bytes(int n_ch, int range_avg, int subarray, int P, bool refine);
constexpr int kMvMaxP = 128;

// apod.hip: learned apodization MLP (256 -> 3 x 256 GELU -> 256, bf16 MFMA).
constexpr int kApodFeatures = 256;
constexpr int kApodLayers = 4;
constexpr int kApodParams = kApodLayers * kApodFeatures * kApodFeatures + kApodLayers * kApodFeatures;
constexpr int kApodChannels = kApodFeatures / 2;
// Largest frame for inference: its channel vectors (1 KiB a pixel, up to the
// last 128-pixel tile) stay below 4 GiB of 32-bit buffer offsets.
constexpr long long kApodMaxPixels = (1ll << 22) - 128;

struct ApodNet {
  float* master;   // kApodParams: W_0..W_3 ([in][out]), then b_0..b_3
  float* m;        // Adam moments
  float* v;
  uint16_t* wf;    // bf16 forward fragments, 4 x 65536
  uint16_t* wb;    // bf16 backward fragments, 4 x 65536
  float* bias;     // = master + 4 * 65536
};
struct ApodData {  // training rows
  uint16_t* x;     // [n][256] bf16 network input
  float* g;        // [n][256] minimum-variance weights (re, im)
  uint32_t* mask;  // [n][4] active channels
  float* sw;       // [n] 1 / n_active, or 0 (excluded)
};
struct ApodStep {
  int batch;       // multiple of 32 (and of 256 above 256)
  long n_data;
  int random;      // 1: rows drawn by hash(seed, step, b); 0: rows 0..n_data-1, then padding
  uint64_t seed;
  int step;
  float grad_scale;  // 2 / rows (rows = batch, or the valid rows of a padded batch)
  float loss_scale;  // 1 / rows
};
struct ApodWork {
  uint16_t* h;     // [4][batch][256] layer inputs (bf16)
  float* z;        // [3][batch][256] pre-activations
  uint16_t* dz;    // [4][batch][256] layer output gradients (bf16)
  float* partial;  // [4][splits][65536] weight-gradient partials
  float* db;       // [4][32][256] bias-gradient partials
  float* loss_part;  // [batch / 16] loss partials
};
struct ApodAdam {
  float lr, beta1, beta2, eps, weight_decay;
  float bc1, bc2;  // 1 - beta^t
};
size_t apod_infer_smem();
int apod_splits(int batch);
void launch_apod_pack(const ApodNet& N, hipStream_t st);
// n_cu: compute units of the device (the kernel is persistent, one workgroup per CU).
void launch_apod_infer(const ApodNet& N, int n_pix, int n_ch, const float2* Y,
                       const uint32_t* mask, float2* frame, float* acc, int cf, int incoherent,
                       int n_cu, hipStream_t st);
void launch_apod_gather(int n_list, const int* pixels, int n_ch, const float2* Y,
                        const uint32_t* mask, const float2* weights, const int* status,
                        const ApodData& D, long dst, hipStream_t st);
// One training step (or, with grads != nullptr, only the gradient of the
// batch into grads[kApodParams]); the step's loss goes to *loss_out (device).
void launch_apod_step(const ApodNet& N, const ApodData& D, const ApodStep& S, const ApodWork& W,
                      const ApodAdam& opt, float* grads, float* loss_out, hipStream_t st);

// sim.hip: point-scatterer forward model.
constexpr int kSimChunk = 512;  // scatterers per partial trace
struct SimArgs {
  int n_elements, n_samples;
  float fs, t0, prf, c, fs_os, wavelength, element_width;
  const float* ex;        // element x, n_elements
  const float* delays;    // [event][element] firing delays
  const int* firing;      // [event][element] firing element indices (first n_firing[event])
  const float* fire_w;    // [event][element] their apodization weights
  const int* n_firing;    // [event]
  const float* pulse;
  int pulse_n;
  float pulse_half;
  float atten;            // dB per metre at the centre frequency, 0 = off
  float max_window;       // s
  float trim_amp;
  float noise_amp;        // 0 = no receive noise
  int inc_cap;            // longest incident waveform (oversampled samples)
  int spreading, directivity, motion, full_diffraction;
};
struct SimIncident {
  int* live;
  float* t0;
  int* n;
  float4* pos;            // position after motion, amplitude
  float* wave;            // [scatterer][inc_cap]
};
size_t sim_incident_smem(const SimArgs& S);
size_t sim_echo_smem(const SimArgs& S);
int sim_chunks(int n_scat);
// One event: rf (n_elements * n_samples) of event ev; z the event's
// standard normals (nullptr: no noise); blockmax: ceil(n_el n_s / 256) floats.
void launch_sim_event(const SimArgs& S, int ev, const float* scat, int n_scat, const SimIncident& I,
                      float* partial, float* blockmax, const float* z, float* rf, hipStream_t st);

// display.hip: envelope, TGC, log compression.
struct DisplayArgs {
  GridView g;
  int apply_tgc;
  float attenuation;   // dB / (cm MHz)
  float f_mhz;         // TGC frequency, MHz
  float dynamic_range_db;
  float gain_db;
};
int display_blocks(const DisplayArgs& D, bool rf);
// env (n_pix) = envelope (|IQ|, or the analytic magnitude along axis1 when rf)
// times the TGC gain; blockmax[display_blocks] its block maxima.
void launch_display_envelope(const DisplayArgs& D, const float2* frame, bool rf, float* env,
                             float* blockmax, hipStream_t st);
// image = log compression of env against *ref (device; <= 0: the maximum of blockmax).
void launch_display_compress(const DisplayArgs& D, const float* env, const float* blockmax,
                             int n_blocks, const float* ref, float* image, hipStream_t st);

// doppler.hip: clutter filter, colour flow, power Doppler, axial displacement
// and strain on beamformed frames. Ensembles are frame-major, n frames of
// n_pix complex pixels; M is the row-major n x n wall-filter projector or
// nullptr (no filter).
struct ColorFlowArgs {
  int n, n0, n1;
  int ka, kl;              // window half sizes (axial, lateral), >= 0
  int loupas;
  float fdem, c, prf, dz;
  float coherence_threshold;
};
struct DisplacementArgs {
  int n0, n1;
  int ka, kl, lag_max;     // ka >= 1, kl >= 0, lag_max >= 0
  float fdem, c, dz;
};
void launch_wall_filter(int n, size_t n_pix, const float2* in, const float* M, float2* out,
                        hipStream_t st);
void launch_power_doppler(int n, size_t n_pix, const float2* in, const float* M, float* p,
                          hipStream_t st);
// lag: 5 * n_pix floats of scratch; pmax_bits: one word of scratch.
void launch_color_flow(const ColorFlowArgs& A, const float2* in, const float* M, float* lag,
                       float* velocity, float* power, float* variance, uint8_t* valid,
                       unsigned int* pmax_bits, float factor, hipStream_t st);
void launch_displacement(const DisplacementArgs& A, const float2* before, const float2* after,
                         float* axial, float* correlation, hipStream_t st);
void launch_strain(int n0, int n1, float dz, int half, const float* axial, float* strain,
                   hipStream_t st);

}  // namespace usxd
