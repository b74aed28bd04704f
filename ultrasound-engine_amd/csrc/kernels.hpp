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
};
// counts: 16 ints; lists: (mv_max_bucket_p / 16) * n_pix ints (device scratch).
void launch_mv(const MvArgs& M, const float2* Y, const uint32_t* mask, float2* frame, float* acc,
               int* counts, int* lists, hipStream_t st);
// Largest covariance size bucket (L rounded up to 16) a configuration can produce.
int mv_max_bucket_p(int n_ch, int subarray);
size_t mv_smem_bytes(int n_ch, int range_avg, int subarray, int P);
constexpr int kMvMaxP = 64;

}  // namespace usxd
