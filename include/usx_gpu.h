/* usx_gpu.h -- C ABI of the MI355X (gfx950) imaging path.
 *
 * This is the drop-in boundary below the reference's C++ API. The reference
 * has no C ABI of its own: its boundary is pybind11 over C++ classes
 * (bindings/pymodule.cpp). This library is what the C++17 host layer of this
 * repository (ultrasound-engine_amd/core) calls from the same three C++
 * entry points the reference exposes, which bindings/pymodule.cpp binds
 * unchanged:
 *
 *   usx_demodulate     replaces  ChannelData demodulate(const ChannelData&, DemodConfig)
 *                                core/include/usx/demodulate.hpp:45,
 *                                core/src/demodulate.cpp:36-121
 *   usx_beamform       replaces  Frame Beamformer::beamform(const ChannelData&,
 *                                const ScanGrid&, int, int) const
 *                                core/include/usx/beamformer.hpp:114-115,
 *                                core/src/beamformer.cpp:297-549
 *   usx_process_frame  replaces  Frame ImagingPipeline::process_frame(const ChannelData&)
 *                                core/include/usx/pipeline.hpp:52,
 *                                core/src/pipeline.cpp:25-46
 *
 * and, downstream of the frames, the estimators of core/include/usx/doppler.hpp
 * (usx_wall_filter, usx_color_flow, usx_power_doppler, usx_axial_displacement,
 * usx_axial_strain; each cites its reference function).
 *
 * Conventions
 *   - Plain pointers and sizes only; no C++ or torch types cross the ABI.
 *   - Channel data is [event][channel][sample], sample fastest
 *     (ChannelData::index, types.hpp:206-209). Frames are [axis0][axis1],
 *     axis1 fastest (Frame::at, types.hpp:278-279).
 *   - Complex values are interleaved float pairs (usx_c32), bit-compatible
 *     with std::complex<float> and numpy complex64.
 *   - `memory` selects the address space of the data pointers: USX_HOST
 *     means caller-owned host buffers and a synchronous call (upload, launch,
 *     download, stream sync); USX_DEVICE means device buffers and an
 *     asynchronous call enqueued on the context's stream.
 *   - Descriptor arrays (taps, rotator, element_x, transmits, axes) are
 *     always host memory, read during the call.
 *   - No exception crosses the ABI. Every call returns a usx_status; the
 *     message of the last failure is usx_ctx_last_error(ctx). The C++ layer
 *     maps USX_ERR_INVALID to std::invalid_argument and the rest to
 *     std::runtime_error, so Python sees ValueError / RuntimeError as with the
 *     reference (pymodule.cpp maps them the same way).
 *   - One context per pipeline or beamformer object. Calls on one context
 *     must not overlap (the reference's ImagingPipeline is not re-entrant
 *     either, pipeline.cpp:27, 43-44); separate contexts are independent.
 */
#ifndef USX_GPU_H
#define USX_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
  USX_OK = 0,
  USX_ERR_INVALID = 1,   /* a descriptor or size is invalid */
  USX_ERR_RUNTIME = 2,   /* HIP runtime failure (message says which call) */
  USX_ERR_NO_DEVICE = 3  /* no gfx950 device, or the requested ordinal is absent */
} usx_status;

enum { USX_HOST = 0, USX_DEVICE = 1 };

typedef struct {
  float re, im;
} usx_c32;

typedef struct usx_ctx usx_ctx;

/* ---- context ----------------------------------------------------------- */

/* Create a context on HIP device `device`: one HIP stream (default
 * priority), pooled device buffers grown on demand, the last error string. */
usx_status usx_ctx_create(int device, usx_ctx** out);
usx_status usx_ctx_destroy(usx_ctx* ctx);
const char* usx_ctx_last_error(const usx_ctx* ctx);
/* The context's hipStream_t, for callers that enqueue their own work. */
void* usx_ctx_stream(usx_ctx* ctx);
usx_status usx_ctx_synchronize(usx_ctx* ctx);
/* Device ordinal, or -1 for a null context. */
int usx_ctx_device(const usx_ctx* ctx);
/* Library build tag (gfx target and compile options), for diagnostics. */
const char* usx_build_info(void);

/* Timing and profiling (no reference counterpart; the reference times stages
 * with std::chrono, pipeline.cpp:6-11). Events are recorded on the context
 * stream, so they bracket exactly the work enqueued between them. */
usx_status usx_ctx_event_record(usx_ctx* ctx, int slot);   /* slot 0..7 */
usx_status usx_ctx_event_elapsed(usx_ctx* ctx, int slot_a, int slot_b, float* ms);
/* When enabled, every kernel launch is bracketed by events; the report
 * (after a synchronize) is JSON: {"kernel": [launches, total_ms], ...}
 * accumulated since profiling was enabled. */
usx_status usx_ctx_profile(usx_ctx* ctx, int enable);
usx_status usx_ctx_profile_report(usx_ctx* ctx, char* buf, size_t len);

/* ---- demodulation (demodulate.cpp:36-121) ----------------------------- */

typedef struct {
  int n_traces;            /* n_events * n_channels */
  int n_in;                /* RF samples per trace */
  int n_out;               /* (n_in + decimation - 1) / decimation */
  int decimation;          /* >= 1 */
  int n_taps;              /* odd, as design_lowpass returns */
  const float* taps;       /* host, n_taps: design_lowpass(...) */
  const usx_c32* rotator;  /* host, n_in: exp(-j 2 pi fdem (t0 + n/fs)) as
                              the reference tabulates it (demodulate.cpp:72-76) */
} usx_demod_desc;

/* rf: n_traces * n_in floats -> iq: n_traces * n_out complex. */
usx_status usx_demodulate(usx_ctx* ctx, const usx_demod_desc* d, const float* rf, usx_c32* iq,
                          int memory);

/* ---- beamforming (beamformer.cpp:297-549) ----------------------------- */

/* One transmit event with everything the per-pixel delay and illumination
 * laws need that does not depend on the pixel, precomputed on the host with
 * the reference's own float expressions (geometry.cpp:82-109,
 * beamformer.cpp:238-288). */
typedef struct {
  int type;              /* 0 Focused, 1 PlaneWave, 2 Diverging */
  float sin_a, cos_a;    /* sin/cos of the steering angle (plane wave) */
  float tan_a;           /* tan of the steering angle (plane wave illumination) */
  float origin_x;        /* beam origin on the face (Transmit::origin.x) */
  float focus[3];        /* focal point / virtual source */
  float axis_len;        /* |focus - origin| (focused) */
  float axis_dir[3];     /* (focus - origin) / axis_len (focused) */
  float origin_focus;    /* |origin - focus| (diverging) */
  int extent_valid;      /* firing aperture not empty */
  float x_lo, x_hi;      /* lateral extent of the firing aperture */
  float aperture;        /* max(x_hi - x_lo, pitch) */
} usx_tx_desc;

typedef struct {
  /* channel data */
  int n_events;            /* events present in the data */
  int n_channels;
  int n_samples;
  int is_iq;               /* 1: data is usx_c32 IQ; 0: data is float RF */
  float sampling_frequency, t0, speed_of_sound;
  float demodulation_frequency;  /* > 0 for IQ; ignored for RF */
  float wavelength;        /* speed_of_sound / center_frequency */
  float pitch;
  const float* element_x;  /* host, n_channels (Probe::element_position) */
  const usx_tx_desc* tx;   /* host, n_events */
  int event_begin, event_end;    /* already clamped, 0 <= begin <= end <= n_events */
  /* grid (ScanGrid::cartesian / sector, types.cpp:66-115) */
  int grid_type;           /* 0 Cartesian, 1 Sector */
  int n_axis0, n_axis1;
  const float* axis0;      /* host, n_axis0 */
  const float* axis1;      /* host, n_axis1 */
  const float* axis0_sin;  /* host, n_axis0, sector only: sin(theta) */
  const float* axis0_cos;  /* host, n_axis0, sector only: cos(theta) */
  float apex_x, apex_z;
  /* BeamformerConfig (beamformer.hpp:66-106) */
  float f_number;
  int rx_window;           /* 0 Rect 1 Hann 2 Hamming 3 Tukey 4 Blackman */
  float tukey_alpha;
  int combiner;            /* 0 DAS 1 DMAS 2 MV; 3 learned apodization (no reference
                              counterpart: needs 128 channels and a network, see usx_apod_*) */
  int compounding;         /* 0 Coherent 1 Incoherent 2 Select */
  int transmit_weighting;
  int coherence_factor;
  int mv_subarray;
  float mv_diagonal_loading;
  int mv_range_average;
  /* additions with no reference counterpart (default 0 = reference behaviour) */
  int mv_precision;        /* 0 FP32 covariance, 1 BF16 inputs, 2 FP16 inputs (FP32 accumulate) */
  int mv_refine;           /* 1: one step of iterative refinement, FP64 residual */
  /* launch-shape tunables (no reference counterpart; 0 = default), as
   * usx_autotune picks them */
  int tile_slots;          /* transmit events per beamforming launch, 1..16 (default 16) */
  int tile_mv_occupancy;   /* persistent MV workgroups per CU (default: as LDS allows) */
} usx_bf_desc;

/* data: n_events * n_channels * n_samples (float RF or usx_c32 IQ);
 * frame: n_axis0 * n_axis1 complex. */
usx_status usx_beamform(usx_ctx* ctx, const usx_bf_desc* b, const void* data, usx_c32* frame,
                        int memory);

/* Stage times of the last call, from HIP events on the context stream
 * (PipelineStats fields, pipeline.hpp:38-44). */
typedef struct {
  double demodulate_ms;
  double beamform_ms;
  double total_ms;
  double postprocess_ms;   /* usx_process_display only */
} usx_stage_times;

/* RF in, frame out: demodulate (d) then beamform (b, whose n_samples,
 * sampling_frequency, demodulation_frequency and is_iq describe the IQ that
 * d produces). times may be NULL. */
usx_status usx_process_frame(usx_ctx* ctx, const usx_demod_desc* d, const usx_bf_desc* b,
                             const float* rf, usx_c32* frame, int memory,
                             usx_stage_times* times);

/* ---- allocation-free frame (no reference counterpart) -------------------
 *
 * For callers that cannot allocate device memory: code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc. The same
 * work as usx_process_frame with memory = USX_DEVICE (RF in device memory ->
 * demodulate -> beamform -> frame in device memory, the same kernels and the
 * same results), but no call allocates: the IQ, the descriptor tables and the
 * beamforming scratch (channel vectors, active masks, minimum-variance work
 * lists) are laid out in one fixed array of USX_NOALLOC_SCRATCH_BYTES per GPU,
 * which the library holds from its first context on that GPU until the
 * process ends. A frame that needs more (the minimum-variance channel vectors
 * of large grids, for one) returns USX_ERR_INVALID with the bytes it needs.
 * Asynchronous on the context stream; the host descriptor arrays are copied
 * before the call returns. Every context on one GPU shares that GPU's array,
 * which keeps the last call's contents: calls that overlap on one GPU (from
 * two contexts, or enqueued on two streams) race. */
#define USX_NOALLOC_SCRATCH_BYTES (64u * 1024u * 1024u)
usx_status usx_process_frame_noalloc(usx_ctx* ctx, const usx_demod_desc* d, const usx_bf_desc* b,
                                     const float* rf, usx_c32* frame);

/* ---- display tail (pipeline.cpp:48-79, postprocess.cpp:12-117) --------- */

typedef struct {
  int apply_tgc;                 /* PipelineConfig.apply_tgc */
  float attenuation_db_cm_mhz;   /* Medium.attenuation */
  float center_frequency_hz;     /* TGC frequency: the frame's demodulation frequency, or 5 MHz */
  float dynamic_range_db, gain_db;
  float reference_smoothing;     /* PipelineConfig.reference_smoothing */
} usx_display_desc;

/* Frame -> display image in [0, 1]: envelope (|IQ|; for an RF frame, rf = 1,
 * the analytic-signal magnitude along axis1), TGC, log compression against
 * the smoothed reference level, which *reference carries between frames
 * (in/out, host; 0 = none yet) and is updated as ImagingPipeline::to_display
 * does (replaces the body of to_display, pipeline.cpp:48-70). b supplies the
 * grid as for usx_beamform. frame, image: n_axis0 * n_axis1. */
usx_status usx_display(usx_ctx* ctx, const usx_bf_desc* b, const usx_display_desc* p,
                       const usx_c32* frame, int rf, float* image, float* reference, int memory);
/* RF in, display image out (ImagingPipeline::process_display,
 * pipeline.cpp:72-79): usx_process_frame then usx_display, with the frame kept
 * in device memory. times may be NULL. */
usx_status usx_process_display(usx_ctx* ctx, const usx_demod_desc* d, const usx_bf_desc* b,
                               const usx_display_desc* p, const float* rf, float* image,
                               float* reference, int memory, usx_stage_times* times);

/* ---- Doppler and speckle tracking (doppler.cpp) ------------------------ *
 *
 * On frames beamformed from IQ. An ensemble is n_frames frames of
 * n_axis0 * n_axis1 complex pixels laid one after the other (frame-major,
 * axis1 fastest within a frame). `reject` is the wall filter as the row-major
 * n_frames x n_frames projector (I - P) that removes the least-squares fit of
 * a low-order polynomial in slow time, built on the host exactly as
 * polynomial_reject_matrix (doppler.cpp:15-49) builds it; NULL means no
 * filter (WallFilter::None). */
typedef struct {
  int n_frames;            /* >= 1 */
  int n_axis0, n_axis1;
  const float* reject;     /* host, n_frames * n_frames, or NULL */
} usx_ensemble_desc;

/* apply_wall_filter (doppler.cpp:51-75): out = reject applied along slow
 * time at every pixel. ensemble, out: n_frames * n_axis0 * n_axis1 complex,
 * separate buffers (they must not overlap). */
usx_status usx_wall_filter(usx_ctx* ctx, const usx_ensemble_desc* e, const usx_c32* ensemble,
                           usx_c32* out, int memory);

/* power_doppler (doppler.cpp:178-186): power[k] = sum over the filtered
 * frames of |z|^2. power: n_axis0 * n_axis1 floats. */
usx_status usx_power_doppler(usx_ctx* ctx, const usx_ensemble_desc* e, const usx_c32* ensemble,
                             float* power, int memory);

/* DopplerConfig (doppler.hpp:41-79) and the frame facts color_flow reads. */
typedef struct {
  float prf, speed_of_sound;
  float demodulation_frequency;  /* of the ensemble's first frame */
  float dz;                      /* |axis1[1] - axis1[0]|, 1e-4 for a single depth */
  int loupas;
  int kernel_axial, kernel_lateral;   /* window half sizes; negative counts as 0 */
  float coherence_threshold;
  float power_factor;            /* 10^(power_threshold_db / 10), as std::pow(10.f, db / 10.f) */
} usx_color_flow_desc;

/* color_flow (doppler.cpp:77-176) after the wall filter of e: per pixel the
 * axial velocity (m/s, + away from the probe), the post-filter power, the
 * normalised spectral width and the valid flag (1: flow), each
 * n_axis0 * n_axis1. Needs n_frames >= 2. */
usx_status usx_color_flow(usx_ctx* ctx, const usx_ensemble_desc* e, const usx_color_flow_desc* f,
                          const usx_c32* ensemble, float* velocity, float* power, float* variance,
                          uint8_t* valid, int memory);

typedef struct {
  int n_axis0, n_axis1;
  float speed_of_sound;
  float demodulation_frequency;  /* of the first frame */
  float dz;                      /* as for color flow */
  int kernel_axial, kernel_lateral, max_lag;  /* as passed; clamped as the reference does */
} usx_displacement_desc;

/* estimate_axial_displacement (doppler.cpp:188-273): axial displacement (m)
 * and |normalised correlation| per pixel from two frames of one grid. */
usx_status usx_axial_displacement(usx_ctx* ctx, const usx_displacement_desc* d,
                                  const usx_c32* before, const usx_c32* after, float* axial,
                                  float* correlation, int memory);

/* axial_strain (doppler.cpp:275-307): least-squares slope of `axial` against
 * depth over fit_length / 2 (at least 1) depths either side. */
usx_status usx_axial_strain(usx_ctx* ctx, int n_axis0, int n_axis1, float dz, int fit_length,
                            const float* axial, float* strain, int memory);

/* ---- minimum-variance channel weights (no reference counterpart) ------- */

/* The minimum-variance frame (b->combiner must be 2, coherent or select
 * compounding) together with each pixel's effective channel weights g, such
 * that, before the coherence factor, out = sum_ch g[ch] * Y[ch] over the
 * rectangular-aperture channel vector Y (beamformer.cpp:202-213 rewritten
 * per channel: g_c = n / (n_sub sum w) * sum over the subarrays covering c
 * of conj(w_a)); inactive channels get 0 and DAS fallbacks get 1.
 * weights: n_axis0 * n_axis1 * n_channels complex; status (may be NULL):
 * n_axis0 * n_axis1 ints, 0 solved, 1 fewer than 4 active channels,
 * 2 solve failed (both DAS), 3 no active channel. */
usx_status usx_mv_weights(usx_ctx* ctx, const usx_bf_desc* b, const void* data, usx_c32* frame,
                          usx_c32* weights, int* status, int memory);

/* ---- learned apodization (no reference counterpart) ---------------------
 *
 * A 256 -> 256 -> 256 -> 256 -> 256 MLP (tanh-form GELU after the first three
 * layers; bf16 weights and activations on MFMA, FP32 accumulation) maps a
 * pixel's rectangular-aperture channel vector (128 channels, normalised to
 * unit mean power and the phase of its sum) to channel weights
 * 1 + o_c; combiner 3 forms out = sum_c (1 + o_c) Y_c over the active
 * channels (DAS below 4 active channels, as minimum variance does). It is
 * trained on the minimum-variance weights of usx_mv_weights. Parameters are
 * FP32, USX_APOD_PARAMS floats: W_0..W_3 (each [in][out], 256 x 256), then
 * b_0..b_3 (256 each). Features are (re, im) interleaved per channel. */
#define USX_APOD_CHANNELS 128
#define USX_APOD_FEATURES 256
#define USX_APOD_PARAMS (4 * 256 * 256 + 4 * 256)

/* Load parameters (host, USX_APOD_PARAMS floats); resets the optimizer. */
usx_status usx_apod_set_params(usx_ctx* ctx, const float* params);
usx_status usx_apod_get_params(usx_ctx* ctx, float* params);

/* Training set held on the device: `capacity` rows, emptied. */
usx_status usx_apod_dataset_reserve(usx_ctx* ctx, size_t capacity);
usx_status usx_apod_dataset_size(usx_ctx* ctx, size_t* n);
/* Append the listed pixels of one frame: run the minimum-variance path on
 * the data (RF through d when d is not NULL, else IQ as b describes it;
 * b->combiner must be 2), then store each listed pixel's network input,
 * weights, active mask and sample weight (1/n_active; 0 for pixels that
 * fell back to DAS). pixels: host, n_pixels indices into the frame
 * (axis0 * n_axis1 + axis1). frame (host, may be NULL) receives the MV frame. */
usx_status usx_apod_dataset_append(usx_ctx* ctx, const usx_demod_desc* d, const usx_bf_desc* b,
                                   const void* data, int memory, const int* pixels, int n_pixels,
                                   usx_c32* frame);

typedef struct {
  int n_steps;
  int batch;               /* multiple of 256 */
  float lr, beta1, beta2, eps, weight_decay;
  unsigned long long seed; /* minibatch rows are drawn by hash(seed, step, b) */
} usx_apod_train_desc;
/* n_steps Adam steps on minibatches of the device training set. losses
 * (host, may be NULL) receives each step's loss. */
usx_status usx_apod_train(usx_ctx* ctx, const usx_apod_train_desc* t, float* losses);

/* Loss and gradient of one batch given on the host, with the current
 * parameters (no update). x: n * 256 bf16 bit patterns; g: n * 256 floats;
 * mask: n * 4 words; sw: n sample weights; grads: USX_APOD_PARAMS floats.
 * n <= 4096. */
usx_status usx_apod_gradient(usx_ctx* ctx, int n, const uint16_t* x, const float* g,
                             const uint32_t* mask, const float* sw, float* loss, float* grads);

/* ---- point-scatterer simulator (simulator.cpp:213-419) ------------------ */

/* Everything Simulator::simulate reads besides the phantom, precomputed on
 * the host with the reference's expressions (pulse(), element positions,
 * transmit delays, attenuation in dB/m at the centre frequency, the trim
 * and noise amplitudes). */
typedef struct {
  int n_elements, n_samples, n_events;
  float sampling_frequency, t0, prf, speed_of_sound, wavelength, element_width;
  int oversampling;
  const float* element_x;         /* host, n_elements */
  const float* delays;            /* host, n_events * n_elements (Transmit::delays) */
  const float* apodization;       /* host, n_events * n_elements; 0 = element does not fire */
  const int* has_apodization;     /* host, n_events; 0: every element fires with weight 1 */
  int pulse_n;
  const float* pulse;             /* host, pulse_n (Simulator::pulse) */
  float pulse_half;               /* its half duration, s */
  float attenuation;              /* dB per metre at the centre frequency; 0 = off */
  float max_window;               /* max_incident_window_us * 1e-6 */
  float trim_amp;                 /* 10^(incident_trim_db / 20) */
  int spreading, directivity, motion, full_diffraction;
  float noise_amp;                /* 10^(noise_db / 20), 0 = no noise */
  const float* noise;             /* host, n_events * n_elements * n_samples standard
                                     normals: std::normal_distribution<float>(0, 1) drawn
                                     from std::mt19937(noise_seed + event); NULL if no noise */
} usx_sim_desc;

/* scatterers: host, n_scatterers * 7 floats (x, y, z, amplitude, vx, vy, vz);
 * rf: n_events * n_elements * n_samples floats. Matches the host simulator
 * to float rounding (device powf/sinf in the amplitudes), deterministically. */
usx_status usx_simulate(usx_ctx* ctx, const usx_sim_desc* s, int n_scatterers,
                        const float* scatterers, float* rf, int memory);

/* ---- device buffers (no reference counterpart) --------------------------
 * Plain device memory for callers without a device allocator of their own
 * (the frame-sequence binding). Copies are synchronous; usx_buffer_zero is a
 * fill kernel on the context stream, so it may be captured in a graph. */
usx_status usx_buffer_alloc(usx_ctx* ctx, size_t bytes, void** dev);
usx_status usx_buffer_free(usx_ctx* ctx, void* dev);
usx_status usx_buffer_upload(usx_ctx* ctx, void* dev, const void* host, size_t bytes);
usx_status usx_buffer_download(usx_ctx* ctx, void* host, const void* dev, size_t bytes);
usx_status usx_buffer_zero(usx_ctx* ctx, void* dev, size_t bytes);

/* ---- frame sequences and launch-shape tuning (no reference counterpart) --- */

/* A fixed sequence of n frames, frame f: RF (device, rf[f]) -> demodulate
 * (d[f]) -> beamform (b[f]) -> frame (device, frames[f]), each with its own
 * probe, grid, transmit count and configuration (the uneven batched stream).
 * Creation uploads every frame's descriptor tables once, sizes the scratch
 * for the largest frame, runs the sequence once and captures it into one
 * HIP graph. The pointers are bound at creation. */
typedef struct usx_plan usx_plan;
usx_status usx_plan_create(usx_ctx* ctx, int n, const usx_demod_desc* d, const usx_bf_desc* b,
                           const float* const* rf, usx_c32* const* frames, usx_plan** out);
/* Enqueue the sequence on the context stream: the graph (use_graph = 1) or
 * the same kernels launched one by one (0). Asynchronous. */
usx_status usx_plan_launch(usx_ctx* ctx, usx_plan* plan, int use_graph);
usx_status usx_plan_destroy(usx_ctx* ctx, usx_plan* plan);

/* Time the candidate launch shapes of one frame (RF and frame in device
 * memory; transmit events per beamforming launch 4/8/16, and for MV the
 * persistent occupancy 0/4/8/16 workgroups per CU), each over `reps`
 * repetitions after a warm-up, and write the fastest into b->tile_slots and
 * b->tile_mv_occupancy (best_ms, may be NULL: its time per frame). */
usx_status usx_autotune(usx_ctx* ctx, const usx_demod_desc* d, usx_bf_desc* b, const float* rf,
                        usx_c32* frame, int reps, float* best_ms);

#ifdef __cplusplus
}
#endif
#endif /* USX_GPU_H */
