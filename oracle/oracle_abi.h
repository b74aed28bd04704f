/* oracle_abi.h -- plain-C descriptors shared by the two CPU checkers under oracle/.
 *
 * TEST INFRASTRUCTURE ONLY. Nothing in the product links this header or the
 * libraries built from it; only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py load them, and only as the checker.
 *
 * Two libraries implement the functions declared here:
 *   liboracle_usx.so      -- this repository's own C++17 restatement of the
 *                            reference's imaging path (usx_oracle.cpp), with an
 *                            FP32 exact-order mode and an FP64 mode for MV;
 *   _ref/libusx_ref.so    -- the reference's own core/src sources compiled from
 *                            /root/reference by oracle/Makefile, wrapped by
 *                            ref_shim.cpp (built only where the sources exist).
 * Both export the same oc_* entry points, so a test can run one case through
 * either and compare.
 *
 * Every struct mirrors a reference type field-for-field, in the reference's
 * units (SI): Probe/Medium/Acquisition (core/include/usx/types.hpp:68-134),
 * Transmit (types.hpp:159-173), ScanGrid factories (types.cpp:66-100),
 * DemodConfig (demodulate.hpp:23-40), BeamformerConfig (beamformer.hpp:66-106),
 * SimulatorConfig (simulator.hpp:77-116).
 */
#ifndef USX_ORACLE_ABI_H
#define USX_ORACLE_ABI_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  /* Probe */
  int n_elements;
  float pitch, element_width, elevation_height, center_frequency, bandwidth;
  /* Medium */
  float speed_of_sound, attenuation, density;
  /* Acquisition */
  float sampling_frequency;
  int n_samples;
  float t0, prf;
} oc_setup;

typedef struct {
  int type;          /* 0 Focused, 1 PlaneWave, 2 Diverging */
  float angle;
  float focus[3];
  float origin[3];
  int n_delays;      /* 0 (empty vector) or n_elements */
  const float* delays;
  int n_apod;        /* 0 (empty vector) or n_elements */
  const float* apodization;
} oc_transmit;

typedef struct {
  int type;          /* 0 Cartesian, 1 Sector */
  float a0_min, a0_max;
  int n0;
  float a1_min, a1_max;
  int n1;
  float apex[3];
  /* A subset of the lateral lines of that grid, with the full grid's axis
   * values: lines line_first, line_first + line_step, ... below n0
   * (line_step 0: every line). Frames then hold that many lines. */
  int line_first, line_step;
} oc_grid;

typedef struct {
  float demodulation_frequency, cutoff_fraction;
  int filter_taps, decimation;
} oc_demod;

typedef struct {
  float f_number;
  int rx_window;     /* Window enum: 0 Rect 1 Hann 2 Hamming 3 Tukey 4 Blackman */
  float tukey_alpha;
  int combiner;      /* 0 DAS 1 DMAS 2 MV */
  int compounding;   /* 0 Coherent 1 Incoherent 2 Select */
  int transmit_weighting;
  int coherence_factor;
  int mv_subarray;
  float mv_diagonal_loading;
  int mv_range_average;
} oc_bf;

typedef struct {
  float pulse_cycles;
  int tx_window;
  float tx_f_number;
  int oversampling;
  int full_transmit_diffraction;
  float incident_trim_db, max_incident_window_us;
  int apply_attenuation, apply_directivity, apply_spreading;
  float noise_db;
  uint32_t noise_seed;
  int enable_motion;
} oc_simcfg;

/* Status: 0 ok, 1 invalid argument (message in oc_last_error), 2 other error. */
const char* oc_last_error(void);
const char* oc_flavour(void);   /* "port" or "reference" */

/* ScanGrid factory: axes (n0 + n1 floats) and points (3*n0*n1 floats). */
int oc_grid_build(const oc_grid* g, float* axis0, float* axis1, float* points);

int oc_design_lowpass(int n_taps, float normalized_cutoff, float* taps_out, int* n_out);
float oc_window_value(int window, float u, float tukey_alpha);

/* Transmit factories of Simulator (simulator.cpp:156-203). kind: 1 plane
 * waves (params = angles), 0 focused scan (params = {x_min, x_max,
 * focal_depth, angle}, n = n_lines), 2 diverging (params = angles, extra =
 * virtual source depth). delays/apod/focus/origin/angle out per transmit. */
int oc_make_transmits(const oc_setup* s, const oc_simcfg* sc, int kind, int n,
                      const float* params, float extra, float* delays_out,
                      float* apod_out, float* focus_out, float* origin_out,
                      float* angle_out);

float oc_transmit_arrival_time(const oc_transmit* tx, const float* point, float c);
float oc_transmit_sensitivity(const oc_setup* s, const oc_transmit* tx,
                              const float* point, float wavelength);

/* Phantom::speckle_box (simulator.cpp:46-65): pos_out 3*n, amp_out n. */
int oc_speckle_box(float x_min, float x_max, float z_min, float z_max, float y_half,
                   int n, uint32_t seed, float sigma, float* pos_out, float* amp_out);

/* Simulator::simulate (simulator.cpp:213-419). vel may be NULL (zero).
 * rf_out: n_tx * n_elements * n_samples. */
int oc_simulate(const oc_setup* s, const oc_simcfg* sc, const oc_transmit* txs, int n_tx,
                int n_scat, const float* pos, const float* amp, const float* vel,
                float* rf_out);

/* demodulate (demodulate.cpp:36-121). iq_out interleaved re,im with
 * n_tx * n_elements * n_out samples; n_out/fs_out/fdem_out receive the output
 * acquisition fields. Pass iq_out = NULL to query n_out only. */
int oc_demodulate(const oc_setup* s, const oc_transmit* txs, int n_tx, const float* rf,
                  const oc_demod* d, float* iq_out, int* n_out, float* fs_out,
                  float* fdem_out);

/* Beamformer::beamform (beamformer.cpp:297-549). `data` is RF (float) when
 * is_iq == 0, else interleaved IQ; s->n_samples / s->sampling_frequency
 * describe `data`. frame_out interleaved, n0*n1 complex. The config is
 * treated as assigned after construction: it is not validated, and
 * out-of-range fields meet the reference's clamps; only minimum variance with
 * a negative range average is refused (returns 1). */
int oc_beamform(const oc_setup* s, const oc_transmit* txs, int n_tx, int is_iq,
                const float* data, float fdem, const oc_grid* g, const oc_bf* bf,
                int event_begin, int event_end, float* frame_out);

/* BeamformerConfig::validate as the Beamformer constructor runs it
 * (beamformer.hpp:110, beamformer.cpp:218-225): 0 or 1 with the message. */
int oc_validate_bf(const oc_bf* bf);

/* ImagingPipeline::process_frame (pipeline.cpp:25-46). stats_out receives
 * {demodulate_ms, beamform_ms, total_ms}; may be NULL. */
int oc_process_frame(const oc_setup* s, const oc_transmit* txs, int n_tx, const float* rf,
                     const oc_grid* g, const oc_demod* d, const oc_bf* bf,
                     float* frame_out, double* stats_out);

/* The display tail: frames through ImagingPipeline::to_display in order
 * (pipeline.cpp:48-70; postprocess.cpp:12-117), so the smoothed reference
 * level carries over from frame to frame. frames: n_frames * n0*n1 complex
 * (interleaved); fdem > 0: IQ frames (envelope |.|, TGC at fdem), 0: RF
 * frames (analytic signal along axis1, TGC at 5 MHz); attenuation from s.
 * images_out: n_frames * n0*n1 floats. */
typedef struct {
  int apply_tgc;
  float dynamic_range_db, gain_db, reference_smoothing;
} oc_display_cfg;
int oc_display(const oc_setup* s, const oc_grid* g, const oc_display_cfg* p, int n_frames,
               const float* frames, float fdem, float* images_out);

/* Doppler and speckle tracking (core/src/doppler.cpp). Frames are n0*n1
 * complex (interleaved) on grid g; an ensemble is n_frames of them, one after
 * the other. wall type: 0 none, 1 polynomial regression, 2 mean subtraction. */
typedef struct {
  int wall_filter, polynomial_order, loupas, kernel_axial, kernel_lateral;
  float power_threshold_db, coherence_threshold;
} oc_doppler;
/* apply_wall_filter (doppler.cpp:51-75): out n_frames * n0*n1 complex. */
int oc_wall_filter(const oc_grid* g, int n_frames, const float* ensemble, int type, int order,
                   float* out);
/* power_doppler (doppler.cpp:178-186): power_out n0*n1. */
int oc_power_doppler(const oc_grid* g, int n_frames, const float* ensemble, int type, int order,
                     float* power_out);
/* color_flow (doppler.cpp:77-176); fdem is the frames' demodulation frequency. */
int oc_color_flow(const oc_grid* g, int n_frames, const float* ensemble, float fdem, float prf,
                  float c, const oc_doppler* d, float* velocity, float* power, float* variance,
                  uint8_t* valid, float* nyquist);
/* estimate_axial_displacement (doppler.cpp:188-273). */
int oc_axial_displacement(const oc_grid* g, const float* before, const float* after, float fdem,
                          float c, int kernel_axial, int kernel_lateral, int max_lag, float* axial,
                          float* correlation);
/* axial_strain (doppler.cpp:275-307). */
int oc_axial_strain(const oc_grid* g, const float* axial, int fit_length, float* strain);

#ifdef __cplusplus
}
#endif
#endif
