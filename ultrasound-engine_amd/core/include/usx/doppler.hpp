// usx/doppler.hpp -- estimators on the phase of beamformed IQ: colour and
// power Doppler across a slow-time ensemble, and axial speckle tracking and
// strain between two frames.
//
// Interface of the reference's core/include/usx/doppler.hpp:27-142 (the same
// enum, structs, defaults and free functions). The per-pixel work runs on the
// GPU through the C ABI (usx_wall_filter, usx_color_flow, usx_power_doppler,
// usx_axial_displacement, usx_axial_strain); the host builds the wall-filter
// projector and validates the inputs.
//
// Sign convention: positive axial velocity and displacement point along +z,
// away from the transducer.
#pragma once

#include <cstdint>
#include <vector>

#include "usx/types.hpp"

namespace usx {

// Clutter (wall) filter applied along slow time at every pixel.
enum class WallFilter : std::uint8_t {
  None = 0,
  PolynomialRegression = 1,  // remove the least-squares polynomial of polynomial_order
  MeanSubtraction = 2,       // remove the ensemble mean (order 0)
};

struct DopplerConfig {
  WallFilter wall_filter = WallFilter::PolynomialRegression;
  int polynomial_order = 1;
  bool loupas = true;              // local centre frequency from the axial phase gradient
  int kernel_axial = 4;            // averaging window half sizes, pixels
  int kernel_lateral = 1;
  real power_threshold_db = -35.0f;  // below the frame maximum: not flow
  real coherence_threshold = 0.35f;  // |R(0,1)| / R(0,0) required for flow
};

struct ColorFlowMap {
  ScanGrid grid;
  std::vector<real> velocity;   // m/s, axial, + away from the probe
  std::vector<real> power;      // post-filter power
  std::vector<real> variance;   // normalised spectral width in [0, 1]
  std::vector<std::uint8_t> valid;
  real nyquist_velocity = 0.0f;  // c * PRF / (4 f0)
};

std::vector<Frame> apply_wall_filter(const std::vector<Frame>& ensemble, WallFilter type,
                                     int polynomial_order);

// `prf`: the ensemble's frame rate (transmits per second over transmits per frame).
ColorFlowMap color_flow(const std::vector<Frame>& ensemble, real prf, real speed_of_sound,
                        DopplerConfig config = {});

std::vector<real> power_doppler(const std::vector<Frame>& ensemble,
                                WallFilter type = WallFilter::PolynomialRegression,
                                int polynomial_order = 1);

struct DisplacementMap {
  ScanGrid grid;
  std::vector<real> axial;        // m, + away from the probe
  std::vector<real> correlation;  // |normalised cross-correlation| in [0, 1]
};

// Coarse integer-sample lag search (|lag| <= max_lag) refined by the phase at
// the winning lag; max_lag = 0 is phase-only.
DisplacementMap estimate_axial_displacement(const Frame& before, const Frame& after,
                                            real speed_of_sound, int kernel_axial = 8,
                                            int kernel_lateral = 2, int max_lag = 8);

// Least-squares slope of the displacement against depth over fit_length depths.
std::vector<real> axial_strain(const DisplacementMap& disp, int fit_length = 15);

// The (I - P) projector of the polynomial wall filter, row-major n x n: P
// projects onto polynomials of degree <= order in slow time (no reference
// public counterpart; doppler.cpp:15-49 builds the same matrix).
std::vector<real> polynomial_reject_matrix(int n, int order);

}  // namespace usx
