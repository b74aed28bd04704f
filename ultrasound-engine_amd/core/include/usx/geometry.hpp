// usx/geometry.hpp -- transmit delay laws, arrival times, receive aperture.
//
// Interface of the reference's core/include/usx/geometry.hpp:14-60.
#pragma once

#include "usx/types.hpp"

namespace usx {

// Per-element fire times (min 0 over firing elements) and apodization of a
// transmit; *t0_offset receives the earliest engine-time delay.
void compute_transmit_delays(const Probe& probe, const Medium& medium, Transmit& tx,
                             Window apod_window, real f_number, real* t0_offset);

// Arrival time of the ideal transmit wavefront at `point`, engine time.
real transmit_arrival_time(const Transmit& tx, const Vec3& point, real c);

inline real receive_arrival_time(const Vec3& element, const Vec3& point, real c) {
  return distance(element, point) / c;
}

// cos(theta) * sinc(pi w sin(theta) / lambda) of a rectangular element.
real element_directivity(real element_width, real wavelength, real sin_theta);

// f-number-limited receive apodization weight of `element` for `point`.
real dynamic_aperture_weight(const Vec3& element, const Vec3& point, real f_number, Window w,
                             real tukey_alpha = 0.5f);

}  // namespace usx
