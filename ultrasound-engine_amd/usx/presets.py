"""Probe and acquisition presets (reference: python/usx/presets.py:18-98).

Values are the reference's, so grids, sample counts and PRFs built from them
match it exactly.
"""

from __future__ import annotations

import numpy as np

from ._usx import Acquisition, Medium, Probe


def _probe(n, pitch, width, elev, f0, bw) -> Probe:
    p = Probe()
    p.n_elements = n
    p.pitch = pitch
    p.element_width = width
    p.elevation_height = elev
    p.center_frequency = f0
    p.bandwidth = bw
    return p


def linear_5mhz() -> Probe:
    """128-element 5 MHz linear array, 300 um pitch."""
    return _probe(128, 300e-6, 270e-6, 5e-3, 5.0e6, 0.65)


def linear_high_freq() -> Probe:
    """128-element 10 MHz linear array, 150 um pitch."""
    return _probe(128, 150e-6, 135e-6, 3e-3, 10.0e6, 0.7)


def phased_3mhz() -> Probe:
    """64-element 3 MHz phased array, half-wavelength pitch."""
    return _probe(64, 250e-6, 220e-6, 12e-3, 3.0e6, 0.6)


def acquisition(probe: Probe, depth: float = 0.06, medium: Medium | None = None,
                samples_per_wavelength: float = 8.0, prf: float | None = None) -> Acquisition:
    """Record length covering the round trip to `depth` plus the aperture
    delay spread and four periods of margin; PRF as fast as the depth allows."""
    c = (medium or Medium()).speed_of_sound
    a = Acquisition()
    a.sampling_frequency = float(samples_per_wavelength * probe.center_frequency)
    margin = probe.aperture_width() / c + 4.0 / probe.center_frequency
    a.n_samples = int(np.ceil((2.0 * depth / c + margin) * a.sampling_frequency))
    a.prf = float(prf) if prf else float(0.95 * c / (2.0 * depth))
    a.t0 = 0.0
    return a


def plane_wave_sequence(sim, half_angle_deg: float = 9.0, n_angles: int = 11):
    """Symmetric steered plane waves over +-half_angle_deg."""
    if n_angles <= 1:
        return sim.make_plane_waves([0.0])
    return sim.make_plane_waves(
        list(np.deg2rad(np.linspace(-half_angle_deg, half_angle_deg, n_angles))))


def focused_sequence(sim, probe: Probe, n_lines: int = 128, focal_depth: float = 0.03):
    """Line-by-line focused scan across 90% of the aperture."""
    half = 0.5 * probe.aperture_width() * 0.9
    return sim.make_focused_scan(n_lines, -half, half, focal_depth)
