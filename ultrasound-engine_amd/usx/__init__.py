"""usx -- the ultrasound imaging path on MI355X.

Drop-in for the reference package's imaging API (python/usx/__init__.py):
the same classes and helpers, backed by hand-written gfx950 HIP kernels
behind a C ABI (include/usx_gpu.h)::

    import usx

    probe = usx.presets.linear_5mhz()
    sim   = usx.Simulator(probe, usx.Medium(), usx.presets.acquisition(probe))
    raw   = sim.simulate(sim.make_plane_waves(usx.angles_deg(-9, 9, 11)), phantom)
    grid  = usx.grid_for(probe, 0.005, 0.050, 256, 640)
    frame = usx.ImagingPipeline(grid).process_frame(raw)   # complex64 (256, 640)

The imaging calls need a gfx950 GPU and raise RuntimeError without one; there
is no CPU fallback. Host-only pieces (types, grids, delay laws, the
simulator, metrics) work anywhere.
"""

from __future__ import annotations

import numpy as np

from ._usx import (  # noqa: F401
    Acquisition,
    Beamformer,
    BeamformerConfig,
    ChannelData,
    ColorFlowMap,
    Combiner,
    Compounding,
    DemodConfig,
    DeviceBuffer,
    DeviceRunner,
    DisplacementMap,
    DopplerConfig,
    DopplerRunner,
    Frame,
    GridType,
    ImagingPipeline,
    Medium,
    MvPrecision,
    Phantom,
    PipelineConfig,
    PipelineStats,
    Probe,
    ScanGrid,
    Scatterer,
    Simulator,
    SimulatorConfig,
    Transmit,
    TransmitType,
    Vec3,
    WallFilter,
    Window,
    __version__,
    apply_wall_filter,
    axial_strain,
    build_info,
    color_flow,
    compute_transmit_delays,
    demodulate,
    design_lowpass,
    dynamic_aperture_weight,
    element_directivity,
    estimate_axial_displacement,
    polynomial_reject_matrix,
    power_doppler,
    transmit_arrival_time,
    transmit_sensitivity,
    window_value,
)

from . import ingest, metrics, presets  # noqa: E402,F401

__all__ = [n for n in dir() if not n.startswith("_")] + [
    "angles_deg", "to_bmode", "grid_for", "extent_mm",
]


def angles_deg(start: float, stop: float, count: int) -> list[float]:
    """Evenly spaced steering angles given in degrees, returned in radians
    (reference: python/usx/__init__.py:81-90)."""
    if count == 1:
        return [np.deg2rad(0.5 * (start + stop))]
    return list(np.deg2rad(np.linspace(start, stop, count)))


def to_bmode(frame: Frame, dynamic_range_db: float = 60.0, gain_db: float = 0.0,
             medium: Medium | None = None, tgc: bool = True) -> np.ndarray:
    """Complex frame to a log-compressed image in [0, 1], shape (axis0, axis1)
    (reference: python/usx/__init__.py:93-105): envelope, TGC at the frame's
    demodulation frequency (5 MHz for RF frames) and log compression against
    the frame's own peak, on the GPU display path (ImagingPipeline.to_display
    of a fresh pipeline, whose first reference level is that peak)."""
    pc = PipelineConfig()
    pc.apply_tgc = tgc
    pc.dynamic_range_db = dynamic_range_db
    pc.gain_db = gain_db
    img = ImagingPipeline(frame.grid, pc).to_display(frame, medium or Medium())
    return np.asarray(img, dtype=np.float32).reshape(frame.grid.n_axis0, frame.grid.n_axis1)


def grid_for(probe: Probe, z_min: float = 0.005, z_max: float = 0.055,
             n_lateral: int = 256, n_axial: int = 768,
             width: float | None = None) -> ScanGrid:
    """Cartesian grid across the probe aperture over a depth range
    (reference: python/usx/__init__.py:108-117)."""
    half = 0.5 * (width if width is not None else probe.aperture_width())
    return ScanGrid.cartesian(-half, half, n_lateral, z_min, z_max, n_axial)


def extent_mm(grid: ScanGrid) -> tuple[float, float, float, float]:
    """matplotlib extent in mm for an image shaped (axis0, axis1), displayed
    depth-down (reference: python/usx/__init__.py:120-128)."""
    a0 = np.asarray(grid.axis0) * 1e3
    a1 = np.asarray(grid.axis1) * 1e3
    return (float(a0[0]), float(a0[-1]), float(a1[-1]), float(a1[0]))
