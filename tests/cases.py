"""Seeded acquisition cases shared by the tests (built with the product's
host simulator, which is bitwise the reference's single-threaded one)."""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np


@dataclass
class Case:
    probe: object
    medium: object
    acq: object
    sim: object
    txs: list
    phantom: object
    raw: object            # ChannelData (RF)
    grid: object           # usx.ScanGrid
    gspec: object          # oracle GridSpec producing the same axes


def make_case(usx, O, n_el=32, fs=20e6, n_samples=768, n_tx=3, half_angle=9.0,
              n_scat=600, seed=0, grid=(24, 0.006, 0.024, 48), noise_db=-60.0,
              cyst=None, targets=None, target_amp=10.0, box=(-0.006, 0.006, 0.005, 0.026)):
    probe = usx.presets.linear_5mhz()
    probe.n_elements = n_el
    med = usx.Medium()
    acq = usx.Acquisition()
    acq.sampling_frequency = fs
    acq.n_samples = n_samples
    acq.prf = 5000.0
    cfg = usx.SimulatorConfig()
    cfg.noise_db = noise_db
    sim = usx.Simulator(probe, med, acq, cfg)
    txs = sim.make_plane_waves(usx.angles_deg(-half_angle, half_angle, n_tx))
    if n_scat > 0:
        ph = usx.Phantom.speckle_box(box[0], box[1], box[2], box[3], 2e-4, n_scat, seed, 1.0)
    else:
        ph = usx.Phantom()
    if cyst is not None:
        ph.scale_amplitude_in_circle(usx.Vec3(cyst[0], 0.0, cyst[1]), cyst[2], 0.0)
    if targets:
        ph.extend(usx.Phantom.point_targets([usx.Vec3(x, 0.0, z) for x, z in targets], target_amp))
    raw = sim.simulate(txs, ph)
    nx, z0, z1, nz = grid
    half = 0.5 * probe.aperture_width()
    g = usx.ScanGrid.cartesian(-half, half, nx, z0, z1, nz)
    gs = O.GridSpec(0, -half, half, nx, z0, z1, nz)
    return Case(probe, med, acq, sim, txs, ph, raw, g, gs)


def bf_config(usx, combiner="DAS", cf=False, compounding="COHERENT", f_number=1.5,
              subarray=16, loading=1e-2, range_avg=2, window="HANN", weighting=True,
              precision="FP32", refine=False):
    b = usx.BeamformerConfig()
    b.f_number = f_number
    b.combiner = getattr(usx.Combiner, combiner)
    b.coherence_factor = cf
    b.compounding = getattr(usx.Compounding, compounding)
    b.mv_subarray = subarray
    b.mv_diagonal_loading = loading
    b.mv_range_average = range_avg
    b.rx_window = getattr(usx.Window, window)
    b.transmit_weighting = weighting
    b.mv_precision = getattr(usx.MvPrecision, precision)
    b.mv_refine = refine
    return b


def oracle_inputs(O, case, data=None):
    s = O.Setup.of(case.probe, case.medium, case.acq)
    txs = [O.Tx.of(t) for t in case.txs]
    return s, txs


def wires(usx, points, amplitude=1.0):
    """Point scatterers at (x, z) positions in the plane y = 0."""
    return usx.Phantom.point_targets([usx.Vec3(float(x), 0.0, float(z)) for x, z in points],
                                     amplitude)


def speckle_with_cysts(usx, probe, medium, cysts, z_min, z_max, per_cell=12.0, seed=3):
    """Gaussian speckle over 1.15 x the aperture width and [z_min, z_max] at
    `per_cell` scatterers per resolution cell (f/2 cell, 0.4 mm thick slab),
    with each (x, z, radius, amplitude scale) circle rescaled."""
    half = 0.575 * probe.aperture_width()
    n = usx.Phantom.recommended_scatterer_count(probe, medium, 2 * half, z_max - z_min, 4e-4,
                                                f_number=2.0, per_cell=per_cell)
    ph = usx.Phantom.speckle_box(-half, half, z_min, z_max, 2e-4, int(min(max(n, 2000), 400000)),
                                 seed)
    for x, z, r, scale in cysts:
        ph.scale_amplitude_in_circle(usx.Vec3(float(x), 0.0, float(z)), float(r), float(scale))
    return ph

