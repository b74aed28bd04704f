"""GPU point-scatterer simulator (sim.hip) against the host simulator.

The host simulator (Simulator.simulate) is the reference's model bit for bit
(pinned in test_host.py by the reference's own RF). The GPU model differs
only by float rounding: device powf/sinf in the path amplitudes and the
chunked (fixed-order) reduction of the echoes. Bars: relative L2 1e-5 over
the record and 1e-4 of the peak in any sample.
"""

from __future__ import annotations

import time

import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu


def _sim(usx, n_el=32, n_samples=1024, noise_db=-60.0, motion=True):
    probe = usx.presets.linear_5mhz()
    probe.n_elements = n_el
    acq = usx.Acquisition()
    acq.sampling_frequency = 20e6
    acq.n_samples = n_samples
    acq.prf = 5000.0
    cfg = usx.SimulatorConfig()
    cfg.noise_db = noise_db
    cfg.enable_motion = motion
    return probe, usx.Simulator(probe, usx.Medium(), acq, cfg)


def _check(a, b):
    a = np.asarray(a.rf, np.float64)
    b = np.asarray(b.rf, np.float64)
    assert a.shape == b.shape
    peak = np.abs(b).max()
    assert rel_l2(a, b) < 1e-5, rel_l2(a, b)
    assert np.abs(a - b).max() < 1e-4 * peak


def test_plane_waves_speckle_and_cyst(usx_mod):
    probe, sim = _sim(usx_mod)
    txs = sim.make_plane_waves(usx_mod.angles_deg(-9, 9, 3))
    ph = usx_mod.Phantom.speckle_box(-0.006, 0.006, 0.005, 0.026, 2e-4, 1500, 3, 1.0)
    ph.scale_amplitude_in_circle(usx_mod.Vec3(0.0, 0.0, 0.015), 0.002, 0.0)
    ph.extend(usx_mod.Phantom.point_targets([usx_mod.Vec3(0.002, 0.0, 0.02)], 10.0))
    _check(sim.simulate_gpu(txs, ph), sim.simulate(txs, ph))


def test_focused_scan_with_apodization_and_motion(usx_mod):
    probe, sim = _sim(usx_mod, n_el=48)
    txs = sim.make_focused_scan(4, -0.004, 0.004, 0.02)
    assert any(len(t.apodization) and min(t.apodization) == 0.0 for t in txs)
    ph = usx_mod.Phantom.speckle_box(-0.008, 0.008, 0.005, 0.03, 2e-4, 800, 5, 1.0)
    ph.add_flow_tube(usx_mod.Vec3(0.0, 0.0, 0.018), 0.002, 0.01, 0.3, 0.5, 300, 9, 0.5)
    _check(sim.simulate_gpu(txs, ph), sim.simulate(txs, ph))


def test_diverging_waves_without_noise(usx_mod):
    probe, sim = _sim(usx_mod, noise_db=-300.0)
    txs = sim.make_diverging_waves(usx_mod.angles_deg(-10, 10, 2), 0.01)
    ph = usx_mod.Phantom.speckle_box(-0.006, 0.006, 0.005, 0.02, 2e-4, 500, 8, 1.0)
    _check(sim.simulate_gpu(txs, ph), sim.simulate(txs, ph))
    # an empty phantom is silence (no peak, so no noise either)
    empty = sim.simulate_gpu(txs, usx_mod.Phantom())
    assert not np.asarray(empty.rf).any()


def test_config0_size_matches_and_is_fast(usx_mod):
    """128 elements, 11 plane waves, 1536 samples, 20k scatterers (the
    benchmark's phantom)."""
    probe = usx_mod.presets.linear_5mhz()
    acq = usx_mod.Acquisition()
    acq.sampling_frequency = 20e6
    acq.n_samples = 1536
    acq.prf = 5000.0
    sim = usx_mod.Simulator(probe, usx_mod.Medium(), acq, usx_mod.SimulatorConfig())
    ph = usx_mod.Phantom.speckle_box(-0.019, 0.019, 0.005, 0.050, 2e-4, 20000, 0, 1.0)
    ph.scale_amplitude_in_circle(usx_mod.Vec3(0.0, 0.0, 0.020), 0.003, 0.0)
    txs = sim.make_plane_waves(usx_mod.angles_deg(-9, 9, 11))
    sim.simulate_gpu(txs, ph)  # warm-up (noise table, buffers)
    t = time.perf_counter()
    g = sim.simulate_gpu(txs, ph)
    dt = time.perf_counter() - t
    print(f"\n[simulate_gpu] config-0 frame: {dt * 1e3:.1f} ms")
    _check(g, sim.simulate(txs, ph))
