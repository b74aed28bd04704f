"""Doppler and speckle tracking on the GPU (usx.apply_wall_filter,
usx.color_flow, usx.power_doppler, usx.estimate_axial_displacement,
usx.axial_strain) against the CPU checker, which equals the reference's
compiled doppler.cpp bit for bit (test_doppler_oracle.py), and the
reference's own physics checks restated.

Bars: the wall filter, power Doppler and strain are bit-exact (the same
float operations in the same order). Colour flow regroups the window and
ensemble sums (per-pixel ensemble sums first) and uses the device atan2f /
hypotf: power, velocity and spectral width within 1e-5 relative L2, and the
valid flags equal except where the checker's power or lag-one coherence lies
within 1e-4 of its threshold. Displacement (device atan2f / hypotf / sqrtf
in the same summation order): within 1e-5 relative L2, identical lags.
"""

from __future__ import annotations

import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu

C = 1540.0


def _standard(usx, n_samples=2200, prf=6000.0):
    """The reference's standard probe and acquisition (core/tests/test_main.cpp:55-72)."""
    probe = usx.Probe()
    probe.n_elements = 96
    probe.pitch = 300e-6
    probe.element_width = 270e-6
    probe.center_frequency = 5.0e6
    probe.bandwidth = 0.6
    med = usx.Medium()
    acq = usx.Acquisition()
    acq.sampling_frequency = 40.0e6
    acq.n_samples = n_samples
    acq.t0 = 0.0
    acq.prf = prf
    cfg = usx.SimulatorConfig()
    cfg.noise_db = -200.0
    return probe, med, acq, usx.Simulator(probe, med, acq, cfg)


def _moving_block(usx, v, seed=5, n=3000, amp=1.0, x=(-0.004, 0.004), z=(0.028, 0.032)):
    ph = usx.Phantom.speckle_box(x[0], x[1], z[0], z[1], 1e-4, n, seed, amp)
    ph.velocities = np.tile(np.asarray(v, np.float32), (len(ph), 1)).astype(np.float32)
    return ph


def _ensemble(usx, sim, ph, n_ens, grid):
    seq = usx.Simulator.repeat(sim.make_plane_waves([0.0]), n_ens)
    iq = usx.demodulate(sim.simulate(seq, ph), usx.DemodConfig())
    return usx.Beamformer().beamform_sequence(iq, grid, 1)


def _spec(O, grid):
    a0, a1 = np.asarray(grid.axis0), np.asarray(grid.axis1)
    return O.GridSpec(0, float(a0[0]), float(a0[-1]), len(a0), float(a1[0]), float(a1[-1]), len(a1))


def _stack(frames):
    return np.stack([np.asarray(f.data) for f in frames]).astype(np.complex64)


@pytest.fixture(scope="module")
def flow(usx_mod):
    """16 frames of a speckle block moving axially at 0.15 m/s, alone and under
    4x stronger static clutter (core/tests/test_main.cpp:431-503)."""
    usx = usx_mod
    probe, med, acq, sim = _standard(usx)
    grid = usx.ScanGrid.cartesian(-0.003, 0.003, 31, 0.029, 0.031, 61)
    moving = _moving_block(usx, (0.0, 0.0, 0.15))
    ens = _ensemble(usx, sim, moving, 16, grid)
    mixed = _moving_block(usx, (0.0, 0.0, 0.15))
    mixed.extend(usx.Phantom.speckle_box(-0.004, 0.004, 0.028, 0.032, 1e-4, 3000, 77, 4.0))
    ens2 = _ensemble(usx, sim, mixed, 16, grid)
    return acq, grid, ens, ens2


def _synthetic(usx_mod, flow, n, seed):
    """n frames on the flow grid with synthetic data: slow-time ramp, clutter, noise."""
    acq, grid, ens, _ = flow
    frames = list(ens) * ((n + len(ens) - 1) // len(ens))
    frames = [usx_mod.apply_wall_filter([f], usx_mod.WallFilter.NONE, 1)[0] for f in frames[:n]]
    rng = np.random.default_rng(seed)
    n0, n1 = grid.n_axis0, grid.n_axis1
    data = (0.3 * (rng.standard_normal((n, n0, n1)) + 1j * rng.standard_normal((n, n0, n1)))
            + np.exp(0.35j * np.arange(n))[:, None, None] * np.exp(1j * np.linspace(0, 9, n1))[None, None, :]
            + 2.0).astype(np.complex64)
    for f, d in zip(frames, data):
        np.asarray(f.data)[...] = d
    return frames, data


@pytest.mark.parametrize("n", [2, 9, 16, 17, 33, 40])
@pytest.mark.parametrize("wall,order", [("POLYNOMIAL", 1), ("POLYNOMIAL", 3), ("MEAN", 1)])
def test_wall_filter_and_power_doppler_are_bit_exact(usx_mod, O, port, flow, n, wall, order):
    frames, data = _synthetic(usx_mod, flow, n, seed=n)
    g = _spec(O, flow[1])
    wt = getattr(usx_mod.WallFilter, wall)
    code = 1 if wall == "POLYNOMIAL" else 2
    got = _stack(usx_mod.apply_wall_filter(frames, wt, order))
    ref = port.wall_filter(g, data, code, order)
    assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))
    pw = np.asarray(usx_mod.power_doppler(frames, wt, order))
    assert np.array_equal(pw.view(np.uint32), port.power_doppler(g, data, code, order).view(np.uint32))


def _near_threshold(var, power, d, pmax):
    coh = 1.0 - var / 2.0
    thr = pmax * 10.0 ** (d.power_threshold_db / 10.0)
    return (np.abs(coh / d.coherence_threshold - 1.0) < 1e-4) | (np.abs(power / thr - 1.0) < 1e-4)


def _compare_flow(usx, O, port, frames, data, grid, prf, dc):
    d = O.Doppler.of(dc)
    m = usx.color_flow(frames, prf, C, dc)
    v, p, w, ok, nyq = port.color_flow(_spec(O, grid), data, frames[0].demodulation_frequency, prf, C, d)
    assert rel_l2(np.asarray(m.power), p) < 1e-5
    assert np.max(np.abs(np.asarray(m.variance) - w)) < 1e-5
    got_ok = np.asarray(m.valid).astype(bool)
    ref_ok = ok.astype(bool)
    diff = got_ok != ref_ok
    assert not np.any(diff & ~_near_threshold(w, p, d, p.max())), int(diff.sum())
    both = got_ok & ref_ok
    if both.any():
        assert rel_l2(np.asarray(m.velocity)[both], v[both]) < 1e-5
    assert m.nyquist_velocity == pytest.approx(nyq, rel=1e-7)
    return m, ref_ok


@pytest.mark.parametrize("wall", ["NONE", "POLYNOMIAL", "MEAN"])
@pytest.mark.parametrize("loupas,ka,kl", [(True, 4, 1), (False, 0, 0), (True, 6, 2)])
def test_color_flow_matches_the_oracle_on_simulated_flow(usx_mod, O, port, flow, wall, loupas, ka, kl):
    acq, grid, _, ens2 = flow
    dc = usx_mod.DopplerConfig()
    dc.wall_filter = getattr(usx_mod.WallFilter, wall)
    dc.loupas = loupas
    dc.kernel_axial = ka
    dc.kernel_lateral = kl
    m, ref_ok = _compare_flow(usx_mod, O, port, ens2, _stack(ens2), grid, acq.prf, dc)
    assert ref_ok.sum() > 0


@pytest.mark.parametrize("n", [2, 17, 40])
@pytest.mark.parametrize("ka,kl", [(4, 1), (20, 1), (2, 6)])
def test_color_flow_matches_the_oracle_on_synthetic_ensembles(usx_mod, O, port, flow, n, ka, kl):
    """Windows beyond the LDS tile (axial > 16 or lateral > 4) take the
    direct-from-memory path."""
    frames, data = _synthetic(usx_mod, flow, n, seed=100 + n)
    dc = usx_mod.DopplerConfig()
    dc.coherence_threshold = 0.2
    dc.power_threshold_db = -10.0
    dc.kernel_axial = ka
    dc.kernel_lateral = kl
    _compare_flow(usx_mod, O, port, frames, data, flow[1], 5000.0, dc)


def test_doppler_recovers_the_axial_velocity(usx_mod, flow):
    """core/tests/test_main.cpp:461-479: mean valid velocity within 0.02 m/s of
    0.15 m/s, below the aliasing limit."""
    acq, grid, ens, _ = flow
    assert len(ens) == 16
    dc = usx_mod.DopplerConfig()
    dc.wall_filter = usx_mod.WallFilter.NONE
    dc.kernel_axial = 4
    dc.kernel_lateral = 1
    m = usx_mod.color_flow(ens, acq.prf, C, dc)
    ok = np.asarray(m.valid).astype(bool)
    assert ok.sum() > 0
    assert float(np.mean(np.asarray(m.velocity)[ok])) == pytest.approx(0.15, abs=0.02)
    assert m.nyquist_velocity > 0.15


def test_wall_filter_improves_the_estimate_under_clutter(usx_mod, flow):
    """core/tests/test_main.cpp:481-503."""
    acq, grid, _, ens2 = flow

    def mean_valid(wall):
        dc = usx_mod.DopplerConfig()
        dc.wall_filter = wall
        dc.polynomial_order = 1
        m = usx_mod.color_flow(ens2, acq.prf, C, dc)
        ok = np.asarray(m.valid).astype(bool)
        return float(np.mean(np.asarray(m.velocity)[ok])) if ok.any() else 0.0

    v_off = mean_valid(usx_mod.WallFilter.NONE)
    v_on = mean_valid(usx_mod.WallFilter.POLYNOMIAL)
    assert abs(v_on - 0.15) < abs(v_off - 0.15)


def test_lateral_flow_produces_no_axial_doppler(usx_mod):
    """tests/test_usx.py:227-244 of the reference: purely lateral motion has
    no axial Doppler."""
    usx = usx_mod
    probe = usx.presets.linear_5mhz()
    probe.n_elements = 64
    med = usx.Medium()
    acq = usx.presets.acquisition(probe, depth=0.04, medium=med)
    cfg = usx.SimulatorConfig()
    cfg.noise_db = -200.0
    sim = usx.Simulator(probe, med, acq, cfg)
    ph = _moving_block(usx, (0.20, 0.0, 0.0), n=3000, x=(-0.006, 0.006), z=(0.022, 0.028))
    grid = usx.ScanGrid.cartesian(-0.003, 0.003, 21, 0.023, 0.027, 61)
    frames = _ensemble(usx, sim, ph, 12, grid)
    dc = usx.DopplerConfig()
    dc.wall_filter = usx.WallFilter.NONE
    m = usx.color_flow(frames, acq.prf, med.speed_of_sound, dc)
    ok = np.asarray(m.valid).astype(bool)
    assert ok.sum() > 0
    assert abs(float(np.median(np.asarray(m.velocity)[ok]))) < 0.02


@pytest.mark.parametrize("ka,kl,lag", [(8, 2, 8), (4, 0, 0), (1, 1, 3)])
def test_displacement_and_strain_match_the_oracle(usx_mod, O, port, flow, ka, kl, lag):
    acq, grid, ens, _ = flow
    before, after = ens[0], ens[5]
    g = _spec(O, grid)
    fdem = before.demodulation_frequency
    m = usx_mod.estimate_axial_displacement(before, after, C, ka, kl, lag)
    ax, co = port.axial_displacement(g, np.asarray(before.data), np.asarray(after.data), fdem, C, ka, kl, lag)
    assert rel_l2(np.asarray(m.axial), ax) < 1e-5
    assert rel_l2(np.asarray(m.correlation), co) < 1e-5
    s = np.asarray(usx_mod.axial_strain(m, 15))
    assert np.array_equal(s.view(np.uint32), port.axial_strain(g, np.asarray(m.axial), 15).view(np.uint32))


def test_displacement_tracks_the_moving_block(usx_mod, flow):
    """Frames 5 apart at PRF 6 kHz: the block moved 0.15 m/s * 5 / 6000 = 125 um
    (+z); the median displacement over well-correlated pixels finds it."""
    acq, grid, ens, _ = flow
    m = usx_mod.estimate_axial_displacement(ens[0], ens[5], C)
    good = np.asarray(m.correlation) > 0.9
    assert good.sum() > 100
    assert float(np.median(np.asarray(m.axial)[good])) == pytest.approx(0.15 * 5 / acq.prf, abs=20e-6)


def test_errors(usx_mod, flow):
    acq, grid, ens, _ = flow
    with pytest.raises(ValueError):
        usx_mod.color_flow(ens[:1], acq.prf, C)
    other = usx_mod.Beamformer().beamform_sequence(
        usx_mod.demodulate(_standard(usx_mod, 600)[3].simulate(
            _standard(usx_mod, 600)[3].make_plane_waves([0.0]), usx_mod.Phantom()), usx_mod.DemodConfig()),
        usx_mod.ScanGrid.cartesian(-0.001, 0.001, 3, 0.005, 0.006, 4), 1)
    with pytest.raises(ValueError):
        usx_mod.color_flow([ens[0], other[0]], acq.prf, C)
    with pytest.raises(ValueError):
        usx_mod.apply_wall_filter([ens[0], other[0]], usx_mod.WallFilter.MEAN, 0)
    with pytest.raises(ValueError):
        usx_mod.estimate_axial_displacement(ens[0], other[0], C)
    assert usx_mod.apply_wall_filter([], usx_mod.WallFilter.MEAN, 0) == []


def test_displacement_and_strain_physics(usx_mod):
    """core/tests/test_main.cpp:506-600 restated: a 20 um translation is
    tracked within 2 um, a 400 um one only with the lag search, a uniform
    translation has no strain and a 0.5 % compression has that strain."""
    usx = usx_mod
    probe, med, acq, _ = _standard(usx, n_samples=2600, prf=4000.0)
    cfg = usx.SimulatorConfig()
    cfg.noise_db = -200.0
    cfg.enable_motion = False
    sim = usx.Simulator(probe, med, acq, cfg)
    before = usx.Phantom.speckle_box(-0.006, 0.006, 0.026, 0.034, 1e-4, 6000, 11)
    txs = sim.make_plane_waves([0.0])
    grid = usx.ScanGrid.cartesian(-0.004, 0.004, 41, 0.028, 0.032, 121)
    bf = usx.Beamformer()
    pos = np.asarray(before.positions, np.float64)

    def frame_of(z_new):
        ph = before.copy()
        p = pos.copy()
        p[:, 2] = z_new
        ph.positions = p.astype(np.float32)
        return bf.beamform(usx.demodulate(sim.simulate(txs, ph), usx.DemodConfig()), grid)

    fa = bf.beamform(usx.demodulate(sim.simulate(txs, before), usx.DemodConfig()), grid)
    fb = frame_of(np.float32(pos[:, 2]) + np.float32(20e-6))
    dm = usx.estimate_axial_displacement(fa, fb, med.speed_of_sound, 8, 2)
    good = np.asarray(dm.correlation) >= 0.9
    assert good.sum() > 100
    assert float(np.mean(np.asarray(dm.axial)[good])) * 1e6 == pytest.approx(20.0, abs=2.0)

    ff = frame_of(np.float32(pos[:, 2]) + np.float32(400e-6))

    def median_of(m):
        ok = np.asarray(m.correlation) > 0.8
        return float(np.median(np.asarray(m.axial)[ok])) if ok.any() else 0.0

    m_phase = median_of(usx.estimate_axial_displacement(fa, ff, med.speed_of_sound, 8, 2, 0))
    m_lag = median_of(usx.estimate_axial_displacement(fa, ff, med.speed_of_sound, 8, 2, 12))
    assert abs(m_lag - 400e-6) < 20e-6
    assert abs(m_phase - 400e-6) > 50e-6

    strain = np.asarray(usx.axial_strain(dm, 15))
    assert abs(float(np.mean(strain[:, 20:-20]))) < 5e-3

    z = pos[:, 2].astype(np.float32)
    fc = frame_of(z + np.float32(-0.005) * (z - np.float32(0.030)))
    dm2 = usx.estimate_axial_displacement(fa, fc, med.speed_of_sound, 8, 2)
    s2 = np.asarray(usx.axial_strain(dm2, 21))[:, 25:-25]
    ok = np.asarray(dm2.correlation)[:, 25:-25] >= 0.9
    assert float(np.mean(s2[ok])) == pytest.approx(-0.005, abs=1.5e-3)
