"""Host-side logic (CPU only): the C-ABI library, the usx:: host layer that
builds every descriptor the kernels read, the simulator that feeds the
tests, the Python helpers and the metrics."""

from __future__ import annotations

import ctypes
import json
import os
import re

import numpy as np
import pytest

from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "usx_gpu.h")
LIB = os.path.join(ROOT, "ultrasound-engine_amd", "usx", "libusx_hip.so")


def _gpu_present():
    return os.path.exists("/dev/kfd") and any(
        os.path.exists(f"/dev/dri/renderD{128 + i}") for i in range(16))


def test_cabi_library_exports_every_declared_symbol():
    text = open(HEADER).read()
    names = sorted(set(re.findall(r"\b(usx_[a-z_0-9]+)\s*\(", text)))
    assert len(names) >= 12
    lib = ctypes.CDLL(LIB)
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    lib.usx_build_info.restype = ctypes.c_char_p
    assert b"gfx950" in lib.usx_build_info()
    # null-safe entry points (no GPU needed)
    lib.usx_ctx_last_error.restype = ctypes.c_char_p
    assert lib.usx_ctx_last_error(None) == b"null context"
    assert lib.usx_ctx_device(None) == -1
    assert lib.usx_ctx_destroy(None) == 0
    assert lib.usx_beamform(None, None, None, None, 0) == 1


@pytest.mark.skipif(_gpu_present(), reason="a GPU is present here")
def test_imaging_calls_fail_loudly_without_a_gpu(usx_mod):
    probe = usx_mod.presets.linear_5mhz()
    probe.n_elements = 8
    acq = usx_mod.Acquisition()
    acq.n_samples = 64
    sim = usx_mod.Simulator(probe, usx_mod.Medium(), acq)
    raw = sim.simulate(sim.make_plane_waves([0.0]), usx_mod.Phantom())
    grid = usx_mod.ScanGrid.cartesian(-0.001, 0.001, 4, 0.005, 0.01, 4)
    with pytest.raises(RuntimeError, match="gfx950"):
        usx_mod.ImagingPipeline(grid).process_frame(raw)
    with pytest.raises(RuntimeError, match="gfx950"):
        usx_mod.demodulate(raw)
    lib = ctypes.CDLL(LIB)
    ctx = ctypes.c_void_p()
    assert lib.usx_ctx_create(0, ctypes.byref(ctx)) == 3  # USX_ERR_NO_DEVICE


# ---- host descriptors vs the reference (bitwise) ---------------------------

@pytest.fixture(scope="module")
def ref(O):
    if not O.available("reference"):
        pytest.skip("oracle/_ref not built here (reference sources absent)")
    return O.load("reference")


def test_grids_match_reference_bitwise(usx_mod, ref, O):
    for args in [(-0.019185, 0.019185, 256, 0.005, 0.050, 640), (-0.01, 0.02, 7, 0.001, 0.04, 1),
                 (0.0, 0.008, 81, 0.021, 0.029, 81)]:
        g = usx_mod.ScanGrid.cartesian(*args)
        a0, a1, pts = ref.grid(O.GridSpec(0, *args))
        assert np.array_equal(np.asarray(g.axis0), a0) and np.array_equal(np.asarray(g.axis1), a1)
        assert np.array_equal(np.asarray(g.points), pts)
    g = usx_mod.ScanGrid.sector(-0.6, 0.6, 11, 0.01, 0.08, 5, usx_mod.Vec3(0.001, 0, -0.002))
    _, _, pts = ref.grid(O.GridSpec(1, -0.6, 0.6, 11, 0.01, 0.08, 5, (0.001, 0, -0.002)))
    assert np.array_equal(np.asarray(g.points), pts)


def test_transmit_factories_match_reference_bitwise(usx_mod, ref, O):
    probe = usx_mod.presets.linear_5mhz()
    med = usx_mod.Medium()
    acq = usx_mod.Acquisition()
    for win in (usx_mod.Window.HANN, usx_mod.Window.RECT, usx_mod.Window.TUKEY):
        cfg = usx_mod.SimulatorConfig()
        cfg.tx_window = win
        sim = usx_mod.Simulator(probe, med, acq, cfg)
        s = O.Setup.of(probe, med, acq)
        sc = O.SimCfg(tx_window=int(win.value))
        ang = np.asarray(usx_mod.angles_deg(-15, 15, 7), np.float32)
        for mine, theirs in [
            (sim.make_plane_waves(list(ang)), ref.make_transmits(s, sc, 1, ang)),
            (sim.make_focused_scan(9, -0.01, 0.01, 0.03, 0.05),
             ref.make_transmits(s, sc, 0, [-0.01, 0.01, 0.03, 0.05], n=9)),
            (sim.make_diverging_waves(list(ang), 0.012),
             ref.make_transmits(s, sc, 2, ang, extra=0.012)),
        ]:
            for a, b in zip(mine, theirs):
                assert np.array_equal(np.asarray(a.delays, np.float32), b.delays)
                assert np.array_equal(np.asarray(a.apodization, np.float32), b.apodization)
                assert (a.focus.x, a.focus.y, a.focus.z) == tuple(np.float32(b.focus))


def test_host_geometry_matches_reference_bitwise(usx_mod, ref, O):
    probe = usx_mod.presets.linear_5mhz()
    probe.n_elements = 48
    s = O.Setup.of(probe, usx_mod.Medium(), usx_mod.Acquisition())
    tx = usx_mod.Transmit()
    tx.type = usx_mod.TransmitType.FOCUSED
    tx.focus = usx_mod.Vec3(0.002, 0, 0.025)
    tx.apodization = [0.0] * 8 + [1.0] * 32 + [0.0] * 8
    tx.delays = [0.0] * 48
    for t in (tx,):
        for p in [(0.0, 0.0, 0.02), (0.004, 0.0, 0.011), (-0.006, 0.0, 0.04)]:
            v = usx_mod.Vec3(*p)
            assert usx_mod.transmit_arrival_time(t, v, 1540.0) == pytest.approx(
                ref.transmit_arrival_time(O.Tx.of(t), p, 1540.0), rel=0, abs=0)
            assert usx_mod.transmit_sensitivity(probe, t, v, 3.08e-4) == pytest.approx(
                ref.transmit_sensitivity(s, O.Tx.of(t), p, 3.08e-4), rel=0, abs=0)
    for w in range(5):
        for u in np.linspace(-1.1, 1.1, 45, dtype=np.float32):
            assert usx_mod.window_value(usx_mod.Window(w), float(u), 0.35) == ref.window_value(
                w, float(u), 0.35)
    for n, fc in [(33, 0.125), (16, 0.4), (65, 0.02)]:
        assert np.array_equal(np.asarray(usx_mod.design_lowpass(n, fc)), ref.design_lowpass(n, fc))


def test_simulator_reproduces_golden_rf_bitwise(usx_mod):
    d = np.load(os.path.join(ROOT, "tests", "golden", "small_case.npz"))
    meta = json.load(open(os.path.join(ROOT, "tests", "golden", "small_case.json")))
    st = meta["setup"]
    probe = usx_mod.Probe()
    for k in ("n_elements", "pitch", "element_width", "elevation_height", "center_frequency",
              "bandwidth"):
        setattr(probe, k, st[k])
    acq = usx_mod.Acquisition()
    acq.sampling_frequency, acq.n_samples, acq.t0, acq.prf = (
        st["sampling_frequency"], st["n_samples"], st["t0"], st["prf"])
    sim = usx_mod.Simulator(probe, usx_mod.Medium(), acq)
    txs = sim.make_plane_waves(meta["angles_rad"])
    assert np.array_equal(np.stack([t.delays for t in txs]).astype(np.float32), d["delays"])
    ph = usx_mod.Phantom.speckle_box(-0.006, 0.006, 0.005, 0.026, 2e-4, 600, 0, 1.0)
    assert np.array_equal(np.asarray(ph.positions), d["positions"][:600])
    assert np.array_equal(np.asarray(ph.amplitudes), d["amplitudes"][:600])
    ph.extend(usx_mod.Phantom.point_targets([usx_mod.Vec3(0.001, 0, 0.012),
                                             usx_mod.Vec3(-0.002, 0, 0.019)], 10.0))
    rf = np.asarray(sim.simulate(txs, ph).rf)
    assert np.array_equal(rf, d["rf"]), "host simulator differs from the reference"


def test_channel_data_views_and_validation(usx_mod):
    probe = usx_mod.presets.linear_5mhz()
    probe.n_elements = 16
    acq = usx_mod.Acquisition()
    acq.n_samples = 100
    cfg = usx_mod.SimulatorConfig()
    cfg.noise_db = -200.0
    sim = usx_mod.Simulator(probe, usx_mod.Medium(), acq, cfg)
    cd = sim.simulate(sim.make_plane_waves([0.0]), usx_mod.Phantom.point_targets([usx_mod.Vec3(0, 0, 0.001)]))
    a = cd.rf
    assert a.shape == (1, 16, 100) and a.dtype == np.float32 and not a.flags.owndata
    a[0, 0, 0] = 12345.0
    assert cd.rf[0, 0, 0] == 12345.0
    with pytest.raises(RuntimeError):
        cd.iq
    with pytest.raises(ValueError):
        usx_mod.ChannelData.from_rf(np.zeros((2, 16, 100), np.float32), probe, usx_mod.Medium(),
                                    acq, sim.make_plane_waves([0.0]))
    bad = usx_mod.BeamformerConfig()
    bad.mv_subarray = 1
    with pytest.raises(ValueError, match="mv_subarray"):
        usx_mod.Beamformer(bad)
    with pytest.raises(ValueError):
        usx_mod.ScanGrid.cartesian(0, 1, 0, 0, 1, 4)
    p = usx_mod.Probe()
    p.element_width = 1.0
    with pytest.raises(ValueError, match="overlap"):
        usx_mod.Simulator(p, usx_mod.Medium(), acq)


def test_python_helpers_match_reference_definitions(usx_mod):
    a = usx_mod.angles_deg(-9, 9, 11)
    assert np.allclose(np.rad2deg(a), np.linspace(-9, 9, 11)) and usx_mod.angles_deg(0, 4, 1) == [
        np.deg2rad(2.0)]
    p = usx_mod.presets.linear_5mhz()
    g = usx_mod.grid_for(p, 0.005, 0.050, 256, 640)
    half = 0.5 * p.aperture_width()
    assert g.shape == (256, 640) and g.axis0[0] == np.float32(-half) and g.axis1[-1] == np.float32(0.05)
    acq = usx_mod.presets.acquisition(p, depth=0.04)
    assert acq.n_samples * 1.0 / acq.sampling_frequency >= 2 * 0.04 / 1540.0
    assert acq.max_depth(1540.0) >= 0.04
    assert usx_mod.presets.phased_3mhz().pitch < 0.55 * usx_mod.presets.phased_3mhz().wavelength(1540)


# ---- metrics ----------------------------------------------------------------

class _G:
    def __init__(self, n0, n1, x=(-0.01, 0.01), z=(0.01, 0.03)):
        self.axis0 = np.linspace(*x, n0)
        self.axis1 = np.linspace(*z, n1)


def test_speckle_snr_of_rayleigh_samples(usx_mod):
    rng = np.random.default_rng(0)
    g = _G(200, 300)
    env = np.abs(rng.normal(size=(200, 300)) + 1j * rng.normal(size=(200, 300)))
    snr = usx_mod.metrics.speckle_snr(env, g, (0.0, 0.02), (0.008, 0.008))
    assert snr == pytest.approx(usx_mod.metrics.RAYLEIGH_SPECKLE_SNR, rel=0.03)
    # closed form: sqrt(pi / (4 - pi)) = 1.9131...
    assert usx_mod.metrics.RAYLEIGH_SPECKLE_SNR == pytest.approx(np.sqrt(np.pi / (4 - np.pi)), abs=1e-3)


def test_gcnr_invariance_and_extremes(usx_mod):
    rng = np.random.default_rng(1)
    g = _G(60, 120, z=(0.018, 0.032))
    x = g.axis0[:, None]
    z = g.axis1[None, :]
    inside = np.hypot(x, z - 0.025) < 0.003
    env = np.abs(rng.normal(size=(60, 120)) + 1j * rng.normal(size=(60, 120)))
    env = np.where(inside, env * 0.15, env)
    g1 = usx_mod.metrics.gcnr(env, g, (0.0, 0.025), 0.003)
    for f in (np.sqrt, lambda v: v ** 2.5, np.log1p):
        assert usx_mod.metrics.gcnr(f(env), g, (0.0, 0.025), 0.003) == pytest.approx(g1, abs=1e-12)
    assert 0.2 < g1 <= 1.0
    sep = np.where(inside, 0.0, 1.0 + env)
    assert usx_mod.metrics.gcnr(sep, g, (0.0, 0.025), 0.003) == pytest.approx(1.0)
    same = np.ones_like(env)
    assert usx_mod.metrics.gcnr(same, g, (0.0, 0.025), 0.003) == pytest.approx(0.0, abs=1e-12)


def test_psf_of_an_analytic_gaussian(usx_mod):
    g = _G(401, 301, x=(-0.004, 0.004), z=(0.027, 0.033))
    sx, sz = 0.4e-3, 0.2e-3
    x = g.axis0[:, None]
    z = g.axis1[None, :]
    env = np.exp(-0.5 * (x / sx) ** 2 - 0.5 * ((z - 0.03) / sz) ** 2)
    psf = usx_mod.metrics.measure_psf(env, g, (0.0, 0.03))
    fwhm = 2 * np.sqrt(2 * np.log(2))
    assert psf.lateral_fwhm == pytest.approx(fwhm * sx, rel=2e-3)
    assert psf.axial_fwhm == pytest.approx(fwhm * sz, rel=2e-3)
    assert (psf.x, psf.z) == pytest.approx((0.0, 0.03), abs=3e-5)
    assert np.isnan(psf.sidelobe_db), "no pixels beyond 3 FWHM inside a 2 mm window"
    assert usx_mod.metrics.measure_psf(env, g, (0.0, 0.03), search_radius=0.004).sidelobe_db < -60


def test_contrast_and_cnr_closed_forms(usx_mod):
    g = _G(80, 80, z=(0.02, 0.04))
    x = g.axis0[:, None]
    z = g.axis1[None, :]
    r = np.hypot(x, z - 0.03)
    env = np.where(r <= 0.7 * 0.003, 0.1, 1.0)
    assert usx_mod.metrics.contrast_db(env, g, (0.0, 0.03), 0.003) == pytest.approx(-20.0, abs=1e-9)
    assert usx_mod.metrics.cnr(env, g, (0.0, 0.03), 0.003) > 1e10
