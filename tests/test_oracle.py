"""Pinning the CPU oracle (runs on the CPU, no GPU).

The oracle (oracle/usx_oracle.cpp) is the checker of the HIP path, so it is
itself checked, each part by something a plausible mistake would fail:
  * bitwise against the golden vectors written by the compiled reference
    (tests/golden/make_golden.py) and, where oracle/_ref is built, live
    against the reference on fresh cases;
  * against closed forms and independent float64 restatements in numpy:
    windows, the FIR design, demodulation of a pure tone and of a delta,
    transmit delay laws, a brute-force DAS on a tiny array, the O(N^2)
    pairwise DMAS, the coherence factor, an MV solved with LAPACK, and the
    MV effective-weight identity.
"""

from __future__ import annotations

import json
import os

import numpy as np
import pytest

from conftest import ROOT, rel_l2

GOLD = os.path.join(ROOT, "tests", "golden")


@pytest.fixture(scope="module")
def golden():
    d = dict(np.load(os.path.join(GOLD, "small_case.npz")))
    meta = json.load(open(os.path.join(GOLD, "small_case.json")))
    return d, meta


def _gold_inputs(O, golden):
    d, meta = golden
    s = O.Setup(**meta["setup"])
    txs = [O.Tx(1, float(a), (0.0, 0.0, 0.03), (0.0, 0.0, 0.0), d["delays"][i], d["apodization"][i])
           for i, a in enumerate(meta["angles_rad"])]
    g = O.GridSpec(**{**meta["grid"], "apex": tuple(meta["grid"]["apex"])})
    return s, txs, g


def bitwise(a, b):
    a = np.ascontiguousarray(a)
    b = np.ascontiguousarray(b)
    return a.shape == b.shape and np.array_equal(a.view(np.uint8), b.view(np.uint8))


# ---- golden vectors (written by the compiled reference) ------------------

def test_port_reproduces_golden_vectors(port, O, golden):
    d, meta = golden
    s, txs, g = _gold_inputs(O, golden)
    iq, fs, fd = port.demodulate(s, txs, d["rf"], O.Demod(decimation=4))
    assert bitwise(iq, d["iq"]) and fs == meta["demod"]["fs_out"] and fd == meta["demod"]["fdem"]
    assert bitwise(port.design_lowpass(33, 0.5 * 5e6 / 20e6), d["taps"])
    a0, a1, _ = port.grid(g)
    assert bitwise(a0, d["axis0"]) and bitwise(a1, d["axis1"])
    s_iq = O.Setup(**{**s.__dict__, "sampling_frequency": fs, "n_samples": iq.shape[2]})
    for k, cfg in meta["configs"].items():
        got = port.beamform(s_iq, txs, d["iq"], g, O.Bf(**cfg), fdem=fd)
        assert bitwise(got, d[f"frame_{k}"]), k
    assert bitwise(port.beamform(s, txs, d["rf"], g, O.Bf(f_number=1.5)), d["frame_das_rf"])


def test_golden_hashes_match_metadata(golden):
    import hashlib
    d, meta = golden
    for k, h in meta["sha256"].items():
        assert hashlib.sha256(np.ascontiguousarray(d[k]).tobytes()).hexdigest() == h, k


# ---- live cross-check against the compiled reference ---------------------

@pytest.fixture(scope="module")
def ref(O):
    if not O.available("reference"):
        pytest.skip("oracle/_ref not built here (reference sources absent)")
    return O.load("reference")


@pytest.mark.parametrize("seed,n_el,n_tx,fnum", [(1, 24, 2, 1.5), (2, 40, 5, 1.0), (3, 16, 1, 2.5)])
def test_port_matches_reference_live(port, ref, O, seed, n_el, n_tx, fnum):
    s = O.Setup(n_elements=n_el, bandwidth=0.65, sampling_frequency=16e6, n_samples=600)
    txs = ref.make_transmits(s, O.SimCfg(), 1,
                             np.deg2rad(np.linspace(-12, 12, n_tx)).astype(np.float32))
    pos, amp = ref.speckle_box(-0.006, 0.006, 0.004, 0.022, 2e-4, 300, seed, 1.0)
    rf = ref.simulate(s, O.SimCfg(), txs, pos, amp)
    g = O.GridSpec(0, -0.004, 0.004, 10, 0.005, 0.02, 30)
    for dec in (1, 3):
        for comb in range(3):
            for cf in (0, 1):
                for comp in range(3):
                    b = O.Bf(f_number=fnum, combiner=comb, coherence_factor=cf, compounding=comp,
                             mv_subarray=6, rx_window=(comb + comp) % 5)
                    a, _ = port.process_frame(s, txs, rf, g, O.Demod(decimation=dec), b)
                    r, _ = ref.process_frame(s, txs, rf, g, O.Demod(decimation=dec), b)
                    assert bitwise(a, r), (dec, comb, cf, comp)


# Fields a BeamformerConfig can hold once assigned after construction, which
# the constructor would reject (beamformer.cpp:218-225) and beamform clamps.
MUTATED = [("f_number", -1.0, 0), ("f_number", 0.0, 1), ("mv_subarray", 1, 2),
           ("mv_subarray", -7, 2), ("mv_diagonal_loading", -1.0, 2), ("f_number", -3.0, 2),
           ("mv_range_average", -1, 0), ("mv_range_average", -2, 1)]


def test_port_matches_reference_on_configs_assigned_after_construction(port, ref, O):
    s = O.Setup(n_elements=24, bandwidth=0.65, sampling_frequency=16e6, n_samples=600)
    txs = ref.make_transmits(s, O.SimCfg(), 1, np.deg2rad(np.linspace(-8, 8, 3)).astype(np.float32))
    pos, amp = ref.speckle_box(-0.006, 0.006, 0.004, 0.022, 2e-4, 300, 4, 1.0)
    rf = ref.simulate(s, O.SimCfg(), txs, pos, amp)
    g = O.GridSpec(0, -0.004, 0.004, 10, 0.005, 0.02, 30)
    for field, value, comb in MUTATED:
        b = O.Bf(**{"combiner": comb, "mv_subarray": 6, field: value})
        for lib in (port, ref):
            with pytest.raises(ValueError, match=field):
                lib.validate_bf(b)
        a, _ = port.process_frame(s, txs, rf, g, O.Demod(decimation=2), b)
        r, _ = ref.process_frame(s, txs, rf, g, O.Demod(decimation=2), b)
        assert bitwise(a, r), (field, value, comb)
        assert np.any(a != 0)
        # the clamps: the same frame as the in-range value they clamp to
        clamp = {"f_number": 1e-3, "mv_subarray": 2, "mv_diagonal_loading": 0.0}.get(field)
        if clamp is not None:
            c, _ = port.process_frame(s, txs, rf, g, O.Demod(decimation=2),
                                      O.Bf(**{"combiner": comb, "mv_subarray": 6, field: clamp}))
            assert bitwise(a, c), (field, value)
    # minimum variance with a negative range average cannot run in the reference
    bad = O.Bf(combiner=2, mv_range_average=-1)
    for lib in (port, ref):
        with pytest.raises(ValueError, match="mv_range_average"):
            lib.process_frame(s, txs, rf, g, O.Demod(decimation=2), bad)
    assert port.validate_bf(O.Bf()) is None and ref.validate_bf(O.Bf()) is None


def test_port_matches_reference_small_functions(port, ref, O):
    for w in range(5):
        for u in np.linspace(-1.2, 1.2, 97, dtype=np.float32):
            assert port.window_value(w, float(u), 0.3) == ref.window_value(w, float(u), 0.3)
    for n, fc in [(33, 0.125), (8, 0.3), (2, 0.6), (65, 1e-6)]:
        assert bitwise(port.design_lowpass(n, fc), ref.design_lowpass(n, fc))
    s = O.Setup(n_elements=48)
    tx_pw = O.Tx(1, 0.2, (0, 0, 0.03), (0.001, 0, 0))
    tx_f = O.Tx(0, 0.0, (0.002, 0, 0.025), (0.0, 0, 0), np.zeros(48, np.float32),
                np.r_[np.zeros(8), np.ones(32), np.zeros(8)].astype(np.float32))
    tx_d = O.Tx(2, 0.1, (-0.001, 0, -0.01), (0, 0, 0))
    for tx in (tx_pw, tx_f, tx_d):
        for p in [(0.0, 0.0, 0.02), (0.004, 0.0, 0.011), (-0.006, 0.0, 0.04), (0.0, 0.0, 0.025)]:
            assert port.transmit_arrival_time(tx, p, 1540.0) == ref.transmit_arrival_time(tx, p, 1540.0)
            assert (port.transmit_sensitivity(s, tx, p, 3.08e-4)
                    == ref.transmit_sensitivity(s, tx, p, 3.08e-4))
    for g in (O.GridSpec(0, -0.01, 0.01, 7, 0.001, 0.04, 9),
              O.GridSpec(1, -0.6, 0.6, 11, 0.01, 0.08, 5, (0.001, 0, -0.002))):
        for x, y in zip(port.grid(g), ref.grid(g)):
            assert bitwise(x, y)


# ---- closed forms and independent restatements ----------------------------

def test_windows_closed_forms(port):
    # reference core/tests/test_main.cpp:125-133
    assert port.window_value(1, 0.0, 0.5) == pytest.approx(1.0, abs=1e-6)
    assert port.window_value(1, -1.0, 0.5) == pytest.approx(0.0, abs=1e-6)
    assert port.window_value(1, 1.0, 0.5) == pytest.approx(0.0, abs=1e-6)
    assert port.window_value(0, 0.5, 0.5) == 1.0 and port.window_value(0, 1.5, 0.5) == 0.0
    assert port.window_value(3, 0.0, 0.5) == 1.0
    u = np.linspace(-1, 1, 41)
    t = 0.5 * (u + 1)
    for w, f in [(1, 0.5 * (1 - np.cos(2 * np.pi * t))), (2, 0.54 - 0.46 * np.cos(2 * np.pi * t)),
                 (4, 0.42 - 0.5 * np.cos(2 * np.pi * t) + 0.08 * np.cos(4 * np.pi * t))]:
        got = np.array([port.window_value(w, float(x), 0.5) for x in u])
        assert np.allclose(got, f, atol=2e-6), w


def test_design_lowpass_is_the_windowed_sinc(port):
    for n, fc in [(33, 0.125), (21, 0.3), (9, 0.05)]:
        h = port.design_lowpass(n, fc).astype(np.float64)
        k = np.arange(n) - n // 2
        ideal = np.where(k == 0, 2 * fc, np.sin(2 * np.pi * fc * k) / (np.pi * np.where(k == 0, 1, k)))
        ref = ideal * (0.54 - 0.46 * np.cos(2 * np.pi * np.arange(n) / (n - 1)))
        ref /= ref.sum()
        assert len(h) == n and np.allclose(h, ref, atol=2e-7)
        assert abs(h.sum() - 1) < 1e-6 and np.allclose(h, h[::-1], atol=1e-7)
    assert len(port.design_lowpass(8, 0.2)) == 9 and len(port.design_lowpass(1, 0.2)) == 3


def test_demodulation_tone_and_delta(port, O):
    # pure tone at f0 -> |IQ| = 1 (reference test_main.cpp:228-250)
    s = O.Setup(n_elements=4, sampling_frequency=40e6, n_samples=512, bandwidth=0.6)
    n = np.arange(512, dtype=np.float32)
    t = (np.float32(0) + n / np.float32(40e6)).astype(np.float32)
    rf = np.tile(np.cos(2 * np.pi * 5e6 * t.astype(np.float64)).astype(np.float32), (1, 4, 1))
    tx = [O.Tx()]
    iq, fs, fd = port.demodulate(s, tx, rf, O.Demod(decimation=1))
    assert np.mean(np.abs(iq[0, 0, 64:-64])) == pytest.approx(1.0, abs=0.02) and fd == 5e6
    iq4, fs4, _ = port.demodulate(s, tx, rf, O.Demod(decimation=4))
    assert iq4.shape[2] == (512 + 3) // 4 and fs4 == pytest.approx(10e6)
    # a delta at sample n0 comes out as 2 * h[k*dec + gd - n0] * exp(-j 2 pi f t_n0)
    d = np.zeros((1, 4, 512), np.float32)
    n0 = 200
    d[0, 2, n0] = 1.0
    for dec in (1, 4):
        iq, _, _ = port.demodulate(s, tx, d, O.Demod(decimation=dec))
        h = port.design_lowpass(33, 0.5 * 5e6 / 40e6).astype(np.float64)
        rot = np.exp(-2j * np.pi * 5e6 * (n0 / 40e6))
        k = np.arange(iq.shape[2])
        j = k * dec + 16 - n0
        want = np.where((j >= 0) & (j < 33), 2 * h[np.clip(j, 0, 32)] * rot, 0)
        assert np.allclose(iq[0, 2], want, atol=2e-6)
        assert not np.any(iq[0, [0, 1, 3]])


def test_transmit_geometry_closed_forms(port, O):
    # plane wave: arrival = projection / c; t = 0 at the origin (test_main.cpp:167-185)
    pw = O.Tx(1, 0.2, (0, 0, 0.03), (0, 0, 0))
    p = (0.004, 0.0, 0.025)
    want = (p[0] * np.sin(np.float32(0.2)) + p[2] * np.cos(np.float32(0.2))) / 1540.0
    assert port.transmit_arrival_time(pw, p, 1540.0) == pytest.approx(want, abs=1e-12)
    assert port.transmit_arrival_time(pw, (0, 0, 0), 1540.0) == 0.0
    # focused: arrival at the focus = |focus - origin| / c, t = 0 at the origin
    f = O.Tx(0, 0.0, (0.003, 0, 0.03), (0.001, 0, 0))
    assert port.transmit_arrival_time(f, (0.003, 0, 0.03), 1540.0) == pytest.approx(
        np.hypot(0.002, 0.03) / 1540.0, rel=1e-6)
    assert port.transmit_arrival_time(f, (0.001, 0, 0), 1540.0) == pytest.approx(0.0, abs=1e-9)
    # diverging: t = 0 at the origin, increasing with depth
    dv = O.Tx(2, 0.0, (0, 0, -0.01), (0, 0, 0))
    assert port.transmit_arrival_time(dv, (0, 0, 0), 1540.0) == pytest.approx(0.0, abs=1e-9)
    assert port.transmit_arrival_time(dv, (0, 0, 0.03), 1540.0) == pytest.approx(0.03 / 1540.0, rel=1e-6)


def _tiny_case(O, port, n_el=4, fnum=1.0):
    rng = np.random.default_rng(7)
    s = O.Setup(n_elements=n_el, sampling_frequency=10e6, n_samples=200, center_frequency=5e6)
    iq = (rng.normal(size=(1, n_el, 200)) + 1j * rng.normal(size=(1, n_el, 200))).astype(np.complex64)
    tx = [O.Tx(1, 0.05, (0, 0, 0.03), (0, 0, 0))]
    g = O.GridSpec(0, -0.0004, 0.0004, 3, 0.004, 0.005, 3)
    return s, iq, tx, g


def test_das_against_float64_restatement(port, O):
    """Tiny array (4 channels, 1 plane wave, 3x3 pixels): an independent
    float64 DAS from the geometry (a sign or index slip fails by >> 1e-3)."""
    s, iq, tx, g = _tiny_case(O, port)
    fd = 5e6
    frame = port.beamform(s, tx, iq, g, O.Bf(f_number=1.0, transmit_weighting=0), fdem=fd)
    x = np.linspace(-0.0004, 0.0004, 3)
    z = np.linspace(0.004, 0.005, 3)
    ex = (np.arange(4) - 1.5) * 300e-6
    want = np.zeros((3, 3), complex)
    for i in range(3):
        for j in range(3):
            ttx = (x[i] * np.sin(0.05) + z[j] * np.cos(0.05)) / 1540.0
            half = 0.5 * z[j] / 1.0
            for ch in range(4):
                u = (ex[ch] - x[i]) / half
                if abs(u) > 1:
                    continue
                a = 0.5 * (1 - np.cos(2 * np.pi * 0.5 * (u + 1)))
                trx = np.hypot(ex[ch] - x[i], z[j]) / 1540.0
                idx = (ttx + trx) * 10e6
                k = int(np.floor(idx))
                f = idx - k
                v = iq[0, ch, k] * (1 - f) + iq[0, ch, k + 1] * f
                want[i, j] += v * np.exp(2j * np.pi * fd * (ttx + trx)) * a
    assert rel_l2(frame, want) < 2e-3


def test_dmas_cf_and_das_from_channel_vectors(port, O, golden):
    """The combiners against their definitions, applied in numpy to the
    oracle's own per-pixel channel vectors: DMAS as the O(N^2) pairwise sum
    of y_i y_j / sqrt(|y_i||y_j|), CF = |sum y|^2 / (N sum |y|^2)."""
    d, meta = golden
    s, txs, g = _gold_inputs(O, golden)
    fs, fd = meta["demod"]["fs_out"], meta["demod"]["fdem"]
    s_iq = O.Setup(**{**s.__dict__, "sampling_frequency": fs, "n_samples": d["iq"].shape[2]})
    das, ex = port.beamform(s_iq, txs, d["iq"], g, O.Bf(f_number=1.5), fdem=fd, ych=True, info=True)
    y = ex["ych"].astype(np.complex128)
    assert rel_l2(das.ravel(), y.sum(1)) < 1e-6
    dmas = port.beamform(s_iq, txs, d["iq"], g, O.Bf(f_number=1.5, combiner=1), fdem=fd).ravel()
    cfr = port.beamform(s_iq, txs, d["iq"], g, O.Bf(f_number=1.5, coherence_factor=1), fdem=fd).ravel()
    info = ex["info"]
    checked = 0
    for p in range(0, y.shape[0], 7):
        n = int(info[p, 0])
        v = y[p][y[p] != 0]
        if n == 0 or v.size == 0:
            continue
        pair = 0j
        for i in range(v.size):
            for j in range(i + 1, v.size):
                pair += v[i] * v[j] / np.sqrt(abs(v[i]) * abs(v[j]))
        scale = 0.5 * max(1, n - 1)
        # float32 evaluation of ((sum v)^2 - sum v^2) / 2 against the pairwise sum
        assert abs(dmas[p] - pair / scale) <= 1e-5 * np.sum(np.abs(v)) ** 2 / scale, p
        cf = abs(v.sum()) ** 2 / (n * np.sum(np.abs(v) ** 2))
        assert abs(cfr[p] - das.ravel()[p] * cf) <= 1e-5 * abs(das.ravel()[p]) + 1e-9, p
        checked += 1
    assert checked > 50


def _mv_numpy(yrows, kw, act, L, loading):
    """Minimum variance of one pixel in float64 with LAPACK (numpy.linalg.solve)."""
    na = len(act)
    ns = na - L + 1
    R = np.zeros((L, L), complex)
    n = 0
    for row in yrows:
        for s0 in range(ns):
            x = row[act[s0:s0 + L]]
            R += np.outer(x, x.conj())
            n += 1
    R /= n
    R += max(loading * np.trace(R).real / L, 1e-20) * np.eye(L)
    w = np.linalg.solve(R, np.ones(L))
    cen = yrows[kw]
    avg = np.array([cen[act[a:a + ns]].mean() for a in range(L)])
    return (w.conj() @ avg) / w.sum() * na


def test_mv_against_lapack_and_weight_identity(port, O, golden):
    d, meta = golden
    s, txs, g = _gold_inputs(O, golden)
    fs, fd = meta["demod"]["fs_out"], meta["demod"]["fdem"]
    s_iq = O.Setup(**{**s.__dict__, "sampling_frequency": fs, "n_samples": d["iq"].shape[2]})
    b = O.Bf(f_number=1.5, combiner=2, mv_subarray=8)
    f32 = port.beamform(s_iq, txs, d["iq"], g, b, fdem=fd)
    f64, ex = port.beamform(s_iq, txs, d["iq"], g, b, fdem=fd, mv_fp64=True, ych=True,
                            weights=True, info=True)
    n0, n1 = g.n0, g.n1
    y = ex["ych"].reshape(n0, n1, -1).astype(np.complex128)
    assert rel_l2(f32, f64) < 1e-4
    checked = 0
    for i in range(0, n0, 3):
        for j in range(0, n1, 5):
            na, L, fb = ex["info"][i * n1 + j]
            if fb:
                continue
            act = np.flatnonzero(y[i, j] != 0)
            if len(act) != na:
                continue  # a zero-valued active channel; covered elsewhere
            rows = [y[i, k] for k in range(max(0, j - 2), min(n1, j + 3))]
            want = _mv_numpy(rows, j - max(0, j - 2), act, L, 1e-2)
            assert abs(f64[i, j] - want) <= 2e-7 * abs(want) + 1e-12, (i, j)  # output rounded to complex64
            checked += 1
    assert checked > 10
    # out == sum_c g_c Y[c] with the effective per-channel weights
    wts = ex["weights"].astype(np.complex128).reshape(n0, n1, -1)
    assert rel_l2((wts * y).sum(-1), f64) < 1e-6


def test_display_matches_reference_and_closed_forms(O):
    """The display tail (pipeline.cpp:48-70 over postprocess.cpp:12-117):
    the restatement equals the compiled reference bit for bit (IQ frames,
    RF frames through the FFT analytic signal, a sector grid, the smoothed
    reference over 3 frames), and obeys closed forms."""
    port = O.load("port")
    s = O.Setup()
    rng = np.random.default_rng(1)
    cases = [(O.GridSpec(0, -0.01, 0.01, 12, 0.005, 0.04, 100), 5e6),
             (O.GridSpec(0, -0.01, 0.01, 9, 0.005, 0.04, 77), 0.0),
             (O.GridSpec(1, -0.5, 0.5, 10, 0.01, 0.06, 64), 3e6)]
    for gs, fdem in cases:
        fr = (rng.normal(size=(3, gs.n0, gs.n1)) + 1j * rng.normal(size=(3, gs.n0, gs.n1))).astype(np.complex64)
        fr *= np.array([1, 5, 2], np.float32)[:, None, None]
        got = port.display(s, gs, fr, fdem)
        if O.available("reference"):
            ref = O.load("reference").display(s, gs, fr, fdem)
            assert np.array_equal(got.view(np.uint32), ref.view(np.uint32)), (gs.type, fdem)
        # first frame: the reference level is its own peak, so the brightest
        # pixel shows 1.0 and a pixel at -dr/2 dB below it 0.5
        assert got[0].max() == 1.0
    # closed forms on an IQ frame without TGC: env = |y|
    gs = O.GridSpec(0, -0.01, 0.01, 4, 0.01, 0.02, 8)
    fr = np.full((1, 4, 8), 1.0, np.complex64)
    fr[0, 0, 0] = 10.0
    fr[0, 1, 1] = 10.0 * 10 ** (-30 / 20)
    fr[0, 2, 2] = 0.0
    img = port.display(s, gs, fr, 5e6, apply_tgc=False, dynamic_range_db=60.0)[0]
    assert img[0, 0] == 1.0
    assert img[1, 1] == pytest.approx(0.5, abs=1e-6)
    assert img[2, 2] == 0.0
    assert img[3, 3] == pytest.approx((20 * np.log10(0.1) + 60) / 60, abs=1e-6)
    # TGC: the gain at depth z is 10^(2 alpha f z_cm / 20) (no smoothing effect on the ratio)
    img_t = port.display(s, gs, np.ones((1, 4, 8), np.complex64), 5e6, apply_tgc=True,
                         dynamic_range_db=200.0)[0]
    z = np.linspace(0.01, 0.02, 8)
    gain_db = 2 * 0.5 * 5.0 * z * 100
    want = (gain_db - gain_db.max() + 200.0) / 200.0
    assert np.allclose(img_t[0], want, atol=2e-6)
    # silence stays black
    assert not port.display(s, gs, np.zeros((2, 4, 8), np.complex64), 5e6).any()
