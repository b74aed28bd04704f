"""Parity of the HIP imaging path with the CPU oracle (gfx950 only).

Each test drives the product API (usx.*), whose C++ host layer calls the C
ABI of include/usx_gpu.h, on seeded RF, and compares with the oracle's
restatement of the reference on the same inputs. Bars: demodulation
bit-exact; DAS, DMAS and coherence factor within 1e-5 relative L2; MV within
1e-4 relative L2 (the reference's own FP32 MV differs from FP64 by ~1e-6).
"""

from __future__ import annotations

import numpy as np
import pytest

from cases import bf_config, make_case, oracle_inputs
from conftest import rel_l2

pytestmark = pytest.mark.gpu

TOL_DAS = 1e-5
TOL_MV = 1e-4


@pytest.fixture(scope="module")
def small(usx_mod, O):
    return make_case(usx_mod, O)


@pytest.fixture(scope="module")
def small_iq(usx_mod, small):
    d = usx_mod.DemodConfig()
    d.decimation = 4
    return usx_mod.demodulate(small.raw, d)


def _oracle_frame(port, O, case, bfc, dec=4, data=None, ev=(0, -1), **kw):
    s, txs = oracle_inputs(O, case)
    rf = np.asarray(case.raw.rf)
    if data is None:
        iq, fs, fd = port.demodulate(s, txs, rf, O.Demod(decimation=dec))
        s2 = O.Setup(**{**s.__dict__, "sampling_frequency": fs, "n_samples": iq.shape[2]})
        return port.beamform(s2, txs, iq, case.gspec, O.Bf.of(bfc), fdem=fd,
                             event_begin=ev[0], event_end=ev[1], **kw)
    return port.beamform(s, txs, data, case.gspec, O.Bf.of(bfc), event_begin=ev[0],
                         event_end=ev[1], **kw)


def test_demodulation_is_bit_exact(usx_mod, port, O, small):
    for dec in (1, 2, 4):
        d = usx_mod.DemodConfig()
        d.decimation = dec
        iq = usx_mod.demodulate(small.raw, d)
        s, txs = oracle_inputs(O, small)
        ref, fs, fd = port.demodulate(s, txs, np.asarray(small.raw.rf), O.Demod(decimation=dec))
        got = np.asarray(iq.iq)
        assert got.shape == ref.shape
        assert np.array_equal(got.view(np.uint32), ref.view(np.uint32)), f"decimation {dec}"
        assert iq.acq.sampling_frequency == pytest.approx(fs, rel=0)
        assert iq.demodulation_frequency == fd


@pytest.mark.parametrize("combiner,cf", [("DAS", False), ("DAS", True), ("DMAS", False),
                                         ("DMAS", True), ("MV", False), ("MV", True)])
@pytest.mark.parametrize("compounding", ["COHERENT", "INCOHERENT", "SELECT"])
def test_process_frame_matches_oracle(usx_mod, port, O, small, combiner, cf, compounding):
    bfc = bf_config(usx_mod, combiner, cf, compounding, subarray=8)
    pc = usx_mod.PipelineConfig()
    pc.demod.decimation = 4
    pc.beamformer = bfc
    frame = usx_mod.ImagingPipeline(small.grid, pc).process_frame(small.raw)
    got = np.asarray(frame.data)
    assert got.dtype == np.complex64 and got.shape == (small.grid.n_axis0, small.grid.n_axis1)
    ref = _oracle_frame(port, O, small, bfc)
    tol = TOL_MV if combiner == "MV" else TOL_DAS
    assert rel_l2(got, ref) < tol, f"{combiner} cf={cf} {compounding}: {rel_l2(got, ref):.3e}"


@pytest.mark.parametrize("window", ["RECT", "HANN", "HAMMING", "TUKEY", "BLACKMAN"])
def test_rx_windows_match_oracle(usx_mod, port, O, small, small_iq, window):
    for combiner in ("DAS", "DMAS", "MV"):
        bfc = bf_config(usx_mod, combiner, window=window, subarray=8)
        got = np.asarray(usx_mod.Beamformer(bfc).beamform(small_iq, small.grid).data)
        s, txs = oracle_inputs(O, small)
        s2 = O.Setup(**{**s.__dict__, "sampling_frequency": small_iq.acq.sampling_frequency,
                        "n_samples": small_iq.n_samples})
        ref = port.beamform(s2, txs, np.asarray(small_iq.iq), small.gspec, O.Bf.of(bfc),
                            fdem=small_iq.demodulation_frequency)
        tol = TOL_MV if combiner == "MV" else TOL_DAS
        assert rel_l2(got, ref) < tol, f"{window} {combiner}: {rel_l2(got, ref):.3e}"


def test_event_ranges_and_rf_input(usx_mod, port, O, small, small_iq):
    bfc = bf_config(usx_mod, "DAS")
    bf = usx_mod.Beamformer(bfc)
    s, txs = oracle_inputs(O, small)
    s2 = O.Setup(**{**s.__dict__, "sampling_frequency": small_iq.acq.sampling_frequency,
                    "n_samples": small_iq.n_samples})
    for ev in [(0, 1), (1, 3), (2, -1), (3, 3)]:
        got = np.asarray(bf.beamform(small_iq, small.grid, ev[0], ev[1]).data)
        ref = port.beamform(s2, txs, np.asarray(small_iq.iq), small.gspec, O.Bf.of(bfc),
                            fdem=small_iq.demodulation_frequency, event_begin=ev[0],
                            event_end=ev[1])
        if not np.any(ref):
            assert not np.any(got)
        else:
            assert rel_l2(got, ref) < TOL_DAS, ev
    # Beamforming RF directly (no demodulation): real-valued interpolation, no phase.
    got = np.asarray(bf.beamform(small.raw, small.grid).data)
    ref = port.beamform(s, txs, np.asarray(small.raw.rf), small.gspec, O.Bf.of(bfc))
    assert rel_l2(got, ref) < TOL_DAS
    assert np.all(got.imag == 0)


def test_many_events_chunking(usx_mod, port, O):
    # 21 transmits: more than one 16-slot chunk on every path.
    case = make_case(usx_mod, O, n_el=24, n_tx=21, n_scat=300, grid=(12, 0.008, 0.02, 40))
    for combiner, cf in [("DAS", False), ("DAS", True), ("DMAS", False), ("MV", False)]:
        bfc = bf_config(usx_mod, combiner, cf, subarray=6)
        pc = usx_mod.PipelineConfig()
        pc.demod.decimation = 2
        pc.beamformer = bfc
        got = np.asarray(usx_mod.ImagingPipeline(case.grid, pc).process_frame(case.raw).data)
        ref = _oracle_frame(port, O, case, bfc, dec=2)
        tol = TOL_MV if combiner == "MV" else TOL_DAS
        assert rel_l2(got, ref) < tol, combiner


def test_focused_scan_select_and_sector_grid(usx_mod, port, O):
    probe = usx_mod.presets.linear_5mhz()
    probe.n_elements = 32
    acq = usx_mod.Acquisition()
    acq.sampling_frequency = 20e6
    acq.n_samples = 900
    cfg = usx_mod.SimulatorConfig()
    cfg.noise_db = -200.0
    sim = usx_mod.Simulator(probe, usx_mod.Medium(), acq, cfg)
    ph = usx_mod.Phantom.point_targets([usx_mod.Vec3(0.001, 0, 0.02), usx_mod.Vec3(-0.002, 0, 0.03)])
    txs = sim.make_focused_scan(12, -0.004, 0.004, 0.025)
    raw = sim.simulate(txs, ph)
    iq = usx_mod.demodulate(raw, usx_mod.DemodConfig())
    s = O.Setup.of(probe, usx_mod.Medium(), iq.acq)
    otx = [O.Tx.of(t) for t in txs]
    grids = [(usx_mod.ScanGrid.cartesian(-0.005, 0.005, 20, 0.012, 0.036, 60),
              O.GridSpec(0, -0.005, 0.005, 20, 0.012, 0.036, 60)),
             (usx_mod.ScanGrid.sector(-0.3, 0.3, 18, 0.012, 0.036, 50, usx_mod.Vec3(0, 0, -0.002)),
              O.GridSpec(1, -0.3, 0.3, 18, 0.012, 0.036, 50, (0.0, 0.0, -0.002)))]
    for g, gs in grids:
        for comp in ("SELECT", "COHERENT", "INCOHERENT"):
            for combiner in ("DAS", "MV"):
                bfc = bf_config(usx_mod, combiner, compounding=comp, subarray=6)
                got = np.asarray(usx_mod.Beamformer(bfc).beamform(iq, g).data)
                ref = port.beamform(s, otx, np.asarray(iq.iq), gs, O.Bf.of(bfc),
                                    fdem=iq.demodulation_frequency)
                tol = TOL_MV if combiner == "MV" else TOL_DAS
                assert rel_l2(got, ref) < tol, (g.type, comp, combiner, rel_l2(got, ref))


def test_diverging_waves(usx_mod, port, O):
    probe = usx_mod.presets.phased_3mhz()
    acq = usx_mod.Acquisition()
    acq.sampling_frequency = 12e6
    acq.n_samples = 700
    sim = usx_mod.Simulator(probe, usx_mod.Medium(), acq, usx_mod.SimulatorConfig())
    ph = usx_mod.Phantom.speckle_box(-0.01, 0.01, 0.01, 0.04, 2e-4, 400, 5)
    txs = sim.make_diverging_waves(usx_mod.angles_deg(-10, 10, 3), 0.01)
    iq = usx_mod.demodulate(sim.simulate(txs, ph), usx_mod.DemodConfig())
    g = usx_mod.ScanGrid.sector(-0.5, 0.5, 24, 0.01, 0.04, 48, usx_mod.Vec3(0, 0, 0))
    gs = O.GridSpec(1, -0.5, 0.5, 24, 0.01, 0.04, 48)
    s = O.Setup.of(probe, usx_mod.Medium(), iq.acq)
    for combiner in ("DAS", "DMAS"):
        bfc = bf_config(usx_mod, combiner)
        got = np.asarray(usx_mod.Beamformer(bfc).beamform(iq, g).data)
        ref = port.beamform(s, [O.Tx.of(t) for t in txs], np.asarray(iq.iq), gs, O.Bf.of(bfc),
                            fdem=iq.demodulation_frequency)
        assert rel_l2(got, ref) < TOL_DAS, combiner


def test_mv_fallbacks(usx_mod, port, O, small, small_iq):
    s, txs = oracle_inputs(O, small)
    s2 = O.Setup(**{**s.__dict__, "sampling_frequency": small_iq.acq.sampling_frequency,
                    "n_samples": small_iq.n_samples})
    # A large f-number leaves shallow pixels with < 4 active channels: those
    # take the DAS sum (beamformer.cpp:142-146), the rest the MV solve.
    bfc = bf_config(usx_mod, "MV", f_number=6.0, subarray=8)
    got = np.asarray(usx_mod.Beamformer(bfc).beamform(small_iq, small.grid).data).ravel()
    ref, ex = port.beamform(s2, txs, np.asarray(small_iq.iq), small.gspec, O.Bf.of(bfc),
                            fdem=small_iq.demodulation_frequency, info=True)
    ref = ref.ravel()
    fb = ex["info"][:, 2] == 1
    assert fb.any() and (~fb).any()
    assert rel_l2(got[fb], ref[fb]) < TOL_DAS
    assert rel_l2(got[~fb], ref[~fb]) < TOL_MV
    # Zero loading on all-zero data: every pivot is 0, so the Cholesky fails
    # and the DAS sum (0) is returned (beamformer.cpp:199).
    zero = usx_mod.ChannelData.from_iq(np.zeros(small_iq.shape, np.complex64), small.probe,
                                       small.medium, small_iq.acq, small.txs,
                                       small_iq.demodulation_frequency)
    bfc = bf_config(usx_mod, "MV", loading=0.0, subarray=8)
    out = np.asarray(usx_mod.Beamformer(bfc).beamform(zero, small.grid).data)
    assert np.all(out == 0) and np.all(np.isfinite(out.view(np.float32)))


def test_point_target_physics(usx_mod):
    """Reference physics checks (core/tests/test_main.cpp:262-396) on the GPU path."""
    probe = usx_mod.Probe()
    probe.n_elements = 96
    probe.bandwidth = 0.6
    acq = usx_mod.Acquisition()
    acq.sampling_frequency = 40e6
    acq.n_samples = 2600
    acq.prf = 4000
    cfg = usx_mod.SimulatorConfig()
    cfg.noise_db = -200.0
    sim = usx_mod.Simulator(probe, usx_mod.Medium(), acq, cfg)
    ph = usx_mod.Phantom.point_targets([usx_mod.Vec3(0.0, 0, 0.030)])
    iq = usx_mod.demodulate(sim.simulate(sim.make_plane_waves([0.0]), ph), usx_mod.DemodConfig())
    grid = usx_mod.ScanGrid.cartesian(-0.004, 0.004, 161, 0.028, 0.032, 81)
    widths = {}
    for combiner in ("DAS", "DMAS", "MV"):
        bfc = bf_config(usx_mod, combiner, subarray=16)
        env = np.abs(np.asarray(usx_mod.Beamformer(bfc).beamform(iq, grid).data))
        i, j = np.unravel_index(env.argmax(), env.shape)
        assert abs(grid.axis0[i]) <= 1.1e-4 and abs(grid.axis1[j] - 0.030) <= 1.1e-4
        widths[combiner] = usx_mod.metrics.measure_psf(env, grid, (0.0, 0.030)).lateral_fwhm
    assert widths["DMAS"] < 0.95 * widths["DAS"]
    assert widths["MV"] < 0.95 * widths["DAS"]


@pytest.mark.parametrize("field,value,combiner", [
    ("f_number", -1.0, "DAS"), ("f_number", 0.0, "DMAS"), ("f_number", -2.0, "MV"),
    ("mv_subarray", 1, "MV"), ("mv_subarray", -3, "MV"), ("mv_diagonal_loading", -1.0, "MV"),
    ("mv_range_average", -1, "DAS")])
def test_config_assigned_after_construction_matches_oracle(usx_mod, port, O, small, small_iq,
                                                           field, value, combiner):
    """The reference validates BeamformerConfig when a Beamformer is built
    (beamformer.hpp:110) and never in beamform (beamformer.cpp:297-299): a
    config assigned afterwards (pymodule.cpp:430-431, 617-619) runs with the
    clamps of geometry.cpp:127 and beamformer.cpp:147, 187. The oracle follows
    the compiled reference bit for bit there (tests/test_oracle.py)."""
    bad = bf_config(usx_mod, combiner, subarray=8)
    setattr(bad, field, value)
    with pytest.raises(ValueError, match=field):
        usx_mod.Beamformer(bad)
    pc = usx_mod.PipelineConfig()
    pc.demod.decimation = 4
    pc.beamformer = bf_config(usx_mod, combiner, subarray=8)
    pipe = usx_mod.ImagingPipeline(small.grid, pc)
    cfg = pipe.config
    cfg.beamformer = bad
    pipe.config = cfg
    got = np.asarray(pipe.process_frame(small.raw).data)
    ref = _oracle_frame(port, O, small, bad)
    tol = TOL_MV if combiner == "MV" else TOL_DAS
    assert np.any(ref != 0) and rel_l2(got, ref) < tol, f"{field}={value}: {rel_l2(got, ref):.3e}"
    bf = usx_mod.Beamformer(bf_config(usx_mod, combiner, subarray=8))
    bf.config = bad
    got = np.asarray(bf.beamform(small_iq, small.grid).data)
    assert rel_l2(got, ref) < tol


def test_mv_with_negative_range_average_raises(usx_mod, port, O, small):
    """Minimum variance with a negative range average is the one mutated
    config the reference cannot run (its 2 r + 1 entry scratch vector,
    beamformer.cpp:357): ValueError, as the oracle reports it."""
    bad = bf_config(usx_mod, "MV", subarray=8)
    bad.mv_range_average = -1
    pc = usx_mod.PipelineConfig()
    pipe = usx_mod.ImagingPipeline(small.grid, pc)
    cfg = pipe.config
    cfg.beamformer = bad
    pipe.config = cfg
    with pytest.raises(ValueError, match="mv_range_average"):
        pipe.process_frame(small.raw)
    with pytest.raises(ValueError, match="mv_range_average"):
        _oracle_frame(port, O, small, bad)


def test_frame_metadata_and_stats(usx_mod, small):
    pc = usx_mod.PipelineConfig()
    pc.demod.decimation = 4
    pipe = usx_mod.ImagingPipeline(small.grid, pc)
    f0 = pipe.process_frame(small.raw)
    f1 = pipe.process_frame(small.raw)
    assert (f0.sequence, f1.sequence) == (0, 1) and pipe.frame_count == 2
    assert f0.demodulation_frequency == pytest.approx(small.probe.center_frequency)
    assert pipe.stats.total_ms > 0 and pipe.stats.beamform_ms > 0
    assert np.array_equal(np.asarray(f0.data), np.asarray(f1.data)), "frames are deterministic"
    assert not np.asarray(f0.data).flags.owndata


@pytest.fixture(scope="module")
def config0(usx_mod, O):
    """The benchmark acquisition (128 elements, 11 plane waves, 1536 samples at
    20 MHz, 20k-scatterer seed-0 speckle with a 3 mm cyst) on a sub-grid of
    8 lateral lines across the full width and the full 5-50 mm depth range."""
    probe = usx_mod.presets.linear_5mhz()
    acq = usx_mod.Acquisition()
    acq.sampling_frequency = 20e6
    acq.n_samples = 1536
    sim = usx_mod.Simulator(probe, usx_mod.Medium(), acq, usx_mod.SimulatorConfig())
    ph = usx_mod.Phantom.speckle_box(-0.019, 0.019, 0.005, 0.050, 2e-4, 20000, 0, 1.0)
    ph.scale_amplitude_in_circle(usx_mod.Vec3(0.0, 0.0, 0.020), 0.003, 0.0)
    txs = sim.make_plane_waves(usx_mod.angles_deg(-9, 9, 11))
    raw = sim.simulate(txs, ph)
    half = 0.5 * probe.aperture_width()
    g = usx_mod.ScanGrid.cartesian(-half, half, 8, 0.005, 0.050, 640)
    gs = O.GridSpec(0, -half, half, 8, 0.005, 0.050, 640)
    return probe, acq, txs, raw, g, gs


def test_config1_sizes_mv_and_das_match_oracle(usx_mod, port, O, config0):
    """Subarray 64: the active aperture reaches 111 channels and L 55, so every
    covariance size bucket (16..64) is exercised."""
    probe, acq, txs, raw, g, gs = config0
    s = O.Setup.of(probe, usx_mod.Medium(), acq)
    otx = [O.Tx.of(t) for t in txs]
    for combiner, cf, tol in [("DAS", False, TOL_DAS), ("DAS", True, TOL_DAS),
                              ("DMAS", False, TOL_DAS), ("MV", False, TOL_MV)]:
        pc = usx_mod.PipelineConfig()
        pc.demod.decimation = 4
        pc.beamformer = bf_config(usx_mod, combiner, cf, subarray=64)
        got = np.asarray(usx_mod.ImagingPipeline(g, pc).process_frame(raw).data)
        ref, _ = port.process_frame(s, otx, np.asarray(raw.rf), gs, O.Demod(decimation=4),
                                    O.Bf.of(pc.beamformer))
        assert rel_l2(got, ref) < tol, (combiner, cf, rel_l2(got, ref))


def test_mv_precision_variants_against_fp64(usx_mod, port, O, config0):
    """FP32 (with and without refinement) meets the 1e-4 bar against the FP64
    oracle; BF16/FP16 covariance inputs (8- and 11-bit mantissas: they cannot meet 1e-4) are reported variants and
    are only bounded loosely here."""
    probe, acq, txs, raw, g, gs = config0
    d = usx_mod.DemodConfig()
    d.decimation = 4
    iq = usx_mod.demodulate(raw, d)
    s = O.Setup.of(probe, usx_mod.Medium(), iq.acq)
    otx = [O.Tx.of(t) for t in txs]
    b64 = bf_config(usx_mod, "MV", subarray=64)
    ref64, _ = port.beamform(s, otx, np.asarray(iq.iq), gs, O.Bf.of(b64),
                             fdem=iq.demodulation_frequency, mv_fp64=True, info=True)
    errs = {}
    for prec, refine, bound in [("FP32", True, 1e-4), ("FP32", False, 1e-4), ("BF16", True, 5e-2),
                                ("FP16", True, 1e-2)]:
        b = bf_config(usx_mod, "MV", subarray=64, precision=prec, refine=refine)
        got = np.asarray(usx_mod.Beamformer(b).beamform(iq, g).data)
        errs[(prec, refine)] = rel_l2(got, ref64)
        assert errs[(prec, refine)] < bound, (prec, refine, errs)
    assert errs[("FP16", True)] < errs[("BF16", True)]
    # the FP64-residual step brings FP32 closer to the FP64 oracle
    print(f"[precision] errors vs FP64: {errs}", flush=True)
    assert errs[("FP32", True)] < errs[("FP32", False)], errs


def test_mv_192_channels_subarray_96(usx_mod, port, O):
    """192-element probe, subarray 96 at f/1: L reaches 96, so the 80- and
    96-row buckets (two matrix rows per lane) run beside the smaller ones."""
    case = make_case(usx_mod, O, n_el=192, n_samples=2048, n_tx=3, n_scat=2500,
                     grid=(12, 0.01, 0.07, 120), box=(-0.028, 0.028, 0.005, 0.072))
    d = usx_mod.DemodConfig()
    d.decimation = 4
    iq = usx_mod.demodulate(case.raw, d)
    s, txs = oracle_inputs(O, case)
    s2 = O.Setup(**{**s.__dict__, "sampling_frequency": iq.acq.sampling_frequency,
                    "n_samples": iq.n_samples})
    bfc = bf_config(usx_mod, "MV", subarray=96, f_number=1.0)
    got = np.asarray(usx_mod.Beamformer(bfc).beamform(iq, case.grid).data)
    ref, ex = port.beamform(s2, txs, np.asarray(iq.iq), case.gspec, O.Bf.of(bfc),
                            fdem=iq.demodulation_frequency, info=True)
    L = ex["info"][:, 1]
    assert L.max() == 96 and (L > 64).sum() > 50 and (L[L > 0] <= 16).any()
    assert rel_l2(got, ref) < TOL_MV, rel_l2(got, ref)
    # the same with the effective weights on (the label path)
    run = usx_mod.DeviceRunner(case.grid, _pc(usx_mod, bfc), case.raw)
    frame, w, st = run.mv_weights(case.raw)
    assert rel_l2(frame, ref) < TOL_MV
    _, exw = port.beamform(s2, txs, np.asarray(iq.iq), case.gspec, O.Bf.of(bfc),
                           fdem=iq.demodulation_frequency, weights=True)
    assert rel_l2(w.reshape(-1, 192), exw["weights"]) < TOL_MV
    # and with the refinement step: the full matrix in LDS (its upper triangle
    # keeps R for the residual) instead of the packed lower triangle
    bfr = bf_config(usx_mod, "MV", subarray=96, f_number=1.0, refine=True)
    got_r = np.asarray(usx_mod.Beamformer(bfr).beamform(iq, case.grid).data)
    assert rel_l2(got_r, ref) < TOL_MV, rel_l2(got_r, ref)


def test_mv_256_channels_subarray_128(usx_mod, port, O):
    """256-element probe, subarray 128 at f/0.5: below about 38 mm every
    element is active and L reaches 128, so the 112- and 128-row buckets run
    (the reference has no upper bound on L, beamformer.cpp:147-148), with and
    without the refinement step (full matrix in LDS)."""
    case = make_case(usx_mod, O, n_el=256, n_samples=2200, n_tx=1, n_scat=1500,
                     grid=(4, 0.030, 0.066, 24), box=(-0.036, 0.036, 0.02, 0.07))
    d = usx_mod.DemodConfig()
    d.decimation = 4
    iq = usx_mod.demodulate(case.raw, d)
    s, txs = oracle_inputs(O, case)
    s2 = O.Setup(**{**s.__dict__, "sampling_frequency": iq.acq.sampling_frequency,
                    "n_samples": iq.n_samples})
    bfc = bf_config(usx_mod, "MV", subarray=128, f_number=0.5)
    ref, ex = port.beamform(s2, txs, np.asarray(iq.iq), case.gspec, O.Bf.of(bfc),
                            fdem=iq.demodulation_frequency, info=True)
    L = ex["info"][:, 1]
    assert L.max() == 128 and ((L > 96) & (L <= 112)).any(), np.unique(L)
    for refine in (False, True):
        b = bf_config(usx_mod, "MV", subarray=128, f_number=0.5, refine=refine)
        got = np.asarray(usx_mod.Beamformer(b).beamform(iq, case.grid).data)
        assert rel_l2(got, ref) < TOL_MV, (refine, rel_l2(got, ref))


def _pc(usx, bfc):
    pc = usx.PipelineConfig()
    pc.demod.decimation = 4
    pc.beamformer = bfc
    return pc
