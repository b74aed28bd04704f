"""Parity at the benchmarked sizes (gfx950 only).

* The benchmark frame itself: 128 elements, 11 plane waves, 1536 samples at
  20 MHz of the seed-0 speckle phantom with its 3 mm cyst, demodulated by 4 and
  beamformed onto the full 256 x 640 grid. DAS, DAS + coherence factor, DMAS
  and DMAS + coherence factor are compared with the oracle on every pixel
  (1e-5 relative L2); minimum variance (subarray 64, loading 1e-2, +-2 depths)
  on 64 lateral lines spread over the width at full depth (1e-4), with and
  without the refinement step.
* The same acquisition beamformed from RF directly (no demodulation; the
  reference's interp_rf path, beamformer.cpp:480-489) through DMAS and MV.
* Full-size frames of the uneven stream (seed 1): the largest 192-element
  frame, the 31-transmit frame, the 4058-sample frame and the 512 x 1280
  frame, recorded into one HIP graph with the packaged launch shapes and
  replayed; sampled lines of each DAS and MV frame against the oracle.

The oracle beamforms a subset of the lateral lines at the full grid's axis
values (GridSpec.line_first / line_step), so its lines are bitwise the same
pixels as the GPU frame's.
"""

from __future__ import annotations

import numpy as np
import pytest

from cases import bf_config
from conftest import rel_l2

pytestmark = pytest.mark.gpu

TOL_DAS = 1e-5
TOL_MV = 1e-4


@pytest.fixture(scope="module")
def bench_case(usx_mod, O):
    probe = usx_mod.presets.linear_5mhz()
    med = usx_mod.Medium()
    acq = usx_mod.Acquisition()
    acq.sampling_frequency = 20e6
    acq.n_samples = 1536
    acq.prf = 5000.0
    sim = usx_mod.Simulator(probe, med, acq, usx_mod.SimulatorConfig())
    ph = usx_mod.Phantom.speckle_box(-0.019, 0.019, 0.005, 0.050, 2e-4, 20000, 0, 1.0)
    ph.scale_amplitude_in_circle(usx_mod.Vec3(0.0, 0.0, 0.020), 0.003, 0.0)
    txs = sim.make_plane_waves(usx_mod.angles_deg(-9, 9, 11))
    raw = sim.simulate(txs, ph)
    grid = usx_mod.grid_for(probe, 0.005, 0.050, 256, 640)
    half = 0.5 * probe.aperture_width()
    s = O.Setup.of(probe, med, acq)
    otx = [O.Tx.of(t) for t in txs]
    print(f"[bench_case] RF {raw.shape}, grid {grid.n_axis0}x{grid.n_axis1}", flush=True)
    return raw, grid, s, otx, half


def _pc(usx, bfc):
    pc = usx.PipelineConfig()
    pc.demod.decimation = 4
    pc.beamformer = bfc
    return pc


@pytest.mark.parametrize("combiner,cf", [("DAS", False), ("DAS", True), ("DMAS", False),
                                         ("DMAS", True)])
def test_benchmark_frame_das_dmas_every_pixel(usx_mod, port, O, bench_case, combiner, cf):
    raw, grid, s, otx, half = bench_case
    bfc = bf_config(usx_mod, combiner, cf, subarray=64)
    got = np.asarray(usx_mod.ImagingPipeline(grid, _pc(usx_mod, bfc)).process_frame(raw).data)
    gs = O.GridSpec(0, -half, half, 256, 0.005, 0.050, 640)
    ref, _ = port.process_frame(s, otx, np.asarray(raw.rf), gs, O.Demod(decimation=4),
                                O.Bf.of(bfc))
    assert got.shape == ref.shape == (256, 640)
    assert rel_l2(got, ref) < TOL_DAS, (combiner, cf, rel_l2(got, ref))


@pytest.mark.parametrize("refine", [False, True])
def test_benchmark_frame_mv_64_lines(usx_mod, port, O, bench_case, refine):
    raw, grid, s, otx, half = bench_case
    bfc = bf_config(usx_mod, "MV", subarray=64, refine=refine)
    got = np.asarray(usx_mod.ImagingPipeline(grid, _pc(usx_mod, bfc)).process_frame(raw).data)
    gs = O.GridSpec(0, -half, half, 256, 0.005, 0.050, 640, line_first=1, line_step=4)
    ref, _ = port.process_frame(s, otx, np.asarray(raw.rf), gs, O.Demod(decimation=4),
                                O.Bf.of(bfc))
    assert ref.shape == (64, 640)
    assert rel_l2(got[1::4], ref) < TOL_MV, (refine, rel_l2(got[1::4], ref))
    # every line is finite and the cyst is dark, also on the lines not compared
    assert np.isfinite(got.view(np.float32)).all()


@pytest.mark.parametrize("combiner,cf,step", [("DMAS", False, 1), ("DMAS", True, 1),
                                              ("MV", False, 8)])
def test_rf_input_dmas_and_mv(usx_mod, port, O, bench_case, combiner, cf, step):
    """Beamformer.beamform on RF channel data: real linear interpolation with
    no phase rotation (interp_rf, beamformer.cpp:34-39, 480-489)."""
    raw, grid, s, otx, half = bench_case
    bfc = bf_config(usx_mod, combiner, cf, subarray=64)
    got = np.asarray(usx_mod.Beamformer(bfc).beamform(raw, grid).data)
    assert np.all(got.imag == 0)
    gs = O.GridSpec(0, -half, half, 256, 0.005, 0.050, 640, line_first=step // 2,
                    line_step=step if step > 1 else 0)
    ref = port.beamform(s, otx, np.asarray(raw.rf), gs, O.Bf.of(bfc))
    tol = TOL_MV if combiner == "MV" else TOL_DAS
    sub = got[step // 2::step]
    assert rel_l2(sub, ref) < tol, (combiner, cf, rel_l2(sub, ref))


# ---- full-size frames of the uneven stream --------------------------------

STREAM_FRAMES = {38: "largest 192-element frame (493 x 1234, 27 transmits)",
                 43: "31 transmits, 192 elements",
                 23: "4058 samples, 192 elements, L = 96",
                 52: "512 x 1280 grid"}


@pytest.fixture(scope="module")
def full_stream(usx_mod):
    from usx import stream as S
    specs = S.draw(seed=1, n=64)
    picked = [specs[i] for i in STREAM_FRAMES]
    print(f"[full_stream] frames {list(STREAM_FRAMES)}", flush=True)
    st = S.Stream(picked, tune=False, cache=S.load_cache())
    st.run(graph=False)
    eager = [f.copy() for f in st.frames()]
    st.seq.clear()
    st.run(graph=True)
    return st, eager


def test_full_stream_graph_equals_eager(full_stream):
    st, eager = full_stream
    for a, b in zip(st.frames(), eager):
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


@pytest.mark.parametrize("k", range(len(STREAM_FRAMES)))
def test_full_stream_frames_match_the_oracle(usx_mod, port, O, full_stream, k):
    st, _ = full_stream
    got = st.frames()
    idx = list(STREAM_FRAMES)[k]
    raw, grid = st.raws[2 * k], st.grids[2 * k]
    s = O.Setup.of(raw.probe, raw.medium, raw.acq)
    txs = [O.Tx.of(t) for t in raw.transmits]
    n0 = grid.n_axis0
    print(f"[stream frame {idx}] {STREAM_FRAMES[idx]}: {raw.shape} -> {n0}x{grid.n_axis1}",
          flush=True)
    for j, (step, tol) in enumerate(((max(1, n0 // 16), TOL_DAS), (max(1, n0 // 8), TOL_MV))):
        gs = O.GridSpec(0, *st.extents[2 * k + j], line_first=step // 2, line_step=step)
        ref, _ = port.process_frame(s, txs, np.asarray(raw.rf), gs, O.Demod(decimation=4),
                                    O.Bf.of(st.cfgs[2 * k + j].beamformer))
        sub = got[2 * k + j][step // 2::step]
        assert sub.shape == ref.shape
        assert np.abs(ref).max() > 0
        assert rel_l2(sub, ref) < tol, (idx, "DAS" if j == 0 else "MV", rel_l2(sub, ref))
