"""The uneven batched stream: one captured HIP graph over frames of different
probe, grid and transmit count (usx.FrameSequence, usx.stream).

* The graph replay equals the same kernels launched one by one, bit for bit,
  on every replay, with other device work and synchronizations between the
  replays.
* Every frame equals ImagingPipeline.process_frame with the same
  configuration (and launch shape), bit for bit.
* Frames of the stream match the oracle (DAS 1e-5, MV 1e-4).
* The autotuner's choices change results by float rounding at most.
* One frame captured by DeviceRunner.capture replays bit for bit (DAS, MV,
  learned apodization).
"""

from __future__ import annotations

import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def small_stream(usx_mod):
    from usx import stream as S
    specs = S.draw(seed=1, n=6, max_scale=1.4, max_samples=1400)
    cache = {"version": S.CACHE_VERSION, "target": "gfx950", "entries": {}}
    return S.Stream(specs, tune=True, cache=cache)


def _bits(a):
    return np.ascontiguousarray(a).view(np.uint32)


def test_graph_replays_equal_eager_launches(small_stream):
    st = small_stream
    st.run(graph=False)
    eager = st.frames()
    assert all(np.isfinite(f.view(np.float32)).all() for f in eager)
    assert any(np.abs(f).max() > 0 for f in eager)
    for _ in range(3):
        st.seq.clear()
        st.seq.synchronize()
        assert not any(np.abs(f).max() > 0 for f in st.frames())
        st.run(graph=True)
        for a, b in zip(st.frames(), eager):
            assert np.array_equal(_bits(a), _bits(b))


def test_stream_frames_equal_the_pipeline(usx_mod, small_stream):
    st = small_stream
    st.run(graph=True)
    got = st.frames()
    for raw, grid, pc, f in zip(st.raws, st.grids, st.cfgs, got):
        ref = np.asarray(usx_mod.ImagingPipeline(grid, pc).process_frame(raw).data)
        assert np.array_equal(_bits(ref), _bits(f))


def test_stream_frames_match_the_oracle(usx_mod, port, O, small_stream):
    st = small_stream
    st.run(graph=True)
    got = st.frames()
    # the first frame of the stream, DAS and MV
    raw, grid = st.raws[0], st.grids[0]
    s = O.Setup.of(raw.probe, raw.medium, raw.acq)
    txs = [O.Tx.of(t) for t in raw.transmits]
    gs = O.GridSpec(0, *st.extents[0])
    for k, tol in ((0, 1e-5), (1, 1e-4)):
        ref, _ = port.process_frame(s, txs, np.asarray(raw.rf), gs, O.Demod(decimation=4),
                                    O.Bf.of(st.cfgs[k].beamformer))
        assert rel_l2(got[k], ref) < tol, (k, rel_l2(got[k], ref))


def test_autotuned_shapes_keep_the_results(usx_mod, small_stream):
    """The MV occupancy only reorders independent pixels (bit-identical
    frames); the event-slot count regroups the sums over transmits (DAS, and
    MV's channel vectors), which moves results by float rounding only."""
    st = small_stream
    assert st.tuned, "the autotuner ran on every shape class"
    for key, ent in st.tuned.items():
        assert ent["slots"] in (4, 8, 16) and ent["mv_occupancy"] in (0, 4, 8, 16) and ent["ms"] > 0
    for raw, grid, pc in zip(st.raws, st.grids, st.cfgs):
        tuned = np.asarray(usx_mod.ImagingPipeline(grid, pc).process_frame(raw).data)
        b = pc.beamformer
        slots = b.tile_slots
        b.tile_mv_occupancy = 0
        pc.beamformer = b
        same_slots = np.asarray(usx_mod.ImagingPipeline(grid, pc).process_frame(raw).data)
        assert np.array_equal(_bits(tuned), _bits(same_slots))
        b.tile_slots = 0
        pc.beamformer = b
        plain = np.asarray(usx_mod.ImagingPipeline(grid, pc).process_frame(raw).data)
        assert rel_l2(tuned, plain) < 1e-5, (slots, rel_l2(tuned, plain))


@pytest.mark.parametrize("combiner", ["DAS", "MV", "Learned"])
def test_device_runner_graph_equals_eager(usx_mod, O, combiner):
    """DeviceRunner.capture_host records one frame into a HIP graph (a
    one-frame plan on the runner's device copy of the RF); every replay, with
    other device work between them, equals ImagingPipeline.process_frame of
    the same data bit for bit."""
    import os
    from cases import bf_config, make_case
    from usx import learned as L
    case = make_case(usx_mod, O, n_el=128, n_samples=1024, n_tx=3, n_scat=800,
                     grid=(16, 0.008, 0.030, 48), box=(-0.02, 0.02, 0.005, 0.032))
    pc = usx_mod.PipelineConfig()
    pc.demod.decimation = 4
    pc.beamformer = bf_config(usx_mod, combiner, subarray=16)
    pipe = usx_mod.ImagingPipeline(case.grid, pc)
    run = usx_mod.DeviceRunner(case.grid, pc, case.raw)
    if combiner == "Learned":
        ckpt = os.path.join(os.path.dirname(usx_mod.__file__), "data", "apod_mlp.npz")
        params = L.load_checkpoint(ckpt)[0]
        pipe.set_apodization(params)
        run.set_apodization(params)
    ref = np.asarray(pipe.process_frame(case.raw).data)
    assert np.abs(ref).max() > 0
    run.capture_host(case.raw)
    for _ in range(3):
        pipe.process_frame(case.raw)  # other device work between the replays
        run.run_graph()
        run.synchronize()
        assert np.array_equal(_bits(run.graph_frame()), _bits(ref))
    if combiner == "DAS":
        # a capture on caller buffers frees capture_host's copies: no stale frame
        rf_host = np.ascontiguousarray(np.asarray(case.raw.rf))
        rf = usx_mod.DeviceBuffer(rf_host.nbytes)
        rf.upload(rf_host)
        shape = [case.grid.n_axis0, case.grid.n_axis1]
        fr = usx_mod.DeviceBuffer(8 * shape[0] * shape[1])
        fr.zero()
        run.capture(rf.ptr, fr.ptr)
        with pytest.raises(Exception, match="capture_host"):
            run.graph_frame()
        run.run_graph()
        run.synchronize()
        assert np.array_equal(_bits(fr.download(np.complex64, shape)), _bits(ref))
