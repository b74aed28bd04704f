"""The allocation-free entry point (usx_process_frame_noalloc, no reference
counterpart) against the entry point it mirrors, usx_process_frame on device
buffers, through the C ABI (DeviceRunner.run_noalloc / DeviceRunner.run).

Two calls in a row on one GPU share the library's fixed scratch array: the
second, on a smaller input, must not see the first's leftovers. Results are
compared bit for bit (the same kernels on the same inputs) and with the
oracle at the parity bars (1e-5 DAS family, 1e-4 minimum variance).
"""

from __future__ import annotations

import numpy as np
import pytest

from cases import bf_config, make_case, oracle_inputs
from conftest import rel_l2

pytestmark = pytest.mark.gpu


def _pc(usx, bfc):
    pc = usx.PipelineConfig()
    pc.demod.decimation = 4
    pc.beamformer = bfc
    return pc


def _frames(usx, case, bfc):
    """(allocation-free frame, usx_process_frame frame) of one case."""
    run = usx.DeviceRunner(case.grid, _pc(usx, bfc), case.raw)
    rf_host = np.ascontiguousarray(np.asarray(case.raw.rf))
    rf = usx.DeviceBuffer(rf_host.nbytes)
    rf.upload(rf_host)
    shape = [case.grid.n_axis0, case.grid.n_axis1]
    a = usx.DeviceBuffer(8 * shape[0] * shape[1])
    b = usx.DeviceBuffer(8 * shape[0] * shape[1])
    a.zero()
    b.zero()
    run.run_noalloc(rf.ptr, a.ptr)
    run.synchronize()
    got = a.download(np.complex64, shape)
    run.run(rf.ptr, b.ptr)
    run.synchronize()
    return got, b.download(np.complex64, shape)


def _oracle(port, O, case, bfc):
    s, txs = oracle_inputs(O, case)
    ref, _ = port.process_frame(s, txs, np.asarray(case.raw.rf), case.gspec,
                                O.Demod(decimation=4), O.Bf.of(bfc))
    return ref


def test_noalloc_equals_process_frame_on_two_calls(usx_mod, port, O):
    big = make_case(usx_mod, O, n_el=128, n_samples=1536, n_tx=11, n_scat=4000,
                    grid=(128, 0.005, 0.050, 320), box=(-0.019, 0.019, 0.005, 0.050))
    small = make_case(usx_mod, O, n_el=32, n_samples=768, n_tx=3, n_scat=600)
    calls = [(big, bf_config(usx_mod, "DMAS", True), 1e-5, "big DMAS + CF"),
             (small, bf_config(usx_mod, "MV", subarray=8), 1e-4, "small MV"),
             (small, bf_config(usx_mod, "DAS"), 1e-5, "small DAS")]
    for case, bfc, tol, name in calls:
        print(f"[noalloc] {name}: RF {case.raw.shape} -> {case.grid.n_axis0}x{case.grid.n_axis1}",
              flush=True)
        got, ref = _frames(usx_mod, case, bfc)
        assert np.abs(ref).max() > 0
        assert np.array_equal(got.view(np.uint32), ref.view(np.uint32)), name
        assert rel_l2(got, _oracle(port, O, case, bfc)) < tol, name


def test_noalloc_rejects_frames_beyond_its_scratch(usx_mod, O):
    """Minimum variance on a 128-channel 256 x 320 grid needs 84 MB of channel
    vectors alone: past the 64 MiB array, so the call is refused (ValueError)
    and the frame buffer is left untouched."""
    case = make_case(usx_mod, O, n_el=128, n_samples=1536, n_tx=11, n_scat=2000,
                     grid=(256, 0.005, 0.050, 320), box=(-0.019, 0.019, 0.005, 0.050))
    run = usx_mod.DeviceRunner(case.grid, _pc(usx_mod, bf_config(usx_mod, "MV", subarray=64)),
                               case.raw)
    rf_host = np.ascontiguousarray(np.asarray(case.raw.rf))
    rf = usx_mod.DeviceBuffer(rf_host.nbytes)
    rf.upload(rf_host)
    shape = [case.grid.n_axis0, case.grid.n_axis1]
    fr = usx_mod.DeviceBuffer(8 * shape[0] * shape[1])
    fr.zero()
    with pytest.raises(ValueError, match="scratch"):
        run.run_noalloc(rf.ptr, fr.ptr)
    run.synchronize()
    assert not np.any(fr.download(np.complex64, shape))
