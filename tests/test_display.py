"""The display tail on the GPU (ImagingPipeline.to_display / process_display)
against the oracle's restatement of the reference's to_display, which equals
the compiled reference bit for bit (test_oracle.py).

Images are in [0, 1]; the bar is 1e-5 absolute (device powf/log10f/hypotf and,
for RF frames, a different FFT twiddle evaluation move the last bits).
"""

from __future__ import annotations

import numpy as np
import pytest

from cases import bf_config, make_case, oracle_inputs

pytestmark = pytest.mark.gpu

TOL = 1e-5


@pytest.fixture(scope="module")
def case(usx_mod, O):
    return make_case(usx_mod, O, n_el=32, n_tx=3, n_scat=600)


def _pc(usx, **kw):
    pc = usx.PipelineConfig()
    pc.demod.decimation = 4
    pc.beamformer = bf_config(usx, "DAS")
    for k, v in kw.items():
        setattr(pc, k, v)
    return pc


def test_to_display_iq_frames_with_smoothing(usx_mod, port, O, case):
    pipe = usx_mod.ImagingPipeline(case.grid, _pc(usx_mod, reference_smoothing=0.6))
    s, _ = oracle_inputs(O, case)
    frames, images = [], []
    for scale in (1.0, 4.0, 0.5):
        f = pipe.process_frame(case.raw)
        d = np.asarray(f.data)  # a view of the frame's own buffer
        d *= np.complex64(scale)
        frames.append(d.copy())
        images.append(np.asarray(pipe.to_display(f, case.medium)))
    want = port.display(s, case.gspec, np.stack(frames), float(f.demodulation_frequency),
                        reference_smoothing=0.6)
    for a, b in zip(images, want):
        assert np.abs(a - b).max() < TOL


def test_to_display_rf_frames_use_the_analytic_signal(usx_mod, port, O, case):
    # beamforming the RF directly gives an RF frame (no demodulation frequency)
    frame = usx_mod.Beamformer(bf_config(usx_mod, "DAS")).beamform(case.raw, case.grid)
    assert frame.demodulation_frequency == 0.0
    pipe = usx_mod.ImagingPipeline(case.grid, _pc(usx_mod))
    img = np.asarray(pipe.to_display(frame, case.medium))
    s, _ = oracle_inputs(O, case)
    want = port.display(s, case.gspec, np.asarray(frame.data)[None], 0.0)[0]
    assert np.abs(img - want).max() < TOL


def test_process_display_keeps_the_frame_on_the_device(usx_mod, port, O, case):
    pc = _pc(usx_mod, gain_db=3.0, dynamic_range_db=50.0)
    pipe = usx_mod.ImagingPipeline(case.grid, pc)
    imgs = [np.asarray(pipe.process_display(case.raw)) for _ in range(2)]
    assert pipe.frame_count == 2 and pipe.stats.postprocess_ms > 0
    s, txs = oracle_inputs(O, case)
    fr, _ = port.process_frame(s, txs, np.asarray(case.raw.rf), case.gspec, O.Demod(decimation=4),
                               O.Bf.of(pc.beamformer))
    want = port.display(s, case.gspec, np.stack([fr, fr]), 5e6, gain_db=3.0, dynamic_range_db=50.0)
    for a, b in zip(imgs, want):
        assert np.abs(a - b).max() < 1e-4  # frame within 1e-5 relative, then the display
