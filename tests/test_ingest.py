"""Raw-data sources and the .npz recording format (usx.ingest; reference
python/usx/ingest.py and its tests in tests/test_usx.py:159-203).

The archive must hold exactly the reference's keys with the reference's
dtypes, so recordings move between the two packages; a round trip must keep
samples and geometry bit for bit, and (GPU) beamform to the same image.
"""

from __future__ import annotations

import numpy as np
import pytest

from cases import wires

REFERENCE_KEYS = {
    "rf": np.float32, "probe_n_elements": np.integer, "probe_pitch": np.floating,
    "probe_element_width": np.floating, "probe_elevation_height": np.floating,
    "probe_center_frequency": np.floating, "probe_bandwidth": np.floating,
    "medium_speed_of_sound": np.floating, "medium_attenuation": np.floating,
    "acq_sampling_frequency": np.floating, "acq_n_samples": np.integer, "acq_t0": np.floating,
    "acq_prf": np.floating, "tx_type": np.int32, "tx_angle": np.float32, "tx_focus": np.float32,
    "tx_origin": np.float32, "tx_delays": np.float32, "tx_apodization": np.float32,
}


@pytest.fixture(scope="module")
def setup(usx_mod):
    usx = usx_mod
    probe = usx.presets.linear_5mhz()
    probe.n_elements = 64
    medium = usx.Medium()
    acq = usx.presets.acquisition(probe, depth=0.04, medium=medium)
    cfg = usx.SimulatorConfig()
    cfg.noise_db = -200.0
    return probe, medium, acq, cfg, usx.Simulator(probe, medium, acq, cfg)


def test_round_trip_keeps_samples_and_geometry(usx_mod, setup, tmp_path):
    usx = usx_mod
    probe, medium, acq, cfg, sim = setup
    ph = wires(usx, [(0.0, 0.025)])
    txs = sim.make_plane_waves([0.0])
    src = usx.ingest.SimulatedSource(probe, medium, acq, ph, txs, config=cfg, n_frames=2)
    frames = list(src.frames())
    assert len(frames) == 2 and src.events_per_frame == 1
    path = usx.ingest.save_npz(tmp_path / "acq.npz", frames[0])
    with np.load(path, allow_pickle=False) as z:
        assert set(z.files) == set(REFERENCE_KEYS)
        for k, t in REFERENCE_KEYS.items():
            assert np.issubdtype(z[k].dtype, t), (k, z[k].dtype)
        assert z["tx_type"].tolist() == [1]
    replayed = next(usx.ingest.FileSource(path).frames())
    assert np.array_equal(np.asarray(frames[0].rf), np.asarray(replayed.rf))
    assert replayed.probe.n_elements == probe.n_elements
    assert replayed.probe.pitch == pytest.approx(probe.pitch)
    assert replayed.acq.sampling_frequency == pytest.approx(acq.sampling_frequency)
    assert replayed.acq.n_samples == acq.n_samples
    assert replayed.medium.speed_of_sound == pytest.approx(medium.speed_of_sound)
    assert len(replayed.transmits) == len(txs)
    assert np.array_equal(np.asarray(replayed.transmits[0].delays, np.float32),
                          np.asarray(txs[0].delays, np.float32))
    assert replayed.transmits[0].type == txs[0].type
    assert replayed.transmits[0].angle == pytest.approx(txs[0].angle)


def test_focused_and_diverging_transmits_survive(usx_mod, setup, tmp_path):
    usx = usx_mod
    probe, medium, acq, cfg, sim = setup
    txs = list(sim.make_focused_scan(4, -0.005, 0.005, 0.03)) + list(
        sim.make_diverging_waves([-0.1, 0.1], -0.01))
    rf = np.random.default_rng(0).standard_normal((len(txs), probe.n_elements, acq.n_samples))
    data = usx.ChannelData.from_rf(rf.astype(np.float32), probe, medium, acq, txs)
    path = usx.ingest.save_npz(tmp_path / "mixed.npz", data)
    back = next(usx.ingest.FileSource(path).frames())
    for a, b in zip(txs, back.transmits):
        assert a.type == b.type
        assert (a.focus.x, a.focus.z) == pytest.approx((b.focus.x, b.focus.z))
        assert np.array_equal(np.asarray(a.apodization, np.float32), np.asarray(b.apodization, np.float32))


def test_multi_frame_archive_and_loop(usx_mod, setup, tmp_path):
    usx = usx_mod
    probe, medium, acq, cfg, sim = setup
    txs = sim.make_plane_waves([-0.05, 0.05])
    rf = np.random.default_rng(1).standard_normal((6, probe.n_elements, acq.n_samples)).astype(np.float32)
    data = usx.ChannelData.from_rf(rf, probe, medium, acq, list(txs) * 3)
    path = usx.ingest.save_npz(tmp_path / "three.npz", data)
    # the archive stores 6 transmits; a source with 2 per frame is built by hand
    with np.load(path, allow_pickle=False) as z:
        assert z["rf"].shape[0] == 6
    src = usx.ingest.FileSource(path)
    assert src.n_frames == 1 and len(src.metadata.transmits) == 6
    arr = usx.ingest.ArraySource([rf[:2], rf[2:4], rf[4:]], probe, medium, acq, txs)
    got = list(arr.frames())
    assert len(got) == 3 and all(np.array_equal(np.asarray(g.rf), rf[2 * i:2 * i + 2])
                                 for i, g in enumerate(got))
    looped = usx.ingest.FileSource(path, loop=True).frames()
    assert [next(looped).n_events for _ in range(3)] == [6, 6, 6]


def test_moving_phantom_advances_between_frames(usx_mod, setup):
    usx = usx_mod
    probe, medium, acq, cfg, sim = setup
    ph = wires(usx, [(0.0, 0.02)])
    ph.velocities = np.array([[0.0, 0.0, 0.5]], np.float32)
    txs = sim.make_plane_waves([0.0])
    frames = list(usx.ingest.SimulatedSource(probe, medium, acq, ph, txs, config=cfg,
                                             n_frames=2).frames())
    a, b = (np.asarray(f.rf)[0, probe.n_elements // 2] for f in frames)
    # the echo moved 0.5 m/s * (1 / PRF) deeper: later peak
    assert np.argmax(np.abs(b)) > np.argmax(np.abs(a))


def test_inconsistent_transmit_vectors_are_rejected(usx_mod, setup, tmp_path):
    usx = usx_mod
    probe, medium, acq, cfg, sim = setup
    txs = list(sim.make_plane_waves([0.0, 0.1]))
    txs[1].apodization = []
    txs[0].apodization = [1.0] * probe.n_elements
    rf = np.zeros((2, probe.n_elements, acq.n_samples), np.float32)
    data = usx.ChannelData.from_rf(rf, probe, medium, acq, txs)
    with pytest.raises(ValueError):
        usx.ingest.save_npz(tmp_path / "bad.npz", data)


@pytest.mark.gpu
def test_replayed_recording_beamforms_like_the_reference(usx_mod, setup, tmp_path, port, O):
    """A recording replayed through FileSource gives the GPU pipeline the same
    frame as the original RF (bit for bit), and that frame matches the oracle's
    restatement of the reference on the original RF (1e-5 relative L2)."""
    from conftest import rel_l2
    usx = usx_mod
    probe, medium, acq, cfg, sim = setup
    ph = wires(usx, [(0.0, 0.025), (0.004, 0.02)], amplitude=5.0)
    ph.extend(usx.Phantom.speckle_box(-0.008, 0.008, 0.01, 0.035, 2e-4, 800, 11, 1.0))
    txs = sim.make_plane_waves(usx.angles_deg(-6, 6, 3))
    raw = sim.simulate(txs, ph)
    path = usx.ingest.save_npz(tmp_path / "acq.npz", raw)
    replayed = next(usx.ingest.FileSource(path).frames())
    grid = usx.ScanGrid.cartesian(-0.006, 0.006, 40, 0.012, 0.034, 90)
    pc = usx.PipelineConfig()
    pc.demod.decimation = 2
    for combiner in ("DAS", "MV"):
        b = pc.beamformer
        b.combiner = getattr(usx.Combiner, combiner)
        b.mv_subarray = 16
        pc.beamformer = b
        a = np.asarray(usx.ImagingPipeline(grid, pc).process_frame(raw).data)
        r = np.asarray(usx.ImagingPipeline(grid, pc).process_frame(replayed).data)
        assert np.array_equal(a.view(np.uint32), r.view(np.uint32)), combiner
        ref, _ = port.process_frame(O.Setup.of(probe, medium, acq), [O.Tx.of(t) for t in txs],
                                    np.asarray(raw.rf), O.GridSpec(0, -0.006, 0.006, 40, 0.012,
                                                                   0.034, 90),
                                    O.Demod(decimation=2), O.Bf.of(b))
        assert rel_l2(r, ref) < (1e-4 if combiner == "MV" else 1e-5), combiner


@pytest.mark.gpu
def test_simulated_source_on_the_device(usx_mod, setup):
    """SimulatedSource(device=True) renders with the GPU simulator: the same
    frames as the host simulator's to float rounding."""
    from conftest import rel_l2
    usx = usx_mod
    probe, medium, acq, cfg, sim = setup
    ph = wires(usx, [(0.0, 0.02), (0.003, 0.03)])
    ph.velocities = np.array([[0.0, 0.0, 0.5], [0.0, 0.0, -0.2]], np.float32)
    txs = sim.make_plane_waves([0.0, 0.1])
    host = list(usx.ingest.SimulatedSource(probe, medium, acq, ph, txs, config=cfg,
                                           n_frames=2).frames())
    dev = list(usx.ingest.SimulatedSource(probe, medium, acq, ph, txs, config=cfg, n_frames=2,
                                          device=True).frames())
    for h, d in zip(host, dev):
        assert rel_l2(np.asarray(d.rf), np.asarray(h.rf)) < 1e-5
