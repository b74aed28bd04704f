"""The reference's imaging-physics checks, restated against the GPU path.

Every frame below is beamformed by the gfx950 kernels (Beamformer.beamform,
ImagingPipeline.process_frame / process_display); the assertions and their
bounds are the reference's own:

* core/tests/test_main.cpp:262-310  three targets, three transmit schemes
* core/tests/test_main.cpp:312-339  RF and IQ paths agree
* core/tests/test_main.cpp:389-394  the coherence factor keeps the target
* core/tests/test_main.cpp:398-426  beam width scales with the f-number
* core/tests/test_main.cpp:619-655  an anechoic lesion is darker
* core/tests/test_main.cpp:681-712  process_display end to end
* tests/test_usx.py:61-156, 255-269 the Python boundary's physics checks

Speckle phantoms go through the GPU simulator (Simulator.simulate_gpu, the
same echo model as the host simulator); point targets through the host one.
"""

from __future__ import annotations

import numpy as np
import pytest

from cases import bf_config, speckle_with_cysts, wires

pytestmark = pytest.mark.gpu


def _standard(usx, n_samples=2600, noise_db=-200.0):
    """The reference's standard probe and acquisition: 96 elements at 300 um
    pitch, 5 MHz, 60 % bandwidth, 40 MHz sampling."""
    probe = usx.Probe()
    probe.n_elements = 96
    probe.pitch = 300e-6
    probe.element_width = 270e-6
    probe.center_frequency = 5.0e6
    probe.bandwidth = 0.6
    med = usx.Medium()
    acq = usx.Acquisition()
    acq.sampling_frequency = 40e6
    acq.n_samples = n_samples
    acq.t0 = 0.0
    acq.prf = 4000.0
    cfg = usx.SimulatorConfig()
    cfg.noise_db = noise_db
    return probe, med, acq, usx.Simulator(probe, med, acq, cfg)


def _peak(env, grid):
    i, j = np.unravel_index(int(np.argmax(env)), env.shape)
    return float(np.asarray(grid.axis0)[i]), float(np.asarray(grid.axis1)[j])


def _lateral_fwhm(env, grid):
    """Width at half maximum of the lateral row through the peak, crossings
    interpolated between pixels (the definition of core/tests/test_main.cpp:90-121)."""
    x = np.asarray(grid.axis0, np.float64)
    i, j = np.unravel_index(int(np.argmax(env)), env.shape)
    row = env[:, j].astype(np.float64)
    half = 0.5 * row[i]
    lo = i
    while lo > 0 and row[lo] > half:
        lo -= 1
    x_lo = x[lo]
    if lo < i and row[lo] <= half:
        x_lo = x[lo] + (half - row[lo]) / (row[lo + 1] - row[lo]) * (x[lo + 1] - x[lo])
    hi = i
    while hi < len(row) - 1 and row[hi] > half:
        hi += 1
    x_hi = x[hi]
    if hi > i and row[hi] <= half:
        x_hi = x[hi] + (half - row[hi]) / (row[hi - 1] - row[hi]) * (x[hi - 1] - x[hi])
    return x_hi - x_lo


@pytest.fixture(scope="module")
def standard(usx_mod):
    return _standard(usx_mod)


@pytest.mark.parametrize("scheme", ["plane wave x1", "plane wave x11", "focused scan"])
def test_three_targets_localize_under_every_transmit_scheme(usx_mod, standard, scheme):
    usx = usx_mod
    probe, med, acq, sim = standard
    targets = [(0.0, 0.015), (0.006, 0.030), (-0.008, 0.045)]
    ph = usx.Phantom.point_targets([usx.Vec3(x, 0.0, z) for x, z in targets])
    if scheme == "plane wave x1":
        txs, comp = sim.make_plane_waves([0.0]), "COHERENT"
    elif scheme == "plane wave x11":
        txs, comp = sim.make_plane_waves([-0.16 + 0.032 * k for k in range(11)]), "COHERENT"
    else:
        txs, comp = sim.make_focused_scan(48, -0.012, 0.012, 0.030), "SELECT"
    iq = usx.demodulate(sim.simulate(txs, ph), usx.DemodConfig())
    bf = usx.Beamformer(bf_config(usx, "DAS", compounding=comp))
    for tx, tz in targets:
        # 0.1 mm pixels: the bound tests the reconstruction, not the grid
        grid = usx.ScanGrid.cartesian(tx - 0.004, tx + 0.004, 81, tz - 0.004, tz + 0.004, 81)
        env = np.abs(np.asarray(bf.beamform(iq, grid).data))
        px, pz = _peak(env, grid)
        assert abs(px - tx) <= 2.1e-4 and abs(pz - tz) <= 2.1e-4, (scheme, tx, tz, px, pz)
        assert env.max() > 0


def test_rf_and_iq_paths_agree(usx_mod, standard):
    usx = usx_mod
    probe, med, acq, sim = standard
    ph = usx.Phantom.point_targets([usx.Vec3(0.003, 0.0, 0.025)])
    raw = sim.simulate(sim.make_plane_waves([0.0]), ph)
    iq = usx.demodulate(raw, usx.DemodConfig())
    grid = usx.ScanGrid.cartesian(-0.004, 0.010, 141, 0.020, 0.030, 201)
    bf = usx.Beamformer(usx.BeamformerConfig())
    f_rf = bf.beamform(raw, grid)
    f_iq = bf.beamform(iq, grid)
    assert f_rf.demodulation_frequency == 0.0 and f_iq.demodulation_frequency > 0.0
    x1, z1 = _peak(np.abs(np.asarray(f_rf.data)), grid)
    x2, z2 = _peak(np.abs(np.asarray(f_iq.data)), grid)
    assert abs(x1 - x2) <= 1.1e-4 and abs(z1 - z2) <= 1.1e-4
    # detection on both: the analytic signal (FFT along depth) for the RF
    # frame, the modulus for the IQ one; both images reach full brightness
    for f in (f_rf, f_iq):  # a fresh pipeline each: no smoothed reference yet
        img = np.asarray(usx.ImagingPipeline(grid, usx.PipelineConfig()).to_display(f, med))
        assert img.max() == pytest.approx(1.0, abs=1e-6) and img.min() >= 0.0


@pytest.fixture(scope="module")
def centre_target(usx_mod, standard):
    usx = usx_mod
    probe, med, acq, sim = standard
    ph = usx.Phantom.point_targets([usx.Vec3(0.0, 0.0, 0.030)])
    return usx.demodulate(sim.simulate(sim.make_plane_waves([0.0]), ph), usx.DemodConfig())


def test_coherence_factor_keeps_the_target(usx_mod, centre_target):
    usx = usx_mod
    grid = usx.ScanGrid.cartesian(-0.004, 0.004, 161, 0.028, 0.032, 81)
    f = usx.Beamformer(bf_config(usx, "DAS", cf=True)).beamform(centre_target, grid)
    px, pz = _peak(np.abs(np.asarray(f.data)), grid)
    assert abs(px) <= 1.1e-4 and abs(pz - 0.030) <= 1.1e-4


def test_beam_width_scales_with_the_f_number(usx_mod, standard, centre_target):
    usx = usx_mod
    probe, med, acq, sim = standard
    grid = usx.ScanGrid.cartesian(-0.005, 0.005, 201, 0.029, 0.031, 41)
    w = {}
    for fn in (3.0, 1.5):
        b = usx.BeamformerConfig()
        b.f_number = fn
        w[fn] = _lateral_fwhm(np.abs(np.asarray(usx.Beamformer(b).beamform(centre_target, grid).data)), grid)
    lam = probe.wavelength(med.speed_of_sound)
    assert w[1.5] < 0.75 * w[3.0]
    assert 1.0 < w[3.0] / lam < 9.0
    assert w[3.0] / (2.0 * w[1.5]) == pytest.approx(1.0, abs=0.25)


def test_anechoic_lesion_is_darker(usx_mod):
    usx = usx_mod
    probe, med, acq, sim = _standard(usx, noise_db=-60.0)
    ph = usx.Phantom.speckle_box(-0.012, 0.012, 0.020, 0.040, 2e-4, 20000, 3)
    cx, cz, radius = 0.0, 0.030, 0.003
    ph.scale_amplitude_in_circle(usx.Vec3(cx, 0.0, cz), radius, 0.0)
    raw = sim.simulate_gpu(sim.make_plane_waves([-0.1, -0.05, 0.0, 0.05, 0.1]), ph)
    iq = usx.demodulate(raw, usx.DemodConfig())
    grid = usx.ScanGrid.cartesian(-0.010, 0.010, 121, 0.024, 0.036, 181)
    env = np.abs(np.asarray(usx.Beamformer().beamform(iq, grid).data)).astype(np.float64)
    x = np.asarray(grid.axis0, np.float64)[:, None]
    z = np.asarray(grid.axis1, np.float64)[None, :]
    r = np.hypot(x - cx, z - cz)
    inside = r < 0.6 * radius
    ring = (r > 1.5 * radius) & (r < 2.5 * radius)
    contrast = 20 * np.log10(np.sqrt(np.mean(env[inside] ** 2)) / np.sqrt(np.mean(env[ring] ** 2)))
    assert contrast < -6.0, contrast


def test_process_display_end_to_end(usx_mod, standard):
    usx = usx_mod
    probe, med, acq, _ = standard
    sim = usx.Simulator(probe, med, acq, usx.SimulatorConfig())
    ph = usx.Phantom.speckle_box(-0.012, 0.012, 0.015, 0.045, 2e-4, 8000, 21)
    bright = usx.Scatterer()
    bright.position = usx.Vec3(0.0, 0.0, 0.030)
    bright.amplitude = 30.0
    ph.add(bright)
    grid = usx.ScanGrid.cartesian(-0.012, 0.012, 96, 0.010, 0.050, 256)
    pc = usx.PipelineConfig()
    pc.demod.decimation = 2
    pipe = usx.ImagingPipeline(grid, pc)
    img = np.asarray(pipe.process_display(sim.simulate_gpu(sim.make_plane_waves([-0.1, 0.0, 0.1]), ph)))
    assert img.size == grid.n_pixels
    assert img.min() >= 0.0 and img.max() <= 1.0 and img.max() > 0.9
    assert pipe.frame_count == 1 and pipe.stats.total_ms > 0


# --- the Python boundary's physics checks (tests/test_usx.py) ---------------


@pytest.fixture(scope="module")
def setup64(usx_mod):
    usx = usx_mod
    probe = usx.presets.linear_5mhz()
    probe.n_elements = 64
    med = usx.Medium()
    acq = usx.presets.acquisition(probe, depth=0.04, medium=med)
    cfg = usx.SimulatorConfig()
    cfg.noise_db = -200.0
    return probe, med, acq, usx.Simulator(probe, med, acq, cfg)


def test_frame_is_complex_and_localizes(usx_mod, setup64):
    usx = usx_mod
    probe, med, acq, sim = setup64
    cd = sim.simulate(sim.make_plane_waves([0.0]), usx.Phantom.point_targets([usx.Vec3(0.004, 0.0, 0.025)]))
    grid = usx.ScanGrid.cartesian(0.0, 0.008, 81, 0.021, 0.029, 81)
    data = np.asarray(usx.ImagingPipeline(grid).process_frame(cd).data)
    assert data.dtype == np.complex64
    px, pz = _peak(np.abs(data), grid)
    assert px == pytest.approx(0.004, abs=2e-4) and pz == pytest.approx(0.025, abs=2e-4)


def test_bmode_is_normalized(usx_mod, setup64):
    usx = usx_mod
    probe, med, acq, sim = setup64
    cd = sim.simulate(sim.make_plane_waves([0.0]), wires(usx, [(0.0, 0.02), (0.0, 0.03)]))
    grid = usx.grid_for(probe, 0.01, 0.04, 64, 160)
    pc = usx.PipelineConfig()
    pc.dynamic_range_db = 50.0
    pipe = usx.ImagingPipeline(grid, pc)
    img = np.asarray(pipe.to_display(pipe.process_frame(cd), med))
    assert img.shape == (64, 160)
    assert img.min() >= 0.0 and img.max() <= 1.0 and img.max() > 0.9


def test_resolution_improves_with_more_angles(usx_mod, setup64):
    usx = usx_mod
    probe, med, acq, sim = setup64
    ph = usx.Phantom.point_targets([usx.Vec3(0.0, 0.0, 0.025)])
    grid = usx.ScanGrid.cartesian(-0.004, 0.004, 161, 0.023, 0.027, 81)
    bf = usx.Beamformer(usx.BeamformerConfig())
    widths = {}
    for n in (1, 11):
        iq = usx.demodulate(sim.simulate(usx.presets.plane_wave_sequence(sim, 12.0, n), ph),
                            usx.DemodConfig())
        env = np.abs(np.asarray(bf.beamform(iq, grid).data))
        widths[n] = usx.metrics.measure_psf(env, grid, (0.0, 0.025)).lateral_fwhm
    assert widths[11] < 0.75 * widths[1], widths


def test_speckle_snr_matches_rayleigh(usx_mod, setup64):
    usx = usx_mod
    probe, med, acq, sim = setup64
    ph = speckle_with_cysts(usx, probe, med, (), 0.010, 0.035, per_cell=15.0)
    iq = usx.demodulate(sim.simulate_gpu(usx.presets.plane_wave_sequence(sim, 9.0, 5), ph),
                        usx.DemodConfig())
    grid = usx.grid_for(probe, 0.015, 0.030, 96, 240)
    env = np.abs(np.asarray(usx.Beamformer().beamform(iq, grid).data))
    snr = usx.metrics.speckle_snr(env, grid, (0.0, 0.0225), (0.005, 0.005))
    assert snr == pytest.approx(usx.metrics.RAYLEIGH_SPECKLE_SNR, rel=0.15), snr


def test_cyst_is_darker_than_background(usx_mod, setup64):
    usx = usx_mod
    probe, med, acq, sim = setup64
    ph = speckle_with_cysts(usx, probe, med, ((0.0, 0.025, 0.003, 0.0),), 0.010, 0.038)
    iq = usx.demodulate(sim.simulate_gpu(usx.presets.plane_wave_sequence(sim, 9.0, 7), ph),
                        usx.DemodConfig())
    grid = usx.grid_for(probe, 0.015, 0.035, 128, 240)
    env = np.abs(np.asarray(usx.Beamformer().beamform(iq, grid).data))
    stats = usx.metrics.summarize(env, grid, lesion_xz=(0.0, 0.025), lesion_radius=0.003)
    assert stats["contrast_db"] < -6.0 and stats["gcnr"] > 0.5, stats


def test_pipeline_reuses_a_smoothed_reference(usx_mod, setup64):
    usx = usx_mod
    probe, med, acq, sim = setup64
    cd = sim.simulate(sim.make_plane_waves([0.0]), wires(usx, [(0.0, 0.025)]))
    pipe = usx.ImagingPipeline(usx.grid_for(probe, 0.015, 0.035, 48, 120))
    a = np.asarray(pipe.process_display(cd))
    b = np.asarray(pipe.process_display(cd))
    assert pipe.frame_count == 2 and pipe.stats.total_ms > 0
    assert np.allclose(a, b, atol=1e-6)
    pipe.reset()
    assert pipe.frame_count == 0
