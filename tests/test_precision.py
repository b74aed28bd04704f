"""MV precision variants on the wire phantom, against the FP64 MV oracle.

Workload: the benchmark acquisition (128 elements, 11 plane waves, 1536
samples at 20 MHz) over the benchmark speckle (seed 0, 20k scatterers,
anechoic cyst at (0, 20 mm)) plus 9 wires of amplitude 10 at x = -10 mm,
z = 10..50 mm in 5 mm steps; MV with subarray 64, loading 1e-2, range +-2,
f/1.5, on the 64 lateral lines of the 256-line grid that span the wires.

For each variant -- FP32 covariance with and without the FP64-residual
refinement step, BF16 and FP16 covariance inputs (FP32 accumulation) --
the test reports IQ relative L2 error and, per wire, lateral FWHM and
sidelobe level, next to the FP64 oracle's. Gates: FP32 within 1e-4 of FP64;
BF16/FP16 (which cannot meet 1e-4) within 2 % of the oracle's FWHM and
1 dB of its sidelobe level on average.
"""

from __future__ import annotations

import json

import numpy as np
import pytest

from cases import bf_config
from conftest import rel_l2

pytestmark = pytest.mark.gpu

WIRES = [(-0.010, 0.010 + 0.005 * k) for k in range(9)]


@pytest.fixture(scope="module")
def wire_case(usx_mod, O):
    probe = usx_mod.presets.linear_5mhz()
    acq = usx_mod.Acquisition()
    acq.sampling_frequency = 20e6
    acq.n_samples = 1536
    acq.prf = 5000.0
    sim = usx_mod.Simulator(probe, usx_mod.Medium(), acq, usx_mod.SimulatorConfig())
    ph = usx_mod.Phantom.speckle_box(-0.019, 0.019, 0.005, 0.050, 2e-4, 20000, 0, 1.0)
    ph.scale_amplitude_in_circle(usx_mod.Vec3(0.0, 0.0, 0.020), 0.003, 0.0)
    ph.extend(usx_mod.Phantom.point_targets([usx_mod.Vec3(x, 0.0, z) for x, z in WIRES], 10.0))
    txs = sim.make_plane_waves(usx_mod.angles_deg(-9, 9, 11))
    raw = sim.simulate(txs, ph)
    full = usx_mod.grid_for(probe, 0.005, 0.052, 256, 640)
    ax0 = np.asarray(full.axis0)
    i0 = int(np.searchsorted(ax0, -0.0148))
    x0, x1 = float(ax0[i0]), float(ax0[i0 + 63])
    g = usx_mod.ScanGrid.cartesian(x0, x1, 64, 0.005, 0.052, 640)
    gs = O.GridSpec(0, x0, x1, 64, 0.005, 0.052, 640)
    d = usx_mod.DemodConfig()
    d.decimation = 4
    iq = usx_mod.demodulate(raw, d)
    return probe, txs, iq, g, gs


def _psf(usx, frame, grid):
    env = np.abs(np.asarray(frame))
    out = [usx.metrics.measure_psf(env, grid, w, search_radius=0.0015) for w in WIRES]
    return (np.array([p.lateral_fwhm for p in out]) * 1e3, np.array([p.sidelobe_db for p in out]))


def test_precision_variants_psf_report(usx_mod, port, O, wire_case):
    probe, txs, iq, g, gs = wire_case
    s = O.Setup.of(probe, usx_mod.Medium(), iq.acq)
    otx = [O.Tx.of(t) for t in txs]
    b64 = bf_config(usx_mod, "MV", subarray=64)
    ref = port.beamform(s, otx, np.asarray(iq.iq), gs, O.Bf.of(b64),
                        fdem=iq.demodulation_frequency, mv_fp64=True, info=True)[0]
    f_ref, sl_ref = _psf(usx_mod, ref, g)
    rows = {"fp64_oracle": {"iq_err": 0.0, "fwhm_mm": float(f_ref.mean()),
                            "sidelobe_db": float(np.nanmean(sl_ref))}}
    for name, prec, refine in [("fp32_refined", "FP32", True), ("fp32", "FP32", False),
                               ("bf16_refined", "BF16", True), ("fp16_refined", "FP16", True)]:
        b = bf_config(usx_mod, "MV", subarray=64, precision=prec, refine=refine)
        got = np.asarray(usx_mod.Beamformer(b).beamform(iq, g).data)
        f, sl = _psf(usx_mod, got, g)
        rows[name] = {"iq_err": rel_l2(got, ref), "fwhm_mm": float(f.mean()),
                      "sidelobe_db": float(np.nanmean(sl)),
                      "fwhm_max_rel_dev": float(np.max(np.abs(f - f_ref) / f_ref)),
                      "sidelobe_max_dev_db": float(np.nanmax(np.abs(sl - sl_ref)))}
    print("\n[precision] " + json.dumps(rows))
    assert rows["fp32_refined"]["iq_err"] < 1e-4 and rows["fp32"]["iq_err"] < 1e-4
    for name in ("bf16_refined", "fp16_refined"):
        r = rows[name]
        assert abs(r["fwhm_mm"] / rows["fp64_oracle"]["fwhm_mm"] - 1.0) < 0.02, (name, r)
        assert abs(r["sidelobe_db"] - rows["fp64_oracle"]["sidelobe_db"]) < 1.0, (name, r)
    assert rows["fp16_refined"]["iq_err"] < rows["bf16_refined"]["iq_err"]
