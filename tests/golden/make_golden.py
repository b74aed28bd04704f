"""Write the golden vectors under tests/golden/ from the compiled reference.

    OMP_NUM_THREADS=1 python tests/golden/make_golden.py

Needs oracle/_ref/libusx_ref.so, the reference's own core compiled from its
sources by oracle/Makefile (only where /root/reference exists). Everything is
produced by the reference's functions: its phantom factory, transmit
factory, simulator (single-threaded, hence deterministic), demodulation and
beamformer. The fixture holds inputs and outputs only.

Case "small": 32-element linear array (linear_5mhz geometry), 3 plane waves
at -9/0/9 deg, fs 20 MHz, 768 samples, 600 speckle scatterers (seed 0) plus
two wires of amplitude 10; demodulation with decimation 4; 24 x 48 grid over
the aperture width, z 6-24 mm; f/1.5. Frames for DAS, DAS+CF, DMAS,
MV (L = 8 and 16, loading 1e-2, +-2 depth averaging), Incoherent DAS and
Select DAS, and the RF-path DAS.
"""

from __future__ import annotations

import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402


def main():
    if os.environ.get("OMP_NUM_THREADS") != "1":
        raise SystemExit("run with OMP_NUM_THREADS=1 (the reference simulator's reduction "
                         "order is only deterministic single-threaded)")
    ref = O.load("reference")
    s = O.Setup(n_elements=32, pitch=300e-6, element_width=270e-6, elevation_height=5e-3,
                center_frequency=5e6, bandwidth=0.65, sampling_frequency=20e6, n_samples=768,
                prf=5000.0)
    sim = O.SimCfg()
    angles = np.deg2rad(np.linspace(-9, 9, 3)).astype(np.float32)
    txs = ref.make_transmits(s, sim, 1, angles)
    pos, amp = ref.speckle_box(-0.006, 0.006, 0.005, 0.026, 2e-4, 600, 0, 1.0)
    pos = np.concatenate([pos, np.array([[0.001, 0, 0.012], [-0.002, 0, 0.019]], np.float32)])
    amp = np.concatenate([amp, np.array([10.0, 10.0], np.float32)])
    rf = ref.simulate(s, sim, txs, pos, amp)
    iq, fs, fd = ref.demodulate(s, txs, rf, O.Demod(decimation=4))
    half = float(np.float32(0.5) * (np.float32(31) * np.float32(300e-6) + np.float32(270e-6)))
    g = O.GridSpec(0, -half, half, 24, 0.006, 0.024, 48)
    s_iq = O.Setup(**{**s.__dict__, "sampling_frequency": fs, "n_samples": iq.shape[2]})
    a0, a1, _ = ref.grid(g)

    frames = {}
    cfgs = {
        "das": O.Bf(f_number=1.5),
        "das_cf": O.Bf(f_number=1.5, coherence_factor=1),
        "dmas": O.Bf(f_number=1.5, combiner=1),
        "mv8": O.Bf(f_number=1.5, combiner=2, mv_subarray=8),
        "mv16": O.Bf(f_number=1.5, combiner=2, mv_subarray=16),
        "das_incoherent": O.Bf(f_number=1.5, compounding=1),
        "das_select": O.Bf(f_number=1.5, compounding=2),
    }
    for k, b in cfgs.items():
        frames[k] = ref.beamform(s_iq, txs, iq, g, b, fdem=fd)
    frames["das_rf"] = ref.beamform(s, txs, rf, g, O.Bf(f_number=1.5))
    taps = ref.design_lowpass(33, 0.5 * 5e6 / 20e6)

    out = os.path.join(HERE, "small_case.npz")
    np.savez_compressed(
        out, rf=rf, iq=iq, taps=taps, axis0=a0, axis1=a1,
        delays=np.stack([t.delays for t in txs]), apodization=np.stack([t.apodization for t in txs]),
        positions=pos, amplitudes=amp, **{f"frame_{k}": v for k, v in frames.items()})
    meta = {
        "generator": "tests/golden/make_golden.py (reference core compiled by oracle/Makefile, "
                     "OMP_NUM_THREADS=1)",
        "setup": s.__dict__, "angles_rad": angles.tolist(), "grid": g.__dict__,
        "demod": {"decimation": 4, "fs_out": fs, "fdem": fd},
        "configs": {k: v.__dict__ for k, v in cfgs.items()},
        "sha256": {k: hashlib.sha256(np.ascontiguousarray(v).tobytes()).hexdigest()
                   for k, v in [("rf", rf), ("iq", iq)] + [(f"frame_{k}", v) for k, v in frames.items()]},
    }
    with open(os.path.join(HERE, "small_case.json"), "w") as f:
        json.dump(meta, f, indent=1, default=float)
    print("wrote", out, os.path.getsize(out), "bytes")


if __name__ == "__main__":
    main()
