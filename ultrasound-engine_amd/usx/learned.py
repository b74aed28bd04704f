"""Learned apodization beside minimum variance (no reference counterpart).

A 256 -> 3 x 256 (GELU) -> 256 MLP maps each pixel's delayed channel vector
(128 channels, rectangular aperture, normalised) to channel weights
1 + o_c; the frame is sum_c (1 + o_c) y_c. It is trained on the GPU, by the
library's own BF16 kernels (csrc/apod.hip), on the effective channel weights
of the minimum-variance beamformer, which the library computes for
simulated phantoms (Simulator.simulate_gpu). Minimum variance stays the
fallback (Combiner.Learned runs MV when no network is loaded) and the judge:
``evaluate`` compares learned, MV and DAS frames by gCNR, speckle SNR,
lateral FWHM and IQ error.

    python -m usx.learned train --out apod.npz [--seeds 100:300] ...
    python -m usx.learned evaluate --checkpoint apod.npz [--seeds 300:320]
"""

from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np

from . import _usx
from . import metrics

FEATURES = 256
LAYERS = 4
N_PARAMS = int(_usx.APOD_PARAMS)
CHECKPOINT_VERSION = 1
ARCH = "mlp-256x(3x256 gelu-tanh)x256-bf16"


def log(*a):
    print("[learned]", *a, flush=True)


# ---------------------------------------------------------------------------
# Acquisition and phantoms
# ---------------------------------------------------------------------------

def acquisition_setup(n_samples: int = 1536, fs: float = 20e6, n_tx: int = 11, half_angle: float = 9.0):
    """The benchmark acquisition: 128-element 5 MHz linear array, plane waves."""
    from . import angles_deg, presets
    probe = presets.linear_5mhz()
    med = _usx.Medium()
    acq = _usx.Acquisition()
    acq.sampling_frequency = fs
    acq.n_samples = n_samples
    acq.prf = 5000.0
    sim = _usx.Simulator(probe, med, acq, _usx.SimulatorConfig())
    txs = sim.make_plane_waves(angles_deg(-half_angle, half_angle, n_tx))
    return probe, med, acq, sim, txs


def phantom_mix(probe, medium, seed: int, x_half: float = 0.019, z_lo: float = 0.005,
                z_hi: float = 0.050, min_cysts: int = 0, min_wires: int = 0):
    """Speckle at 10 scatterers per resolution cell, 0-3 anechoic or
    hyperechoic cysts of radius 2-5 mm, 0-10 wires of amplitude 5-20, all
    drawn from `seed`. Returns (phantom, description)."""
    rng = np.random.default_rng(seed)
    n = _usx.Phantom.recommended_scatterer_count(probe, medium, 2 * x_half, z_hi - z_lo, 4e-4,
                                                 f_number=2.0, per_cell=10.0)
    ph = _usx.Phantom.speckle_box(-x_half, x_half, z_lo, z_hi, 2e-4, int(n), int(seed) & 0xFFFFFFFF, 1.0)
    cysts, wires = [], []
    for _ in range(max(min_cysts, int(rng.integers(0, 4)))):
        r = float(rng.uniform(0.002, 0.005))
        cx = float(rng.uniform(-x_half + r + 0.002, x_half - r - 0.002))
        cz = float(rng.uniform(z_lo + r + 0.003, z_hi - r - 0.003))
        # the first forced cyst is anechoic, so every evaluation phantom has one
        anechoic = (len(cysts) < min_cysts) or rng.random() < 0.5
        scale = 0.0 if anechoic else float(rng.uniform(2.0, 4.0))
        ph.scale_amplitude_in_circle(_usx.Vec3(cx, 0.0, cz), r, scale)
        cysts.append((cx, cz, r, scale))
    for _ in range(max(min_wires, int(rng.integers(0, 11)))):
        wx = float(rng.uniform(-x_half + 0.002, x_half - 0.002))
        wz = float(rng.uniform(z_lo + 0.003, z_hi - 0.002))
        amp = float(rng.uniform(5.0, 20.0))
        ph.extend(_usx.Phantom.point_targets([_usx.Vec3(wx, 0.0, wz)], amp))
        wires.append((wx, wz, amp))
    return ph, {"seed": int(seed), "scatterers": int(n), "cysts": cysts, "wires": wires}


def mv_config(subarray: int = 64, loading: float = 1e-2, range_avg: int = 2, f_number: float = 1.5):
    pc = _usx.PipelineConfig()
    pc.demod.decimation = 4
    b = pc.beamformer
    b.f_number = f_number
    b.combiner = _usx.Combiner.MV
    b.mv_subarray = subarray
    b.mv_diagonal_loading = loading
    b.mv_range_average = range_avg
    pc.beamformer = b
    return pc


# ---------------------------------------------------------------------------
# Parameters and checkpoints
# ---------------------------------------------------------------------------

def init_params(seed: int = 0) -> np.ndarray:
    """He-normal hidden layers, zero output layer and biases: the untrained
    network is DAS with the rectangular aperture."""
    rng = np.random.default_rng(seed)
    p = np.zeros(N_PARAMS, np.float32)
    for l in range(LAYERS - 1):
        p[l * FEATURES * FEATURES:(l + 1) * FEATURES * FEATURES] = rng.normal(
            0.0, np.sqrt(2.0 / FEATURES), FEATURES * FEATURES).astype(np.float32)
    return p


def _bf16_bits(w: np.ndarray) -> np.ndarray:
    u = np.ascontiguousarray(w, np.float32).view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) & 0xFFFF
    return r.astype(np.uint16)


def save_checkpoint(path: str, params: np.ndarray, meta: dict) -> None:
    """Weights as bf16 bit patterns (what the kernels compute with), biases
    in FP32, plus the version, architecture and training record."""
    p = np.asarray(params, np.float32)
    nw = LAYERS * FEATURES * FEATURES
    arrays = {f"w{l}_bf16": _bf16_bits(p[l * FEATURES * FEATURES:(l + 1) * FEATURES * FEATURES])
              for l in range(LAYERS)}
    arrays.update({f"b{l}": p[nw + l * FEATURES:nw + (l + 1) * FEATURES] for l in range(LAYERS)})
    np.savez_compressed(path, version=np.int32(CHECKPOINT_VERSION), arch=np.array(ARCH),
                        meta=np.array(json.dumps(meta)), **arrays)


def load_checkpoint(path: str):
    """(params FP32 vector, meta dict). Rejects other versions/architectures."""
    with np.load(path, allow_pickle=False) as z:
        if int(z["version"]) != CHECKPOINT_VERSION or str(z["arch"]) != ARCH:
            raise ValueError(f"{path}: unsupported checkpoint {int(z['version'])} / {z['arch']}")
        p = np.zeros(N_PARAMS, np.float32)
        nw = LAYERS * FEATURES * FEATURES
        for l in range(LAYERS):
            bits = z[f"w{l}_bf16"].astype(np.uint32) << 16
            p[l * FEATURES * FEATURES:(l + 1) * FEATURES * FEATURES] = bits.view(np.float32)
            p[nw + l * FEATURES:nw + (l + 1) * FEATURES] = z[f"b{l}"]
        return p, json.loads(str(z["meta"]))


# ---------------------------------------------------------------------------
# Training
# ---------------------------------------------------------------------------

def build_dataset(runner, sim, txs, probe, medium, seeds, per_frame: int, rng, log_every: int = 10):
    """Append `per_frame` random pixels of every seed's phantom (simulated on
    the GPU) with their MV weights to the runner's device training set."""
    n_pix = runner.frame_pixels
    t0 = time.time()
    for i, seed in enumerate(seeds):
        ph, _ = phantom_mix(probe, medium, seed)
        raw = sim.simulate_gpu(txs, ph)
        pick = rng.choice(n_pix, size=min(per_frame, n_pix), replace=False).astype(np.int32)
        runner.dataset_append(raw, np.sort(pick))
        if (i + 1) % log_every == 0 or i + 1 == len(seeds):
            log(f"dataset: {i + 1}/{len(seeds)} phantoms, {runner.dataset_size} rows, "
                f"{time.time() - t0:.1f} s")


def train(runner, epochs: float, batch: int, lr: float, seed: int, log_every_s: float = 20.0):
    """Adam with a cosine learning-rate schedule; returns the loss history."""
    n = runner.dataset_size
    steps = max(1, int(epochs * n / batch))
    chunk = 100
    hist = []
    t0 = time.time()
    last = t0
    done = 0
    while done < steps:
        k = min(chunk, steps - done)
        frac = (done + 0.5 * k) / steps
        cur = lr * 0.5 * (1.0 + np.cos(np.pi * frac))
        hist.append(runner.train(k, batch=batch, lr=float(cur), seed=seed))
        done += k
        if time.time() - last > log_every_s or done == steps:
            last = time.time()
            log(f"train: step {done}/{steps} lr {cur:.2e} loss {np.mean(hist[-1]):.5f} "
                f"({time.time() - t0:.1f} s)")
    return np.concatenate(hist)


# ---------------------------------------------------------------------------
# Evaluation
# ---------------------------------------------------------------------------

def _clear(desc, x, z, r):
    for cx, cz, cr, _ in desc["cysts"]:
        if np.hypot(x - cx, z - cz) < cr + r + 0.001:
            return False
    for wx, wz, _ in desc["wires"]:
        if np.hypot(x - wx, z - wz) < r + 0.003:
            return False
    return True


def frame_metrics(frame, grid, desc, ref=None) -> dict:
    env = np.abs(np.asarray(frame))
    out = {}
    an = [c for c in desc["cysts"] if c[3] == 0.0]
    if an:
        out["gcnr"] = float(np.mean([metrics.gcnr(env, grid, (c[0], c[1]), c[2]) for c in an]))
    if desc["wires"]:
        w = []
        for wx, wz, _ in desc["wires"]:
            if not _clear({"cysts": desc["cysts"], "wires": []}, wx, wz, 0.002):
                continue
            psf = metrics.measure_psf(env, grid, (wx, wz))
            if np.isfinite(psf.lateral_fwhm):
                w.append(psf.lateral_fwhm)
        if w:
            out["fwhm_mm"] = float(np.mean(w) * 1e3)
    for bx in (-0.012, 0.0, 0.012, -0.006, 0.006):
        for bz in (0.035, 0.025, 0.042, 0.018):
            if _clear(desc, bx, bz, 0.004 * 1.5):
                out["speckle_snr"] = float(metrics.speckle_snr(env, grid, (bx, bz), (0.004, 0.004)))
                break
        if "speckle_snr" in out:
            break
    if ref is not None:
        r = np.asarray(ref)
        out["iq_err"] = float(np.linalg.norm(np.asarray(frame) - r) / np.linalg.norm(r))
    return out


def evaluate(params, seeds, subarray: int = 64, verbose: bool = True) -> dict:
    """Learned vs MV vs DAS (rectangular aperture, the untrained network)
    on phantoms of the training mix with at least one anechoic cyst and one
    wire. Returns per-method means and per-seed records."""
    from . import grid_for
    probe, med, acq, sim, txs = acquisition_setup()
    grid = grid_for(probe, 0.005, 0.050, 256, 640)
    pc_mv = mv_config(subarray)
    pc_l = mv_config(subarray)
    pc_l.beamformer.combiner = _usx.Combiner.Learned
    learned = _usx.ImagingPipeline(grid, pc_l)
    learned.set_apodization(params)
    zero = _usx.ImagingPipeline(grid, pc_l)
    zero.set_apodization(np.zeros(N_PARAMS, np.float32))
    mv = _usx.ImagingPipeline(grid, pc_mv)
    rec = []
    for seed in seeds:
        ph, desc = phantom_mix(probe, med, seed, min_cysts=1, min_wires=1)
        raw = sim.simulate_gpu(txs, ph)
        f_mv = np.asarray(mv.process_frame(raw).data)
        f_l = np.asarray(learned.process_frame(raw).data)
        f_d = np.asarray(zero.process_frame(raw).data)
        r = {"seed": int(seed), "mv": frame_metrics(f_mv, grid, desc),
             "learned": frame_metrics(f_l, grid, desc, f_mv), "das_rect": frame_metrics(f_d, grid, desc, f_mv)}
        rec.append(r)
        if verbose:
            log(f"eval seed {seed}: " + "; ".join(
                f"{k} " + " ".join(f"{m}={v:.3f}" for m, v in r[k].items()) for k in ("mv", "learned", "das_rect")))
    summary = {}
    for k in ("mv", "learned", "das_rect"):
        keys = sorted({m for r in rec for m in r[k]})
        summary[k] = {m: float(np.mean([r[k][m] for r in rec if m in r[k]])) for m in keys}
    # per-seed agreement with minimum variance, the judge
    d = [abs(r["learned"]["gcnr"] - r["mv"]["gcnr"]) for r in rec
         if "gcnr" in r["learned"] and "gcnr" in r["mv"]]
    if d:
        summary["gcnr_delta"] = {"mean_abs": float(np.mean(d)), "max_abs": float(np.max(d)),
                                 "of_means": abs(summary["learned"]["gcnr"] - summary["mv"]["gcnr"]),
                                 "seeds": len(d)}
    return {"summary": summary, "records": rec}


# ---------------------------------------------------------------------------
# Command line
# ---------------------------------------------------------------------------

def _seeds(s: str):
    a, b = s.split(":")
    return list(range(int(a), int(b)))


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(prog="python -m usx.learned")
    sub = ap.add_subparsers(dest="cmd", required=True)
    t = sub.add_parser("train")
    t.add_argument("--out", required=True)
    t.add_argument("--seeds", default="100:300")
    t.add_argument("--per-frame", type=int, default=100_000)
    t.add_argument("--epochs", type=float, default=4.0)
    t.add_argument("--batch", type=int, default=4096)
    t.add_argument("--lr", type=float, default=1e-3)
    t.add_argument("--init-seed", type=int, default=0)
    t.add_argument("--eval-seeds", default="300:320")
    e = sub.add_parser("evaluate")
    e.add_argument("--checkpoint", required=True)
    e.add_argument("--seeds", default="300:320")
    e.add_argument("--json", default=None)
    a = ap.parse_args(argv)
    if a.cmd == "train":
        from . import grid_for
        probe, med, acq, sim, txs = acquisition_setup()
        grid = grid_for(probe, 0.005, 0.050, 256, 640)
        ph, _ = phantom_mix(probe, med, 0)
        like = sim.simulate_gpu(txs, ph)
        runner = _usx.DeviceRunner(grid, mv_config(), like)
        seeds = _seeds(a.seeds)
        runner.dataset_reserve(len(seeds) * a.per_frame)
        rng = np.random.default_rng(12345)
        t0 = time.time()
        build_dataset(runner, sim, txs, probe, med, seeds, a.per_frame, rng)
        t_data = time.time() - t0
        runner.set_apodization(init_params(a.init_seed))
        t1 = time.time()
        hist = train(runner, a.epochs, a.batch, a.lr, seed=7)
        t_train = time.time() - t1
        params = runner.get_apodization()
        meta = {"seeds": a.seeds, "per_frame": a.per_frame, "rows": int(runner.dataset_size),
                "epochs": a.epochs, "batch": a.batch, "lr": a.lr, "steps": int(hist.size),
                "loss_first": float(hist[:50].mean()), "loss_last": float(hist[-50:].mean()),
                "data_s": t_data, "train_s": t_train}
        save_checkpoint(a.out, params, meta)
        log("saved", a.out, json.dumps(meta))
        if a.eval_seeds:
            res = evaluate(params, _seeds(a.eval_seeds))
            log("summary", json.dumps(res["summary"]))
        return 0
    params, meta = load_checkpoint(a.checkpoint)
    res = evaluate(params, _seeds(a.seeds))
    log("summary", json.dumps(res["summary"]))
    if a.json:
        with open(a.json, "w") as f:
            json.dump({"meta": meta, **res}, f, indent=1)
    return 0


if __name__ == "__main__":
    sys.exit(main())
