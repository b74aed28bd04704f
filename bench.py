"""Benchmark of the MI355X imaging path (one GPU).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--skip-cpu]

Workload (one step = one frame through the hot path, RF already in HBM):
    128-element 5 MHz linear array, 11 plane waves steered -9..9 deg,
    RF float32 (11, 128, 1536) at 20 MHz from the point-scatterer simulator
    (seed 0: 20k scatterers in x [-19, 19] mm, z [5, 50] mm, anechoic 3 mm
    cyst at (0, 20 mm)); demodulation with decimation 4; minimum-variance
    beamforming (subarray 64, diagonal loading 1e-2, +-2 depth averaging,
    f/1.5) onto a 256 x 640 grid; complex64 frame out.
The headline `value` is MV frames/s. DAS, DAS + coherence factor and DMAS
frame rates on the same RF, the end-to-end (host RF in, host frame out)
MV rate, the learned-apodization rate (trained network shipped in
usx/data, with its gCNR on the cyst beside MV's, its BF16 MFMA roofline and
gCNR, speckle SNR, lateral FWHM and IQ error against MV on the 20 held-out
phantoms, seeds 300-319)
the uneven batched stream (64 frames of seed 1, DAS + MV each, one HIP
graph) and colour flow / power Doppler / wall filter on a 16-frame ensemble
of the benchmark grid are reported under "extras", with the MV, DAS and
learned rates of the same frame captured into a HIP graph and replayed
(`*_fps_graph`, each checked equal to the eager frame).

Frame rates come from a timed region of eager launches without per-kernel
events; a second timed region of as many frames, with HIP events recorded
around every launch, gives the per-kernel times.

roofline: the dominant kernel of the MV step (`mv`, covariance + solve) is
priced against the FP32 MFMA peak with the reference's FLOP count
(8 sum n_snap L(L+1)/2 + 8 sum (L^3/6 + L^2) over pixels, the algorithmic
work the measurement contract fixes), divided by the kernel's mean duration
from HIP events recorded on the library's stream over the profiled timed
region; `effective_tflops` repeats that rate under its own name. Beside it:
the FLOP the kernel actually executes (shift recursion: 6x fewer) as
`executed_tflops` / `executed_frac`, and the stage's VALU and MFMA busy
fractions from the committed rocprofv3 counter summary.

cpu_baseline: the reference's own core (oracle/_ref, built from its sources)
or else this repository's restatement (oracle/), timed on this host's cores
(OMP_NUM_THREADS, else the CPUs this process may use) on a bounded sample:
MV on 64 of the 256 lateral lines at full depth, one warm-up then the median
of 3, scaled x4; beside it a 1-thread MV sample, DAS / DAS+CF / DMAS full
frames (median of 5) and the CPU model and SMT state.

parity: the relative L2 of the timed GPU frames (DAS, DAS + CF, DMAS on every
pixel; MV on the 64 lines the CPU computed) against the reference's frames of
the same RF. Above 1e-5 (DAS family) or 1e-4 (MV) the bench exits with 3.
"""

from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "ultrasound-engine_amd"))

METRIC = "frames/s @256x640x11PW: DAS, MV, learned-MV at matched gCNR; % HBM roofline"
PEAK_FP32_MFMA = 157.3e12
PEAK_BF16_MFMA = 2.5e15
PEAK_HBM = 8.0e12


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def make_workload(usx):
    """The seed-0 RF of the benchmark configuration and its grid."""
    probe = usx.presets.linear_5mhz()
    med = usx.Medium()
    acq = usx.Acquisition()
    acq.sampling_frequency = 20e6
    acq.n_samples = 1536
    acq.prf = 5000.0
    sim = usx.Simulator(probe, med, acq, usx.SimulatorConfig())
    ph = usx.Phantom.speckle_box(-0.019, 0.019, 0.005, 0.050, 2e-4, 20000, 0, 1.0)
    ph.scale_amplitude_in_circle(usx.Vec3(0.0, 0.0, 0.020), 0.003, 0.0)
    txs = sim.make_plane_waves(usx.angles_deg(-9, 9, 11))
    raw = sim.simulate(txs, ph)
    grid = usx.grid_for(probe, 0.005, 0.050, 256, 640)
    return probe, med, acq, txs, ph, raw, grid


def pipeline_config(usx, combiner, cf=False, subarray=64):
    pc = usx.PipelineConfig()
    pc.demod.decimation = 4
    b = pc.beamformer
    b.f_number = 1.5
    b.combiner = getattr(usx.Combiner, combiner)
    b.coherence_factor = cf
    b.mv_subarray = subarray
    b.mv_diagonal_loading = 1e-2
    b.mv_range_average = 2
    pc.beamformer = b
    return pc


def mv_reference_flops(probe, grid, f_number=1.5, subarray=64, range_avg=2):
    """The reference's MV work per frame: covariance upper triangle
    8 * n_snap * L(L+1)/2 plus Cholesky and solves 8 * (L^3/6 + L^2), per
    pixel, with the reference's active-set, subarray and depth-window rules
    (beamformer.cpp:142-173). Active channels: |x_e - x_p| <= z / (2 F#)."""
    ex = np.asarray(probe.element_x, dtype=np.float64)
    x = np.asarray(grid.axis0, dtype=np.float64)
    z = np.asarray(grid.axis1, dtype=np.float64)
    half = 0.5 * z / f_number
    na = (np.abs(ex[None, None, :] - x[:, None, None]) < half[None, :, None]).sum(-1)
    L = np.clip(np.minimum(subarray, np.maximum(2, na // 2)), 2, None)
    ns = na - L + 1
    n1 = z.size
    j = np.arange(n1)
    rows = (np.minimum(j + range_avg, n1 - 1) - np.maximum(j - range_avg, 0) + 1)[None, :]
    live = na >= 4
    cov = 8.0 * rows * ns * L * (L + 1) / 2.0
    sol = 8.0 * (L ** 3 / 6.0 + L ** 2)
    return float(np.sum(np.where(live, cov + sol, 0.0)))


def mv_executed_flops(probe, grid, f_number=1.5, subarray=64, range_avg=2):
    """FLOP the default MV kernel actually executes per frame (8 per complex
    multiply-add): the first covariance column 8 * rows * n_sub * L, the
    sub-diagonal walks 8 * 2 rows * L(L-1)/2, the Cholesky 8 * L^3/6 and the
    two folded forward substitutions 8 * L^2. Far below the reference's count:
    the shift recursion replaces the O(L^2 n_snap) covariance."""
    ex = np.asarray(probe.element_x, dtype=np.float64)
    x = np.asarray(grid.axis0, dtype=np.float64)
    z = np.asarray(grid.axis1, dtype=np.float64)
    half = 0.5 * z / f_number
    na = (np.abs(ex[None, None, :] - x[:, None, None]) < half[None, :, None]).sum(-1)
    L = np.clip(np.minimum(subarray, np.maximum(2, na // 2)), 2, None)
    ns = na - L + 1
    n1 = z.size
    j = np.arange(n1)
    rows = (np.minimum(j + range_avg, n1 - 1) - np.maximum(j - range_avg, 0) + 1)[None, :]
    work = 8.0 * (rows * ns * L + rows * L * (L - 1) + L ** 3 / 6.0 + L ** 2)
    return float(np.sum(np.where(na >= 4, work, 0.0)))


def das_gathers(probe, grid, txs, f_number=1.5, c=1540.0):
    """Active (pixel, channel, transmit) gathers of the DAS frame: receive
    apodization non-zero (|x_e - x_p| < z / 2F#, dynamic_aperture_weight,
    geometry.cpp:123-130) times transmits whose Tukey illumination is
    non-zero (|x_p - centre| < half of the steered, diffraction-widened band,
    tx_sensitivity, beamformer.cpp:249-288)."""
    ex = np.asarray(probe.element_x, dtype=np.float64)
    x = np.asarray(grid.axis0, dtype=np.float64)[:, None]
    z = np.asarray(grid.axis1, dtype=np.float64)[None, :]
    half_rx = 0.5 * z / f_number
    n_rx = (np.abs(ex[None, None, :] - x[..., None]) < half_rx[..., None]).sum(-1)
    lo, hi = ex.min(), ex.max()
    aperture = max(hi - lo, probe.pitch)
    lam = c / probe.center_frequency
    diff = np.clip(z / aperture, 0.5, 20.0) * lam
    n_tx = np.zeros_like(n_rx)
    for t in txs:
        shift = z * np.tan(t.angle)
        a, b = lo + shift - diff, hi + shift + diff
        n_tx += (np.abs(x - 0.5 * (a + b)) < 0.5 * (b - a)).astype(n_rx.dtype)
    return int(np.sum(n_rx * n_tx))


def stage_busy():
    """(VALU busy, MFMA busy) of the MV stage from the committed counter
    summary, as fractions weighted by each solve kernel's busy cycles, or
    (None, None)."""
    path = os.path.join(ROOT, "profiles", "mv_frame_256x640_11pw_counters.md")
    if not os.path.exists(path):
        return None, None
    for line in open(path):
        if line.startswith("mv_stage_busy:"):
            v, m = line.split(":")[1].split(",")
            return float(v), float(m)
    return None, None


def _timed(runner, torch, steps, launch):
    """Wall ms and event ms per frame over `steps` launches, the timed region
    bracketed by synchronizations on both sides."""
    torch.cuda.synchronize()
    runner.synchronize()
    t0 = time.perf_counter()
    runner.record(0)
    for _ in range(steps):
        launch()
    runner.record(1)
    runner.synchronize()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    return wall * 1e3 / steps, runner.elapsed(0, 1) / steps


def time_device(usx, torch, raw, grid, pc, steps, warmup, profile=False, apod=None, graph=False):
    """Device-resident frames/s: RF in HBM before the timed regions.

    Returns (ms, ev_ms, kernels, frame, graph_ms): ms / ev_ms per frame over
    `steps` eager frames (host clock / HIP events); with `profile`, a second
    timed region of `steps` frames with HIP events recorded around every
    launch on the library's stream gives the per-kernel ms (kept apart so the
    events do not slow the throughput figure); with `graph`, the frame
    captured into a HIP graph and replayed `steps` times (graph_ms, else None).
    """
    runner = usx.DeviceRunner(grid, pc, raw)
    if apod is not None:
        runner.set_apodization(apod)
    rf = torch.from_numpy(np.ascontiguousarray(np.asarray(raw.rf))).to("cuda:0")
    frame = torch.empty((grid.n_axis0, grid.n_axis1), dtype=torch.complex64, device="cuda:0")
    torch.cuda.synchronize()
    for _ in range(warmup):
        runner.run(rf.data_ptr(), frame.data_ptr())
    eager = lambda: runner.run(rf.data_ptr(), frame.data_ptr())
    ms, ev_ms = _timed(runner, torch, steps, eager)
    kern = {}
    if profile:
        runner.profile(True)
        _timed(runner, torch, steps, eager)
        kern = json.loads(runner.profile_report())
        runner.profile(False)
    out = frame.cpu().numpy()
    graph_ms = None
    if graph:
        gframe = torch.empty_like(frame)
        runner.capture(rf.data_ptr(), gframe.data_ptr())
        for _ in range(warmup):
            runner.run_graph()
        graph_ms, _ = _timed(runner, torch, steps, runner.run_graph)
        if not torch.equal(gframe, frame):
            raise RuntimeError("graph replay differs from the eager frame")
    return ms, ev_ms, kern, out, graph_ms


def doppler_leg(usx, torch, frame, grid, steps, warmup, n_frames=16, prf=5000.0):
    """Colour flow, power Doppler and the wall filter on a 16-frame ensemble of
    the benchmark grid, device-resident: the DAS frame under a slow-time phase
    ramp (0.4 rad per frame) and a slight amplitude trend. Colour flow uses the
    reference CLI's settings (polynomial wall filter of order 1, Loupas, 6 x 1
    window). HBM roofline on compulsory bytes: the ensemble read once, the four
    output maps written once (13 B per pixel)."""
    n_pix = grid.n_axis0 * grid.n_axis1
    base = np.asarray(frame, np.complex64)
    ens = np.stack([base * np.complex64(np.exp(0.4j * e) * (1.0 + 0.01 * e))
                    for e in range(n_frames)]).astype(np.complex64)
    dc = usx.DopplerConfig()
    dc.wall_filter = usx.WallFilter.POLYNOMIAL
    dc.polynomial_order = 1
    dc.kernel_axial = 6
    dc.kernel_lateral = 1
    runner = usx.DopplerRunner(grid, n_frames, prf, 1540.0, 5.0e6, dc)
    d_ens = torch.from_numpy(ens).to("cuda:0")
    shape = (grid.n_axis0, grid.n_axis1)
    v, pw, w = (torch.empty(shape, dtype=torch.float32, device="cuda:0") for _ in range(3))
    ok = torch.empty(shape, dtype=torch.uint8, device="cuda:0")
    filt = torch.empty_like(d_ens)
    calls = {
        "color_flow": lambda: runner.color_flow(d_ens.data_ptr(), v.data_ptr(), pw.data_ptr(),
                                                w.data_ptr(), ok.data_ptr()),
        "power_doppler": lambda: runner.power_doppler(d_ens.data_ptr(), pw.data_ptr()),
        "wall_filter": lambda: runner.wall_filter(d_ens.data_ptr(), filt.data_ptr()),
    }
    out = {"frames": n_frames, "grid": list(shape)}
    torch.cuda.synchronize()
    for name, fn in calls.items():
        for _ in range(warmup):
            fn()
        runner.synchronize()
        runner.profile(True)
        runner.record(0)
        for _ in range(steps):
            fn()
        runner.record(1)
        runner.synchronize()
        k = json.loads(runner.profile_report())
        runner.profile(False)
        ms = runner.elapsed(0, 1) / steps
        kms = k[name][1] / k[name][0]
        out[f"{name}_ms"] = ms
        out[f"{name}_kernel_ms"] = kms
    cf_bytes = n_frames * n_pix * 8 + n_pix * 13
    out["color_flow_hbm"] = {"bound": "hbm", "bytes_per_call": cf_bytes,
                             "achieved_GBps": cf_bytes / (out["color_flow_kernel_ms"] * 1e-3) / 1e9,
                             "frac": cf_bytes / (out["color_flow_kernel_ms"] * 1e-3) / PEAK_HBM}
    wf_bytes = 2 * n_frames * n_pix * 8
    out["wall_filter_hbm_frac"] = wf_bytes / (out["wall_filter_kernel_ms"] * 1e-3) / PEAK_HBM
    pd_bytes = n_frames * n_pix * 8 + n_pix * 4
    out["power_doppler_hbm_frac"] = pd_bytes / (out["power_doppler_kernel_ms"] * 1e-3) / PEAK_HBM
    out["valid_fraction"] = float(ok.float().mean().item())
    return out, ens, dc


def cpu_info():
    """Host CPU facts for the baseline: model, logical CPUs usable by this
    process, and whether SMT is on."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    smt = None
    try:
        with open("/sys/devices/system/cpu/smt/active") as f:
            smt = f.read().strip() == "1"
    except OSError:
        pass
    usable = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    return {"model": model, "logical_cpus": os.cpu_count(), "usable_cpus": usable, "smt": smt}


def _omp_threads(n):
    """Set the OpenMP thread count of the CPU checkers' parallel regions
    (they share the process's libgomp)."""
    import ctypes
    ctypes.CDLL("libgomp.so.1").omp_set_num_threads(int(n))


def _median_s(fn, reps):
    """One untimed warm-up call, then the median wall time of `reps` calls;
    returns (median seconds, every time, the last result)."""
    fn()
    times, out = [], None
    for _ in range(reps):
        t0 = time.perf_counter()
        out = fn()
        times.append(time.perf_counter() - t0)
    return float(np.median(times)), times, out


def cpu_legs(usx, probe, med, acq, txs, raw, ens, dc, threads, prf=5000.0):
    """The reference's CPU pipeline (its own core built from its sources,
    oracle/_ref, else this repository's restatement) on the same RF:

    * MV process_frame on 64 of the 256 lateral lines (lines 2, 6, ..., 254 of
      the benchmark grid, at its axis values), warm-up + median of 3, at
      `threads` OpenMP threads, scaled x4 to the full frame; and on 4 lines
      (34, 98, 162, 226) with 1 thread, median of 3, scaled x64;
    * DAS, DAS + coherence factor and DMAS process_frame on the full frame,
      warm-up + median of 5 at `threads`, and DAS with 1 thread, median of 3;
    * colour flow on the bench's 16-frame ensemble (single-threaded, as the
      reference's colour flow has no OpenMP loop), warm-up + median of 3.

    Returns the MV baseline, the DAS baseline, the colour-flow time and the
    reference frames (DAS, DAS + CF, DMAS full; MV on its 64 lines)."""
    from oracle import oracle as O

    flavour = "reference" if O.available("reference") else "port"
    orc = O.load(flavour)
    s = O.Setup.of(probe, med, acq)
    otx = [O.Tx.of(t) for t in txs]
    half = 0.5 * probe.aperture_width()
    rf = np.asarray(raw.rf)
    dm = O.Demod(decimation=4)

    def frame(comb, cf=0, first=0, step=0):
        gs = O.GridSpec(0, -half, half, 256, 0.005, 0.050, 640, line_first=first, line_step=step)
        bf = O.Bf(f_number=1.5, combiner=comb, coherence_factor=cf, mv_subarray=64,
                  mv_diagonal_loading=1e-2, mv_range_average=2)
        return lambda: orc.process_frame(s, otx, rf, gs, dm, bf)[0]

    refs = {}
    _omp_threads(threads)
    mv_s, mv_all, refs["mv"] = _median_s(frame(2, 0, 2, 4), 3)
    das_s, das_all, refs["das"] = _median_s(frame(0), 5)
    refs["das_cf"] = frame(0, 1)()
    refs["dmas"] = frame(1)()
    _omp_threads(1)
    mv1_s, mv1_all, _ = _median_s(frame(2, 0, 34, 64), 3)
    das1_s, das1_all, _ = _median_s(frame(0), 3)
    d = O.Doppler.of(dc)
    cf_s, _, _ = _median_s(lambda: orc.color_flow(O.GridSpec(0, -half, half, 256, 0.005, 0.050, 640),
                                                  ens, 5.0e6, prf, 1540.0, d), 3)
    _omp_threads(threads)
    host = cpu_info()
    mv = {"value": 1.0 / (mv_s * 4.0), "unit": "frames/s", "cores": threads, "kind": flavour,
          "sample": "MV process_frame on 64 of the 256 lateral lines (every 4th) x 640 depths, "
                    "one warm-up then the median of 3 runs, scaled x4 to the full frame",
          "sample_seconds": [round(t, 3) for t in mv_all],
          "one_thread": {"value": 1.0 / (mv1_s * 64.0), "unit": "frames/s", "cores": 1,
                         "sample": "4 lateral lines x 640 depths, median of 3, scaled x64",
                         "sample_seconds": [round(t, 3) for t in mv1_all]},
          "host": host}
    das = {"value": 1.0 / das_s, "unit": "frames/s", "cores": threads, "kind": flavour,
           "sample": "DAS process_frame, full 256 x 640 frame, one warm-up then the median of 5",
           "sample_seconds": [round(t, 4) for t in das_all],
           "one_thread": {"value": 1.0 / das1_s, "unit": "frames/s", "cores": 1,
                          "sample_seconds": [round(t, 4) for t in das1_all]},
           "host": host}
    cf = {"ms": cf_s * 1e3, "cores": 1, "kind": flavour,
          "sample": f"color_flow on the {ens.shape[0]}-frame 256 x 640 ensemble, median of 3"}
    return mv, das, cf, refs


def parity(frames, refs):
    """Relative L2 of the timed GPU frames against the reference's CPU frames
    of the same RF: DAS, DAS + CF and DMAS on every pixel (bar 1e-5), MV with
    and without the refinement step on the reference's 64 lines (bar 1e-4)."""
    def rel(a, b):
        a = np.asarray(a, np.complex128)
        b = np.asarray(b, np.complex128)
        return float(np.linalg.norm(a - b) / np.linalg.norm(b))
    out = {"das": rel(frames["das"], refs["das"]), "das_cf": rel(frames["das_cf"], refs["das_cf"]),
           "dmas": rel(frames["dmas"], refs["dmas"]), "mv": rel(frames["mv"][2::4], refs["mv"]),
           "mv_refined": rel(frames["mv_refined"][2::4], refs["mv"]),
           "tolerance": {"das": 1e-5, "das_cf": 1e-5, "dmas": 1e-5, "mv": 1e-4, "mv_refined": 1e-4},
           "mv_lines": "2, 6, ..., 254 of 256"}
    out["pass"] = all(out[k] < out["tolerance"][k] for k in out["tolerance"])
    return out


def time_host(usx, raw, grid, pc, steps):
    """End-to-end: host ChannelData in, host Frame out (PCIe included)."""
    pipe = usx.ImagingPipeline(grid, pc)
    pipe.process_frame(raw)
    t0 = time.perf_counter()
    for _ in range(steps):
        pipe.process_frame(raw)
    return (time.perf_counter() - t0) * 1e3 / steps


def run_stream(n_frames, cache_path=None):
    """The uneven batched stream (seed 1): n frames, each beamformed with DAS
    and MV (subarray = elements / 2), recorded into one HIP graph; launch
    shapes from the autotuner's versioned cache (tuned where missing)."""
    from usx import stream as S
    specs = S.draw(seed=1, n=n_frames)
    cache = S.load_cache(cache_path or S.CACHE_PATH)
    t = time.perf_counter()
    st = S.Stream(specs, tune=True, cache=cache, log=log)
    setup_s = time.perf_counter() - t
    if st.tuned:
        S.save_cache(st.cache, cache_path or S.CACHE_PATH)
    st.run(graph=True)
    eager_ms = st.run(graph=False, reps=2)
    graph_ms = st.run(graph=True, reps=3)
    pixels = sum(sp["grid"][0] * sp["grid"][1] for sp in specs)
    res = {"frames": n_frames, "jobs": 2 * n_frames, "graph_ms": graph_ms, "eager_ms": eager_ms,
           "frames_per_s": n_frames / (graph_ms * 1e-3), "mpixels": pixels / 1e6,
           "elements": sorted({sp["elements"] for sp in specs}),
           "transmits": [min(len(sp["angles_deg"]) for sp in specs),
                         max(len(sp["angles_deg"]) for sp in specs)],
           "setup_s": setup_s, "classes_tuned_now": len(st.tuned),
           "classes": len({S.shape_class(sp, c) for sp in specs for c in ("DAS", "MV")})}
    log(f"  stream: {res}")
    return res


def pmc_workload(n_frames, combiner="MV"):
    """The workload of the counter passes: n frames of the benchmark through
    `combiner` (MV, DAS, DMAS or Learned; device-resident RF) and, for
    calibration, one device copy of exactly 256 MiB. Under rocprofv3 --pmc
    FETCH_SIZE (then WRITE_SIZE), traffic per frame = counter sum over a
    frame's dispatches, times (copy bytes / the copy's counter)."""
    import torch
    import usx
    probe, med, acq, txs, ph, raw, grid = make_workload(usx)
    pc = pipeline_config(usx, combiner)
    runner = usx.DeviceRunner(grid, pc, raw)
    if combiner == "Learned":
        from usx import learned as L
        runner.set_apodization(L.load_checkpoint(os.path.join(
            ROOT, "ultrasound-engine_amd", "usx", "data", "apod_mlp.npz"))[0])
    rf = torch.from_numpy(np.ascontiguousarray(np.asarray(raw.rf))).to("cuda:0")
    frame = torch.empty((grid.n_axis0, grid.n_axis1), dtype=torch.complex64, device="cuda:0")
    for _ in range(n_frames):
        runner.run(rf.data_ptr(), frame.data_ptr())
    runner.synchronize()
    src = torch.ones(64 << 20, dtype=torch.float32, device="cuda:0")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    dst.copy_(src)
    torch.cuda.synchronize()
    log(f"pmc workload: {n_frames} {combiner} frames, one {src.numel() * 4} B device copy")


def measured_traffic():
    """HBM bytes per MV frame from the committed counter summary, or None."""
    path = os.path.join(ROOT, "profiles", "mv_stage_256x640_hbm_traffic.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_frame:"):
            return float(line.split(":")[1])
    return None


def main():
    ap = argparse.ArgumentParser(description="MI355X imaging-path benchmark")
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--skip-cpu", action="store_true", help="omit the CPU baseline leg")
    ap.add_argument("--pmc-workload", type=int, default=0, metavar="N",
                    help="run only N MV frames and a 256 MiB calibration copy (for rocprofv3 --pmc)")
    ap.add_argument("--pmc-combiner", default="MV", choices=["MV", "DAS", "DMAS", "Learned"],
                    help="combiner of the --pmc-workload frames")
    ap.add_argument("--skip-stream", action="store_true", help="omit the uneven-stream leg")
    ap.add_argument("--skip-eval", action="store_true",
                    help="omit the 20-phantom learned-vs-MV image-quality evaluation")
    ap.add_argument("--stream-only", action="store_true",
                    help="run only the uneven-stream leg and print its result (for profiling)")
    ap.add_argument("--stream-frames", type=int, default=64)
    ap.add_argument("--tile-cache", default=None,
                    help="autotuner cache file (default: the package's usx/data/tile_cache.json)")
    a = ap.parse_args()
    if a.pmc_workload:
        pmc_workload(a.pmc_workload, a.pmc_combiner)
        return 0
    if a.gpus != 1:
        print(json.dumps({"error": "this benchmark runs on one GPU; the imaging path is "
                                   "data-parallel over frames and is measured per GPU",
                          "n_gpus": a.gpus}))
        return 2
    if a.stream_only:
        print(json.dumps(run_stream(a.stream_frames, a.tile_cache)))
        return 0

    import torch
    import usx

    log("simulating the benchmark RF (host) ...")
    t = time.perf_counter()
    probe, med, acq, txs, ph, raw, grid = make_workload(usx)
    log(f"  RF {raw.shape} in {time.perf_counter() - t:.1f} s")

    pc_mv = pipeline_config(usx, "MV")
    log("MV device-resident ...")
    mv_ms, mv_ev_ms, kern, mv_frame, mv_graph_ms = time_device(usx, torch, raw, grid, pc_mv,
                                                               a.steps, a.warmup, profile=True,
                                                               graph=True)
    log(f"  {mv_ms:.3f} ms/frame (events {mv_ev_ms:.3f}, graph {mv_graph_ms:.3f}), kernels {kern}")
    extras = {"mv_fps": 1e3 / mv_ms, "mv_fps_graph": 1e3 / mv_graph_ms}
    # the same frame with one step of iterative refinement (FP64 residual)
    pc_ref = pipeline_config(usx, "MV")
    b = pc_ref.beamformer
    b.mv_refine = True
    pc_ref.beamformer = b
    r_ms, _, r_kern, r_frame, _ = time_device(usx, torch, raw, grid, pc_ref, a.steps, a.warmup,
                                              profile=True)
    extras["mv_fps_refined"] = 1e3 / r_ms
    extras["mv_refined_kernels_ms"] = {n: v[1] / v[0] for n, v in r_kern.items()}
    log(f"  refined: {r_ms:.3f} ms/frame, kernels {r_kern}")
    das_frame = None
    timed_frames = {"mv": mv_frame, "mv_refined": r_frame}
    for name, comb, cf in [("das", "DAS", False), ("das_cf", "DAS", True), ("dmas", "DMAS", False)]:
        ms, ev_ms, k, fr, g_ms = time_device(usx, torch, raw, grid, pipeline_config(usx, comb, cf),
                                             max(a.steps, 100), a.warmup, profile=True,
                                             graph=name == "das")
        timed_frames[name] = fr
        if name == "das":
            das_frame = fr
            extras["das_fps_graph"] = 1e3 / g_ms
        extras[f"{name}_fps"] = 1e3 / ms
        extras[f"{name}_kernels_ms"] = {n: v[1] / v[0] for n, v in k.items()}
        log(f"  {name}: {ms:.4f} ms/frame, kernels {k}")
    extras["mv_end_to_end_fps"] = 1e3 / time_host(usx, raw, grid, pc_mv, max(3, a.steps // 4))

    # learned apodization (trained checkpoint shipped with the package)
    from usx import learned as L
    apod, _ = L.load_checkpoint(os.path.join(ROOT, "ultrasound-engine_amd", "usx", "data",
                                             "apod_mlp.npz"))
    pc_l = pipeline_config(usx, "Learned")
    l_ms, _, lk, l_frame, l_graph_ms = time_device(usx, torch, raw, grid, pc_l, max(a.steps, 100),
                                                   a.warmup, profile=True, apod=apod, graph=True)
    extras["learned_fps"] = 1e3 / l_ms
    extras["learned_fps_graph"] = 1e3 / l_graph_ms
    extras["learned_speedup_vs_mv"] = mv_ms / l_ms
    extras["learned_speedup_vs_mv_graph"] = mv_graph_ms / l_graph_ms
    extras["learned_kernels_ms"] = {n: v[1] / v[0] for n, v in lk.items()}
    cyst = ((0.0, 0.020), 0.003)
    extras["gcnr_mv"] = usx.metrics.gcnr(np.abs(mv_frame), grid, *cyst)
    extras["gcnr_learned"] = usx.metrics.gcnr(np.abs(l_frame), grid, *cyst)
    extras["learned_iq_err_vs_mv"] = float(np.linalg.norm(l_frame - mv_frame) / np.linalg.norm(mv_frame))
    infer_ms = extras["learned_kernels_ms"]["apod_infer"]
    infer_flops = 2.0 * grid.n_pixels * 4 * 256 * 256
    extras["apod_infer_roofline"] = {"bound": "mfma", "achieved": infer_flops / (infer_ms * 1e-3) / 1e12,
                                     "peak": PEAK_BF16_MFMA / 1e12, "unit": "TFLOP/s",
                                     "frac": infer_flops / (infer_ms * 1e-3) / PEAK_BF16_MFMA,
                                     "flops_per_launch": infer_flops, "kernel_ms": infer_ms}
    log(f"  learned: {l_ms:.4f} ms/frame ({mv_ms / l_ms:.1f}x MV), kernels {lk}, "
        f"gCNR {extras['gcnr_learned']:.3f} vs MV {extras['gcnr_mv']:.3f}")
    if not a.skip_eval:
        log("learned vs MV on the 20 held-out phantoms (seeds 300-319) ...")
        ev = L.evaluate(apod, list(range(300, 320)), verbose=False)["summary"]
        extras["learned_eval_20_seeds"] = ev
        log(f"  {ev}")
    extras["mv_kernels_ms"] = {n: v[1] / v[0] for n, v in kern.items()}

    flops = mv_reference_flops(probe, grid)
    mv_kernel_ms = kern["mv"][1] / kern["mv"][0]
    achieved = flops / (mv_kernel_ms * 1e-3)
    roof = {"bound": "mfma", "achieved": achieved / 1e12, "peak": PEAK_FP32_MFMA / 1e12,
            "unit": "TFLOP/s", "frac": achieved / PEAK_FP32_MFMA, "traffic": measured_traffic(),
            "kernel": "mv", "flops_per_launch": flops, "kernel_ms": mv_kernel_ms,
            "executed_flops_per_launch": mv_executed_flops(probe, grid),
            "executed_tflops": mv_executed_flops(probe, grid) / (mv_kernel_ms * 1e-3) / 1e12}
    vb, mb = stage_busy()
    roof["executed_frac"] = roof["executed_tflops"] * 1e12 / PEAK_FP32_MFMA
    roof["effective_tflops"] = roof["achieved"]
    roof["stage_valu_busy"] = vb
    roof["stage_mfma_busy"] = mb
    # DAS is priced on compulsory HBM bytes (RF in + frame out), and on the
    # reference's 24 FLOP per active gather against the FP32 peak.
    das_bytes = raw.rf.nbytes + grid.n_pixels * 8
    das_total_ms = sum(extras["das_kernels_ms"].values())
    extras["das_hbm_frac"] = das_bytes / (das_total_ms * 1e-3) / PEAK_HBM
    n_g = das_gathers(probe, grid, txs)
    extras["das_gathers"] = n_g
    extras["das_valu_frac"] = 24.0 * n_g / (das_total_ms * 1e-3) / PEAK_FP32_MFMA

    log("Doppler (16-frame ensemble, device-resident) ...")
    extras["doppler"], d_ens, d_cfg = doppler_leg(usx, torch, das_frame, grid, max(a.steps, 50),
                                                  a.warmup)
    log(f"  {extras['doppler']}")

    if not a.skip_stream:
        extras["stream"] = run_stream(a.stream_frames, a.tile_cache)

    cpu = None
    par = None
    if not a.skip_cpu:
        threads = int(os.environ.get("OMP_NUM_THREADS", 0)) or len(os.sched_getaffinity(0))
        log(f"CPU baseline: the reference's pipeline on {threads} threads and on 1 ...")
        cpu, cpu_das, cpu_cf, refs = cpu_legs(usx, probe, med, acq, txs, raw, d_ens, d_cfg, threads)
        log(f"  MV {cpu}")
        log(f"  DAS {cpu_das}, colour flow {cpu_cf}")
        extras["mv_speedup_vs_cpu"] = (1e3 / mv_ms) / cpu["value"]
        extras["mv_speedup_vs_cpu_1thread"] = (1e3 / mv_ms) / cpu["one_thread"]["value"]
        extras["cpu_das"] = cpu_das
        extras["das_speedup_vs_cpu"] = extras["das_fps"] / cpu_das["value"]
        extras["doppler"]["cpu_color_flow"] = cpu_cf
        extras["doppler"]["color_flow_speedup_vs_cpu"] = cpu_cf["ms"] / extras["doppler"]["color_flow_ms"]
        par = parity(timed_frames, refs)
        log(f"  parity of the timed frames vs the reference: {par}")

    line = {
        "metric": METRIC, "value": 1e3 / mv_ms, "unit": "frames/s", "n_gpus": 1,
        "steps": a.steps, "warmup": a.warmup, "ms_per_step": mv_ms, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seed-0 point-scatterer RF from the host simulator)",
        "config": {"workload": "MV process_frame: RF f32 (11,128,1536) @20 MHz, demod dec 4, "
                               "MV L=64 loading 1e-2 range +-2, f/1.5, 256x640 grid",
                   "grid": [256, 640], "channels": 128, "transmits": 11,
                   "parallelism": "single GPU"},
        "roofline": roof, "cpu_baseline": cpu, "parity": par, "extras": extras,
    }
    print(json.dumps(line))
    if par is not None and not par["pass"]:
        log(f"FAIL: timed frames differ from the reference beyond the bars: {par}")
        return 3
    return 0


if __name__ == "__main__":
    sys.exit(main())
