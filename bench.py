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
usx/data, with its gCNR on the cyst beside MV's and its BF16 MFMA roofline)
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
(8 sum n_snap L(L+1)/2 + 8 sum (L^3/6 + L^2) over pixels), divided by the
kernel's mean duration from HIP events recorded on the library's stream
over the profiled timed region.

cpu_baseline: the reference's own core (oracle/_ref, built from its sources)
or else this repository's restatement (oracle/), timed on this host's cores
on a bounded sample (64 lateral lines of the same depth range), scaled to a
full frame.
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


def cpu_das_and_doppler(usx, probe, med, acq, txs, raw, grid, ens, dc, prf=5000.0):
    """The reference's CPU DAS process_frame on the full frame (its OpenMP
    threads) and its colour flow on the same 16-frame ensemble as the GPU leg
    (single-threaded in the reference), each timed once after a warm-up of
    the smaller one."""
    from oracle import oracle as O

    flavour = "reference" if O.available("reference") else "port"
    orc = O.load(flavour)
    s = O.Setup.of(probe, med, acq)
    otx = [O.Tx.of(t) for t in txs]
    half = 0.5 * probe.aperture_width()
    gs = O.GridSpec(0, -half, half, 256, 0.005, 0.050, 640)
    bf = O.Bf(f_number=1.5, combiner=0)
    rf = np.asarray(raw.rf)
    t0 = time.perf_counter()
    orc.process_frame(s, otx, rf, gs, O.Demod(decimation=4), bf)
    das_s = time.perf_counter() - t0
    d = O.Doppler.of(dc)
    t0 = time.perf_counter()
    orc.color_flow(gs, ens, 5.0e6, prf, 1540.0, d)
    cf_s = time.perf_counter() - t0
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return ({"value": 1.0 / das_s, "unit": "frames/s", "cores": threads, "kind": flavour,
             "sample": "DAS process_frame, full 256 x 640 frame"},
            {"ms": cf_s * 1e3, "cores": 1, "kind": flavour,
             "sample": f"color_flow on the {ens.shape[0]}-frame 256 x 640 ensemble"})


def time_host(usx, raw, grid, pc, steps):
    """End-to-end: host ChannelData in, host Frame out (PCIe included)."""
    pipe = usx.ImagingPipeline(grid, pc)
    pipe.process_frame(raw)
    t0 = time.perf_counter()
    for _ in range(steps):
        pipe.process_frame(raw)
    return (time.perf_counter() - t0) * 1e3 / steps


def cpu_baseline(usx, probe, med, acq, txs, raw, n_lines=128):
    """Reference CPU MV on a bounded sample (n_lines lateral lines over the
    full width and depth range), scaled to the 256-line frame."""
    from oracle import oracle as O

    flavour = "reference" if O.available("reference") else "port"
    orc = O.load(flavour)
    s = O.Setup.of(probe, med, acq)
    otx = [O.Tx.of(t) for t in txs]
    half = 0.5 * probe.aperture_width()
    gs = O.GridSpec(0, -half, half, n_lines, 0.005, 0.050, 640)
    bf = O.Bf(f_number=1.5, combiner=2, mv_subarray=64, mv_diagonal_loading=1e-2,
              mv_range_average=2)
    rf = np.asarray(raw.rf)
    t0 = time.perf_counter()
    _, st = orc.process_frame(s, otx, rf, gs, O.Demod(decimation=4), bf)
    sample_s = time.perf_counter() - t0
    frame_s = sample_s * (256.0 / n_lines)
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {"value": 1.0 / frame_s, "unit": "frames/s", "cores": threads, "kind": flavour,
            "sample": f"MV frame on {n_lines} of 256 lateral lines x 640 depths "
                      f"({sample_s:.1f} s), scaled x{256 // n_lines} to the full frame",
            "sample_seconds": round(sample_s, 3)}


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


def pmc_workload(n_frames):
    """The workload of the HBM-traffic counters: n MV frames of the benchmark
    (device-resident RF) and, for calibration, one device copy of exactly
    256 MiB. Under rocprofv3 --pmc FETCH_SIZE (then WRITE_SIZE), traffic per
    MV frame = counter sum over the MV-stage dispatches of a frame, times
    (copy bytes / the copy's counter)."""
    import torch
    import usx
    probe, med, acq, txs, ph, raw, grid = make_workload(usx)
    pc = pipeline_config(usx, "MV")
    runner = usx.DeviceRunner(grid, pc, raw)
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
    log(f"pmc workload: {n_frames} MV frames, one {src.numel() * 4} B device copy")


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
    ap.add_argument("--skip-stream", action="store_true", help="omit the uneven-stream leg")
    ap.add_argument("--stream-only", action="store_true",
                    help="run only the uneven-stream leg and print its result (for profiling)")
    ap.add_argument("--stream-frames", type=int, default=64)
    ap.add_argument("--tile-cache", default=None,
                    help="autotuner cache file (default: the package's usx/data/tile_cache.json)")
    a = ap.parse_args()
    if a.pmc_workload:
        pmc_workload(a.pmc_workload)
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
    das_frame = None
    for name, comb, cf in [("das", "DAS", False), ("das_cf", "DAS", True), ("dmas", "DMAS", False)]:
        ms, ev_ms, k, fr, g_ms = time_device(usx, torch, raw, grid, pipeline_config(usx, comb, cf),
                                             max(a.steps, 100), a.warmup, profile=True,
                                             graph=name == "das")
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
    extras["mv_kernels_ms"] = {n: v[1] / v[0] for n, v in kern.items()}

    flops = mv_reference_flops(probe, grid)
    mv_kernel_ms = kern["mv"][1] / kern["mv"][0]
    achieved = flops / (mv_kernel_ms * 1e-3)
    roof = {"bound": "mfma", "achieved": achieved / 1e12, "peak": PEAK_FP32_MFMA / 1e12,
            "unit": "TFLOP/s", "frac": achieved / PEAK_FP32_MFMA, "traffic": measured_traffic(),
            "kernel": "mv", "flops_per_launch": flops, "kernel_ms": mv_kernel_ms,
            "executed_flops_per_launch": mv_executed_flops(probe, grid),
            "executed_tflops": mv_executed_flops(probe, grid) / (mv_kernel_ms * 1e-3) / 1e12}
    # DAS is priced on compulsory HBM bytes (RF in + frame out) for reference.
    das_bytes = raw.rf.nbytes + grid.n_pixels * 8
    das_total_ms = sum(extras["das_kernels_ms"].values())
    extras["das_hbm_frac"] = das_bytes / (das_total_ms * 1e-3) / PEAK_HBM

    log("Doppler (16-frame ensemble, device-resident) ...")
    extras["doppler"], d_ens, d_cfg = doppler_leg(usx, torch, das_frame, grid, max(a.steps, 50),
                                                  a.warmup)
    log(f"  {extras['doppler']}")

    if not a.skip_stream:
        extras["stream"] = run_stream(a.stream_frames, a.tile_cache)

    cpu = None
    if not a.skip_cpu:
        log("CPU baseline (bounded sample) ...")
        cpu = cpu_baseline(usx, probe, med, acq, txs, raw)
        log(f"  {cpu}")
        extras["mv_speedup_vs_cpu"] = (1e3 / mv_ms) / cpu["value"]
        log("CPU DAS frame and colour flow (reference) ...")
        cpu_das, cpu_cf = cpu_das_and_doppler(usx, probe, med, acq, txs, raw, grid, d_ens, d_cfg)
        extras["cpu_das"] = cpu_das
        extras["das_speedup_vs_cpu"] = extras["das_fps"] / cpu_das["value"]
        extras["doppler"]["cpu_color_flow"] = cpu_cf
        extras["doppler"]["color_flow_speedup_vs_cpu"] = cpu_cf["ms"] / extras["doppler"]["color_flow_ms"]
        log(f"  DAS {cpu_das}, colour flow {cpu_cf}")

    line = {
        "metric": METRIC, "value": 1e3 / mv_ms, "unit": "frames/s", "n_gpus": 1,
        "steps": a.steps, "warmup": a.warmup, "ms_per_step": mv_ms, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seed-0 point-scatterer RF from the host simulator)",
        "config": {"workload": "MV process_frame: RF f32 (11,128,1536) @20 MHz, demod dec 4, "
                               "MV L=64 loading 1e-2 range +-2, f/1.5, 256x640 grid",
                   "grid": [256, 640], "channels": 128, "transmits": 11,
                   "parallelism": "single GPU"},
        "roofline": roof, "cpu_baseline": cpu, "extras": extras,
    }
    print(json.dumps(line))
    return 0


if __name__ == "__main__":
    sys.exit(main())
