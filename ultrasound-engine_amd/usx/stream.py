"""The uneven batched stream: frames of different probe, grid and transmit
count beamformed as one captured launch sequence (no reference counterpart).

Each frame is drawn from ``np.random.default_rng(seed)``: a 64-, 128- or
192-element 5 MHz linear array, 1 to 31 plane waves at angles uniform in
[-15, 15] degrees, a grid of 128x320 up to 512x1280 pixels (one scale factor
for both axes) over the aperture and from 5 mm to 90 % of the record depth,
and 1024 to 4096 RF samples at 20 MHz of a Gaussian speckle phantom at 10
scatterers per resolution cell (simulated on the GPU). Every frame is
beamformed twice, DAS (f/1.5, Hann) and MV (subarray = elements / 2), each
with the launch shape the autotuner picked for its shape class; all
2 x n frames are recorded into one HIP graph (usx.FrameSequence) and
replayed.
"""

from __future__ import annotations

import json
import os
import time

import numpy as np

from . import _usx

CACHE_VERSION = 1
CACHE_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "tile_cache.json")


def draw(seed: int = 1, n: int = 64, max_scale: float = 4.0, max_samples: int = 4096) -> list[dict]:
    """Frame parameters of the stream."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        n_el = int(rng.choice([64, 128, 192]))
        n_tx = int(rng.integers(1, 32))
        angles = np.sort(rng.uniform(-15.0, 15.0, n_tx)).tolist()
        scale = float(rng.uniform(1.0, max_scale))
        n_samples = int(rng.integers(1024, max_samples + 1))
        out.append({"elements": n_el, "angles_deg": angles, "grid": [int(round(128 * scale)),
                                                                    int(round(320 * scale))],
                    "samples": n_samples, "seed": int(rng.integers(0, 2 ** 31 - 1))})
    return out


def simulate(spec: dict):
    """(raw ChannelData, grid, (x_min, x_max, n_axis0, z_min, z_max, n_axis1))
    of one frame, RF simulated on the GPU."""
    from . import grid_for, presets
    probe = presets.linear_5mhz()
    probe.n_elements = spec["elements"]
    med = _usx.Medium()
    acq = _usx.Acquisition()
    acq.sampling_frequency = 20e6
    acq.n_samples = spec["samples"]
    acq.prf = 5000.0
    sim = _usx.Simulator(probe, med, acq, _usx.SimulatorConfig())
    txs = sim.make_plane_waves([float(a) for a in spec["angles_deg"]])
    z_max = 0.9 * med.speed_of_sound * spec["samples"] / (2.0 * acq.sampling_frequency)
    half = 0.5 * probe.aperture_width()
    n = _usx.Phantom.recommended_scatterer_count(probe, med, 2 * half, z_max - 0.005, 4e-4,
                                                 f_number=2.0, per_cell=10.0)
    ph = _usx.Phantom.speckle_box(-half, half, 0.005, z_max, 2e-4, int(n), spec["seed"], 1.0)
    raw = sim.simulate_gpu(txs, ph)
    grid = grid_for(probe, 0.005, z_max, spec["grid"][0], spec["grid"][1])
    return raw, grid, (-half, half, spec["grid"][0], 0.005, z_max, spec["grid"][1])


def configs(spec: dict):
    """(DAS config, MV config) of a frame."""
    das = _usx.PipelineConfig()
    das.demod.decimation = 4
    b = das.beamformer
    b.f_number = 1.5
    b.combiner = _usx.Combiner.DAS
    das.beamformer = b
    mv = _usx.PipelineConfig()
    mv.demod.decimation = 4
    b = mv.beamformer
    b.f_number = 1.5
    b.combiner = _usx.Combiner.MV
    b.mv_subarray = spec["elements"] // 2
    b.mv_diagonal_loading = 1e-2
    b.mv_range_average = 2
    mv.beamformer = b
    return das, mv


def shape_class(spec: dict, combiner: str) -> str:
    """Autotuner cache key. The event-slot choice depends on the transmit
    count, the MV occupancy on the covariance size (elements) and the pixel
    count, so DAS classes are (elements, transmit class) and MV classes add
    a pixel-count class."""
    n_tx = len(spec["angles_deg"])
    tx_c = 4 if n_tx <= 4 else (8 if n_tx <= 8 else (16 if n_tx <= 16 else 32))
    key = f"{combiner}/e{spec['elements']}/t{tx_c}"
    if combiner == "MV":
        pix = spec["grid"][0] * spec["grid"][1]
        key += "/p" + ("s" if pix <= 2 ** 17 else ("m" if pix <= 2 ** 19 else "l"))
    return key


def load_cache(path: str = CACHE_PATH) -> dict:
    if os.path.exists(path):
        with open(path) as f:
            c = json.load(f)
        if c.get("version") == CACHE_VERSION and c.get("target") == "gfx950":
            return c
    return {"version": CACHE_VERSION, "target": "gfx950", "entries": {}}


def save_cache(cache: dict, path: str = CACHE_PATH) -> None:
    with open(path, "w") as f:
        json.dump(cache, f, indent=1, sort_keys=True)


class Stream:
    """Every frame's RF resident in HBM, tuned configurations, and the
    captured sequence (2 x n entries: DAS then MV of each frame)."""

    def __init__(self, specs, tune: bool = True, cache: dict | None = None, log=None):
        self.specs = specs
        self.cache = cache if cache is not None else load_cache()
        self.raws, self.grids, self.cfgs, self.extents = [], [], [], []
        self.tuned = {}
        t0 = time.time()
        for i, spec in enumerate(specs):
            raw, grid, extent = simulate(spec)
            das, mv = configs(spec)
            for name, pc in (("DAS", das), ("MV", mv)):
                key = shape_class(spec, name)
                ent = self.cache["entries"].get(key)
                if ent is None and tune:
                    slots, occ, ms = _usx.DeviceRunner(grid, pc, raw).autotune_host(raw, 2)
                    ent = {"slots": int(slots), "mv_occupancy": int(occ), "ms": float(ms)}
                    self.cache["entries"][key] = ent
                    self.tuned[key] = ent
                if ent is not None:
                    b = pc.beamformer
                    b.tile_slots = ent["slots"]
                    b.tile_mv_occupancy = ent["mv_occupancy"]
                    pc.beamformer = b
                self.raws.append(raw)
                self.grids.append(grid)
                self.cfgs.append(pc)
                self.extents.append(extent)
            if log and ((i + 1) % 8 == 0 or i + 1 == len(specs)):
                log(f"stream: {i + 1}/{len(specs)} frames simulated and tuned ({time.time() - t0:.1f} s)")
        self.seq = _usx.FrameSequence(self.raws, self.grids, self.cfgs)

    def run(self, graph: bool = True, reps: int = 1) -> float:
        """ms per pass of the whole sequence (HIP events on the sequence's stream)."""
        self.seq.synchronize()
        self.seq.record(0)
        for _ in range(reps):
            self.seq.launch(graph)
        self.seq.record(1)
        self.seq.synchronize()
        return self.seq.elapsed(0, 1) / reps

    def frames(self):
        return self.seq.frames()
