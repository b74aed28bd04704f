"""Raw channel data in: sources and the .npz recording format.

A source is anything with ``metadata`` (probe, medium, acquisition,
transmits) and ``frames()`` (an iterator of ChannelData, one frame's events
each), the interface of the reference's python/usx/ingest.py:36-62, so a
pipeline is written once whether its RF was simulated, replayed from a
recording or handed over by acquisition hardware as arrays.

The recording format is the reference's (python/usx/ingest.py:201-233): one
``np.savez_compressed`` archive with the RF samples and every parameter needed
to beamform them, under the key names and dtypes of ``ARCHIVE`` below. That
table is the only part kept alike, so that recordings move between the two
packages; saving and loading are both driven by it.

Sources here add two things: ``SimulatedSource(..., device=True)`` renders the
frames with the GPU simulator, and ``FileSource`` keeps the archive's samples
as one (frames, events, channels, samples) view and hands out frames without
copying them again.
"""

from __future__ import annotations

from dataclasses import dataclass, field
from pathlib import Path
from typing import Iterator

import numpy as np

from ._usx import (Acquisition, ChannelData, Medium, Phantom, Probe, Simulator, SimulatorConfig,
                   Transmit, TransmitType, Vec3)

# Transmit type codes of the archive's tx_type column.
TX_CODES = {TransmitType.FOCUSED: 0, TransmitType.PLANE_WAVE: 1, TransmitType.DIVERGING: 2}
TX_TYPES = {code: kind for kind, code in TX_CODES.items()}

# key -> (owner, attribute, dtype). Scalars are stored as 0-d arrays of the
# dtype numpy gives the Python value (int64 / float64), per-transmit columns
# with the dtype named here, as the reference writes them.
ARCHIVE = {
    "probe_n_elements": ("probe", "n_elements", int),
    "probe_pitch": ("probe", "pitch", float),
    "probe_element_width": ("probe", "element_width", float),
    "probe_elevation_height": ("probe", "elevation_height", float),
    "probe_center_frequency": ("probe", "center_frequency", float),
    "probe_bandwidth": ("probe", "bandwidth", float),
    "medium_speed_of_sound": ("medium", "speed_of_sound", float),
    "medium_attenuation": ("medium", "attenuation", float),
    "acq_sampling_frequency": ("acq", "sampling_frequency", float),
    "acq_n_samples": ("acq", "n_samples", int),
    "acq_t0": ("acq", "t0", float),
    "acq_prf": ("acq", "prf", float),
}
TX_COLUMNS = {
    "tx_type": (np.int32, lambda t: TX_CODES[t.type]),
    "tx_angle": (np.float32, lambda t: t.angle),
    "tx_focus": (np.float32, lambda t: (t.focus.x, t.focus.y, t.focus.z)),
    "tx_origin": (np.float32, lambda t: (t.origin.x, t.origin.y, t.origin.z)),
    "tx_delays": (np.float32, lambda t: list(t.delays)),
    "tx_apodization": (np.float32, lambda t: list(t.apodization)),
}


@dataclass
class SourceMetadata:
    probe: Probe
    medium: Medium
    acquisition: Acquisition
    transmits: list = field(default_factory=list)


class RawSource:
    """Base of every source: subclasses provide ``metadata`` and ``_blocks()``
    (float32 arrays of shape (events, channels, samples), one per frame)."""

    metadata: SourceMetadata

    def _blocks(self) -> Iterator[np.ndarray]:
        raise NotImplementedError

    def frames(self) -> Iterator[ChannelData]:
        m = self.metadata
        for block in self._blocks():
            yield ChannelData.from_rf(block, m.probe, m.medium, m.acquisition, m.transmits)

    def __iter__(self):
        return self.frames()

    @property
    def events_per_frame(self) -> int:
        return len(self.metadata.transmits)

    def close(self) -> None:
        pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


class SimulatedSource(RawSource):
    """Frames rendered from a phantom by the point-scatterer simulator,
    ``n_frames`` of them (None: without end). Scatterers with a velocity
    keep moving between frames: frame i starts i x (events / prf) after
    frame 0, and the simulator moves them within each frame's events itself.
    ``device``: render with the GPU simulator (Simulator.simulate_gpu) instead
    of the host one, whose RF is the reference's bit for bit."""

    def __init__(self, probe: Probe, medium: Medium, acquisition: Acquisition, phantom: Phantom,
                 transmits: list, config: SimulatorConfig | None = None,
                 n_frames: int | None = 1, device: bool = False):
        self.metadata = SourceMetadata(probe, medium, acquisition, list(transmits))
        self.simulator = Simulator(probe, medium, acquisition, config or SimulatorConfig())
        self._phantom = phantom
        self._n_frames = n_frames
        self._render = self.simulator.simulate_gpu if device else self.simulator.simulate

    def _phantom_at(self, i: int) -> Phantom:
        """The phantom as it stands when frame i starts."""
        vel = np.asarray(self._phantom.velocities, dtype=np.float64)
        if i == 0 or not vel.any():
            return self._phantom
        start = i * len(self.metadata.transmits) / self.metadata.acquisition.prf
        moved = self._phantom.copy()
        moved.positions = (np.asarray(self._phantom.positions, dtype=np.float64)
                           + vel * start).astype(np.float32)
        return moved

    def frames(self) -> Iterator[ChannelData]:
        i = 0
        while self._n_frames is None or i < self._n_frames:
            yield self._render(self.metadata.transmits, self._phantom_at(i))
            i += 1


class ArraySource(RawSource):
    """Frames handed over as arrays of shape (events, channels, samples), e.g.
    by acquisition hardware, with the geometry to beamform them."""

    def __init__(self, arrays, probe: Probe, medium: Medium, acquisition: Acquisition,
                 transmits: list):
        self.metadata = SourceMetadata(probe, medium, acquisition, list(transmits))
        self._arrays = arrays

    def _blocks(self):
        for a in self._arrays:
            yield np.asarray(a, dtype=np.float32)


class FileSource(RawSource):
    """Frames replayed from an archive written by ``save_npz``: the samples are
    (events, channels, samples) with the events of successive frames one after
    the other; a trailing partial frame is dropped. ``loop``: start over at
    the end, without end."""

    def __init__(self, path, loop: bool = False):
        self.path = Path(path)
        self._loop = loop
        with np.load(self.path, allow_pickle=False) as z:
            self.metadata = _metadata_from(z)
            rf = np.asarray(z["rf"], dtype=np.float32)
        n_ev = max(1, self.events_per_frame)
        if rf.ndim == 3:
            n = max(1, rf.shape[0] // n_ev)
            self._stack = rf[: n * n_ev].reshape(n, -1, *rf.shape[1:])
        else:
            self._stack = rf[None]

    @property
    def n_frames(self) -> int:
        return self._stack.shape[0]

    def _blocks(self):
        while True:
            yield from self._stack
            if not self._loop:
                return


def _metadata_from(z) -> SourceMetadata:
    objs = {"probe": Probe(), "medium": Medium(), "acq": Acquisition()}
    for key, (owner, attr, kind) in ARCHIVE.items():
        setattr(objs[owner], attr, kind(z[key]))
    n = len(z["tx_type"])
    transmits = []
    for i in range(n):
        t = Transmit()
        t.type = TX_TYPES[int(z["tx_type"][i])]
        t.angle = float(z["tx_angle"][i])
        t.focus = Vec3(*map(float, z["tx_focus"][i]))
        t.origin = Vec3(*map(float, z["tx_origin"][i]))
        t.delays = [float(v) for v in z["tx_delays"][i]]
        t.apodization = [float(v) for v in z["tx_apodization"][i]]
        transmits.append(t)
    return SourceMetadata(objs["probe"], objs["medium"], objs["acq"], transmits)


def save_npz(path, data: ChannelData) -> Path:
    """Write one ChannelData (RF) with everything needed to beamform it, in the
    reference's archive layout. Transmits must agree on the lengths of their
    delay and apodization lists (the columns are 2-D arrays)."""
    path = Path(path)
    objs = {"probe": data.probe, "medium": data.medium, "acq": data.acq}
    arrays = {"rf": np.asarray(data.rf, dtype=np.float32)}
    for key, (owner, attr, _) in ARCHIVE.items():
        arrays[key] = getattr(objs[owner], attr)
    txs = list(data.transmits)
    for key, (dtype, get) in TX_COLUMNS.items():
        values = [get(t) for t in txs]
        if key in ("tx_delays", "tx_apodization") and len({len(v) for v in values}) > 1:
            raise ValueError(f"save_npz: transmits disagree on the length of {key[3:]}")
        arrays[key] = np.array(values, dtype=dtype)
    np.savez_compressed(path, **arrays)
    return path
