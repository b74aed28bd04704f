"""Raw channel-data sources (reference: python/usx/ingest.py:1-233).

Every source answers the same two questions -- what acquired this data
(``metadata``: probe, medium, acquisition, transmits) and what are its frames
(``frames()``: an iterator of ``ChannelData``) -- so the pipeline does not care
whether RF was simulated, replayed from a file or handed over by a scanner
driver. The archive format of ``save_npz`` / ``FileSource`` is the
reference's, key for key, so recordings made with either package load in the
other.

Loading never unpickles: archives are read with ``numpy.load(...,
allow_pickle=False)``.
"""

from __future__ import annotations

import abc
from dataclasses import dataclass, field
from pathlib import Path

import numpy as np

from ._usx import (Acquisition, ChannelData, Medium, Phantom, Probe, Simulator,
                   SimulatorConfig, Transmit, TransmitType, Vec3)

# Archive code of each transmit type (the reference's order).
_TYPE_CODES = {TransmitType.FOCUSED: 0, TransmitType.PLANE_WAVE: 1, TransmitType.DIVERGING: 2}
_CODE_TYPES = {v: k for k, v in _TYPE_CODES.items()}


@dataclass
class SourceMetadata:
    probe: Probe
    medium: Medium
    acquisition: Acquisition
    transmits: list = field(default_factory=list)


class RawSource(abc.ABC):
    """A source of pre-beamforming channel data."""

    @property
    @abc.abstractmethod
    def metadata(self) -> SourceMetadata:
        ...

    @abc.abstractmethod
    def frames(self):
        """Yield one ``ChannelData`` per image frame."""

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
    """Frames simulated from a phantom on demand; ``n_frames=None`` streams
    forever. Moving scatterers advance by one frame period (transmits / PRF)
    per frame, so a stream shows continuous motion."""

    def __init__(self, probe: Probe, medium: Medium, acquisition: Acquisition, phantom: Phantom,
                 transmits: list, config: SimulatorConfig | None = None,
                 n_frames: int | None = 1, gpu: bool = False):
        self._sim = Simulator(probe, medium, acquisition, config or SimulatorConfig())
        self._meta = SourceMetadata(probe, medium, acquisition, list(transmits))
        self._phantom = phantom
        self._n_frames = n_frames
        self._gpu = gpu

    @property
    def metadata(self) -> SourceMetadata:
        return self._meta

    @property
    def simulator(self) -> Simulator:
        return self._sim

    def _simulate(self, phantom: Phantom) -> ChannelData:
        if self._gpu:
            return self._sim.simulate_gpu(self._meta.transmits, phantom)
        return self._sim.simulate(self._meta.transmits, phantom)

    def frames(self):
        period = len(self._meta.transmits) / self._meta.acquisition.prf
        vel = np.asarray(self._phantom.velocities, dtype=np.float64)
        moving = bool(vel.size and np.any(vel != 0.0))
        start = np.asarray(self._phantom.positions, dtype=np.float64) if moving else None
        f = 0
        while self._n_frames is None or f < self._n_frames:
            if moving and f > 0:
                p = self._phantom.copy()
                p.positions = (start + vel * (f * period)).astype(np.float32)
                yield self._simulate(p)
            else:
                yield self._simulate(self._phantom)
            f += 1


class FileSource(RawSource):
    """Frames replayed from an archive written by :func:`save_npz`; an RF
    array of more events than transmits holds several frames back to back."""

    def __init__(self, path, loop: bool = False):
        self.path = Path(path)
        self._loop = loop
        with np.load(self.path, allow_pickle=False) as z:
            self._rf = np.asarray(z["rf"], dtype=np.float32)
            probe = Probe()
            probe.n_elements = int(z["probe_n_elements"])
            probe.pitch = float(z["probe_pitch"])
            probe.element_width = float(z["probe_element_width"])
            probe.elevation_height = float(z["probe_elevation_height"])
            probe.center_frequency = float(z["probe_center_frequency"])
            probe.bandwidth = float(z["probe_bandwidth"])
            medium = Medium()
            medium.speed_of_sound = float(z["medium_speed_of_sound"])
            medium.attenuation = float(z["medium_attenuation"])
            acq = Acquisition()
            acq.sampling_frequency = float(z["acq_sampling_frequency"])
            acq.n_samples = int(z["acq_n_samples"])
            acq.t0 = float(z["acq_t0"])
            acq.prf = float(z["acq_prf"])
            transmits = transmits_from_arrays(z["tx_type"], z["tx_angle"], z["tx_focus"],
                                              z["tx_origin"], z["tx_delays"], z["tx_apodization"])
        self._meta = SourceMetadata(probe, medium, acq, transmits)

    @property
    def metadata(self) -> SourceMetadata:
        return self._meta

    @property
    def n_frames(self) -> int:
        n_ev = max(1, len(self._meta.transmits))
        return max(1, self._rf.shape[0] // n_ev) if self._rf.ndim == 3 else 1

    def frames(self):
        n_ev = len(self._meta.transmits)
        m = self._meta
        while True:
            for f in range(self.n_frames):
                block = self._rf[f * n_ev:(f + 1) * n_ev] if self._rf.ndim == 3 else self._rf
                yield ChannelData.from_rf(block, m.probe, m.medium, m.acquisition, m.transmits)
            if not self._loop:
                return


class ArraySource(RawSource):
    """In-memory (events, channels, samples) RF blocks -- the adapter a scanner
    driver feeds."""

    def __init__(self, arrays, probe: Probe, medium: Medium, acquisition: Acquisition,
                 transmits: list):
        self._arrays = arrays
        self._meta = SourceMetadata(probe, medium, acquisition, list(transmits))

    @property
    def metadata(self) -> SourceMetadata:
        return self._meta

    def frames(self):
        m = self._meta
        for a in self._arrays:
            yield ChannelData.from_rf(np.asarray(a, dtype=np.float32), m.probe, m.medium,
                                      m.acquisition, m.transmits)


def transmits_from_arrays(types, angles, focus, origin, delays, apod) -> list:
    """Transmit objects from the archive's per-transmit arrays."""
    out = []
    for i in range(len(types)):
        t = Transmit()
        t.type = _CODE_TYPES[int(types[i])]
        t.angle = float(angles[i])
        t.focus = Vec3(*(float(v) for v in focus[i]))
        t.origin = Vec3(*(float(v) for v in origin[i]))
        t.delays = [float(v) for v in np.ravel(delays[i])]
        t.apodization = [float(v) for v in np.ravel(apod[i])]
        out.append(t)
    return out


def _per_transmit(rows, what: str) -> np.ndarray:
    """Per-transmit vectors as one (n_transmits, width) float32 array; all
    empty gives width 0 (an empty vector means "none given", which zeros
    would not preserve, so empty and filled vectors cannot be mixed)."""
    rows = [np.asarray(r, np.float32).ravel() for r in rows]
    sizes = {r.size for r in rows}
    if len(sizes) > 1:
        raise ValueError(f"save_npz: transmits have {what} vectors of different lengths {sorted(sizes)}")
    width = sizes.pop() if sizes else 0
    return np.stack(rows) if rows else np.zeros((0, width), np.float32)


def save_npz(path, data: ChannelData) -> Path:
    """Write RF channel data with everything needed to interpret it (probe,
    medium, acquisition, transmits) as a compressed .npz archive."""
    path = Path(path)
    txs = list(data.transmits)
    p, med, acq = data.probe, data.medium, data.acq
    np.savez_compressed(
        path,
        rf=np.asarray(data.rf, dtype=np.float32),
        probe_n_elements=p.n_elements,
        probe_pitch=p.pitch,
        probe_element_width=p.element_width,
        probe_elevation_height=p.elevation_height,
        probe_center_frequency=p.center_frequency,
        probe_bandwidth=p.bandwidth,
        medium_speed_of_sound=med.speed_of_sound,
        medium_attenuation=med.attenuation,
        acq_sampling_frequency=acq.sampling_frequency,
        acq_n_samples=acq.n_samples,
        acq_t0=acq.t0,
        acq_prf=acq.prf,
        tx_type=np.array([_TYPE_CODES[t.type] for t in txs], dtype=np.int32),
        tx_angle=np.array([t.angle for t in txs], dtype=np.float32),
        tx_focus=np.array([[t.focus.x, t.focus.y, t.focus.z] for t in txs], dtype=np.float32),
        tx_origin=np.array([[t.origin.x, t.origin.y, t.origin.z] for t in txs], dtype=np.float32),
        tx_delays=_per_transmit([t.delays for t in txs], "delay"),
        tx_apodization=_per_transmit([t.apodization for t in txs], "apodization"),
    )
    return path
