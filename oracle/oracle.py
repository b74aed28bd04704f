"""ctypes front end for the CPU checkers (test infrastructure only).

Only tests/, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
bench.py import this module, and only as the checker: it never stands in for
the HIP path. Two libraries export the same ``oc_*`` C entry points
(oracle_abi.h):

* ``load("port")``      -> oracle/liboracle_usx.so, this repository's own
  C++17 restatement of the reference imaging path (usx_oracle.cpp);
* ``load("reference")`` -> oracle/_ref/libusx_ref.so, the reference's own
  core sources compiled by oracle/Makefile (present where it was built).

Inputs are plain Python/numpy values or any objects carrying the reference's
attribute names (``probe.n_elements``, ``transmit.delays`` ...), so tests can
hand over objects from the product's ``usx`` package directly.
"""

from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBS = {
    "port": os.path.join(HERE, "liboracle_usx.so"),
    "reference": os.path.join(HERE, "_ref", "libusx_ref.so"),
}

_f = C.c_float
_i = C.c_int
_fp = C.POINTER(C.c_float)
_ip = C.POINTER(C.c_int)


class OcDisplay(C.Structure):
    _fields_ = [("apply_tgc", _i), ("dynamic_range_db", _f), ("gain_db", _f),
                ("reference_smoothing", _f)]


class OcSetup(C.Structure):
    _fields_ = [("n_elements", _i), ("pitch", _f), ("element_width", _f),
                ("elevation_height", _f), ("center_frequency", _f), ("bandwidth", _f),
                ("speed_of_sound", _f), ("attenuation", _f), ("density", _f),
                ("sampling_frequency", _f), ("n_samples", _i), ("t0", _f), ("prf", _f)]


class OcTransmit(C.Structure):
    _fields_ = [("type", _i), ("angle", _f), ("focus", _f * 3), ("origin", _f * 3),
                ("n_delays", _i), ("delays", _fp), ("n_apod", _i), ("apodization", _fp)]


class OcGrid(C.Structure):
    _fields_ = [("type", _i), ("a0_min", _f), ("a0_max", _f), ("n0", _i),
                ("a1_min", _f), ("a1_max", _f), ("n1", _i), ("apex", _f * 3),
                ("line_first", _i), ("line_step", _i)]


class OcDemod(C.Structure):
    _fields_ = [("demodulation_frequency", _f), ("cutoff_fraction", _f),
                ("filter_taps", _i), ("decimation", _i)]


class OcBf(C.Structure):
    _fields_ = [("f_number", _f), ("rx_window", _i), ("tukey_alpha", _f), ("combiner", _i),
                ("compounding", _i), ("transmit_weighting", _i), ("coherence_factor", _i),
                ("mv_subarray", _i), ("mv_diagonal_loading", _f), ("mv_range_average", _i)]


class OcSimCfg(C.Structure):
    _fields_ = [("pulse_cycles", _f), ("tx_window", _i), ("tx_f_number", _f),
                ("oversampling", _i), ("full_transmit_diffraction", _i),
                ("incident_trim_db", _f), ("max_incident_window_us", _f),
                ("apply_attenuation", _i), ("apply_directivity", _i), ("apply_spreading", _i),
                ("noise_db", _f), ("noise_seed", C.c_uint32), ("enable_motion", _i)]


# --------------------------------------------------------------------------
# Plain-value descriptions (reference defaults: types.hpp, *.hpp configs)
# --------------------------------------------------------------------------

@dataclass
class Setup:
    n_elements: int = 128
    pitch: float = 300e-6
    element_width: float = 270e-6
    elevation_height: float = 5e-3
    center_frequency: float = 5.0e6
    bandwidth: float = 0.6
    speed_of_sound: float = 1540.0
    attenuation: float = 0.5
    density: float = 1000.0
    sampling_frequency: float = 40.0e6
    n_samples: int = 2048
    t0: float = 0.0
    prf: float = 5000.0

    @staticmethod
    def of(probe, medium, acq) -> "Setup":
        return Setup(probe.n_elements, probe.pitch, probe.element_width,
                     probe.elevation_height, probe.center_frequency, probe.bandwidth,
                     medium.speed_of_sound, medium.attenuation, medium.density,
                     acq.sampling_frequency, acq.n_samples, acq.t0, acq.prf)

    def c(self) -> OcSetup:
        return OcSetup(*[getattr(self, n) for n, _ in OcSetup._fields_])


@dataclass
class Tx:
    type: int = 1
    angle: float = 0.0
    focus: tuple = (0.0, 0.0, 0.03)
    origin: tuple = (0.0, 0.0, 0.0)
    delays: np.ndarray = field(default_factory=lambda: np.zeros(0, np.float32))
    apodization: np.ndarray = field(default_factory=lambda: np.zeros(0, np.float32))

    @staticmethod
    def of(t) -> "Tx":
        def v3(v):
            return (float(v.x), float(v.y), float(v.z)) if hasattr(v, "x") else tuple(v)
        ty = t.type
        ty = int(ty.value) if hasattr(ty, "value") else int(ty)
        return Tx(ty, float(t.angle), v3(t.focus), v3(t.origin),
                  np.ascontiguousarray(t.delays, np.float32),
                  np.ascontiguousarray(t.apodization, np.float32))


@dataclass
class GridSpec:
    type: int = 0
    a0_min: float = -0.01
    a0_max: float = 0.01
    n0: int = 16
    a1_min: float = 0.01
    a1_max: float = 0.03
    n1: int = 32
    apex: tuple = (0.0, 0.0, 0.0)
    line_first: int = 0      # with line_step > 0: only lines line_first + k line_step,
    line_step: int = 0       # at the full grid's axis values

    @property
    def lines(self) -> int:
        """Lateral lines the frames of this spec hold."""
        if self.line_step <= 0:
            return self.n0
        return len(range(self.line_first, self.n0, self.line_step))

    def c(self) -> OcGrid:
        return OcGrid(self.type, self.a0_min, self.a0_max, self.n0, self.a1_min, self.a1_max,
                      self.n1, (_f * 3)(*self.apex), self.line_first, self.line_step)


@dataclass
class Demod:
    demodulation_frequency: float = 0.0
    cutoff_fraction: float = 0.5
    filter_taps: int = 33
    decimation: int = 1

    @staticmethod
    def of(d) -> "Demod":
        return Demod(d.demodulation_frequency, d.cutoff_fraction, d.filter_taps, d.decimation)

    def c(self) -> OcDemod:
        return OcDemod(self.demodulation_frequency, self.cutoff_fraction, self.filter_taps,
                       self.decimation)


@dataclass
class Bf:
    f_number: float = 1.75
    rx_window: int = 1
    tukey_alpha: float = 0.5
    combiner: int = 0
    compounding: int = 0
    transmit_weighting: int = 1
    coherence_factor: int = 0
    mv_subarray: int = 16
    mv_diagonal_loading: float = 1e-2
    mv_range_average: int = 2

    @staticmethod
    def of(b) -> "Bf":
        def e(v):
            return int(v.value) if hasattr(v, "value") else int(v)
        return Bf(b.f_number, e(b.rx_window), b.tukey_alpha, e(b.combiner), e(b.compounding),
                  int(bool(b.transmit_weighting)), int(bool(b.coherence_factor)),
                  b.mv_subarray, b.mv_diagonal_loading, b.mv_range_average)

    def c(self) -> OcBf:
        return OcBf(*[getattr(self, n) for n, _ in OcBf._fields_])


@dataclass
class SimCfg:
    pulse_cycles: float = 2.5
    tx_window: int = 1
    tx_f_number: float = 2.0
    oversampling: int = 4
    full_transmit_diffraction: int = 1
    incident_trim_db: float = -40.0
    max_incident_window_us: float = 30.0
    apply_attenuation: int = 1
    apply_directivity: int = 1
    apply_spreading: int = 1
    noise_db: float = -60.0
    noise_seed: int = 7
    enable_motion: int = 1

    def c(self) -> OcSimCfg:
        return OcSimCfg(*[getattr(self, n) for n, _ in OcSimCfg._fields_])


class OcDoppler(C.Structure):
    _fields_ = [("wall_filter", _i), ("polynomial_order", _i), ("loupas", _i),
                ("kernel_axial", _i), ("kernel_lateral", _i), ("power_threshold_db", _f),
                ("coherence_threshold", _f)]


@dataclass
class Doppler:
    """DopplerConfig (reference doppler.hpp:41-79); wall_filter 0 none,
    1 polynomial regression, 2 mean subtraction."""
    wall_filter: int = 1
    polynomial_order: int = 1
    loupas: int = 1
    kernel_axial: int = 4
    kernel_lateral: int = 1
    power_threshold_db: float = -35.0
    coherence_threshold: float = 0.35

    @staticmethod
    def of(d) -> "Doppler":
        def e(v):
            return int(v.value) if hasattr(v, "value") else int(v)
        return Doppler(e(d.wall_filter), d.polynomial_order, int(bool(d.loupas)), d.kernel_axial,
                       d.kernel_lateral, d.power_threshold_db, d.coherence_threshold)

    def c(self) -> OcDoppler:
        return OcDoppler(*[getattr(self, n) for n, _ in OcDoppler._fields_])


class OracleError(RuntimeError):
    pass


def _ptr(a):
    return a.ctypes.data_as(_fp)


class Oracle:
    """One loaded checker library."""

    def __init__(self, path: str):
        self.path = path
        self.lib = C.CDLL(path)
        L = self.lib
        L.oc_last_error.restype = C.c_char_p
        L.oc_flavour.restype = C.c_char_p
        L.oc_window_value.restype = _f
        L.oc_window_value.argtypes = [_i, _f, _f]
        L.oc_transmit_arrival_time.restype = _f
        L.oc_transmit_sensitivity.restype = _f
        self.flavour = L.oc_flavour().decode()
        self.has_ex = hasattr(L, "oc_beamform_ex")

    def _check(self, rc):
        if rc != 0:
            msg = self.lib.oc_last_error().decode()
            raise (ValueError if rc == 1 else OracleError)(msg)

    @staticmethod
    def _txs(txs):
        arr = (OcTransmit * len(txs))()
        keep = []
        for i, t in enumerate(txs):
            t = t if isinstance(t, Tx) else Tx.of(t)
            d = np.ascontiguousarray(t.delays, np.float32)
            a = np.ascontiguousarray(t.apodization, np.float32)
            keep += [d, a]
            arr[i].type = t.type
            arr[i].angle = t.angle
            arr[i].focus = (_f * 3)(*t.focus)
            arr[i].origin = (_f * 3)(*t.origin)
            arr[i].n_delays = d.size
            arr[i].delays = _ptr(d)
            arr[i].n_apod = a.size
            arr[i].apodization = _ptr(a)
        return arr, keep

    # ---- small functions -------------------------------------------------
    def grid(self, g: GridSpec):
        a0 = np.zeros(g.n0, np.float32)
        a1 = np.zeros(g.n1, np.float32)
        pts = np.zeros((g.n0 * g.n1, 3), np.float32)
        gc = g.c()
        self._check(self.lib.oc_grid_build(C.byref(gc), _ptr(a0), _ptr(a1), _ptr(pts)))
        return a0, a1, pts

    def design_lowpass(self, n_taps: int, fc: float) -> np.ndarray:
        n = _i(0)
        self._check(self.lib.oc_design_lowpass(n_taps, _f(fc), None, C.byref(n)))
        h = np.zeros(n.value, np.float32)
        self._check(self.lib.oc_design_lowpass(n_taps, _f(fc), _ptr(h), C.byref(n)))
        return h

    def window_value(self, w: int, u: float, alpha: float = 0.5) -> float:
        return float(self.lib.oc_window_value(w, u, alpha))

    def transmit_arrival_time(self, tx, point, c: float) -> float:
        arr, keep = self._txs([tx])
        p = (_f * 3)(*point)
        self.lib.oc_transmit_arrival_time.argtypes = [C.POINTER(OcTransmit), _f * 3, _f]
        return float(self.lib.oc_transmit_arrival_time(arr, p, c))

    def transmit_sensitivity(self, setup: Setup, tx, point, wavelength: float) -> float:
        arr, keep = self._txs([tx])
        s = setup.c()
        p = (_f * 3)(*point)
        self.lib.oc_transmit_sensitivity.argtypes = [C.POINTER(OcSetup), C.POINTER(OcTransmit),
                                                     _f * 3, _f]
        return float(self.lib.oc_transmit_sensitivity(C.byref(s), arr, p, wavelength))

    # ---- simulator-side factories (reference library only) --------------
    def make_transmits(self, setup: Setup, sim: SimCfg, kind: int, params, n=None,
                       extra: float = 0.0):
        params = np.ascontiguousarray(params, np.float32)
        n = len(params) if n is None else n
        ne = setup.n_elements
        d = np.zeros((n, ne), np.float32)
        a = np.zeros((n, ne), np.float32)
        fo = np.zeros((n, 3), np.float32)
        og = np.zeros((n, 3), np.float32)
        an = np.zeros(n, np.float32)
        s, sc = setup.c(), sim.c()
        self._check(self.lib.oc_make_transmits(C.byref(s), C.byref(sc), kind, n, _ptr(params),
                                               _f(extra), _ptr(d), _ptr(a), _ptr(fo), _ptr(og),
                                               _ptr(an)))
        ty = {1: 1, 0: 0, 2: 2}[kind]
        return [Tx(ty, float(an[i]), tuple(map(float, fo[i])), tuple(map(float, og[i])),
                   d[i].copy(), a[i].copy()) for i in range(n)]

    def speckle_box(self, x_min, x_max, z_min, z_max, y_half, n, seed=1234, sigma=1.0):
        pos = np.zeros((n, 3), np.float32)
        amp = np.zeros(n, np.float32)
        self._check(self.lib.oc_speckle_box(_f(x_min), _f(x_max), _f(z_min), _f(z_max),
                                            _f(y_half), n, C.c_uint32(seed), _f(sigma),
                                            _ptr(pos), _ptr(amp)))
        return pos, amp

    def simulate(self, setup: Setup, sim: SimCfg, txs, pos, amp, vel=None) -> np.ndarray:
        arr, keep = self._txs(txs)
        pos = np.ascontiguousarray(pos, np.float32)
        amp = np.ascontiguousarray(amp, np.float32)
        velp = None
        if vel is not None:
            vel = np.ascontiguousarray(vel, np.float32)
            velp = _ptr(vel)
        rf = np.zeros((len(txs), setup.n_elements, setup.n_samples), np.float32)
        s, sc = setup.c(), sim.c()
        self._check(self.lib.oc_simulate(C.byref(s), C.byref(sc), arr, len(txs), len(amp),
                                         _ptr(pos), _ptr(amp), velp, _ptr(rf)))
        return rf

    # ---- hot path ----------------------------------------------------------
    def demodulate(self, setup: Setup, txs, rf, demod: Demod):
        arr, keep = self._txs(txs)
        rf = np.ascontiguousarray(rf, np.float32)
        s, d = setup.c(), demod.c()
        n_out, fs, fdem = _i(0), _f(0), _f(0)
        self._check(self.lib.oc_demodulate(C.byref(s), arr, len(txs), _ptr(rf), C.byref(d), None,
                                           C.byref(n_out), C.byref(fs), C.byref(fdem)))
        iq = np.zeros((len(txs), setup.n_elements, n_out.value), np.complex64)
        self._check(self.lib.oc_demodulate(C.byref(s), arr, len(txs), _ptr(rf), C.byref(d),
                                           iq.ctypes.data_as(_fp), C.byref(n_out), C.byref(fs),
                                           C.byref(fdem)))
        return iq, float(fs.value), float(fdem.value)

    def beamform(self, setup: Setup, txs, data, grid: GridSpec, bf: Bf, fdem: float = 0.0,
                 event_begin: int = 0, event_end: int = -1, mv_fp64: bool = False,
                 weights: bool = False, ych: bool = False, info: bool = False):
        """Beamform RF (float32) or IQ (complex64) channel data; returns the
        frame (n0, n1) complex64, plus a dict of optional per-pixel outputs."""
        arr, keep = self._txs(txs)
        data = np.ascontiguousarray(data)
        is_iq = int(np.iscomplexobj(data))
        data = data.astype(np.complex64 if is_iq else np.float32, copy=False)
        s, g, b = setup.c(), grid.c(), bf.c()
        frame = np.zeros((grid.lines, grid.n1), np.complex64)
        extra = {}
        if mv_fp64 or weights or ych or info:
            if not self.has_ex:
                raise OracleError(f"{self.flavour} library has no oc_beamform_ex")
            npix, nch = grid.lines * grid.n1, setup.n_elements
            w = np.zeros((npix, nch), np.complex64) if weights else None
            y = np.zeros((npix, nch), np.complex64) if ych else None
            inf = np.zeros((npix, 3), np.int32) if info else None
            self._check(self.lib.oc_beamform_ex(
                C.byref(s), arr, len(txs), is_iq, data.ctypes.data_as(_fp), _f(fdem),
                C.byref(g), C.byref(b), event_begin, event_end, frame.ctypes.data_as(_fp),
                int(mv_fp64), w.ctypes.data_as(_fp) if w is not None else None,
                y.ctypes.data_as(_fp) if y is not None else None,
                inf.ctypes.data_as(_ip) if inf is not None else None))
            extra = {k: v for k, v in (("weights", w), ("ych", y), ("info", inf)) if v is not None}
        else:
            self._check(self.lib.oc_beamform(C.byref(s), arr, len(txs), is_iq,
                                             data.ctypes.data_as(_fp), _f(fdem), C.byref(g),
                                             C.byref(b), event_begin, event_end,
                                             frame.ctypes.data_as(_fp)))
        return (frame, extra) if extra else frame

    def display(self, setup: Setup, grid: GridSpec, frames, fdem: float, apply_tgc: bool = True,
                dynamic_range_db: float = 60.0, gain_db: float = 0.0,
                reference_smoothing: float = 0.85):
        """Frames (k, n0, n1) complex64 through ImagingPipeline::to_display in
        order; returns the display images (k, n0, n1) float32."""
        fr = np.ascontiguousarray(np.asarray(frames, np.complex64).reshape(-1, grid.n0, grid.n1))
        out = np.zeros(fr.shape, np.float32)
        cfg = OcDisplay(int(apply_tgc), dynamic_range_db, gain_db, reference_smoothing)
        s = setup.c()
        self._check(self.lib.oc_display(C.byref(s), C.byref(grid.c()), C.byref(cfg), fr.shape[0],
                                        fr.ctypes.data_as(_fp), _f(fdem), out.ctypes.data_as(_fp)))
        return out

    def validate_bf(self, bf: Bf) -> None:
        """BeamformerConfig::validate as the constructor runs it; raises
        ValueError with the reference's message."""
        b = bf.c()
        self._check(self.lib.oc_validate_bf(C.byref(b)))

    def process_frame(self, setup: Setup, txs, rf, grid: GridSpec, demod: Demod, bf: Bf):
        arr, keep = self._txs(txs)
        rf = np.ascontiguousarray(rf, np.float32)
        s, g, d, b = setup.c(), grid.c(), demod.c(), bf.c()
        frame = np.zeros((grid.lines, grid.n1), np.complex64)
        stats = (C.c_double * 3)()
        self._check(self.lib.oc_process_frame(C.byref(s), arr, len(txs), _ptr(rf), C.byref(g),
                                              C.byref(d), C.byref(b), frame.ctypes.data_as(_fp),
                                              stats))
        return frame, {"demodulate_ms": stats[0], "beamform_ms": stats[1],
                       "total_ms": stats[2]}

    # ---- Doppler and speckle tracking (doppler.cpp) ------------------------

    @staticmethod
    def _ensemble(grid: GridSpec, frames):
        return np.ascontiguousarray(np.asarray(frames, np.complex64).reshape(-1, grid.n0, grid.n1))

    def wall_filter(self, grid: GridSpec, frames, wall: int, order: int = 1):
        ens = self._ensemble(grid, frames)
        out = np.zeros_like(ens)
        self._check(self.lib.oc_wall_filter(C.byref(grid.c()), ens.shape[0], ens.ctypes.data_as(_fp),
                                            wall, order, out.ctypes.data_as(_fp)))
        return out

    def power_doppler(self, grid: GridSpec, frames, wall: int = 1, order: int = 1):
        ens = self._ensemble(grid, frames)
        out = np.zeros((grid.n0, grid.n1), np.float32)
        self._check(self.lib.oc_power_doppler(C.byref(grid.c()), ens.shape[0],
                                              ens.ctypes.data_as(_fp), wall, order, _ptr(out)))
        return out

    def color_flow(self, grid: GridSpec, frames, fdem: float, prf: float, c: float, d: Doppler):
        """Returns (velocity, power, variance, valid, nyquist_velocity)."""
        ens = self._ensemble(grid, frames)
        shape = (grid.n0, grid.n1)
        v, p, w = (np.zeros(shape, np.float32) for _ in range(3))
        ok = np.zeros(shape, np.uint8)
        nyq = C.c_float(0.0)
        self._check(self.lib.oc_color_flow(C.byref(grid.c()), ens.shape[0], ens.ctypes.data_as(_fp),
                                           _f(fdem), _f(prf), _f(c), C.byref(d.c()), _ptr(v),
                                           _ptr(p), _ptr(w), ok.ctypes.data_as(C.POINTER(C.c_uint8)),
                                           C.byref(nyq)))
        return v, p, w, ok, nyq.value

    def axial_displacement(self, grid: GridSpec, before, after, fdem: float, c: float,
                           kernel_axial: int = 8, kernel_lateral: int = 2, max_lag: int = 8):
        a = self._ensemble(grid, before)
        b = self._ensemble(grid, after)
        ax = np.zeros((grid.n0, grid.n1), np.float32)
        co = np.zeros_like(ax)
        self._check(self.lib.oc_axial_displacement(C.byref(grid.c()), a.ctypes.data_as(_fp),
                                                   b.ctypes.data_as(_fp), _f(fdem), _f(c),
                                                   kernel_axial, kernel_lateral, max_lag,
                                                   _ptr(ax), _ptr(co)))
        return ax, co

    def axial_strain(self, grid: GridSpec, axial, fit_length: int = 15):
        ax = np.ascontiguousarray(np.asarray(axial, np.float32).reshape(grid.n0, grid.n1))
        out = np.zeros_like(ax)
        self._check(self.lib.oc_axial_strain(C.byref(grid.c()), _ptr(ax), fit_length, _ptr(out)))
        return out


_cache: dict = {}


def available(flavour: str) -> bool:
    return os.path.exists(LIBS[flavour])


def load(flavour: str = "port") -> Oracle:
    if flavour not in _cache:
        path = LIBS[flavour]
        if not os.path.exists(path):
            target = "port" if flavour == "port" else "ref"
            raise OracleError(f"{path} is not built (run `make -C oracle {target}`)")
        _cache[flavour] = Oracle(path)
    return _cache[flavour]
