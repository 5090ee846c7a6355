"""ctypes wrapper of the CPU oracle (fluid_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, by ``__graft_entry__.smoke()``
as the checker, and by bench.py's CPU-baseline leg.  The product library
never calls it.

Arrays are numpy arrays of shape (H, W) in the reference's memory order
(row r holds j = H - 1 - r), the same order the solver's C ABI returns.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "libfluid_oracle.so"

_lib = None


def build() -> Path:
    """Builds libfluid_oracle.so with the Makefile beside this file."""
    subprocess.run(["make", "-s", "-C", os.fspath(HERE)], check=True)
    return LIB_PATH


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        _lib = ctypes.CDLL(os.fspath(LIB_PATH))
        assert _lib.orc_abi_marker() == 0x05A7A1
        _declare(_lib)
    return _lib


_P = ctypes.c_void_p


def _declare(L):
    L.orc_build_masks.restype = ctypes.c_int
    L.orc_build_masks.argtypes = [_P, _P, _P]
    for sfx, real in (("_f32", ctypes.c_float), ("_f64", ctypes.c_double)):
        f = getattr(L, "orc_step" + sfx)
        f.restype = None
        f.argtypes = [_P, _P, _P, _P, _P, _P, _P, _P, _P, _P, real, ctypes.c_int, ctypes.c_int]
        f = getattr(L, "orc_phase_forces" + sfx)
        f.restype = None
        f.argtypes = [_P, _P, _P, _P, _P, _P, _P, _P, real]
        f = getattr(L, "orc_phase_projection" + sfx)
        f.restype = None
        f.argtypes = [_P, _P, _P, _P, _P, _P, real, ctypes.c_int, ctypes.c_int]
        f = getattr(L, "orc_phase_colour" + sfx)
        f.restype = None
        f.argtypes = [_P, _P, _P, _P, _P, _P, real, ctypes.c_int]
        f = getattr(L, "orc_phase_extrapolation" + sfx)
        f.restype = None
        f.argtypes = [_P, _P, _P, _P, _P]
        f = getattr(L, "orc_phase_advect_velocity" + sfx)
        f.restype = None
        f.argtypes = [_P, _P, _P, _P, _P, _P, _P, real]
        f = getattr(L, "orc_phase_advect_smoke" + sfx)
        f.restype = None
        f.argtypes = [_P, _P, _P, _P, _P, _P, _P, real]
        f = getattr(L, "orc_phase_decay" + sfx)
        f.restype = None
        f.argtypes = [_P, _P, _P, _P, real, real]
        for name in ("orc_point_gx", "orc_point_gy", "orc_point_idw"):
            f = getattr(L, name + sfx)
            f.restype = real
            f.argtypes = [_P, _P, _P, _P, real, real]
        f = getattr(L, "orc_divergence" + sfx)
        f.restype = None
        f.argtypes = [_P, _P, _P, _P, _P, _P]


    for sfx, real in (("_f32", ctypes.c_float), ("_f64", ctypes.c_double)):
        f = getattr(L, "orc_phase_diffusion" + sfx)
        f.restype = None
        f.argtypes = [_P, _P, _P, _P, _P, real, ctypes.c_int]
    L.orc_round16_array.restype = None
    L.orc_round16_array.argtypes = [_P, ctypes.c_long, ctypes.c_int]
    L.orc_pixel.restype = ctypes.c_uint32
    L.orc_pixel.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_float,
                            ctypes.c_float, ctypes.c_float, ctypes.c_float]
    L.orc_render_rgba.restype = None
    L.orc_render_rgba.argtypes = [_P, _P, _P, _P, ctypes.c_float, ctypes.c_float, _P]
    L.orc_trace_grid.restype = ctypes.c_int
    L.orc_trace_grid.argtypes = [_P, _P, _P, _P, _P, ctypes.c_int, ctypes.c_int, ctypes.c_float,
                                 _P, _P]


def _ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def round16(a: np.ndarray, storage: int) -> np.ndarray:
    """float32 values rounded to binary16 (storage 1) or bfloat16 (2), ties
    to even, widened back; the oracle's rounding for 16-bit storage."""
    out = np.ascontiguousarray(a, dtype=np.float32).copy()
    lib().orc_round16_array(_ptr(out), out.size, storage)
    return out


def pixel(solid: bool, pressure_on: bool, smoke_on: bool, smoke: float, p: float,
          pmin: float, pmax: float) -> int:
    """One RGBA8888 pixel of update_fluid_pixels_kernel
    (src/graphics_handler.cu:258-285)."""
    return int(lib().orc_pixel(int(solid), int(pressure_on), int(smoke_on), smoke, p, pmin, pmax))


def render_rgba(cfg, solid, smoke, p, pmin: float, pmax: float) -> np.ndarray:
    """The frame of update_fluid_pixels_kernel as (H, W) uint32 RGBA8888
    (r << 24 | g << 16 | b << 8 | a), row 0 the top of the screen, from
    float32 smoke and pressure in the reference's memory order."""
    H, W = cfg.height, cfg.width
    solid = np.ascontiguousarray(solid, np.int32)
    smoke = np.ascontiguousarray(smoke, np.float32)
    p = np.ascontiguousarray(p, np.float32)
    assert solid.shape == smoke.shape == p.shape == (H, W)
    out = np.zeros((H, W), np.uint32)
    lib().orc_render_rgba(ctypes.byref(cfg), _ptr(solid), _ptr(smoke), _ptr(p), pmin, pmax,
                          _ptr(out))
    return out


def trace_grid(cfg, solid, total_s, u, v, distance: int, length: int, d_t: float):
    """Path lines of update_traces_kernel (src/graphics_handler.cu:384-396):
    (traces_x, traces_y), each (H // distance, W // distance, length) int32
    in indx_traces order (row 0 holds the top row of seeds)."""
    if distance < 1 or length < 1:
        raise ValueError("distance and length must be at least 1")
    th, tw = cfg.height // distance, cfg.width // distance
    tx = np.zeros((th, tw, length), np.int32)
    ty = np.zeros((th, tw, length), np.int32)
    u = np.ascontiguousarray(u, np.float32)
    v = np.ascontiguousarray(v, np.float32)
    rc = lib().orc_trace_grid(ctypes.byref(cfg), _ptr(np.ascontiguousarray(solid, np.int32)),
                              _ptr(np.ascontiguousarray(total_s, np.int32)), _ptr(u), _ptr(v),
                              int(distance), int(length), float(d_t), _ptr(tx), _ptr(ty))
    if rc != 0:
        raise ValueError("distance and length must be at least 1")
    return tx, ty


def build_masks(cfg):
    """(is_solid, total_s) as (H, W) int32 arrays (src/fluid.cu:111-142)."""
    H, W = cfg.height, cfg.width
    solid = np.zeros((H, W), np.int32)
    ts = np.zeros((H, W), np.int32)
    rc = lib().orc_build_masks(ctypes.byref(cfg), _ptr(solid), _ptr(ts))
    if rc != 0:
        raise ValueError("grid too small")
    return solid, ts


class OracleFluid:
    """CPU restatement of one Fluid: state arrays plus the update rule.

    ``real`` is np.float32 (the reference arithmetic) or np.float64 (the
    FP64 run)."""

    def __init__(self, cfg, real=np.float32, nthreads: int = 1):
        self.cfg = cfg.copy() if hasattr(cfg, "copy") else cfg
        self.real = np.dtype(real)
        self.sfx = "_f32" if self.real == np.float32 else "_f64"
        self.ctype = ctypes.c_float if self.real == np.float32 else ctypes.c_double
        self.nthreads = nthreads
        H, W = cfg.height, cfg.width
        self.solid, self.total_s = build_masks(cfg)
        self.u = np.zeros((H, W), self.real)
        self.v = np.zeros((H, W), self.real)
        self.smoke = np.zeros((H, W), self.real)
        self.p = np.zeros((H, W), self.real)
        self._scratch = np.zeros((3, H, W), self.real)
        self._prange = np.zeros(2, self.real)

    def _fn(self, name):
        return getattr(lib(), name + self.sfx)

    def _c(self):
        return ctypes.byref(self.cfg)

    def set_state(self, u=None, v=None, smoke=None, p=None):
        for name, val in (("u", u), ("v", v), ("smoke", smoke), ("p", p)):
            if val is not None:
                getattr(self, name)[...] = np.asarray(val, dtype=self.real)

    @property
    def pressure_range(self):
        return float(self._prange[0]), float(self._prange[1])

    def update(self, source=None, d_t=None, steps: int = 1):
        dt = self.cfg.d_t if d_t is None else d_t
        src = ctypes.byref(source) if source is not None else None
        self._fn("orc_step")(self._c(), _ptr(self.solid), _ptr(self.total_s), _ptr(self.u),
                             _ptr(self.v), _ptr(self.smoke), _ptr(self.p), _ptr(self._scratch),
                             _ptr(self._prange), src, self.ctype(dt), int(steps),
                             int(self.nthreads))

    # -- single phases, for unit tests --------------------------------------
    def forces(self, source=None, d_t=None):
        dt = self.cfg.d_t if d_t is None else d_t
        src = ctypes.byref(source) if source is not None else None
        self._fn("orc_phase_forces")(self._c(), _ptr(self.solid), _ptr(self.total_s),
                                     _ptr(self.u), _ptr(self.v), _ptr(self.smoke), _ptr(self.p),
                                     src, self.ctype(dt))

    def projection(self, sweeps=None, d_t=None, nthreads=None):
        dt = self.cfg.d_t if d_t is None else d_t
        n = self.cfg.projection_n if sweeps is None else sweeps
        self._fn("orc_phase_projection")(self._c(), _ptr(self.solid), _ptr(self.total_s),
                                         _ptr(self.u), _ptr(self.v), _ptr(self.p),
                                         self.ctype(dt), int(n),
                                         int(nthreads or self.nthreads))

    def colour_phase(self, colour, d_t=None):
        dt = self.cfg.d_t if d_t is None else d_t
        self._fn("orc_phase_colour")(self._c(), _ptr(self.solid), _ptr(self.total_s),
                                     _ptr(self.u), _ptr(self.v), _ptr(self.p), self.ctype(dt),
                                     int(colour))

    def extrapolation(self):
        self._fn("orc_phase_extrapolation")(self._c(), _ptr(self.solid), _ptr(self.total_s),
                                            _ptr(self.u), _ptr(self.v))

    def advect_velocity(self, d_t=None):
        dt = self.cfg.d_t if d_t is None else d_t
        u2 = np.empty_like(self.u)
        v2 = np.empty_like(self.v)
        self._fn("orc_phase_advect_velocity")(self._c(), _ptr(self.solid), _ptr(self.total_s),
                                              _ptr(self.u), _ptr(self.v), _ptr(u2), _ptr(v2),
                                              self.ctype(dt))
        self.u, self.v = u2, v2

    def advect_smoke(self, d_t=None):
        dt = self.cfg.d_t if d_t is None else d_t
        s2 = np.empty_like(self.smoke)
        self._fn("orc_phase_advect_smoke")(self._c(), _ptr(self.solid), _ptr(self.total_s),
                                           _ptr(self.u), _ptr(self.v), _ptr(self.smoke),
                                           _ptr(s2), self.ctype(dt))
        self.smoke = s2

    def decay(self, rate=None, d_t=None):
        dt = self.cfg.d_t if d_t is None else d_t
        r = self.cfg.decay_rate if rate is None else rate
        self._fn("orc_phase_decay")(self._c(), _ptr(self.solid), _ptr(self.total_s),
                                    _ptr(self.smoke), self.ctype(r), self.ctype(dt))

    def diffusion(self, sweeps=None, d_t=None):
        """The diffusion pass alone (Jacobi rounds on u)."""
        dt = self.cfg.d_t if d_t is None else d_t
        n = self.cfg.projection_n if sweeps is None else sweeps
        self._fn("orc_phase_diffusion")(self._c(), _ptr(self.solid), _ptr(self.total_s),
                                        _ptr(self.u), _ptr(self._scratch), self.ctype(dt),
                                        int(n))

    def gx(self, x, y):
        return self._fn("orc_point_gx")(self._c(), _ptr(self.solid), _ptr(self.total_s),
                                        _ptr(self.u), self.ctype(x), self.ctype(y))

    def gy(self, x, y):
        return self._fn("orc_point_gy")(self._c(), _ptr(self.solid), _ptr(self.total_s),
                                        _ptr(self.v), self.ctype(x), self.ctype(y))

    def idw(self, x, y):
        return self._fn("orc_point_idw")(self._c(), _ptr(self.solid), _ptr(self.total_s),
                                         _ptr(self.smoke), self.ctype(x), self.ctype(y))

    def render_rgba(self):
        """The frame for the current state, with the pressure range of the
        last update (as the reference passes fluid.min_pressure / max_pressure)."""
        pmin, pmax = self.pressure_range
        return render_rgba(self.cfg, self.solid, self.smoke, self.p, pmin, pmax)

    def traces(self, distance: int, length: int, d_t=None):
        dt = self.cfg.d_t if d_t is None else d_t
        return trace_grid(self.cfg, self.solid, self.total_s, self.u, self.v, distance, length,
                          dt)

    def divergence_norm(self):
        out = np.zeros(2, np.float64)
        self._fn("orc_divergence")(self._c(), _ptr(self.solid), _ptr(self.total_s),
                                   _ptr(self.u), _ptr(self.v), _ptr(out))
        return float(out[0]), float(out[1])
