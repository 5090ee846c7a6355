"""OpenSayal per-step fluid solver for AMD Instinct MI355X (gfx950).

Python mirror of the reference's solver interface over the C ABI of
``libopensayal_amd.so`` (declared in include/opensayal.h):

* ``ConfigParser(path).parse()`` -> ``Config`` (src/config_parser.cpp:11-184)
* ``Fluid(config)`` (src/fluid.cu:41-71), ``Fluid.update(source, d_t)``
  (src/fluid.cu:770-795), ``min_pressure`` / ``max_pressure`` (inc/fluid.cuh:61-62)
* ``Source`` (inc/fluid.cuh:8-13)

Fields come back as numpy arrays of shape (H, W) in the reference's memory
order: ``a[r, i]`` is cell ``(i, j = H - 1 - r)`` (Fluid::indx,
src/fluid.cu:163-165).  Every compute call runs the HIP kernels; there is no
CPU fallback, and a missing library raises ``OpenSayalError``.
"""
from __future__ import annotations

import ctypes
import os
from typing import Optional

import numpy as np

from ._abi import (FIELDS, STORAGE, OpenSayalError, OsyBatchInfo, OsyBatchItemTime,
                   OsyBatchTuning, OsyConfig, OsyFluidInfo, OsySource, OsyTuning, check, lib)

__all__ = ["Batch", "Config", "ConfigParser", "Event", "Fluid", "Mouse", "Source", "Stream",
           "OpenSayalError", "default_config", "device_bytes", "FIELDS", "STORAGE"]

Config = OsyConfig
Source = OsySource


def _stream_ptr(stream) -> ctypes.c_void_p:
    """A ``Stream``, a raw hipStream_t value of the library's runtime, or
    None (the handle's own stream)."""
    if isinstance(stream, Stream):
        return ctypes.c_void_p(stream.handle)
    return ctypes.c_void_p(stream or None)


class Stream:
    """A HIP stream of the library's own runtime (non-blocking, default
    priority).  ``handle`` is what the ``stream`` arguments below take."""

    def __init__(self):
        h = ctypes.c_void_p()
        check(lib().osy_stream_create(ctypes.byref(h)))
        self.handle = h.value

    def synchronize(self) -> None:
        check(lib().osy_stream_synchronize(ctypes.c_void_p(self.handle)))

    def close(self) -> None:
        if self.handle:
            check(lib().osy_stream_destroy(ctypes.c_void_p(self.handle)))
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Event:
    """A HIP timing event of the library's runtime."""

    def __init__(self):
        h = ctypes.c_void_p()
        check(lib().osy_event_create(ctypes.byref(h)))
        self.handle = h.value

    def record(self, stream: "Stream | None" = None) -> None:
        s = stream.handle if stream is not None else None
        check(lib().osy_event_record(ctypes.c_void_p(self.handle), ctypes.c_void_p(s)))

    def synchronize(self) -> None:
        check(lib().osy_event_synchronize(ctypes.c_void_p(self.handle)))

    def elapsed_ms(self, end: "Event") -> float:
        ms = ctypes.c_float()
        check(lib().osy_event_elapsed_ms(ctypes.c_void_p(self.handle),
                                         ctypes.c_void_p(end.handle), ctypes.byref(ms)))
        return ms.value

    def close(self) -> None:
        if self.handle:
            check(lib().osy_event_destroy(ctypes.c_void_p(self.handle)))
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def default_config(width: int = 1920, height: int = 1080, **overrides) -> Config:
    """The reference's defaults (src/config_parser.cpp:21-181) for a W x H grid.

    Defaults that depend on the grid size (pipe and smoke heights, obstacle
    centre and radius) follow ``width``/``height``.  ``viscosity`` is 0.
    ``overrides`` are ``Config`` field names.
    """
    text = '{"sim": {"width": %d, "height": %d}, "fluid": {"viscosity": 0}}' % (width, height)
    c = Config()
    check(lib().osy_config_parse_string(text.encode(), ctypes.byref(c)))
    for k, v in overrides.items():
        if k == "storage" and isinstance(v, str):
            v = STORAGE[v]
        if not hasattr(c, k):
            raise AttributeError(f"Config has no field {k!r}")
        setattr(c, k, v)
    return c


class ConfigParser:
    """ConfigParser (src/config_parser.cpp:11-13): reads OpenSayal.conf.json."""

    def __init__(self, config_file_name: str = "OpenSayal.conf.json"):
        self.config_file_name = config_file_name

    def parse(self) -> Config:
        c = Config()
        check(lib().osy_config_parse_file(os.fsencode(self.config_file_name), ctypes.byref(c)))
        return c

    @staticmethod
    def parse_string(text: str) -> Config:
        c = Config()
        check(lib().osy_config_parse_string(text.encode(), ctypes.byref(c)))
        return c


class Fluid:
    """The solver state of one grid on the current HIP device (class Fluid,
    inc/fluid.cuh:15-116).  Owns its device memory until ``close()``."""

    def __init__(self, config: Config):
        self._h = ctypes.c_void_p()
        check(lib().osy_fluid_create(ctypes.byref(config), ctypes.byref(self._h)))
        self.config = config.copy()
        self._refresh_info()

    def _refresh_info(self) -> None:
        info = OsyFluidInfo()
        check(lib().osy_fluid_get_info(self._h, ctypes.byref(info)))
        self.width, self.height = info.width, info.height
        self.pitch = info.pitch
        self.device_bytes = info.device_bytes
        self.projection_kernel = info.projection_kernel
        self.sweeps_per_launch = info.sweeps_per_launch

    @property
    def tuning(self) -> dict:
        t = OsyTuning()
        check(lib().osy_fluid_get_tuning(self._h, ctypes.byref(t)))
        return {name: getattr(t, name) for name, _ in t._fields_}

    def set_tuning(self, projection_kernel: int = 1, tile_width: int = 240,
                   tile_height: int = 32, sweeps_per_launch: int = 4) -> None:
        """Projection launch shape; results are bit-identical for every shape."""
        t = OsyTuning(projection_kernel, tile_width, tile_height, sweeps_per_launch)
        check(lib().osy_fluid_set_tuning(self._h, ctypes.byref(t)))
        self._refresh_info()

    def set_wait_limit(self, microseconds: int) -> None:
        """Bound of one neighbour wait of the resident projection (default
        2 s); a step whose wait exceeds it raises OSY_E_STATE."""
        check(lib().osy_fluid_set_wait_limit(self._h, int(microseconds)))

    def autotune(self) -> tuple:
        """Time every projection launch shape on this grid and keep the
        fastest; returns (tuning dict, ms per projection)."""
        t = OsyTuning()
        ms = ctypes.c_double()
        check(lib().osy_fluid_autotune(self._h, ctypes.byref(t), ctypes.byref(ms)))
        self._refresh_info()
        return {name: getattr(t, name) for name, _ in t._fields_}, ms.value

    def autotune_cached(self) -> bool:
        """Adopt a cached autotuner choice for this grid shape, if any."""
        st = lib().osy_fluid_autotune_cached(self._h)
        if st < 0:
            check(st)
        self._refresh_info()
        return st == 1

    @property
    def projection_kernel_name(self) -> str:
        return {0: "projection_colour", 1: "projection_lds_tiled",
                2: "projection_register_tiled",
                3: "projection_register_resident"}.get(self.projection_kernel, "?")

    def launches_per_call(self, sweeps: int) -> int:
        """Projection kernel launches one ``project(sweeps)`` call makes."""
        if self.projection_kernel == 0:
            return 2 * sweeps
        if self.projection_kernel == 3:
            return 1 if sweeps > 0 else 0
        k = self.sweeps_per_launch
        return -(-sweeps // k)

    @classmethod
    def from_json(cls, path: str) -> "Fluid":
        return cls(ConfigParser(path).parse())

    # -- lifetime ----------------------------------------------------------
    def close(self) -> None:
        if self._h:
            check(lib().osy_fluid_destroy(self._h))
            self._h = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- stepping ----------------------------------------------------------
    def update(self, source: Optional[Source] = None, d_t: Optional[float] = None) -> None:
        """Fluid::update(Source, d_t): one synchronous step."""
        dt = self.config.d_t if d_t is None else d_t
        src = ctypes.byref(source) if source is not None else None
        check(lib().osy_fluid_update(self._h, src, ctypes.c_float(dt)))

    def step(self, steps: int, d_t: Optional[float] = None) -> None:
        """``steps`` updates with an inactive source, replayed from a HIP graph."""
        dt = self.config.d_t if d_t is None else d_t
        check(lib().osy_fluid_step(self._h, int(steps), ctypes.c_float(dt)))

    def step_async(self, steps: int, d_t: Optional[float] = None, stream=None) -> None:
        dt = self.config.d_t if d_t is None else d_t
        check(lib().osy_fluid_step_async(self._h, int(steps), ctypes.c_float(dt),
                                         _stream_ptr(stream)))

    def synchronize(self) -> None:
        check(lib().osy_fluid_synchronize(self._h))

    def project(self, sweeps: Optional[int] = None, stream=None, sync: bool = True) -> None:
        """Fluid::apply_projection alone (``sweeps`` defaults to projection n)."""
        n = self.config.projection_n if sweeps is None else sweeps
        check(lib().osy_fluid_project(self._h, int(n), _stream_ptr(stream)))
        if sync:
            self.synchronize()

    # -- state access ------------------------------------------------------
    def get_field(self, name: str) -> np.ndarray:
        fid = FIELDS[name]
        dtype = np.int32 if name in ("is_solid", "total_s") else np.float32
        out = np.empty((self.height, self.width), dtype=dtype)
        check(lib().osy_fluid_get_field(self._h, fid, out.ctypes.data_as(ctypes.c_void_p),
                                        out.nbytes))
        return out

    def set_field(self, name: str, values: np.ndarray) -> None:
        a = np.ascontiguousarray(values, dtype=np.float32)
        if a.shape != (self.height, self.width):
            raise ValueError(f"expected shape {(self.height, self.width)}, got {a.shape}")
        check(lib().osy_fluid_set_field(self._h, FIELDS[name], a.ctypes.data_as(ctypes.c_void_p),
                                        a.nbytes))

    @property
    def pressure_range(self) -> tuple:
        lo, hi = ctypes.c_float(), ctypes.c_float()
        check(lib().osy_fluid_pressure_range(self._h, ctypes.byref(lo), ctypes.byref(hi)))
        return lo.value, hi.value

    @property
    def min_pressure(self) -> float:
        return self.pressure_range[0]

    @property
    def max_pressure(self) -> float:
        return self.pressure_range[1]

    def divergence_norm(self) -> tuple:
        """(L2, max-abs) of the divergence over interior fluid cells, in FP64."""
        l2, mx = ctypes.c_double(), ctypes.c_double()
        check(lib().osy_fluid_divergence_norm(self._h, ctypes.byref(l2), ctypes.byref(mx)))
        return l2.value, mx.value

    # -- what GraphicsHandler reads each frame ------------------------------
    def render_rgba(self) -> np.ndarray:
        """The frame of update_fluid_pixels_kernel
        (src/graphics_handler.cu:258-285): (H, W) uint32 RGBA8888 pixels
        (r << 24 | g << 16 | b << 8 | a), row 0 the top of the screen."""
        out = np.empty((self.height, self.width), dtype=np.uint32)
        check(lib().osy_fluid_render_rgba(self._h, out.ctypes.data_as(ctypes.c_void_p),
                                          out.nbytes))
        return out

    def traces(self, distance: int = 20, length: int = 20,
               d_t: Optional[float] = None) -> tuple:
        """Path lines of update_traces_kernel (src/graphics_handler.cu:384-396)
        for seeds every ``distance`` cells (visual.path_line.distance / length
        default to 20): (traces_x, traces_y), each an int32 array of shape
        (H // distance, W // distance, length) in indx_traces order (row 0 =
        the top row of seeds); a solid seed's entries are -1."""
        if distance < 1:
            raise ValueError("distance must be at least 1")
        dt = self.config.d_t if d_t is None else d_t
        shape = (self.height // distance, self.width // distance, int(length))
        tx = np.empty(shape, dtype=np.int32)
        ty = np.empty(shape, dtype=np.int32)
        check(lib().osy_fluid_traces(self._h, int(distance), int(length), float(dt),
                                     tx.ctypes.data_as(ctypes.c_void_p),
                                     ty.ctypes.data_as(ctypes.c_void_p), tx.nbytes))
        return tx, ty


INTERACTIVE_STARTING_SPEED = 200  # inc/config.hpp:10


def _trunc_div(a: int, b: int) -> int:
    """C integer division (truncates toward zero)."""
    q = abs(a) // abs(b)
    return q if (a >= 0) == (b >= 0) else -q


class Mouse:
    """Mouse (src/mouse.cu:4-45) without SDL: a window's event loop feeds it
    button, wheel and motion events, and each frame asks for the Source to
    pass to Fluid.update (src/main.cu:85-88)."""

    def __init__(self):
        self.is_down = False
        self.button = 0
        self.wheel_value = INTERACTIVE_STARTING_SPEED
        self.x = 0
        self.y = 0

    def button_down(self, button: int) -> None:  # SDL_MOUSEBUTTONDOWN
        self.is_down = True
        self.button = int(button)

    def button_up(self) -> None:  # SDL_MOUSEBUTTONUP
        self.is_down = False

    def wheel(self, y: int) -> None:  # SDL_MOUSEWHEEL: 10 per notch, never below 0
        value = self.wheel_value + int(y) * 10
        if value >= 0:
            self.wheel_value = value

    def motion(self, x: int, y: int) -> None:  # SDL_MOUSEMOTION, window pixels
        self.x, self.y = int(x), int(y)

    def make_source(self, fluid_height: int, cell_size: int) -> Source:
        """Mouse::make_source: the cell under the pointer (window y grows
        downwards, fluid j upwards), smoke with the left button, the wheel's
        speed, pointing backwards with the middle button."""
        src = Source()
        src.active = int(self.is_down)
        src.smoke = 1.0 if self.button == 1 else 0.0
        src.velocity = float(-self.wheel_value if self.button == 2 else self.wheel_value)
        src.position_x = _trunc_div(self.x, cell_size)
        src.position_y = fluid_height - _trunc_div(self.y, cell_size)
        return src


def device_bytes(config: Config) -> int:
    """Device bytes a Fluid of this configuration allocates."""
    n = ctypes.c_uint64()
    check(lib().osy_fluid_device_bytes(ctypes.byref(config), ctypes.byref(n)))
    return n.value


class Batch:
    """Many Fluid handles stepped together by one persistent kernel launch
    per step (per storage/pressure group), replayed from a HIP graph.  Each
    member's fields after k batch steps are bit-identical to k ``step``
    calls on that Fluid alone.  ``budget_bytes`` bounds the members' device
    memory (0: none).  Members are stepped with their own config's d_t."""

    def __init__(self, budget_bytes: int = 0):
        self._h = ctypes.c_void_p()
        check(lib().osy_batch_create(ctypes.c_uint64(budget_bytes), ctypes.byref(self._h)))
        self.members = []

    def add(self, fluid: Fluid) -> None:
        check(lib().osy_batch_add(self._h, fluid._h))
        self.members.append(fluid)

    def remove(self, fluid: Fluid) -> None:
        check(lib().osy_batch_remove(self._h, fluid._h))
        self.members.remove(fluid)

    def step(self, steps: int = 1) -> None:
        check(lib().osy_batch_step(self._h, steps))

    def step_async(self, steps: int = 1, stream=None) -> None:
        check(lib().osy_batch_step_async(self._h, steps, _stream_ptr(stream)))

    def synchronize(self) -> None:
        check(lib().osy_batch_synchronize(self._h))

    @property
    def tuning(self) -> dict:
        t = OsyBatchTuning()
        check(lib().osy_batch_get_tuning(self._h, ctypes.byref(t)))
        return {"rows": t.rows, "sweeps_per_tile": t.sweeps_per_tile, "occupancy": t.occupancy}

    def set_tuning(self, rows: int = 12, sweeps_per_tile: int = 10, occupancy: int = 2) -> None:
        t = OsyBatchTuning(rows, sweeps_per_tile, occupancy)
        check(lib().osy_batch_set_tuning(self._h, ctypes.byref(t)))

    def autotune(self) -> tuple:
        """Time every supported shape on the members (state restored after);
        keeps the fastest; returns (tuning dict, ms per batch step)."""
        t = OsyBatchTuning()
        ms = ctypes.c_double()
        check(lib().osy_batch_autotune(self._h, ctypes.byref(t), ctypes.byref(ms)))
        return {"rows": t.rows, "sweeps_per_tile": t.sweeps_per_tile,
                "occupancy": t.occupancy}, ms.value

    def set_per_grid_cells(self, cells: int) -> None:
        """Members of at least `cells` cells step on their own per-grid
        graphs ahead of the persistent launch (default 2^20; 0 = none)."""
        check(lib().osy_batch_set_per_grid_cells(self._h, int(cells)))

    def set_timing(self, enable: bool = True) -> None:
        """Stamp every work item of every launch (claim, dependencies
        complete, done; results unchanged)."""
        check(lib().osy_batch_set_timing(self._h, 1 if enable else 0))

    def item_times(self) -> dict:
        """The last step's work items as arrays: member, kind (0 forces, 1
        projection, 2 velocity advection, 3 smoke advection), stage, tile,
        hw_id, and claim / ready / done in nanoseconds from the step's first
        claim."""
        n = ctypes.c_int64()
        tick = ctypes.c_double()
        lib().osy_batch_item_times(self._h, None, 0, ctypes.byref(n), ctypes.byref(tick))
        arr = (OsyBatchItemTime * max(1, n.value))()
        check(lib().osy_batch_item_times(self._h, arr, n.value, ctypes.byref(n),
                                         ctypes.byref(tick)))
        rec = np.ctypeslib.as_array(arr)[: n.value]
        out = {f: rec[f].copy() for f in ("member", "kind", "stage", "tile", "hw_id")}
        t0 = int(rec["claim"].min()) if n.value else 0
        for f in ("claim", "ready", "done"):
            out[f] = (rec[f].astype(np.int64) - t0) * tick.value
        out["tick_ns"] = tick.value
        return out

    @property
    def info(self) -> dict:
        i = OsyBatchInfo()
        check(lib().osy_batch_get_info(self._h, ctypes.byref(i)))
        return {name: getattr(i, name) for name, _ in i._fields_}

    def close(self) -> None:
        if self._h:
            check(lib().osy_batch_destroy(self._h))
            self._h = ctypes.c_void_p()
            self.members = []

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
