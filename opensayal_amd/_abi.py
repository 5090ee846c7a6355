"""ctypes declarations of the C ABI in include/opensayal.h.

The library is loaded lazily and from this package's directory only, so the
HIP solver that runs is always the in-tree build.  Loading fails loudly when
it is missing: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import os
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libopensayal_amd.so"
# the headless program (cli.cpp), linked against LIB_PATH by build.py
CLI_PATH = Path(__file__).resolve().parent / "fluid-simulator"

OSY_OK = 0
OSY_E_INVALID = -1
OSY_E_UNSUPPORTED = -2
OSY_E_HIP = -3
OSY_E_NOMEM = -4
OSY_E_CONFIG = -5
OSY_E_BUDGET = -6
OSY_E_STATE = -7

STATUS_NAMES = {
    OSY_OK: "OSY_OK",
    OSY_E_INVALID: "OSY_E_INVALID",
    OSY_E_UNSUPPORTED: "OSY_E_UNSUPPORTED",
    OSY_E_HIP: "OSY_E_HIP",
    OSY_E_NOMEM: "OSY_E_NOMEM",
    OSY_E_CONFIG: "OSY_E_CONFIG",
    OSY_E_BUDGET: "OSY_E_BUDGET",
    OSY_E_STATE: "OSY_E_STATE",
}

STORAGE = {"f32": 0, "f16": 1, "bf16": 2}
FIELDS = {"vel_x": 0, "vel_y": 1, "smoke": 2, "pressure": 3, "is_solid": 4, "total_s": 5}

_i32 = ctypes.c_int32
_f32 = ctypes.c_float


class OsyConfig(ctypes.Structure):
    """struct osy_config; field names follow OpenSayal.conf.json keys."""

    _fields_ = [
        ("width", _i32), ("height", _i32), ("cell_size", _f32), ("cell_pixel_size", _i32),
        ("enable_drain", _i32), ("enable_pressure", _i32), ("enable_smoke", _i32),
        ("enable_interactive", _i32), ("projection_n", _i32), ("projection_o", _f32),
        ("pipe_height", _i32), ("pipe_length", _i32), ("smoke_length", _i32),
        ("smoke_height", _i32), ("smoke_count", _i32), ("wind_speed", _f32),
        ("wind_smoke", _f32), ("g", _f32), ("d_t", _f32), ("enable_real_time", _i32),
        ("real_time_multiplier", _f32), ("enable_decay", _i32), ("decay_rate", _f32),
        ("obstacle_enable", _i32), ("obstacle_center_x", _i32), ("obstacle_center_y", _i32),
        ("obstacle_radius", _f32), ("density", _f32), ("drag_coeff", _f32),
        ("viscosity", _f32), ("block_size_x", _i32), ("block_size_y", _i32),
        ("omp_thread_count", _i32), ("storage", _i32),
    ]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}

    def copy(self) -> "OsyConfig":
        c = OsyConfig()
        ctypes.memmove(ctypes.byref(c), ctypes.byref(self), ctypes.sizeof(OsyConfig))
        return c


class OsySource(ctypes.Structure):
    """struct osy_source: the reference's Source (inc/fluid.cuh:8-13)."""

    _fields_ = [("active", _i32), ("smoke", _f32), ("velocity", _f32),
                ("position_x", _i32), ("position_y", _i32)]


class OsyTuning(ctypes.Structure):
    """struct osy_tuning: projection launch shape."""

    _fields_ = [("projection_kernel", _i32), ("tile_width", _i32), ("tile_height", _i32),
                ("sweeps_per_launch", _i32)]


class OsyFluidInfo(ctypes.Structure):
    _fields_ = [("width", _i32), ("height", _i32), ("cell_size", _i32),
                ("projection_n", _i32), ("pitch", _i32), ("storage", _i32),
                ("projection_kernel", _i32), ("sweeps_per_launch", _i32),
                ("device_bytes", ctypes.c_size_t)]


class OsyBatchTuning(ctypes.Structure):
    """struct osy_batch_tuning: register-tile rows per wavefront, sweeps per tile."""

    _fields_ = [("rows", _i32), ("sweeps_per_tile", _i32), ("occupancy", _i32)]


class OsyBatchInfo(ctypes.Structure):
    _fields_ = [("count", _i32), ("groups", _i32), ("work_items", ctypes.c_int64),
                ("cells", ctypes.c_int64), ("budget_bytes", ctypes.c_uint64),
                ("used_bytes", ctypes.c_uint64)]


class OsyBatchItemTime(ctypes.Structure):
    """struct osy_batch_item_time: one work item of a batch's last step."""

    _fields_ = [("member", _i32), ("kind", _i32), ("stage", _i32), ("tile", _i32),
                ("claim", ctypes.c_uint64), ("ready", ctypes.c_uint64),
                ("done", ctypes.c_uint64), ("hw_id", ctypes.c_uint64)]


# name -> (restype, argtypes); every function of include/opensayal.h.
PROTOTYPES = {
    "osy_abi_version": (ctypes.c_int, []),
    "osy_last_error": (ctypes.c_char_p, []),
    "osy_config_defaults": (ctypes.c_int, [ctypes.POINTER(OsyConfig)]),
    "osy_config_parse_file": (ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(OsyConfig)]),
    "osy_config_parse_string": (ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(OsyConfig)]),
    "osy_fluid_create": (ctypes.c_int, [ctypes.POINTER(OsyConfig), ctypes.POINTER(ctypes.c_void_p)]),
    "osy_fluid_create_from_json": (ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(ctypes.c_void_p)]),
    "osy_fluid_destroy": (ctypes.c_int, [ctypes.c_void_p]),
    "osy_fluid_update": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(OsySource), _f32]),
    "osy_fluid_step": (ctypes.c_int, [ctypes.c_void_p, _i32, _f32]),
    "osy_fluid_step_async": (ctypes.c_int, [ctypes.c_void_p, _i32, _f32, ctypes.c_void_p]),
    "osy_fluid_synchronize": (ctypes.c_int, [ctypes.c_void_p]),
    "osy_fluid_project": (ctypes.c_int, [ctypes.c_void_p, _i32, ctypes.c_void_p]),
    "osy_fluid_get_field": (ctypes.c_int, [ctypes.c_void_p, _i32, ctypes.c_void_p, ctypes.c_size_t]),
    "osy_fluid_set_field": (ctypes.c_int, [ctypes.c_void_p, _i32, ctypes.c_void_p, ctypes.c_size_t]),
    "osy_fluid_pressure_range": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(_f32), ctypes.POINTER(_f32)]),
    "osy_fluid_divergence_norm": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]),
    "osy_fluid_get_info": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(OsyFluidInfo)]),
    "osy_fluid_render_rgba": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t]),
    "osy_fluid_traces": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32,
                                        ctypes.c_float, ctypes.c_void_p, ctypes.c_void_p,
                                        ctypes.c_size_t]),
    "osy_fluid_traces_noalloc": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32,
                                                ctypes.c_int32, ctypes.c_float, ctypes.c_void_p,
                                                ctypes.c_void_p, ctypes.c_size_t]),
    "osy_fluid_set_tuning": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(OsyTuning)]),
    "osy_fluid_get_tuning": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(OsyTuning)]),
    "osy_fluid_set_wait_limit": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int64]),
    "osy_fluid_autotune": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(OsyTuning),
                                          ctypes.POINTER(ctypes.c_double)]),
    "osy_fluid_autotune_cached": (ctypes.c_int, [ctypes.c_void_p]),
    "osy_fluid_device_bytes": (ctypes.c_int, [ctypes.POINTER(OsyConfig),
                                              ctypes.POINTER(ctypes.c_uint64)]),
    "osy_batch_create": (ctypes.c_int, [ctypes.c_uint64, ctypes.POINTER(ctypes.c_void_p)]),
    "osy_batch_destroy": (ctypes.c_int, [ctypes.c_void_p]),
    "osy_batch_add": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    "osy_batch_remove": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    "osy_batch_step": (ctypes.c_int, [ctypes.c_void_p, _i32]),
    "osy_batch_step_async": (ctypes.c_int, [ctypes.c_void_p, _i32, ctypes.c_void_p]),
    "osy_batch_synchronize": (ctypes.c_int, [ctypes.c_void_p]),
    "osy_batch_set_tuning": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(OsyBatchTuning)]),
    "osy_batch_get_tuning": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(OsyBatchTuning)]),
    "osy_batch_autotune": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(OsyBatchTuning),
                                          ctypes.POINTER(ctypes.c_double)]),
    "osy_batch_get_info": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(OsyBatchInfo)]),
    "osy_batch_set_timing": (ctypes.c_int, [ctypes.c_void_p, _i32]),
    "osy_batch_set_per_grid_cells": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int64]),
    "osy_arith_check": (ctypes.c_int, [_i32, ctypes.POINTER(ctypes.c_uint64),
                                       ctypes.POINTER(ctypes.c_uint32)]),
    "osy_batch_item_times": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(OsyBatchItemTime),
                                            ctypes.c_int64, ctypes.POINTER(ctypes.c_int64),
                                            ctypes.POINTER(ctypes.c_double)]),
    "osy_stream_create": (ctypes.c_int, [ctypes.POINTER(ctypes.c_void_p)]),
    "osy_stream_destroy": (ctypes.c_int, [ctypes.c_void_p]),
    "osy_stream_synchronize": (ctypes.c_int, [ctypes.c_void_p]),
    "osy_event_create": (ctypes.c_int, [ctypes.POINTER(ctypes.c_void_p)]),
    "osy_event_destroy": (ctypes.c_int, [ctypes.c_void_p]),
    "osy_event_record": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    "osy_event_synchronize": (ctypes.c_int, [ctypes.c_void_p]),
    "osy_event_elapsed_ms": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p,
                                            ctypes.POINTER(_f32)]),
    "osy_calibration_copy": (ctypes.c_int, [ctypes.c_uint64, _i32]),
}

_lib = None


class OpenSayalError(RuntimeError):
    """A C ABI call returned an error status."""

    def __init__(self, status: int, message: str):
        super().__init__(f"{STATUS_NAMES.get(status, status)}: {message}")
        self.status = status


def lib() -> ctypes.CDLL:
    """Loads the in-tree HIP library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise OpenSayalError(
                OSY_E_STATE,
                f"{LIB_PATH} is missing: build it with `python -m opensayal_amd.build`"
                " (the solver has no CPU fallback)")
        handle = ctypes.CDLL(os.fspath(LIB_PATH), mode=ctypes.RTLD_LOCAL)
        for name, (res, args) in PROTOTYPES.items():
            fn = getattr(handle, name)
            fn.restype = res
            fn.argtypes = args
        _lib = handle
    return _lib


def check(status: int) -> None:
    if status != OSY_OK:
        msg = lib().osy_last_error()
        raise OpenSayalError(status, msg.decode() if msg else "")
