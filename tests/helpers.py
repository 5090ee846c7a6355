"""Test helpers: configurations built without the solver library.

``make_config`` fills the ABI struct from the defaults documented for
OpenSayal.conf.json (docs/configuration.md, src/config_parser.cpp:21-181) so
oracle tests do not depend on the product's parser; test_config.py checks the
parser against the same table.
"""
from __future__ import annotations

import numpy as np

from opensayal_amd._abi import OsyConfig


def make_config(width: int, height: int, **kw) -> OsyConfig:
    c = OsyConfig()
    c.width, c.height = width, height
    c.cell_size = 1.0
    c.cell_pixel_size = 1
    c.enable_drain = 1
    c.enable_pressure = 0
    c.enable_smoke = 1
    c.enable_interactive = 0
    c.projection_n = 50
    c.projection_o = 1.9
    c.pipe_height = height // 4
    c.pipe_length = 0
    c.smoke_length = 1
    c.smoke_height = height // 4
    c.smoke_count = 1
    c.wind_speed = 0.0
    c.wind_smoke = 1.0
    c.g = 0.0
    c.d_t = 0.05
    c.enable_real_time = 0
    c.real_time_multiplier = 1.0
    c.enable_decay = 0
    c.decay_rate = 0.05
    c.obstacle_enable = 1
    c.obstacle_center_x = width // 2
    c.obstacle_center_y = height // 2
    c.obstacle_radius = float(np.float32(min(width, height)) / np.float32(30.0))
    c.density = 1.0
    c.drag_coeff = 0.0
    c.viscosity = 0.0
    c.block_size_x = 64
    c.block_size_y = 1
    c.omp_thread_count = 0
    c.storage = 0
    for k, v in kw.items():
        if not hasattr(c, k):
            raise AttributeError(k)
        setattr(c, k, v)
    return c


def wind_small(**kw) -> OsyConfig:
    """The 256x144 wind-tunnel benchmark configuration."""
    base = dict(pipe_height=36, obstacle_center_x=128, obstacle_center_y=72, obstacle_radius=10.0,
                wind_speed=2.0, g=-5.0, d_t=0.05, projection_n=50, projection_o=1.9)
    base.update(kw)
    return make_config(256, 144, **base)


def cell(a: np.ndarray, i: int, j: int):
    """Element (i, j) of an (H, W) array in the reference order."""
    return a[a.shape[0] - 1 - j, i]


def set_cell(a: np.ndarray, i: int, j: int, val) -> None:
    a[a.shape[0] - 1 - j, i] = val


def noise(shape, seed, amp=0.01, dtype=np.float32):
    rng = np.random.Generator(np.random.PCG64(seed))
    return (rng.uniform(-amp, amp, size=shape)).astype(dtype)


def rel_linf(a: np.ndarray, b: np.ndarray) -> float:
    """max |a - b| / max |b| (0 when both are 0)."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    den = np.max(np.abs(b))
    num = np.max(np.abs(a - b))
    return float(num / den) if den > 0 else float(num)
