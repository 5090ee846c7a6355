"""Seeded synthetic workloads of the wind-tunnel benchmark configurations.

Everything here is deterministic from a seed: a SplitMix64 counter hash of
(seed, field, linear index) gives the initial velocity noise, so the same
state can be rebuilt on any host without shipping data.  Arrays are (H, W) in
the reference's memory order (row r holds j = H - 1 - r).
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

from . import Config, default_config

_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def splitmix64(x: np.ndarray) -> np.ndarray:
    """SplitMix64 finaliser on uint64 arrays (wrapping arithmetic)."""
    with np.errstate(over="ignore"):
        z = (x + np.uint64(0x9E3779B97F4A7C15)) & _M64
        z = ((z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)) & _M64
        z = ((z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)) & _M64
        return z ^ (z >> np.uint64(31))


def uniform_field(seed: int, field_id: int, shape, lo: float, hi: float) -> np.ndarray:
    """float32 values uniform in [lo, hi): 24 hash bits per element."""
    n = int(np.prod(shape))
    with np.errstate(over="ignore"):
        key = (np.uint64(seed) * np.uint64(0x100000001B3)) ^ (np.uint64(field_id) << np.uint64(56))
        idx = np.arange(n, dtype=np.uint64) ^ key
    bits = splitmix64(idx) >> np.uint64(40)
    u01 = bits.astype(np.float64) / float(1 << 24)
    return (lo + (hi - lo) * u01).astype(np.float32).reshape(shape)


def inflow_smoke(cfg: Config) -> np.ndarray:
    """Initial smoke: 1 on the inflow cells (1 <= i <= smoke_length,
    H/2 - ph/2 <= j <= H/2 + ph/2, the band of src/fluid.cu:317-319), else 0."""
    H, W = cfg.height, cfg.width
    s = np.zeros((H, W), np.float32)
    lo = H // 2 - _cdiv(cfg.pipe_height, 2)
    hi = H // 2 + _cdiv(cfg.pipe_height, 2)
    for j in range(max(lo, 0), min(hi, H - 1) + 1):
        s[H - 1 - j, 1:cfg.smoke_length + 1] = 1.0
    return s


def _cdiv(a: int, b: int) -> int:
    """C integer division (truncates toward zero)."""
    q = abs(a) // abs(b)
    return q if (a >= 0) == (b >= 0) else -q


def initial_state(cfg: Config, seed: int):
    """(u, v, smoke): velocity noise uniform in +-0.01, smoke on the inflow."""
    shape = (cfg.height, cfg.width)
    u = uniform_field(seed, 1, shape, -0.01, 0.01)
    v = uniform_field(seed, 2, shape, -0.01, 0.01)
    return u, v, inflow_smoke(cfg)


def wind_tunnel(width: int, height: int, *, pipe_height: int, cx: int, cy: int, radius: float,
                speed: float = 2.0, n: int = 50, o: float = 1.9, g: float = -5.0,
                d_t: float = 0.05, pressure: bool = False, storage: str = "f32",
                viscosity: float = 0.0) -> Config:
    """A wind-tunnel configuration with the benchmark's fixed choices: drain
    open (solid walls top, bottom and left), smoke 1 in a single band of the
    default height, no decay, no drag, cell size 1; viscosity 0 (no diffusion
    pass) unless given."""
    return default_config(
        width, height, pipe_height=pipe_height, obstacle_center_x=cx, obstacle_center_y=cy,
        obstacle_radius=radius, wind_speed=speed, projection_n=n, projection_o=o, g=g, d_t=d_t,
        enable_pressure=int(pressure), enable_drain=1, enable_smoke=1, wind_smoke=1.0,
        smoke_count=1, smoke_length=1, enable_decay=0, drag_coeff=0.0, viscosity=viscosity,
        density=1.0, storage=storage)


def plain_small(**kw) -> Config:
    """Plain method, small grid: 256x144, pipe 36, obstacle (128,72) r=10."""
    return wind_tunnel(256, 144, pipe_height=36, cx=128, cy=72, radius=10.0, **kw)


def plain_full(**kw) -> Config:
    """Plain method, full size: 1920x1080, pipe 270, obstacle (960,540) r=36."""
    return wind_tunnel(1920, 1080, pipe_height=270, cx=960, cy=540, radius=36.0, **kw)


def mixed_precision(storage: str = "f16", **kw) -> Config:
    """3840x2160, pipe 540, obstacle at the centre r=72, n=100, pressure on."""
    kw.setdefault("n", 100)
    return wind_tunnel(3840, 2160, pipe_height=540, cx=1920, cy=1080, radius=72.0,
                       pressure=True, storage=storage, **kw)


@dataclass
class SimSpec:
    """One simulation of a batched or serving workload."""
    cfg: Config
    seed: int
    steps: int


def batched_specs(count: int = 256, seed: int = 2024, steps: int = 500):
    """Uneven grids: width from {128, 192, ..., 1024}, height = round(9W/16),
    obstacle centre uniform in the middle 60 % of the domain, radius uniform in
    [0.02, 0.06] * W, inflow speed uniform [0.5, 3.0], pipe height H/4,
    n from {20, 50, 100}, o = 1.9, d_t = 0.05, g = -5."""
    rng = np.random.Generator(np.random.PCG64(seed))
    out = []
    for k in range(count):
        W = int(rng.integers(0, 15)) * 64 + 128
        H = int(round(W * 9 / 16))
        cx = int(round(rng.uniform(0.2, 0.8) * W))
        cy = int(round(rng.uniform(0.2, 0.8) * H))
        r = float(rng.uniform(0.02, 0.06) * W)
        speed = float(rng.uniform(0.5, 3.0))
        n = int(rng.choice([20, 50, 100]))
        cfg = wind_tunnel(W, H, pipe_height=H // 4, cx=cx, cy=cy, radius=r, speed=speed, n=n)
        out.append(SimSpec(cfg, seed * 1000003 + k, steps))
    return out


def small_specs(count: int = 256, seed: int = 99, steps: int = 200):
    """Mixed small grids: width from {64, 96, ..., 384}, height = round(9W/16),
    otherwise drawn as in ``batched_specs``."""
    rng = np.random.Generator(np.random.PCG64(seed))
    out = []
    for k in range(count):
        W = int(rng.integers(0, 11)) * 32 + 64
        H = int(round(W * 9 / 16))
        cx = int(round(rng.uniform(0.2, 0.8) * W))
        cy = int(round(rng.uniform(0.2, 0.8) * H))
        r = float(rng.uniform(0.02, 0.06) * W)
        speed = float(rng.uniform(0.5, 3.0))
        n = int(rng.choice([20, 50, 100]))
        cfg = wind_tunnel(W, H, pipe_height=H // 4, cx=cx, cy=cy, radius=r, speed=speed, n=n)
        out.append(SimSpec(cfg, seed * 1000003 + k, steps))
    return out


def start(spec: SimSpec):
    """A Fluid of ``spec.cfg`` holding ``initial_state(spec.cfg, spec.seed)``."""
    from . import Fluid

    fl = Fluid(spec.cfg)
    u, v, s = initial_state(spec.cfg, spec.seed)
    fl.set_field("vel_x", u)
    fl.set_field("vel_y", v)
    fl.set_field("smoke", s)
    return fl


def serving_specs(count: int, seed: int = 7, rate_hz: float = 200.0):
    """Poisson arrivals of requests: width log-uniform in [64, 2048] (16:9),
    steps uniform in [50, 2000], n from {20, 50, 100}, obstacle and inflow as
    in ``batched_specs``, FP32 or FP16 storage with equal odds.  Returns a
    list of (arrival_time_s, SimSpec)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    t = 0.0
    out = []
    for k in range(count):
        t += float(rng.exponential(1.0 / rate_hz))
        W = int(round(math.exp(rng.uniform(math.log(64), math.log(2048)))))
        W = max(W, 16)
        H = max(int(round(W * 9 / 16)), 9)
        cx = int(round(rng.uniform(0.2, 0.8) * W))
        cy = int(round(rng.uniform(0.2, 0.8) * H))
        r = float(rng.uniform(0.02, 0.06) * W)
        speed = float(rng.uniform(0.5, 3.0))
        n = int(rng.choice([20, 50, 100]))
        steps = int(rng.integers(50, 2001))
        storage = "f32" if rng.uniform() < 0.5 else "f16"
        cfg = wind_tunnel(W, H, pipe_height=max(H // 4, 1), cx=cx, cy=cy, radius=r, speed=speed,
                          n=n, storage=storage)
        out.append((t, SimSpec(cfg, seed * 1000003 + k, steps)))
    return out
