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


# ---- which wavefronts of the advection kernels leave their LDS stage -------
# The advection kernels (advect_device.hpp) stage a 64 x 32 tile plus a
# 4-cell halo; a wavefront (one row of 64 cells of a tile) samples from LDS
# when every lane's sample cell and its neighbours are staged, and otherwise
# takes the guarded samplers that read HBM.  These functions restate that
# test on the CPU, in float32 as the kernels compute the sample positions.
ADV_TX, ADV_TY, ADV_HALO = 64, 32, 4


def _grid_view(a):
    """(H, W) reference order -> (j, i) indexing."""
    return np.flipud(np.asarray(a))


def _shift(a, di, dj, fill):
    """b[j, i] = a[j + dj, i + di], `fill` outside the grid."""
    H, W = a.shape
    out = np.full_like(a, fill)
    js, ie = slice(max(0, -dj), min(H, H - dj)), slice(max(0, -di), min(W, W - di))
    jd, idd = slice(max(0, dj), min(H, H + dj)), slice(max(0, di), min(W, W + di))
    out[js, ie] = a[jd, idd]
    return out


def _to_index(q):
    """int(q) after clamping to +-2^24 (NaN to -2^24), as to_index()."""
    q = np.where(np.isnan(q), np.float32(-16777216.0), q)
    q = np.clip(q, np.float32(-16777216.0), np.float32(16777216.0))
    return np.trunc(q).astype(np.int64)


def _inner(ci, cj, i, j):
    """Sample cell (ci, cj) of the lane at (i, j) and its neighbours are in
    the lane's staged region (AdvRegion::inner)."""
    x0 = (i // ADV_TX) * ADV_TX - ADV_HALO
    y0 = (j // ADV_TY) * ADV_TY - ADV_HALO
    rx, ry = ADV_TX + 2 * ADV_HALO, ADV_TY + 2 * ADV_HALO
    return ((ci - x0 - 1 >= 0) & (ci - x0 - 1 < rx - 2) &
            (cj - y0 - 1 >= 0) & (cj - y0 - 1 < ry - 2))


def _fallback_rows(staged):
    """Number of wavefronts (64-cell row segments of a tile) with a lane
    outside its stage, and the number of wavefronts."""
    H, W = staged.shape
    nb = (W + ADV_TX - 1) // ADV_TX
    pad = np.ones((H, nb * ADV_TX), bool)
    pad[:, :W] = staged
    rows = pad.reshape(H, nb, ADV_TX).all(axis=2)
    return int((~rows).sum()), int(rows.size)


def velocity_advection_fallback(cfg, solid, u_ext, v_ext):
    """(wavefronts that take the HBM path, wavefronts) of the velocity
    advection reading the extrapolated field (u_ext, v_ext)."""
    F = _grid_view(solid) == 0
    U = _grid_view(u_ext).astype(np.float32)
    V = _grid_view(v_ext).astype(np.float32)
    H, W = U.shape
    cs = int(cfg.cell_size)
    dt = np.float32(cfg.d_t)
    j, i = np.mgrid[0:H, 0:W]
    nz = np.float32(-0.0)

    def avg(base, A, offs):
        acc, cnt = base.copy(), np.ones_like(base)
        for di, dj in offs:
            ok = _shift(F, di, dj, False)
            acc = acc + np.where(ok, _shift(A, di, dj, np.float32(0)), nz)
            cnt += ok
        return (acc / cnt).astype(np.float32)

    avg_v = avg(V, V, [(-1, 1), (0, 1), (-1, 0)])
    avg_u = avg(U, U, [(1, 0), (0, -1), (1, -1)])
    csf = np.float32(cs)
    xu = (i * cs).astype(np.float32) - U * dt
    yu = ((j + 0.5) * cs).astype(np.float32) - avg_v * dt
    xv = ((i + 0.5) * cs).astype(np.float32) - avg_u * dt
    yv = (j * cs).astype(np.float32) - V * dt
    staged = (_inner(_to_index(xu / csf), _to_index(yu / csf), i, j) &
              _inner(_to_index(xv / csf), _to_index(yv / csf), i, j))
    return _fallback_rows(staged)


def smoke_advection_fallback(cfg, solid, u_new, v_new):
    """(wavefronts that take the HBM path, wavefronts) of the smoke
    advection reading the advected velocities (u_new, v_new)."""
    F = _grid_view(solid) == 0
    U = _grid_view(u_new).astype(np.float32)
    V = _grid_view(v_new).astype(np.float32)
    H, W = U.shape
    cs = int(cfg.cell_size)
    dt = np.float32(cfg.d_t)
    j, i = np.mgrid[0:H, 0:W]
    half, zero, nz = np.float32(0.5), np.float32(0.0), np.float32(-0.0)
    # the velocity at a cell centre: two half-weighted faces, two zero-weighted
    vx = zero + half * U
    vx = vx + np.where(_shift(F, 1, 0, False), half * _shift(U, 1, 0, zero), nz)
    vx = np.where(F, vx, zero)
    vy = zero + half * V
    vy = vy + np.where(_shift(F, 0, 1, False), half * _shift(V, 0, 1, zero), nz)
    vy = np.where(F, vy, zero)
    csf = np.float32(cs)
    x = ((i + 0.5) * cs).astype(np.float32) - vx * dt
    y = ((j + 0.5) * cs).astype(np.float32) - vy * dt
    staged = _inner(_to_index(x / csf), _to_index(y / csf), i, j)
    return _fallback_rows(staged)


def advection_fallback_counts(o):
    """From an OracleFluid's current state, the fallback wavefront counts of
    the next step's velocity and smoke advections: ((vel, total), (smoke,
    total)).  Runs the step's phases on a copy."""
    from oracle.oracle import OracleFluid

    c = OracleFluid(o.cfg, o.real, nthreads=o.nthreads)
    c.set_state(u=o.u, v=o.v, smoke=o.smoke, p=o.p)
    c.forces()
    c.projection()
    c.extrapolation()
    vel = velocity_advection_fallback(c.cfg, c.solid, c.u, c.v)
    c.advect_velocity()
    smk = smoke_advection_fallback(c.cfg, c.solid, c.u, c.v)
    return vel, smk
