"""Pins the oracle's diffusion pass (apply_diffusion, src/fluid.cu:167-190),
restated as Jacobi rounds on u: every interior cell becomes
(u + a * (sum of its 4 neighbours)) / (1 + 4a) of the previous round.

Checks a plausible mistake fails: linear (discrete harmonic) fields are
fixed points; a unit impulse spreads to exactly 1/2 and four 1/8 in one
round with a = 1/4 (a Gauss-Seidel sweep would not give four equal
neighbours); mass moves only through the borders; borders and v are never
written; solid cells are diffused as the reference's kernel does; and the
FP64 rounds converge to the discrete Laplace solution for the fixed border
(solved directly).
"""
import numpy as np
import pytest
import scipy.sparse as sps
import scipy.sparse.linalg as spl

from oracle.oracle import OracleFluid
from tests.helpers import make_config


def grid(W=24, H=18, viscosity=5.0, d_t=0.05, **kw):
    # viscosity 5 with d_t 0.05 gives a = 0.25 exactly in float and double
    return make_config(W, H, viscosity=viscosity, d_t=d_t, obstacle_enable=0, **kw)


@pytest.mark.parametrize("real", [np.float32, np.float64])
def test_linear_field_is_a_fixed_point(real):
    o = OracleFluid(grid(), real)
    H, W = o.u.shape
    j = (H - 1 - np.arange(H))[:, None]
    i = np.arange(W)[None, :]
    o.u[...] = (0.25 * i - 0.5 * j + 3.0).astype(real)  # dyadic: every sum exact
    before = o.u.copy()
    o.diffusion(sweeps=7)
    assert np.array_equal(o.u, before)


def test_impulse_spreads_as_one_jacobi_round():
    o = OracleFluid(grid())
    H, W = o.u.shape
    r, c = H - 1 - 9, 11  # cell (11, 9)
    o.u[r, c] = 1.0
    o.diffusion(sweeps=1)
    want = np.zeros_like(o.u)
    want[r, c] = 0.5          # (1 + a*0) / (1 + 4a), a = 1/4
    for dr, dc in ((0, 1), (0, -1), (1, 0), (-1, 0)):
        want[r + dr, c + dc] = 0.125  # (0 + a*1) / 2
    assert np.array_equal(o.u, want)
    o.diffusion(sweeps=1)
    assert o.u.sum() == pytest.approx(1.0, abs=1e-7)  # far from the borders


def test_borders_v_and_solids():
    cfg = make_config(30, 20, viscosity=5.0, d_t=0.05, obstacle_center_x=15,
                      obstacle_center_y=10, obstacle_radius=4.0)
    o = OracleFluid(cfg)
    rng = np.random.Generator(np.random.PCG64(1))
    o.u[...] = rng.uniform(-1, 1, o.u.shape).astype(np.float32)
    o.v[...] = rng.uniform(-1, 1, o.v.shape).astype(np.float32)
    u0, v0 = o.u.copy(), o.v.copy()
    o.diffusion(sweeps=3)
    assert np.array_equal(o.v, v0)
    for edge in (np.s_[0, :], np.s_[-1, :], np.s_[:, 0], np.s_[:, -1]):
        assert np.array_equal(o.u[edge], u0[edge])
    inner = o.solid[1:-1, 1:-1] != 0
    assert inner.any()
    assert (o.u[1:-1, 1:-1][inner] != u0[1:-1, 1:-1][inner]).any()  # solids diffuse too


def test_fp64_rounds_converge_to_laplace_solution():
    """The fixed point of the rounds is the harmonic field with the initial
    border; a wrong sign, neighbour or divisor moves or destroys it."""
    W, H = 14, 11
    o = OracleFluid(grid(W, H, viscosity=20.0), np.float64)
    rng = np.random.Generator(np.random.PCG64(2))
    o.u[...] = rng.uniform(-1, 1, o.u.shape)
    u0 = o.u.copy()
    o.diffusion(sweeps=4000)
    # direct solve: 4 u(i,j) - sum of neighbours = 0 on the interior
    ni, nj = W - 2, H - 2
    idx = lambda a, b: (b - 1) * ni + (a - 1)
    A = sps.lil_matrix((ni * nj, ni * nj))
    rhs = np.zeros(ni * nj)
    val = lambda a, b: u0[H - 1 - b, a]
    for b in range(1, H - 1):
        for a in range(1, W - 1):
            k = idx(a, b)
            A[k, k] = 4.0
            for da, db in ((1, 0), (-1, 0), (0, 1), (0, -1)):
                x, y = a + da, b + db
                if 1 <= x <= W - 2 and 1 <= y <= H - 2:
                    A[k, idx(x, y)] = -1.0
                else:
                    rhs[k] += val(x, y)
    sol = spl.spsolve(A.tocsr(), rhs)
    got = np.array([o.u[H - 1 - b, a] for b in range(1, H - 1) for a in range(1, W - 1)])
    assert np.max(np.abs(got - sol)) < 1e-9


def test_update_runs_diffusion_only_with_viscosity():
    """With viscosity 0 the update is the plain one; with viscosity the
    result differs and equals forces, diffusion, projection, ... composed."""
    base = dict(projection_n=6, wind_speed=1.0, g=-2.0)
    a = OracleFluid(make_config(40, 26, viscosity=0.0, **base))
    b = OracleFluid(make_config(40, 26, viscosity=0.7, **base))
    rng = np.random.Generator(np.random.PCG64(3))
    u = rng.uniform(-1, 1, a.u.shape).astype(np.float32)
    a.set_state(u=u)
    b.set_state(u=u)
    a.update()
    b.update()
    assert not np.array_equal(a.u, b.u)
    c = OracleFluid(make_config(40, 26, viscosity=0.7, **base))
    c.set_state(u=u)
    c.forces()
    c.diffusion()
    c.projection()
    c.extrapolation()
    c.advect_velocity()
    assert np.array_equal(c.u, b.u) and np.array_equal(c.v, b.v)
