"""Pins the oracle's semi-Lagrangian advection (src/fluid.cu:364-716):
bilinear interpolation reproduces linear fields, the edge-velocity averages
sample the right four faces, zero displacement is the identity, and the
inverse-distance smoke gather uses the four nearest centres with weights
1/(d + 1e-6)."""
import math

import numpy as np
import pytest

from oracle.oracle import OracleFluid
from tests.helpers import cell, make_config, set_cell


def open_grid(W=40, H=30, **kw):
    """No obstacle; fluid everywhere except the left, bottom and top walls."""
    return make_config(W, H, obstacle_enable=0, **kw)


def fill(o, arr, fn):
    H, W = arr.shape
    for j in range(H):
        for i in range(W):
            set_cell(arr, i, j, fn(i, j))


@pytest.mark.parametrize("real", [np.float32, np.float64])
def test_gx_gy_reproduce_linear_fields(real):
    """u samples sit at (i, j+0.5) and v samples at (i+0.5, j); bilinear
    interpolation of samples of f = a x + b y + c returns f(x, y) wherever
    all four samples are fluid.  A swapped weight or neighbour breaks it."""
    o = OracleFluid(open_grid(), real)
    f = lambda x, y: 0.3 * x - 0.7 * y + 2.0
    fill(o, o.u, lambda i, j: f(i, j + 0.5))
    fill(o, o.v, lambda i, j: f(i + 0.5, j))
    rng = np.random.Generator(np.random.PCG64(1))
    tol = 2e-5 if real == np.float32 else 1e-12
    for x, y in rng.uniform([3.0, 3.0], [35.0, 25.0], size=(400, 2)):
        x, y = float(real(x)), float(real(y))
        assert o.gx(x, y) == pytest.approx(f(x, y), abs=tol * 30)
        assert o.gy(x, y) == pytest.approx(f(x, y), abs=tol * 30)


def test_gx_returns_zero_in_solid_and_outside():
    o = OracleFluid(open_grid())
    o.u[...] = 1.0
    o.v[...] = 1.0
    for x, y in ((0.5, 10.0), (10.0, 0.2), (10.0, 29.5), (-3.0, 5.0), (5.0, 1e9),
                 (float("nan"), 3.0), (float("inf"), 4.0)):
        assert o.gx(x, y) == 0.0 and o.gy(x, y) == 0.0


def test_own_face_sample_is_exact():
    """At a u face (i, j+0.5) the weights are exactly 1 and 0, so GX returns
    u(i, j) bit for bit; likewise GY at (i+0.5, j)."""
    o = OracleFluid(open_grid())
    rng = np.random.Generator(np.random.PCG64(2))
    o.u[...] = rng.uniform(-1, 1, o.u.shape).astype(np.float32)
    o.v[...] = rng.uniform(-1, 1, o.v.shape).astype(np.float32)
    for i, j in ((5, 5), (17, 11), (30, 20)):
        assert o.gx(float(i), j + 0.5) == cell(o.u, i, j)
        assert o.gy(i + 0.5, float(j)) == cell(o.v, i, j)


def test_zero_time_step_advection_is_identity_on_fluid():
    o = OracleFluid(open_grid())
    rng = np.random.Generator(np.random.PCG64(3))
    u0 = rng.uniform(-1, 1, o.u.shape).astype(np.float32)
    v0 = rng.uniform(-1, 1, o.v.shape).astype(np.float32)
    o.set_state(u=u0, v=v0)
    o.advect_velocity(d_t=0.0)
    fluid = o.solid == 0
    assert np.array_equal(o.u[fluid], u0[fluid]) and np.array_equal(o.v[fluid], v0[fluid])
    assert np.all(o.u[~fluid] == 0) and np.all(o.v[~fluid] == 0)


def test_edge_velocity_averages_sample_the_four_surrounding_faces():
    """u depends on y only and v is linear, so the back-traced u is
    alpha * (y0 - avg_v * dt) and reveals avg_v; for an interior u face the
    four v faces around (i, j+0.5) average to the linear v field there.
    The same with the roles swapped reveals avg_u at (i+0.5, j)."""
    W, H, dt = 40, 30, 0.5
    alpha, beta, gamma = 0.1, 0.02, 0.03
    o = OracleFluid(open_grid(W, H), np.float64)
    fill(o, o.u, lambda i, j: alpha * (j + 0.5))
    fill(o, o.v, lambda i, j: beta * (i + 0.5) + gamma * j)
    u_in = o.u.copy()
    o.advect_velocity(d_t=dt)
    for i, j in ((10, 10), (20, 5), (30, 22)):
        avg_v = (alpha * (j + 0.5) - cell(o.u, i, j)) / (alpha * dt)
        assert avg_v == pytest.approx(beta * i + gamma * (j + 0.5), abs=1e-9)
    kappa, a2, d2 = 0.05, 0.04, 0.01
    o2 = OracleFluid(open_grid(W, H), np.float64)
    fill(o2, o2.u, lambda i, j: a2 * i + d2 * (j + 0.5))
    fill(o2, o2.v, lambda i, j: kappa * (i + 0.5))
    o2.advect_velocity(d_t=dt)
    for i, j in ((10, 10), (20, 5), (30, 22)):
        avg_u = (kappa * (i + 0.5) - cell(o2.v, i, j)) / (kappa * dt)
        assert avg_u == pytest.approx(a2 * (i + 0.5) + d2 * j, abs=1e-9)
    assert u_in.shape == o.u.shape


def test_idw_centre_weight_closed_form():
    """At a cell centre the sample cell has distance 0 (weight 1/1e-6) and
    the gather takes the right/top quadrant: w1 = 1e6 / (1e6 + 2/(1+1e-6)
    + 1/(sqrt 2 + 1e-6)) = 0.9999973."""
    o = OracleFluid(open_grid())
    i, j = 12, 9
    set_cell(o.smoke, i, j, 1.0)
    w1 = o.idw(i + 0.5, j + 0.5)
    exact = 1e6 / (1e6 + 2 / (1 + 1e-6) + 1 / (math.sqrt(2) + 1e-6))
    assert w1 == pytest.approx(exact, abs=2e-7)
    assert abs(w1 - 0.9999973) < 1e-7


def test_idw_uses_the_four_surrounding_centres_with_inverse_distance_weights():
    """The gather takes the four cell centres around the sample point (the
    corners of the dual cell holding it), weighted by 1/(d + 1e-6)."""
    o = OracleFluid(open_grid())
    rng = np.random.Generator(np.random.PCG64(4))
    for _ in range(40):
        i, j = int(rng.integers(4, 34)), int(rng.integers(4, 24))
        x, y = float(np.float32(i + rng.uniform())), float(np.float32(j + rng.uniform()))
        cands = [(a, b) for a in range(i - 1, i + 2) for b in range(j - 1, j + 2)]
        dist = {c: math.hypot(x - (c[0] + 0.5), y - (c[1] + 0.5)) for c in cands}
        a0, b0 = math.floor(x - 0.5), math.floor(y - 0.5)
        nearest = [(a0, b0), (a0 + 1, b0), (a0, b0 + 1), (a0 + 1, b0 + 1)]
        inv = {c: 1.0 / (dist[c] + 1e-6) for c in nearest}
        tot = sum(inv.values())
        for c in cands:
            o.smoke[...] = 0
            set_cell(o.smoke, c[0], c[1], 1.0)
            w = o.idw(x, y)
            if c in nearest:
                assert w == pytest.approx(inv[c] / tot, rel=1e-5, abs=1e-6)
            else:
                assert w == 0.0


def test_idw_drops_solid_neighbours_without_renormalising():
    """Weights of solid cells are dropped, the rest are not rescaled: a
    constant field next to the bottom wall loses the wall cells' share."""
    o = OracleFluid(open_grid())
    o.smoke[...] = 1.0
    x, y = 10.7, 1.2  # quadrant right/low: (10,1), (11,1), (10,0), (11,0)
    d = lambda a, b: math.hypot(x - (a + 0.5), y - (b + 0.5))
    inv = [1 / (d(10, 1) + 1e-6), 1 / (d(11, 1) + 1e-6), 1 / (d(10, 0) + 1e-6),
           1 / (d(11, 0) + 1e-6)]
    assert o.idw(x, y) == pytest.approx((inv[0] + inv[1]) / sum(inv), rel=1e-5)


def test_smoke_advection_uses_new_velocity_and_translates_uniform_flow():
    """Uniform u = 1 over the fluid (v = 0): the centre back-trace moves by
    exactly dt = 1 cell to the left, so smoke shifts one cell right up to
    the IDW centre weight."""
    W, H = 40, 30
    o = OracleFluid(open_grid(W, H))
    fluid = o.solid == 0
    o.u[...] = 0
    o.u[:, 2:] = 1.0
    o.u[~fluid] = 0
    rng = np.random.Generator(np.random.PCG64(5))
    s0 = rng.uniform(0, 1, (H, W)).astype(np.float32)
    o.set_state(smoke=s0)
    o.advect_smoke(d_t=1.0)
    for i, j in ((10, 10), (20, 15), (30, 5)):
        assert cell(o.smoke, i, j) == pytest.approx(cell(s0, i - 1, j), abs=5e-6)
