"""Pins the oracle's mask build (src/fluid.cu:111-142) by brute force and by
lattice-point counts, and pins the integer restatement of the obstacle test
that the GPU mask kernel uses."""
import math

import numpy as np
import pytest

from oracle.oracle import build_masks
from tests.helpers import cell, make_config


def brute_solid(c):
    H, W = c.height, c.width
    s = np.zeros((H, W), np.int32)
    for j in range(H):
        for i in range(W):
            d = math.sqrt((i - c.obstacle_center_x) ** 2 + (j - c.obstacle_center_y) ** 2)
            solid = (i == 0 or j == 0 or j == H - 1 or (not c.enable_drain and i == W - 1)
                     or (c.obstacle_enable and d < c.obstacle_radius))
            s[H - 1 - j, i] = int(solid)
    return s


def brute_total_s(solid):
    H, W = solid.shape
    ts = np.zeros_like(solid)
    for r in range(H):
        for i in range(W):
            n = 0
            for dr, di in ((0, -1), (0, 1), (1, 0), (-1, 0)):
                rr, ii = r + dr, i + di
                if 0 <= rr < H and 0 <= ii < W and not solid[rr, ii]:
                    n += 1
            ts[r, i] = n
    return ts


@pytest.mark.parametrize("kw", [
    dict(width=40, height=30),
    dict(width=40, height=30, enable_drain=0),
    dict(width=41, height=23, obstacle_center_x=5, obstacle_center_y=3, obstacle_radius=7.5),
    dict(width=33, height=17, obstacle_enable=0),
    dict(width=50, height=40, obstacle_radius=0.0),
    dict(width=50, height=40, obstacle_radius=-3.0),
    dict(width=64, height=36, obstacle_center_x=-4, obstacle_center_y=100, obstacle_radius=70.25),
    dict(width=20, height=20, obstacle_radius=5.0),
])
def test_masks_match_brute_force(kw):
    W, H = kw.pop("width"), kw.pop("height")
    c = make_config(W, H, **kw)
    solid, ts = build_masks(c)
    exp = brute_solid(c)
    assert np.array_equal(solid, exp)
    assert np.array_equal(ts, brute_total_s(exp))


def test_obstacle_cell_count_is_gauss_circle():
    """r = 10 strictly inside: 317 lattice points with x^2+y^2 <= 100, minus
    the 12 on the circle (0,+-10), (+-10,0), (+-6,+-8), (+-8,+-6)."""
    c = make_config(256, 144, obstacle_center_x=128, obstacle_center_y=72, obstacle_radius=10.0)
    solid, _ = build_masks(c)
    borders = 144 + 256 + 256 - 2
    assert solid.sum() == borders + 305


def test_cells_at_exact_radius_are_fluid():
    c = make_config(30, 30, obstacle_center_x=15, obstacle_center_y=15, obstacle_radius=5.0)
    solid, _ = build_masks(c)
    assert cell(solid, 20, 15) == 0 and cell(solid, 19, 15) == 1
    assert cell(solid, 18, 19) == 0 and cell(solid, 18, 18) == 1  # (3,4): distance 5


def test_right_column_follows_drain():
    open_, _ = build_masks(make_config(16, 9, obstacle_enable=0, enable_drain=1))
    closed, _ = build_masks(make_config(16, 9, obstacle_enable=0, enable_drain=0))
    assert open_[1:-1, -1].sum() == 0
    assert closed[:, -1].sum() == 9


def test_total_s_closed_form_on_empty_box():
    """No obstacle, drain closed: interior corners see 2 fluid neighbours,
    interior edges 3, the rest 4; border cells see their interior neighbour."""
    c = make_config(10, 8, obstacle_enable=0, enable_drain=0)
    _, ts = build_masks(c)
    assert cell(ts, 1, 1) == 2 and cell(ts, 8, 6) == 2
    assert cell(ts, 1, 3) == 3 and cell(ts, 4, 1) == 3
    assert cell(ts, 4, 4) == 4
    assert cell(ts, 0, 3) == 1 and cell(ts, 0, 0) == 0


def test_integer_obstacle_predicate_equals_sqrt_form():
    """The GPU mask kernel tests d2 < r*r (r > 0) instead of sqrt(d2) < r.
    For integer d2 and a float radius both agree; checked on every d2 below
    2^17 against radii near perfect squares, integers and random floats."""
    d2 = np.arange(0, 1 << 17, dtype=np.float64)
    rng = np.random.Generator(np.random.PCG64(5))
    radii = [np.float32(x) for x in rng.uniform(0.0, 370.0, 300)]
    for k in range(1, 362):
        radii += [np.float32(k), np.nextafter(np.float32(k), np.float32(0)),
                  np.nextafter(np.float32(k), np.float32(1e9)),
                  np.float32(math.sqrt(k)), np.nextafter(np.float32(math.sqrt(k)), np.float32(0))]
    radii += [np.float32(0.0), np.float32(-2.0)]
    for r32 in radii:
        r = float(r32)
        a = np.sqrt(d2) < r
        b = (r > 0) & (d2 < r * r)
        assert np.array_equal(a, b), r
