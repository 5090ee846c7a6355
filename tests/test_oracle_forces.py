"""Pins the oracle's apply_external_forces_at (src/fluid.cu:308-349) and
decay_smoke_at (src/fluid.cu:758-762) with closed forms."""
import math

import numpy as np

from oracle.oracle import OracleFluid
from opensayal_amd._abi import OsySource
from tests.helpers import cell, make_config, noise


def test_gravity_is_exact_quarter_on_every_cell():
    """g = -5, d_t = 0.05f: g*d_t rounds to exactly -0.25 in binary32, and the
    reference adds it to v on all W*H cells, solid ones included."""
    assert np.float32(-5.0) * np.float32(0.05) == np.float32(-0.25)
    c = make_config(24, 16, g=-5.0, d_t=0.05, wind_smoke=0.0)
    o = OracleFluid(c)
    o.forces()
    assert np.all(o.v == np.float32(-0.25))
    assert np.all(o.u == 0) and np.all(o.smoke == 0)


def test_inflow_band_sets_speed_and_smoke():
    """u = speed and smoke = wind_tunnel.smoke on 1 <= i <= smoke_length and
    H/2 - ph/2 <= j <= H/2 + ph/2; other cells keep their values."""
    W, H, ph, sl = 40, 31, 9, 3
    c = make_config(W, H, pipe_height=ph, smoke_length=sl, smoke_height=5, wind_speed=2.5,
                    wind_smoke=0.75)
    o = OracleFluid(c)
    u0 = noise((H, W), 1)
    o.set_state(u=u0)
    o.forces()
    lo, hi = H // 2 - ph // 2, H // 2 + ph // 2
    sband = (H // 2 - 5 // 2, H // 2 + 5 // 2)
    for j in range(H):
        for i in range(W):
            in_band = 1 <= i <= sl and lo <= j <= hi
            exp_u = np.float32(2.5) if in_band else cell(u0, i, j)
            assert cell(o.u, i, j) == exp_u
            in_smoke = in_band and sband[0] <= j <= sband[1]
            assert cell(o.smoke, i, j) == (np.float32(0.75) if in_smoke else 0)


def test_smoke_stripes_have_count_height_and_spacing():
    """smoke_count = 3, smoke_height = 4, pipe 30: spacing (30 - 12) / 2 = 9,
    so three 4-row stripes separated by 9-row gaps, counted from the top of
    the pipe (src/fluid.cu:312-327)."""
    W, H = 12, 60
    c = make_config(W, H, pipe_height=30, smoke_count=3, smoke_height=4, wind_smoke=1.0)
    o = OracleFluid(c)
    o.forces()
    col = [cell(o.smoke, 1, j) for j in range(H)]
    rows = [j for j in range(H) if col[j] == 1.0]
    top = H // 2 + 15
    runs = []
    for j in sorted(rows, reverse=True):
        if runs and runs[-1][-1] == j + 1:
            runs[-1].append(j)
        else:
            runs.append([j])
    assert [len(r) for r in runs] == [4, 4, 4]
    assert runs[0][0] == top
    assert runs[0][-1] - runs[1][0] - 1 == 9 and runs[1][-1] - runs[2][0] - 1 == 9


def test_drag_scales_by_exponential():
    c = make_config(16, 10, drag_coeff=0.5, d_t=0.05)
    o = OracleFluid(c)
    u0, v0 = noise((10, 16), 2), noise((10, 16), 3)
    o.set_state(u=u0, v=v0)
    o.forces()
    damp = math.exp(-0.5 * float(np.float32(0.05)))
    inflow = np.zeros((10, 16), bool)  # speed 0 sets u = 0 in the band before drag
    for j in range(10 // 2 - 10 // 4 // 2, 10 // 2 + 10 // 4 // 2 + 1):
        inflow[10 - 1 - j, 1] = True
    np.testing.assert_allclose(o.u[~inflow], u0[~inflow] * damp, rtol=2e-7)
    np.testing.assert_allclose(o.v, v0 * damp, rtol=2e-7)


def test_active_source_disc():
    """A source adds velocity * (i - x, j - y) and sets smoke inside the open
    disc of radius 40 around its position (src/fluid.cu:336-346)."""
    W, H = 120, 100
    c = make_config(W, H)
    o = OracleFluid(c)
    src = OsySource(active=1, smoke=0.5, velocity=0.25, position_x=60, position_y=50)
    o.forces(source=src)
    for (i, j) in ((60, 50), (99, 50), (100, 50), (60, 89), (60, 90), (84, 82), (20, 20)):
        inside = (i - 60) ** 2 + (j - 50) ** 2 < 1600
        assert cell(o.smoke, i, j) == (0.5 if inside else 0.0)
        if inside:
            assert cell(o.u, i, j) == np.float32(0.25) * np.float32(i - 60)
            assert cell(o.v, i, j) == np.float32(0.25) * np.float32(j - 50)


def test_pressure_zeroed_only_when_enabled():
    for on in (0, 1):
        c = make_config(10, 8, enable_pressure=on)
        o = OracleFluid(c)
        o.set_state(p=np.full((8, 10), 3.0, np.float32))
        o.forces()
        assert np.all(o.p == (0.0 if on else 3.0))


def test_decay_clamps_at_zero():
    c = make_config(10, 8)
    o = OracleFluid(c)
    s0 = np.linspace(-0.5, 1.0, 80, dtype=np.float32).reshape(8, 10)
    o.set_state(smoke=s0)
    o.decay(rate=2.0, d_t=0.05)
    exp = np.maximum(s0 - np.float32(2.0) * np.float32(0.05), 0).astype(np.float32)
    assert np.array_equal(o.smoke, exp)
