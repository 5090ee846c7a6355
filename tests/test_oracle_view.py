"""Pins the oracle's frame export and path lines.

Frame export, update_fluid_pixels_kernel (src/graphics_handler.cu:214-285)
with hsv_to_rgb / map_rgba / clamp (src/helper.cu:3-50): solid cells are
(80, 80, 80, 255); smoke-only pixels are (255, c, c, 255) with
c = 255 - trunc(255 smoke); pressure colours follow the hue wheel
(1 - norm) * 120 degrees, checked against Python's colorsys (an independent
HSV implementation in double) and at the sector corners in closed form.

Path lines, Fluid::trace (src/fluid.cu:16-36) over the seed grid of
update_traces_kernel (src/graphics_handler.cu:384-396): a still field gives
the rounded cell centre at every point, a uniform field a straight line
whose points are known exactly (quarter-cell steps are exact in float), and
the rounding is half away from zero.
"""
import colorsys

import numpy as np
import pytest

from oracle import oracle as orc
from oracle.oracle import OracleFluid
from tests.helpers import make_config

SOLID = 0x505050FF


def rgba(r, g, b, a=255):
    return (r << 24) | (g << 16) | (b << 8) | a


def unpack(px):
    return (px >> 24) & 255, (px >> 16) & 255, (px >> 8) & 255, px & 255


def test_solid_pixel_wins_over_every_mode():
    for pressure_on in (0, 1):
        for smoke_on in (0, 1):
            assert orc.pixel(1, pressure_on, smoke_on, 0.7, -3.0, -5.0, 5.0) == SOLID


def test_no_mode_leaves_pixel_zero():
    assert orc.pixel(0, 0, 0, 0.7, 1.0, -1.0, 1.0) == 0


def test_smoke_only_closed_form():
    smoke = np.concatenate([np.linspace(0, 1, 4097, dtype=np.float32),
                            np.float32([0.5, 1 / 255, 254.5 / 255, 1e-9])])
    for s in smoke:
        c = 255 - int(np.float32(s) * np.float32(255))
        assert orc.pixel(0, 0, 1, float(s), 0.0, 0.0, 0.0) == rgba(255, c, c)


@pytest.mark.parametrize("norm,rgb", [(1.0, (255, 0, 0)),     # hue 0: red
                                      (0.5, (255, 255, 0)),   # hue 60: yellow
                                      (0.0, (0, 255, 0)),     # hue 120: green
                                      (-0.5, (0, 255, 255)),  # hue 180: cyan
                                      (-1.0, (0, 0, 255))])   # hue 240: blue
def test_pressure_hue_corners(norm, rgb):
    pmin, pmax = -8.0, 4.0
    p = norm * (pmax if norm >= 0 else -pmin)
    assert orc.pixel(0, 1, 0, 0.0, p, pmin, pmax) == rgba(*rgb)
    # with smoke on, the value channel is the smoke: full smoke is the same colour
    assert orc.pixel(0, 1, 1, 1.0, p, pmin, pmax) == rgba(*rgb)


def test_pressure_clamps_beyond_range():
    assert orc.pixel(0, 1, 0, 0.0, 100.0, -1.0, 1.0) == rgba(255, 0, 0)
    assert orc.pixel(0, 1, 0, 0.0, -100.0, -1.0, 1.0) == rgba(0, 0, 255)


def test_pressure_colours_match_colorsys():
    """Hue (1 - norm) * 120 with saturation 1 and value v; each channel is
    trunc(255 * channel).  colorsys computes in double with another formula,
    so a channel may differ by one where 255 * channel lies within rounding of
    an integer; everything else must agree exactly."""
    rng = np.random.Generator(np.random.PCG64(3))
    pmin, pmax = -2.5, 3.0
    n, exact = 3000, 0
    for p, v in zip(rng.uniform(-2.5, 3.0, n).astype(np.float32),
                    rng.uniform(0, 1, n).astype(np.float32)):
        norm = float(np.float32(-p) / np.float32(pmin) if p < 0 else np.float32(p) / np.float32(pmax))
        hue = (1.0 - norm) * 120.0
        want = [int(c * 255) for c in colorsys.hsv_to_rgb(hue / 360.0, 1.0, float(v))]
        got = unpack(orc.pixel(0, 1, 1, float(v), float(p), pmin, pmax))
        assert got[3] == 255
        assert all(abs(a - b) <= 1 for a, b in zip(got[:3], want)), (p, v, got, want)
        exact += tuple(got[:3]) == tuple(want)
    assert exact >= 0.98 * n


def test_smoke_and_pressure_zero_range_guards():
    """With smoke on, a zero min (or max) pressure skips that division
    (src/graphics_handler.cu:227-232): p < 0 with min 0 falls through to
    p / max; p >= 0 with max 0 leaves norm 0 (green)."""
    assert orc.pixel(0, 1, 1, 1.0, -0.5, 0.0, 1.0) == rgba(0, 255, 255)  # norm -0.5
    assert orc.pixel(0, 1, 1, 1.0, 0.5, -1.0, 0.0) == rgba(0, 255, 0)    # norm 0
    assert orc.pixel(0, 1, 1, 0.5, 0.0, 0.0, 0.0) == rgba(0, 127, 0)     # value 0.5


def test_pressure_only_zero_range_is_red_not_undefined():
    """Without smoke the reference divides unguarded: 0 / 0 is NaN, the
    clamp keeps it, every hue test fails and the last sector gives
    (c, 0, NaN): red with the NaN channel converted to 0."""
    assert orc.pixel(0, 1, 0, 0.0, 0.0, 0.0, 0.0) == rgba(255, 0, 0)


def smoke_frame_closed_form(solid, smoke):
    c = (255 - np.trunc(smoke.astype(np.float32) * np.float32(255)).astype(np.int64)) & 255
    px = (255 << 24) | (c << 16) | (c << 8) | 255
    return np.where(solid != 0, SOLID, px).astype(np.uint32)


def test_frame_layout_and_smoke_mode():
    """The frame is in the reference's indx order (row 0 = top of the
    screen = j = H - 1), one pixel per cell."""
    cfg = make_config(61, 37, enable_pressure=0, enable_smoke=1)
    o = OracleFluid(cfg)
    rng = np.random.Generator(np.random.PCG64(4))
    o.smoke[...] = rng.uniform(0, 1, o.smoke.shape).astype(np.float32)
    img = o.render_rgba()
    assert img.shape == (37, 61) and img.dtype == np.uint32
    assert np.array_equal(img, smoke_frame_closed_form(o.solid, o.smoke))
    assert (img[0] == SOLID).all() and (img[-1] == SOLID).all() and (img[:, 0] == SOLID).all()


def test_frame_pressure_mode_uses_range_of_last_update():
    cfg = make_config(48, 30, enable_pressure=1, enable_smoke=0, wind_speed=2.0, g=-5.0,
                      projection_n=10)
    o = OracleFluid(cfg)
    o.update(steps=2)
    pmin, pmax = o.pressure_range
    assert pmin < 0 < pmax
    img = o.render_rgba()
    for r, c in ((5, 5), (15, 30), (28, 40), (10, 3)):
        want = SOLID if o.solid[r, c] else orc.pixel(0, 1, 0, 0.0, float(o.p[r, c]), pmin, pmax)
        assert img[r, c] == want


def test_frame_no_mode():
    cfg = make_config(20, 12, enable_pressure=0, enable_smoke=0)
    o = OracleFluid(cfg)
    img = o.render_rgba()
    assert np.array_equal(img, np.where(o.solid != 0, SOLID, 0).astype(np.uint32))


# ---- path lines -----------------------------------------------------------

def open_grid(W, H, **kw):
    return make_config(W, H, obstacle_enable=0, **kw)


@pytest.mark.parametrize("cs", [1, 2])
def test_trace_still_field_stays_at_rounded_centre(cs):
    cfg = open_grid(50, 30, cell_size=float(cs))
    o = OracleFluid(cfg)
    d, L = 7, 5
    tx, ty = o.traces(d, L)
    th, tw = 30 // d, 50 // d
    assert tx.shape == ty.shape == (th, tw, L)
    for tj in range(th):
        for ti in range(tw):
            i, j = ti * d, tj * d
            r = th - 1 - tj  # indx_traces puts the top row of seeds first
            if o.solid[30 - 1 - j, i]:
                assert (tx[r, ti] == -1).all() and (ty[r, ti] == -1).all()
                continue
            # centre (a + 0.5) * cs, rounded half away from zero
            x = int(np.floor((i + 0.5) * cs + 0.5))
            y = int(np.floor((j + 0.5) * cs + 0.5))
            assert (tx[r, ti] == x).all() and (ty[r, ti] == 30 - 1 - y).all()


def test_trace_uniform_flow_is_exact_straight_line():
    """u = 1, v = 0 everywhere: inside the fluid every bilinear weight is a
    quarter, so GX is exactly 1 and x advances by d_t = 0.25 per point;
    round half away from zero turns i + 0.5 into i + 1 and i + 2.5 into
    i + 3 (half to even would give i + 2)."""
    cfg = open_grid(64, 24)
    o = OracleFluid(cfg)
    o.u[...] = 1.0
    d, L = 8, 12
    tx, ty = o.traces(d, L, d_t=0.25)
    th, tw = 24 // d, 64 // d
    for tj in range(1, th):
        for ti in range(1, tw - 1):
            i, j = ti * d, tj * d
            r = th - 1 - tj
            xs = [i + 0.5 + 0.25 * k for k in range(L)]
            want_x = [int(np.floor(x + 0.5)) for x in xs]
            assert tx[r, ti].tolist() == want_x
            assert ty[r, ti].tolist() == [24 - 1 - (j + 1)] * L


def test_trace_vertical_flow_flips_y():
    cfg = open_grid(40, 64)
    o = OracleFluid(cfg)
    o.v[...] = 1.0
    d, L = 10, 9
    tx, ty = o.traces(d, L, d_t=0.5)
    th = 64 // d
    for tj in range(1, th - 1):
        j = tj * d
        r = th - 1 - tj
        ys = [j + 0.5 + 0.5 * k for k in range(L)]
        assert ty[r, 2].tolist() == [64 - 1 - int(np.floor(y + 0.5)) for y in ys]
        assert (tx[r, 2] == 21).all()


def test_trace_stops_in_solid_and_composes_samplers():
    """A path that enters the obstacle stops there (GX = GY = 0 in a solid
    cell), and on a random field every point is the previous one plus
    (GX, GY) * d_t in float32, the samplers being the oracle's GX / GY."""
    cfg = make_config(48, 32, obstacle_center_x=30, obstacle_center_y=16, obstacle_radius=5.0)
    o = OracleFluid(cfg)
    o.u[...] = 2.0
    tx, ty = o.traces(4, 30, d_t=0.25)
    row = 8 - 1 - 4  # seeds at j = 16
    xs = tx[row, 5]  # seed i = 20, heading into the disc at x ~ 25
    assert xs[-1] == xs[-2] and 24 <= xs[-1] <= 27

    rng = np.random.Generator(np.random.PCG64(5))
    o.u[...] = rng.uniform(-3, 3, o.u.shape).astype(np.float32)
    o.v[...] = rng.uniform(-3, 3, o.v.shape).astype(np.float32)
    dt = np.float32(0.3)
    tx, ty = o.traces(6, 8, d_t=float(dt))
    for tj, ti in ((1, 1), (2, 3), (4, 6)):
        i, j = ti * 6, tj * 6
        x, y = np.float32(i + 0.5), np.float32(j + 0.5)
        r = 32 // 6 - 1 - tj
        for k in range(8):
            if k:
                vx, vy = np.float32(o.gx(float(x), float(y))), np.float32(o.gy(float(x), float(y)))
                x, y = np.float32(x + vx * dt), np.float32(y + vy * dt)
            assert tx[r, ti, k] == int(np.sign(x) * np.floor(abs(x) + 0.5))
            assert ty[r, ti, k] == 31 - int(np.sign(y) * np.floor(abs(y) + 0.5))


def test_trace_rejects_bad_arguments():
    o = OracleFluid(open_grid(20, 12))
    with pytest.raises(ValueError):
        o.traces(0, 4)
    with pytest.raises(ValueError):
        o.traces(3, 0)
