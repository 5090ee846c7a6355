"""Frame export and path lines on the GPU, bit-exact against the oracle.

The frame (update_fluid_pixels_kernel) and the path lines (Fluid::trace over
update_traces_kernel's seed grid) are integer outputs: the GPU's must equal
the oracle's on the same state, which the oracle computes from the fields the
GPU exports (so the comparison isolates the kernels) and, end to end, from
its own run of the same steps.
"""
import numpy as np
import pytest

from opensayal_amd import Batch, Fluid, OpenSayalError, workloads as wl
from oracle import oracle as orc
from oracle.oracle import OracleFluid
from tests.helpers import make_config

pytestmark = pytest.mark.gpu


def started(cfg, seed, steps=0):
    fl = wl.start(wl.SimSpec(cfg, seed, 0))
    if steps:
        fl.step(steps)
    return fl


def oracle_frame(fl):
    cfg = fl.config
    solid = fl.get_field("is_solid")
    pmin, pmax = fl.pressure_range
    return orc.render_rgba(cfg, solid, fl.get_field("smoke"), fl.get_field("pressure"), pmin,
                           pmax)


def oracle_traces(fl, distance, length, d_t):
    return orc.trace_grid(fl.config, fl.get_field("is_solid"), fl.get_field("total_s"),
                          fl.get_field("vel_x"), fl.get_field("vel_y"), distance, length, d_t)


def assert_frames_equal(got, want, what):
    if not np.array_equal(got, want):
        bad = np.argwhere(got != want)
        r, c = bad[0]
        pytest.fail(f"{what}: {len(bad)} pixels differ, first at {(r, c)}: "
                    f"{got[r, c]:#010x} != {want[r, c]:#010x}")


FRAME_CONFIGS = [
    ("smoke", dict(enable_pressure=0, enable_smoke=1)),
    ("pressure+smoke", dict(enable_pressure=1, enable_smoke=1)),
    ("pressure", dict(enable_pressure=1, enable_smoke=0)),
    ("none", dict(enable_pressure=0, enable_smoke=0)),
    ("f16 pressure+smoke", dict(enable_pressure=1, enable_smoke=1, storage=1)),
    ("bf16 smoke", dict(enable_pressure=0, enable_smoke=1, storage=2)),
    ("cs2 pressure+smoke", dict(enable_pressure=1, enable_smoke=1, cell_size=2.0)),
]


@pytest.mark.parametrize("name,kw", FRAME_CONFIGS, ids=[n for n, _ in FRAME_CONFIGS])
def test_frame_matches_oracle_on_same_state(name, kw):
    cfg = make_config(203, 117, projection_n=20, wind_speed=2.0, g=-5.0, decay_rate=0.3,
                      enable_decay=1, obstacle_center_x=70, obstacle_center_y=60,
                      obstacle_radius=12.0, **kw)
    fl = started(cfg, 41, steps=5)
    assert_frames_equal(fl.render_rgba(), oracle_frame(fl), name)


def test_frame_full_size():
    fl = started(wl.plain_full(), 7, steps=3)
    img = fl.render_rgba()
    assert img.shape == (1080, 1920)
    assert_frames_equal(img, oracle_frame(fl), "1920x1080")


def test_frame_end_to_end_with_oracle_run():
    cfg = make_config(96, 64, projection_n=15, enable_pressure=1, wind_speed=2.0, g=-5.0)
    fl = Fluid(cfg)
    o = OracleFluid(cfg)
    fl.step(4)
    o.update(steps=4)
    assert_frames_equal(fl.render_rgba(), o.render_rgba(), "after 4 steps")


def test_frame_saturates_out_of_range_smoke():
    """Smoke outside [0, 1] and NaN: the byte conversions saturate (NaN to 0)
    on both sides."""
    cfg = make_config(64, 40, enable_pressure=0, enable_smoke=1)
    fl = Fluid(cfg)
    s = np.linspace(-2, 3, 64 * 40, dtype=np.float32).reshape(40, 64)
    s[10, 10:20] = np.nan
    s[11, 10:20] = np.inf
    fl.set_field("smoke", s)
    assert_frames_equal(fl.render_rgba(), oracle_frame(fl), "saturation")


def test_frame_of_batch_member_equals_alone():
    cfg = make_config(150, 70, projection_n=17, enable_pressure=1, wind_speed=2.0, g=-5.0)
    member, alone = started(cfg, 3), started(cfg, 3)
    with Batch() as b:
        b.add(member)
        b.add(started(make_config(64, 36, projection_n=5), 4))
        b.step(3)
        alone.step(3)
        assert_frames_equal(member.render_rgba(), alone.render_rgba(), "batch member")


TRACE_CASES = [
    # (cfg kwargs, distance, length, d_t)
    (dict(), 20, 20, None),
    (dict(), 7, 33, 0.37),
    (dict(), 1, 3, 0.05),
    (dict(cell_size=2.0), 9, 25, 0.2),
    (dict(storage=1), 11, 16, 0.5),
    (dict(storage=2), 13, 16, 0.5),
]


@pytest.mark.parametrize("kw,distance,length,d_t", TRACE_CASES)
def test_traces_match_oracle_on_same_state(kw, distance, length, d_t):
    cfg = make_config(211, 97, projection_n=20, wind_speed=3.0, g=-5.0,
                      obstacle_center_x=80, obstacle_center_y=50, obstacle_radius=15.0, **kw)
    fl = started(cfg, 17, steps=6)
    dt = cfg.d_t if d_t is None else d_t
    tx, ty = fl.traces(distance, length, d_t=dt)
    wx, wy = oracle_traces(fl, distance, length, dt)
    assert tx.shape == wx.shape == (97 // distance, 211 // distance, length)
    assert np.array_equal(tx, wx), np.argwhere(tx != wx)[:3]
    assert np.array_equal(ty, wy), np.argwhere(ty != wy)[:3]
    if distance <= 20:
        assert (tx[:, 0, :] == -1).all()  # seeds on the left wall are solid


def test_traces_full_size_defaults():
    fl = started(wl.plain_full(), 9, steps=2)
    tx, ty = fl.traces()
    wx, wy = oracle_traces(fl, 20, 20, fl.config.d_t)
    assert tx.shape == (54, 96, 20)
    assert np.array_equal(tx, wx) and np.array_equal(ty, wy)


def test_traces_extreme_velocities():
    """Huge, infinite and NaN velocities: positions leave the grid, become
    infinite or NaN; the int conversion saturates (NaN to 0) on both sides."""
    cfg = make_config(80, 50, obstacle_enable=0)
    fl = Fluid(cfg)
    rng = np.random.Generator(np.random.PCG64(8))
    u = rng.uniform(-5, 5, (50, 80)).astype(np.float32)
    v = rng.uniform(-5, 5, (50, 80)).astype(np.float32)
    u[5:10, 10:30] = 1e30
    v[20:25, 40:60] = -1e30
    u[30:35, 5:20] = np.inf
    v[40:45, 50:70] = np.nan
    fl.set_field("vel_x", u)
    fl.set_field("vel_y", v)
    for distance, length, dt in ((3, 12, 0.5), (5, 8, 1e10)):
        tx, ty = fl.traces(distance, length, d_t=dt)
        wx, wy = oracle_traces(fl, distance, length, dt)
        assert np.array_equal(tx, wx) and np.array_equal(ty, wy), (distance, length, dt)


def test_traces_of_batch_member_equal_alone():
    cfg = make_config(120, 90, projection_n=12, wind_speed=2.5, g=-1.0)
    member, alone = started(cfg, 5), started(cfg, 5)
    with Batch() as b:
        b.add(member)
        b.step(3)
        alone.step(3)
        a, c = member.traces(6, 10), alone.traces(6, 10)
        assert np.array_equal(a[0], c[0]) and np.array_equal(a[1], c[1])


def test_view_argument_errors():
    fl = Fluid(make_config(40, 30))
    with pytest.raises(ValueError):
        fl.traces(0, 5)
    with pytest.raises(OpenSayalError):
        fl.traces(4, 0)
    with pytest.raises(OpenSayalError):
        fl.traces(4, 65537)
    with pytest.raises(OpenSayalError):
        fl.traces(4, 5, d_t=float("nan"))
    tx, ty = fl.traces(50, 5)  # no seed fits: empty arrays, nothing launched
    assert tx.size == 0 and ty.size == 0
