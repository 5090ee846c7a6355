"""The resident projection (one launch per projection: every tile's
workgroup keeps its region in registers for all n sweeps and exchanges only
its halo with the adjacent tiles between stages of k sweeps) against the
oracle, bit for bit: multi-tile grids in both directions, n not a multiple
of k, one stage and many, pressure sums, 16-bit storage, the full-size
1920x1080 workload, and the shapes it must refuse."""
import numpy as np
import pytest

from opensayal_amd import Fluid, OpenSayalError, workloads as wl
from tests.helpers import make_config
from tests.test_gpu_parity import assert_bitwise, pair

pytestmark = pytest.mark.gpu

RESIDENT = [
    dict(projection_kernel=3, tile_width=108, tile_height=44, sweeps_per_launch=5),  # 8 rows
    dict(projection_kernel=3, tile_width=120, tile_height=24, sweeps_per_launch=2),  # 4 rows
    dict(projection_kernel=3, tile_width=124, tile_height=28, sweeps_per_launch=1),  # 4 rows
    dict(projection_kernel=3, tile_width=116, tile_height=52, sweeps_per_launch=3),  # 8 rows
    dict(projection_kernel=3, tile_width=104, tile_height=40, sweeps_per_launch=6),  # 8 rows
]
TALL = [
    dict(projection_kernel=3, tile_width=108, tile_height=60, sweeps_per_launch=5),  # 10 rows
    dict(projection_kernel=3, tile_width=100, tile_height=68, sweeps_per_launch=7),  # 12 rows
    dict(projection_kernel=3, tile_width=108, tile_height=108, sweeps_per_launch=5),  # 16 rows
]


def obstacle_grid(W, H, **kw):
    return make_config(W, H, wind_speed=2.0, g=-5.0, obstacle_center_x=W * 2 // 5,
                       obstacle_center_y=H * 9 // 20, obstacle_radius=min(W, H) * 0.12, **kw)


@pytest.mark.parametrize("t", range(len(RESIDENT)))
@pytest.mark.parametrize("pressure", [0, 1])
def test_resident_shapes_bit_exact(t, pressure):
    """3 x 6 tiles or more (tiles clipped on the right and top edges),
    n = 13 (a short last stage), projection alone and then whole steps."""
    cfg = obstacle_grid(250, 230, projection_n=13, enable_pressure=pressure)
    fl, o = pair(cfg, seed=21)
    fl.set_tuning(**RESIDENT[t])
    assert fl.projection_kernel == 3
    if pressure:
        p0 = np.random.default_rng(4).uniform(-1, 1, (230, 250)).astype(np.float32)
        fl.set_field("pressure", p0)
        o.set_state(p=p0)
    fl.project()
    o.projection()
    assert_bitwise(fl, o, f"projection {RESIDENT[t]}")
    fl.step(3)
    o.update(steps=3)
    assert_bitwise(fl, o, f"steps {RESIDENT[t]}")


@pytest.mark.parametrize("t", range(len(TALL)))
def test_resident_tall_tiles_bit_exact(t):
    cfg = obstacle_grid(250, 230, projection_n=13)
    fl, o = pair(cfg, seed=22)
    fl.set_tuning(**TALL[t])
    fl.step(3)
    o.update(steps=3)
    assert_bitwise(fl, o, f"steps {TALL[t]}")


@pytest.mark.parametrize("n", [1, 2, 5, 50])
def test_resident_stage_counts(n):
    """One stage (n <= k), a stage and a short one, and 50 one-sweep stages
    (49 halo exchanges per launch, flags cleared between launches)."""
    cfg = obstacle_grid(300, 170, projection_n=n)
    fl, o = pair(cfg, seed=23)
    fl.set_tuning(**(RESIDENT[2] if n == 50 else RESIDENT[0]))
    fl.step(4)
    o.update(steps=4)
    assert_bitwise(fl, o, f"n={n}")


def test_resident_single_tile_and_ragged():
    """A grid smaller than one tile, and one a few cells over a tile edge."""
    for W, H in ((64, 36), (110, 45)):
        cfg = obstacle_grid(W, H, projection_n=11)
        fl, o = pair(cfg, seed=24)
        fl.set_tuning(**RESIDENT[0])
        fl.step(3)
        o.update(steps=3)
        assert_bitwise(fl, o, f"{W}x{H}")


def test_resident_full_size_bit_exact():
    """The 1920x1080 n = 50 workload: 450 tiles, 10 stages per step."""
    cfg = wl.plain_full()
    fl, o = pair(cfg, seed=25)
    fl.set_tuning(**RESIDENT[0])
    fl.step(3)
    o.update(steps=3)
    assert_bitwise(fl, o, "1920x1080 resident")


@pytest.mark.parametrize("storage", ["f16", "bf16"])
def test_resident_half_storage_bit_exact(storage):
    cfg = wl.plain_small(storage=storage, pressure=True)
    fl, o = pair(cfg, seed=26)
    fl.set_tuning(**RESIDENT[1])
    fl.step(4)
    o.update(steps=4)
    assert_bitwise(fl, o, f"{storage} resident")


@pytest.mark.parametrize("storage", [0, 1, 2])
def test_resident_odd_width_bit_exact(storage):
    """An odd width: the last column's cells are stored (and their rings
    exchanged) one element at a time, in every storage type."""
    cfg = obstacle_grid(251, 131, projection_n=12, storage=storage)
    fl, o = pair(cfg, seed=29)
    fl.set_tuning(**RESIDENT[3])
    fl.step(3)
    o.update(steps=3)
    assert_bitwise(fl, o, f"251x131 storage {storage}")


def test_resident_agrees_with_launch_per_stage():
    """The same shape as one resident launch and as launches of k sweeps
    give the same bits over many steps (no oracle: 40 steps at 640x360)."""
    cfg = obstacle_grid(640, 360, projection_n=20)
    a, _ = pair(cfg, seed=27)
    b, _ = pair(cfg, seed=27)
    a.set_tuning(**RESIDENT[0])
    b.set_tuning(**dict(RESIDENT[0], projection_kernel=2))
    a.step(40)
    b.step(40)
    for name in ("vel_x", "vel_y", "smoke"):
        assert np.array_equal(a.get_field(name).view(np.uint32),
                              b.get_field(name).view(np.uint32)), name


def test_resident_refuses_shapes_that_do_not_fit():
    # tile sides below 2k: the halo would reach past the adjacent tiles
    with pytest.raises(OpenSayalError):
        Fluid(make_config(256, 128)).set_tuning(projection_kernel=3, tile_width=80,
                                                tile_height=16, sweeps_per_launch=12)
    Fluid(make_config(256, 128)).set_tuning(projection_kernel=2, tile_width=80,
                                            tile_height=16, sweeps_per_launch=12)
    # more tiles than workgroups the GPU holds at once (36 x 74 = 2664 tiles
    # of 4 rows per wavefront, at most 3 workgroups on each of 256 CUs)
    fl = Fluid(make_config(4400, 2070))
    with pytest.raises(OpenSayalError, match="fit"):
        fl.set_tuning(projection_kernel=3, tile_width=124, tile_height=28, sweeps_per_launch=1)
    fl.set_tuning(projection_kernel=2, tile_width=124, tile_height=28, sweeps_per_launch=1)


def test_resident_wait_timeout_is_reported_and_recovers():
    """A neighbour wait bounded at 0 times out: the launch drains without
    waiting, the step reports OSY_E_STATE, and with the bound restored the
    handle steps exactly again (the flags were cleared, the error word
    read and cleared)."""
    cfg = wl.plain_full()
    fl, o = pair(cfg, seed=28)
    fl.set_tuning(**RESIDENT[0])
    fl.set_wait_limit(0)
    with pytest.raises(OpenSayalError, match="timed out"):
        fl.step(1)
    fl.set_wait_limit(2_000_000)
    u, v, s = wl.initial_state(cfg, 28)
    for name, arr in (("vel_x", u), ("vel_y", v), ("smoke", s)):
        fl.set_field(name, arr)
    fl.step(2)
    o.update(steps=2)
    assert_bitwise(fl, o, "after a timed-out launch")
    with pytest.raises(OpenSayalError):
        fl.set_wait_limit(-1)
    # a timeout of an asynchronous step is reported by the next synchronize
    # even when the launch shape changed in between
    fl.set_wait_limit(0)
    fl.step_async(1)
    fl.set_tuning(**dict(RESIDENT[0], projection_kernel=2))
    with pytest.raises(OpenSayalError, match="timed out"):
        fl.synchronize()
    fl.synchronize()
