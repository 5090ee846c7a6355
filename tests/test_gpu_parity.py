"""Parity of the HIP solver with the CPU oracle, through the C ABI.

FP32 storage is held to bit equality (the kernels evaluate the reference's
update rule in the same IEEE order as the oracle); the stated tolerances,
1e-6 relative after one step and 1e-4 relative L-inf after 100 steps, are
checked as well so a failure says which bar broke.
"""
import os

import numpy as np
import pytest

from opensayal_amd import Fluid, OpenSayalError, Source, workloads as wl
from oracle.oracle import OracleFluid
from tests.helpers import make_config, rel_linf

pytestmark = pytest.mark.gpu

NT = min(os.cpu_count() or 1, 16)
FIELDS = (("vel_x", "u"), ("vel_y", "v"), ("smoke", "smoke"))


def pair(cfg, seed=1):
    fl = Fluid(cfg)
    o = OracleFluid(cfg, nthreads=NT)
    u, v, s = wl.initial_state(cfg, seed)
    for name, arr in (("vel_x", u), ("vel_y", v), ("smoke", s)):
        fl.set_field(name, arr)
    o.set_state(u=u, v=v, smoke=s)
    return fl, o


def assert_bitwise(fl, o, what=""):
    for name, attr in FIELDS + ((("pressure", "p"),) if o.cfg.enable_pressure else ()):
        g = fl.get_field(name)
        ref = getattr(o, attr)
        if not np.array_equal(g.view(np.uint32), ref.view(np.uint32)):
            bad = np.argwhere(g.view(np.uint32) != ref.view(np.uint32))
            r, i = bad[0]
            pytest.fail(f"{what} {name}: {len(bad)} cells differ; first at i={i} "
                        f"j={o.cfg.height - 1 - r}: gpu={g[r, i]!r} oracle={ref[r, i]!r}; "
                        f"rel_linf={rel_linf(g, ref):.3e}")


MASK_CASES = [
    wl.plain_small(),
    wl.plain_full(),
    make_config(77, 53, obstacle_center_x=20, obstacle_center_y=30, obstacle_radius=12.5),
    make_config(130, 40, enable_drain=0, obstacle_radius=0.0),
    make_config(64, 64, obstacle_center_x=-5, obstacle_center_y=70, obstacle_radius=40.0),
    make_config(200, 120, obstacle_radius=17.0, cell_size=2.0),
]


@pytest.mark.parametrize("k", range(len(MASK_CASES)))
def test_masks_bit_exact(k):
    cfg = MASK_CASES[k]
    fl = Fluid(cfg)
    o = OracleFluid(cfg)
    assert np.array_equal(fl.get_field("is_solid"), o.solid)
    assert np.array_equal(fl.get_field("total_s"), o.total_s)


@pytest.mark.parametrize("pressure", [0, 1])
def test_small_one_step_bit_exact(pressure):
    cfg = wl.plain_small(pressure=bool(pressure))
    fl, o = pair(cfg)
    fl.update()
    o.update()
    for name, attr in FIELDS:
        assert rel_linf(fl.get_field(name), getattr(o, attr)) <= 1e-6
    assert_bitwise(fl, o, "1 step")
    if pressure:
        assert fl.pressure_range == o.pressure_range


def test_small_100_steps():
    cfg = wl.plain_small()
    fl, o = pair(cfg, seed=7)
    fl.step(100)
    o.update(steps=100)
    for name, attr in FIELDS:
        assert rel_linf(fl.get_field(name), getattr(o, attr)) <= 1e-4
    assert_bitwise(fl, o, "100 steps")


def test_full_size_steps_bit_exact():
    cfg = wl.plain_full()
    fl, o = pair(cfg, seed=3)
    fl.update()
    o.update()
    assert_bitwise(fl, o, "1 step")
    fl.step(4)
    o.update(steps=4)
    assert_bitwise(fl, o, "5 steps")


@pytest.mark.parametrize("case", ["drag_decay", "stripes", "odd_cs2", "narrow", "drain_closed"])
def test_config_variants_bit_exact(case):
    kw = {
        "drag_decay": dict(width=96, height=64, drag_coeff=0.3, enable_decay=1, decay_rate=0.4,
                           wind_speed=1.5, g=-2.0, enable_pressure=1),
        "stripes": dict(width=128, height=90, pipe_height=60, smoke_count=4, smoke_height=6,
                        smoke_length=3, wind_speed=2.5, g=-1.0),
        "odd_cs2": dict(width=77, height=53, cell_size=2.0, wind_speed=3.0, g=-5.0,
                        obstacle_center_x=30, obstacle_center_y=25, obstacle_radius=9.0),
        "narrow": dict(width=9, height=200, wind_speed=1.0, g=-5.0, projection_n=20),
        "drain_closed": dict(width=150, height=70, enable_drain=0, wind_speed=2.0, g=-5.0,
                             projection_o=1.2, projection_n=37),
    }[case]
    W, H = kw.pop("width"), kw.pop("height")
    cfg = make_config(W, H, **kw)
    fl, o = pair(cfg, seed=11)
    fl.step(6)
    o.update(steps=6)
    assert_bitwise(fl, o, case)


@pytest.mark.parametrize("shape", [(3, 3), (3, 67), (67, 3), (65, 33), (129, 4)])
def test_smallest_and_ragged_grids_bit_exact(shape):
    """Grids of one interior row or column, and grids one cell past a tile
    (64 wide, 32 tall): the advections' staged halo is mostly outside the
    grid, and the extrapolated borders are read from rows and columns 1 and
    W-2 / H-2, which here are the same cells."""
    W, H = shape
    cfg = make_config(W, H, wind_speed=1.5, g=-3.0, enable_pressure=1, projection_n=7,
                      pipe_height=max(1, H // 2), smoke_height=max(1, H // 4),
                      obstacle_radius=0.0)
    fl, o = pair(cfg, seed=13)
    fl.step(5)
    o.update(steps=5)
    assert_bitwise(fl, o, f"{W}x{H}")


def test_active_source_step_bit_exact():
    cfg = make_config(160, 100, wind_speed=1.0, g=-1.0)
    fl, o = pair(cfg, seed=5)
    src = Source(active=1, smoke=0.8, velocity=0.05, position_x=70, position_y=45)
    for _ in range(3):
        fl.update(src)
        o.update(source=src)
    fl.update()
    o.update()
    assert_bitwise(fl, o, "source")


def test_projection_only_matches_oracle():
    cfg = wl.plain_small(pressure=True)
    fl, o = pair(cfg, seed=9)
    fl.set_field("pressure", np.zeros((cfg.height, cfg.width), np.float32))
    fl.project(sweeps=13)
    o.projection(sweeps=13)
    assert_bitwise(fl, o, "projection")


def test_snapshot_restore_continues_identically():
    cfg = wl.plain_small()
    a, _ = pair(cfg, seed=2)
    a.step(5)
    snap = {n: a.get_field(n) for n, _ in FIELDS}
    a.step(7)
    b = Fluid(cfg)
    for n, arr in snap.items():
        b.set_field(n, arr)
    b.step(7)
    for n, _ in FIELDS:
        assert np.array_equal(a.get_field(n), b.get_field(n))


def test_update_and_graph_step_agree():
    cfg = wl.plain_small()
    a, _ = pair(cfg, seed=4)
    b, _ = pair(cfg, seed=4)
    for _ in range(5):
        a.update()
    b.step(5)
    for n, _ in FIELDS:
        assert np.array_equal(a.get_field(n), b.get_field(n))


def test_divergence_norm_matches_oracle():
    cfg = wl.plain_small()
    fl, o = pair(cfg, seed=6)
    fl.step(3)
    o.update(steps=3)
    g = fl.divergence_norm()
    r = o.divergence_norm()
    assert g[0] == pytest.approx(r[0], rel=1e-12)
    assert g[1] == r[1]


@pytest.mark.parametrize("storage", ["f16", "bf16"])
def test_half_storage_tracks_fp64(storage):
    """u, v, smoke stored in 16 bits, arithmetic in FP32: the error against
    the FP64 oracle stays small after 10 steps (bound stated in the test)."""
    cfg = wl.plain_small(storage=storage)
    fl, _ = pair(cfg, seed=8)
    cfg64 = wl.plain_small()
    o = OracleFluid(cfg64, np.float64, nthreads=NT)
    u, v, s = wl.initial_state(cfg64, 8)
    o.set_state(u=u, v=v, smoke=s)
    fl.step(10)
    o.update(steps=10)
    bound = {"f16": 2e-2, "bf16": 1e-1}[storage]
    for name, attr in FIELDS:
        err = rel_linf(fl.get_field(name), getattr(o, attr))
        assert err < bound, (name, err)


TUNINGS = [
    dict(projection_kernel=0),
    dict(projection_kernel=1, tile_width=240, tile_height=24, sweeps_per_launch=4),
    dict(projection_kernel=1, tile_width=56, tile_height=16, sweeps_per_launch=2),
    dict(projection_kernel=1, tile_width=84, tile_height=28, sweeps_per_launch=3),
    dict(projection_kernel=1, tile_width=160, tile_height=12, sweeps_per_launch=8),
    dict(projection_kernel=1, tile_width=2, tile_height=1, sweeps_per_launch=1),
    dict(projection_kernel=1, tile_width=480, tile_height=8, sweeps_per_launch=2),
    dict(projection_kernel=2, tile_width=120, tile_height=24, sweeps_per_launch=2),
    dict(projection_kernel=2, tile_width=112, tile_height=48, sweeps_per_launch=4),
    dict(projection_kernel=2, tile_width=116, tile_height=52, sweeps_per_launch=3),
    dict(projection_kernel=2, tile_width=124, tile_height=28, sweeps_per_launch=1),
]


@pytest.mark.parametrize("t", range(len(TUNINGS)))
@pytest.mark.parametrize("pressure", [0, 1])
def test_projection_shapes_bit_exact(t, pressure):
    """Every launch shape gives the oracle's bits (n = 13 is not a multiple
    of any k above; the 77x53 grid clips tiles on both edges)."""
    cfg = make_config(77, 53, projection_n=13, enable_pressure=pressure, wind_speed=2.0,
                      g=-5.0, obstacle_center_x=30, obstacle_center_y=20, obstacle_radius=8.5)
    fl, o = pair(cfg, seed=12)
    fl.set_tuning(**TUNINGS[t])
    if pressure:
        p0 = np.random.default_rng(3).uniform(-1, 1, (53, 77)).astype(np.float32)
        fl.set_field("pressure", p0)
        o.set_state(p=p0)
    fl.project()
    o.projection()
    assert_bitwise(fl, o, f"projection {TUNINGS[t]}")
    fl.step(3)
    o.update(steps=3)
    assert_bitwise(fl, o, f"steps {TUNINGS[t]}")


# register tiles taller than 8 rows per wavefront (pressure off only: the
# pressure sums would not fit the registers)
TALL_TUNINGS = [
    dict(projection_kernel=2, tile_width=104, tile_height=56, sweeps_per_launch=6),  # 10 rows
    dict(projection_kernel=2, tile_width=108, tile_height=60, sweeps_per_launch=5),  # 10 rows
    dict(projection_kernel=2, tile_width=100, tile_height=68, sweeps_per_launch=7),  # 12 rows
    dict(projection_kernel=2, tile_width=108, tile_height=108, sweeps_per_launch=5),  # 16 rows
]


@pytest.mark.parametrize("t", range(len(TALL_TUNINGS)))
def test_tall_register_tiles_bit_exact(t):
    """Tall register tiles on a grid several tiles tall and wide, n = 13 (not
    a multiple of k), and at full size."""
    cfg = make_config(250, 230, projection_n=13, wind_speed=2.0, g=-5.0, obstacle_center_x=90,
                      obstacle_center_y=110, obstacle_radius=20.0)
    fl, o = pair(cfg, seed=14)
    fl.set_tuning(**TALL_TUNINGS[t])
    fl.step(3)
    o.update(steps=3)
    assert_bitwise(fl, o, f"steps {TALL_TUNINGS[t]}")
    with pytest.raises(OpenSayalError):
        Fluid(make_config(64, 36, enable_pressure=1)).set_tuning(**TALL_TUNINGS[t])


def test_tall_register_tile_full_size():
    cfg = wl.plain_full()
    fl, o = pair(cfg, seed=15)
    fl.set_tuning(**TALL_TUNINGS[0])
    fl.step(2)
    o.update(steps=2)
    assert_bitwise(fl, o, "1920x1080, 10 rows per wavefront")


@pytest.mark.parametrize("t", [1, 4, 8])
def test_full_size_tiled_steps_bit_exact(t):
    cfg = wl.plain_full()
    fl, o = pair(cfg, seed=13)
    fl.set_tuning(**TUNINGS[t])
    fl.step(3)
    o.update(steps=3)
    assert_bitwise(fl, o, f"full {TUNINGS[t]}")


def test_autotune_keeps_state_and_results():
    """The autotuner projects a copy of the state: the fields are unchanged
    afterwards, and stepping with the chosen shape still matches the oracle;
    a second handle of the same shape adopts the cached choice."""
    cfg = make_config(300, 170, wind_speed=2.0, g=-5.0, projection_n=12,
                      obstacle_center_x=150, obstacle_center_y=85, obstacle_radius=9.0)
    fl, o = pair(cfg, seed=14)
    before = {n: fl.get_field(n) for n, _ in FIELDS}
    tuning, ms = fl.autotune()
    assert ms > 0 and tuning["projection_kernel"] in (0, 1, 2, 3)
    for n, _ in FIELDS:
        assert np.array_equal(fl.get_field(n), before[n])
    fl.step(3)
    o.update(steps=3)
    assert_bitwise(fl, o, f"after autotune {tuning}")
    other = Fluid(cfg)
    assert other.autotune_cached()
    assert other.tuning == tuning


def test_invalid_tuning_rejected():
    fl = Fluid(make_config(64, 36))
    with pytest.raises(OpenSayalError):
        fl.set_tuning(projection_kernel=1, tile_width=15, tile_height=8, sweeps_per_launch=4)
    with pytest.raises(OpenSayalError):
        fl.set_tuning(projection_kernel=1, tile_width=400, tile_height=200, sweeps_per_launch=2)
    with pytest.raises(OpenSayalError):
        fl.set_tuning(projection_kernel=1, tile_width=1100, tile_height=4, sweeps_per_launch=2)


def test_errors_are_reported():
    with pytest.raises(OpenSayalError) as e:
        Fluid(make_config(64, 36, viscosity=float("inf")))
    assert "viscosity" in str(e.value)
    fl = Fluid(make_config(64, 36))
    with pytest.raises(OpenSayalError):
        fl.set_field("is_solid", np.zeros((36, 64), np.float32))
    with pytest.raises(ValueError):
        fl.set_field("vel_x", np.zeros((10, 10), np.float32))


def test_library_stream_and_events():
    """Steps enqueued on a library Stream give the same fields as synchronous
    steps, and events recorded around them time a positive interval."""
    from opensayal_amd import Event, Stream

    cfg = wl.plain_small()
    a, _ = pair(cfg, seed=15)
    b, _ = pair(cfg, seed=15)
    s = Stream()
    e0, e1 = Event(), Event()
    e0.record(s)
    a.step_async(4, stream=s)
    e1.record(s)
    s.synchronize()
    assert e0.elapsed_ms(e1) > 0
    b.step(4)
    for n, _ in FIELDS:
        assert np.array_equal(a.get_field(n), b.get_field(n))
    s.close()


@pytest.mark.parametrize("storage", ["f16", "bf16"])
def test_half_storage_bit_exact_with_oracle(storage):
    """16-bit storage: every phase's outputs rounded to the storage type,
    FP32 arithmetic; the oracle's storage mode gives the same bits."""
    cfg = wl.plain_small(storage=storage, pressure=True)
    fl, o = pair(cfg, seed=16)
    fl.step(5)
    o.update(steps=5)
    assert_bitwise(fl, o, f"{storage} 5 steps")


@pytest.mark.parametrize("t", [0, 1, 2, 7, 8])
@pytest.mark.parametrize("storage", [1, 2])
def test_half_storage_shapes_bit_exact(t, storage):
    """Rounding per phase makes 16-bit results independent of the launch
    shape (sweeps per launch, tile), so every shape matches the oracle."""
    cfg = make_config(77, 53, projection_n=13, wind_speed=2.0, g=-5.0, obstacle_center_x=30,
                      obstacle_center_y=20, obstacle_radius=8.5, storage=storage,
                      enable_decay=1, decay_rate=0.2)
    fl, o = pair(cfg, seed=17)
    fl.set_tuning(**TUNINGS[t])
    fl.step(3)
    o.update(steps=3)
    assert_bitwise(fl, o, f"storage {storage} {TUNINGS[t]}")
