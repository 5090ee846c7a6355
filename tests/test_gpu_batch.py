"""The batched persistent step: every member's fields after k batch steps
are bit-identical to k steps of the same handle alone (which the parity
tests tie to the oracle), for uneven sizes, every stage kind, storage and
pressure groups, odd step counts, membership changes and every launch shape.
"""
import numpy as np
import pytest

from opensayal_amd import Batch, Fluid, OpenSayalError, device_bytes, workloads as wl
from oracle.oracle import OracleFluid
from tests.helpers import make_config

pytestmark = pytest.mark.gpu

FIELDS = ("vel_x", "vel_y", "smoke")


def mixed_configs():
    """Uneven grids covering every stage chain: n = 0 (forces stage), n < k,
    n not a multiple of k, smoke off (forces stage for the smoke), decay and
    drag, cell size 2, pressure, 16-bit storage, odd and narrow shapes."""
    return [
        make_config(77, 53, projection_n=13, wind_speed=2.0, g=-5.0, obstacle_center_x=30,
                    obstacle_center_y=20, obstacle_radius=8.5),
        make_config(256, 144, projection_n=50, wind_speed=2.0, g=-5.0, pipe_height=36,
                    obstacle_center_x=128, obstacle_center_y=72, obstacle_radius=10.0),
        make_config(130, 40, projection_n=0, wind_speed=1.0, g=-2.0),
        make_config(64, 36, projection_n=3, wind_speed=3.0, g=-5.0),
        make_config(96, 64, projection_n=20, drag_coeff=0.3, enable_decay=1, decay_rate=0.4,
                    wind_speed=1.5, g=-2.0),
        make_config(120, 90, projection_n=17, wind_smoke=0.0, wind_speed=2.5, g=-1.0),
        make_config(101, 61, projection_n=11, cell_size=2.0, wind_speed=3.0, g=-5.0,
                    obstacle_center_x=40, obstacle_center_y=30, obstacle_radius=9.0),
        make_config(9, 200, projection_n=20, wind_speed=1.0, g=-5.0),
        make_config(150, 70, projection_n=37, enable_pressure=1, enable_drain=0, wind_speed=2.0,
                    g=-5.0, projection_o=1.2),
        make_config(200, 113, projection_n=25, storage=1, wind_speed=2.0, g=-5.0),
        make_config(160, 90, projection_n=12, storage=2, enable_pressure=1, wind_speed=2.0,
                    g=-5.0),
        make_config(331, 187, projection_n=100, wind_speed=0.7, g=-5.0, obstacle_center_x=100,
                    obstacle_center_y=90, obstacle_radius=14.0),
    ]


def started(cfg, seed):
    return wl.start(wl.SimSpec(cfg, seed, 0))


def state(fl):
    out = {n: fl.get_field(n) for n in FIELDS}
    if fl.config.enable_pressure:
        out["pressure"] = fl.get_field("pressure")
    return out


def assert_same(a, b, what):
    for n in a:
        if not np.array_equal(a[n].view(np.uint32), b[n].view(np.uint32)):
            bad = np.argwhere(a[n].view(np.uint32) != b[n].view(np.uint32))
            pytest.fail(f"{what} {n}: {len(bad)} cells differ, first at {tuple(bad[0])}")


def pairs(cfgs, seed0=100):
    batched = [started(c, seed0 + i) for i, c in enumerate(cfgs)]
    alone = [started(c, seed0 + i) for i, c in enumerate(cfgs)]
    return batched, alone


@pytest.mark.parametrize("steps", [(1,), (2, 3)])
def test_batch_equals_alone(steps):
    cfgs = mixed_configs()
    batched, alone = pairs(cfgs)
    with Batch() as b:
        for fl in batched:
            b.add(fl)
        for s in steps:
            b.step(s)
        for fl in alone:
            fl.step(sum(steps))
        for i, (x, y) in enumerate(zip(batched, alone)):
            assert_same(state(x), state(y), f"sim {i} ({cfgs[i].width}x{cfgs[i].height})")
            if cfgs[i].enable_pressure:
                assert x.pressure_range == y.pressure_range


def test_batch_member_matches_oracle():
    cfg = make_config(77, 53, projection_n=13, wind_speed=2.0, g=-5.0)
    fl = started(cfg, 5)
    o = OracleFluid(cfg)
    u, v, s = wl.initial_state(cfg, 5)
    o.set_state(u=u, v=v, smoke=s)
    with Batch() as b:
        b.add(fl)
        b.add(started(make_config(64, 36, projection_n=7), 6))
        b.step(4)
    o.update(steps=4)
    for name, attr in (("vel_x", "u"), ("vel_y", "v"), ("smoke", "smoke")):
        assert np.array_equal(fl.get_field(name).view(np.uint32), getattr(o, attr).view(np.uint32))


SHAPES = [(8, 2, 1), (8, 5, 1), (8, 9, 1), (8, 12, 1), (8, 3, 2), (8, 6, 2), (12, 1, 2),
          (12, 7, 2), (12, 10, 2), (12, 17, 2)]


@pytest.mark.parametrize("rows,k,occ", SHAPES)
def test_batch_shapes_equal_alone(rows, k, occ):
    cfgs = mixed_configs()[:7] + mixed_configs()[8:11]
    batched, alone = pairs(cfgs, 300)
    with Batch() as b:
        b.set_tuning(rows=rows, sweeps_per_tile=k, occupancy=occ)
        for fl in batched:
            b.add(fl)
        b.step(3)
        for fl in alone:
            fl.step(3)
        for i, (x, y) in enumerate(zip(batched, alone)):
            assert_same(state(x), state(y), f"rows={rows} k={k} occupancy={occ} sim {i}")


def test_membership_changes_keep_results():
    cfgs = mixed_configs()
    batched, alone = pairs(cfgs, 500)
    with Batch() as b:
        for fl in batched[:6]:
            b.add(fl)
        b.step(2)
        b.remove(batched[2])
        for fl in batched[6:]:
            b.add(fl)
        b.step(3)
        batched[2].step(1)   # stepped alone again once it left the batch
        b.add(batched[2])
        b.step(1)
    steps = [6, 6, 4, 6, 6, 6, 4, 4, 4, 4, 4, 4]
    for fl, s in zip(alone, steps):
        fl.step(s)
    for i, (x, y) in enumerate(zip(batched, alone)):
        assert_same(state(x), state(y), f"sim {i}")


def test_batch_scale_uneven_grids():
    """A larger mix from the batched workload generator, 64 sims."""
    specs = wl.batched_specs(count=64, seed=11)[:64]
    specs = [s for s in specs if s.cfg.width <= 512][:24]
    batched = [wl.start(s) for s in specs]
    alone = [wl.start(s) for s in specs]
    with Batch() as b:
        for fl in batched:
            b.add(fl)
        b.step(2)
        info = b.info
        assert info["count"] == len(specs) and info["work_items"] > 0
    for fl in alone:
        fl.step(2)
    for i, (x, y) in enumerate(zip(batched, alone)):
        assert_same(state(x), state(y), f"spec {i}")


def test_autotune_restores_state():
    cfgs = mixed_configs()[:5]
    batched, _ = pairs(cfgs, 700)
    with Batch() as b:
        for fl in batched:
            b.add(fl)
        before = [state(fl) for fl in batched]
        tuning, ms = b.autotune()
        assert ms > 0 and tuning["rows"] in (8, 12) and tuning["occupancy"] in (1, 2)
        for fl, s in zip(batched, before):
            assert_same(state(fl), s, "after autotune")
        b.step(2)
    ref = [started(c, 700 + i) for i, c in enumerate(cfgs)]
    for fl in ref:
        fl.step(2)
    for x, y in zip(batched, ref):
        assert_same(state(x), state(y), f"after autotune {tuning}")


def test_budget_and_state_errors():
    small = make_config(64, 36)
    need = device_bytes(small)
    with Batch(budget_bytes=2 * need + 1) as b:
        f1, f2, f3 = Fluid(small), Fluid(small), Fluid(small)
        b.add(f1)
        b.add(f2)
        with pytest.raises(OpenSayalError) as e:
            b.add(f3)
        assert "BUDGET" in str(e.value)
        with pytest.raises(OpenSayalError):
            b.add(f1)  # already a member
        with pytest.raises(OpenSayalError):
            f1.step(1)  # stepped by the batch
        with pytest.raises(OpenSayalError):
            b.set_tuning(rows=16, sweeps_per_tile=2)
        with pytest.raises(OpenSayalError):
            b.set_tuning(rows=12, sweeps_per_tile=2, occupancy=1)
        b.remove(f2)
        b.add(f3)
        b.step(1)
        f2.step(1)  # a former member steps alone again
    f1.step(1)  # the batch is gone


def test_destroyed_member_leaves_batch():
    with Batch() as b:
        keep = started(make_config(80, 45, projection_n=9), 1)
        drop = started(make_config(70, 40, projection_n=9), 2)
        b.add(keep)
        b.add(drop)
        b.step(1)
        drop.close()
        assert b.info["count"] == 1
        b.step(1)
    ref = started(make_config(80, 45, projection_n=9), 1)
    ref.step(2)
    assert_same(state(keep), state(ref), "after a member was destroyed")


def test_batched_workload_256_grids():
    """The batched workload itself: 256 grids 128..1024 wide (height 9W/16),
    n from {20, 50, 100}, obstacle and inflow drawn per grid, stepped by one
    persistent launch per step.  After 3 batch steps every member equals the
    same grid stepped alone; the smallest grid, the largest and the smallest
    n = 100 grid equal the oracle."""
    specs = wl.batched_specs()
    assert len(specs) == 256
    batched = [wl.start(s) for s in specs]
    alone = [wl.start(s) for s in specs]
    with Batch() as b:
        for fl in batched:
            b.add(fl)
        b.step(3)
        assert b.info["count"] == 256
        for fl in alone:
            fl.step(3)
        for i, (x, y) in enumerate(zip(batched, alone)):
            assert_same(state(x), state(y), f"grid {i} ({specs[i].cfg.width}x{specs[i].cfg.height})")
    cells = lambda k: specs[k].cfg.width * specs[k].cfg.height
    hundred = [k for k in range(256) if specs[k].cfg.projection_n == 100]
    picks = {min(range(256), key=cells), max(range(256), key=cells), min(hundred, key=cells)}
    for k in sorted(picks):
        cfg = specs[k].cfg
        o = OracleFluid(cfg, nthreads=16)
        u, v, s = wl.initial_state(cfg, specs[k].seed)
        o.set_state(u=u, v=v, smoke=s)
        o.update(steps=3)
        for name, attr in (("vel_x", "u"), ("vel_y", "v"), ("smoke", "smoke")):
            got = batched[k].get_field(name)
            assert np.array_equal(got.view(np.uint32), getattr(o, attr).view(np.uint32)), \
                (k, cfg.width, cfg.height, cfg.projection_n, name)
    for fl in batched + alone:
        fl.close()


def test_item_timing_records_every_item_and_changes_nothing():
    """With per-item timing on, each work item of the last step has one
    record (member, stage kind, stage, tile; claim <= ready <= done), the
    items per member and kind are the member's stage chain times its tiles,
    and every member's fields equal those of the same batch untimed."""
    cfgs = mixed_configs()
    timed, plain = pairs(cfgs)
    with Batch() as bt, Batch() as bp:
        for fl in timed:
            bt.add(fl)
        for fl in plain:
            bp.add(fl)
        bt.set_timing(True)
        with pytest.raises(OpenSayalError):
            bt.item_times()  # no timed step yet
        bt.step(2)
        bp.step(2)
        rec = bt.item_times()
        tun = bt.tuning
        # a record array one short is refused, with the count needed
        import ctypes
        from opensayal_amd import _abi
        need = ctypes.c_int64()
        short = (_abi.OsyBatchItemTime * max(1, len(rec["kind"]) - 1))()
        assert _abi.lib().osy_batch_item_times(bt._h, short, len(rec["kind"]) - 1,
                                               ctypes.byref(need), None) == _abi.OSY_E_INVALID
        assert need.value == len(rec["kind"])
        for i, (x, y) in enumerate(zip(timed, plain)):
            assert_same(state(x), state(y), f"timed batch member {i}")
        n = len(rec["kind"])
        assert n == bt.info["work_items"]
        assert np.all(rec["claim"] <= rec["ready"]) and np.all(rec["ready"] <= rec["done"])
        assert rec["claim"].min() == 0 and rec["done"].max() > 0
        k = tun["sweeps_per_tile"]
        tw, th = 128 - 4 * k, 8 * tun["rows"] - 4 * k
        for m, c in enumerate(cfgs):
            mine = rec["member"] == m
            nproj = -(-c.projection_n // k) if c.projection_n else 0
            tiles_p = -(-c.width // tw) * -(-c.height // th)
            tiles_a = -(-c.width // 64) * -(-c.height // 64)
            assert np.sum(mine & (rec["kind"] == 1)) == nproj * tiles_p, m
            assert np.sum(mine & (rec["kind"] == 2)) == tiles_a, m
            smoke_on = c.enable_smoke and c.wind_smoke != 0
            assert np.sum(mine & (rec["kind"] == 3)) == (tiles_a if smoke_on else 0), m
            forces = c.projection_n == 0 or not smoke_on
            assert np.sum(mine & (rec["kind"] == 0)) == (tiles_a if forces else 0), m
        bt.set_timing(False)
        bt.step(1)
        with pytest.raises(OpenSayalError):
            bt.item_times()


def test_large_members_step_on_their_own_graphs():
    """Members of at least per_grid_cells cells step through their per-grid
    graphs ahead of the persistent launch; with them in the mix, alone in
    the batch, or forced into the persistent launch (0), every member's
    fields equal its run alone."""
    big = make_config(1100, 620, projection_n=23, wind_speed=2.0, g=-5.0, pipe_height=150,
                      obstacle_center_x=400, obstacle_center_y=300, obstacle_radius=30.0)
    small = mixed_configs()[:4]
    for cells, cfgs in (((1 << 20) // 2, [big] + small), ((1 << 20) // 2, [big]),
                        (0, [big] + small)):
        batched, alone = pairs(cfgs, seed0=300)
        with Batch() as b:
            b.set_per_grid_cells(cells)
            for fl in batched:
                b.add(fl)
            b.step(2)
            b.step(1)
            info = b.info
            for fl in alone:
                fl.step(3)
            for i, (x, y) in enumerate(zip(batched, alone)):
                assert_same(state(x), state(y), f"cells={cells} member {i}")
            if cells and len(cfgs) == 1:
                assert info["work_items"] == 0 and info["groups"] == 0
        for fl in batched + alone:
            fl.close()
    with Batch() as b, pytest.raises(OpenSayalError):
        b.set_per_grid_cells(-1)


def test_batch_autotune_tunes_large_members():
    """The batch autotuner also picks the projection shape of the members
    that step on their own graphs (the per-grid autotuner's choice, which a
    fresh handle of the same shape then adopts from its cache), and the
    fields still equal the runs alone."""
    big = make_config(1100, 620, projection_n=23, wind_speed=2.0, g=-5.0, pipe_height=150,
                      obstacle_center_x=400, obstacle_center_y=300, obstacle_radius=30.0)
    cfgs = [big] + mixed_configs()[:2]
    batched, alone = pairs(cfgs, seed0=310)
    with Batch() as b:
        b.set_per_grid_cells((1 << 20) // 2)
        for fl in batched:
            b.add(fl)
        b.autotune()
        member = batched[0]
        other = Fluid(big)
        assert other.autotune_cached()
        assert member.tuning == other.tuning
        other.close()
        b.step(3)
        for fl in alone:
            fl.step(3)
        for i, (x, y) in enumerate(zip(batched, alone)):
            assert_same(state(x), state(y), f"member {i}")
    for fl in batched + alone:
        fl.close()
