"""Parity at the benchmark's stated length and velocities, through the C ABI.

* The 1920x1080, n = 50 wind tunnel for its stated 1000 steps, bit-exact
  against the oracle, with the stated tolerances (1e-4 relative L-inf after
  100 steps) asserted as well.  As the flow develops, gravity drives some
  back-traces of the advections out of the kernels' LDS stage, so the run
  takes the HBM fallback of the samplers (get_general_velocity_x/_y and
  interpolate_smoke, src/fluid.cu:418-539, 644-716) on some wavefronts.
* High-velocity states (|v| d_t of several cells) that send many wavefronts
  of both advections through that fallback (up to 468 of the velocity
  advection's 576 and 76 of the smoke advection's), at cell sizes 1, 2 and
  3; the
  CPU restatement of the kernels' staged test (tests/helpers.py) shows from
  the oracle's state that the fallback is taken before the step runs.
"""
import os

import numpy as np
import pytest

from opensayal_amd import Fluid, workloads as wl
from oracle.oracle import OracleFluid
from tests.helpers import advection_fallback_counts, make_config, rel_linf

pytestmark = pytest.mark.gpu

NT = min(os.cpu_count() or 1, 16)
FIELDS = (("vel_x", "u"), ("vel_y", "v"), ("smoke", "smoke"))


def pair(cfg, u, v, s):
    fl = Fluid(cfg)
    o = OracleFluid(cfg, nthreads=NT)
    for name, arr in (("vel_x", u), ("vel_y", v), ("smoke", s)):
        fl.set_field(name, arr)
    o.set_state(u=u, v=v, smoke=s)
    return fl, o


def assert_bitwise(fl, o, what):
    for name, attr in FIELDS:
        g = fl.get_field(name)
        ref = getattr(o, attr)
        if not np.array_equal(g.view(np.uint32), ref.view(np.uint32)):
            bad = np.argwhere(g.view(np.uint32) != ref.view(np.uint32))
            r, i = bad[0]
            pytest.fail(f"{what} {name}: {len(bad)} cells differ; first at i={i} "
                        f"j={o.cfg.height - 1 - r}: gpu={g[r, i]!r} oracle={ref[r, i]!r}; "
                        f"rel_linf={rel_linf(g, ref):.3e}")


def test_full_size_1000_steps_bit_exact():
    """The plain wind-tunnel workload at full size: 1920x1080, n = 50, o = 1.9, 1000 steps."""
    cfg = wl.plain_full()
    fl, o = pair(cfg, *wl.initial_state(cfg, 3))
    fl.step(100)
    o.update(steps=100)
    for name, attr in FIELDS:
        err = rel_linf(fl.get_field(name), getattr(o, attr))
        assert err <= 1e-4, (name, err)
    assert_bitwise(fl, o, "100 steps")
    counts = []
    for done in range(100, 1000, 300):
        fl.step(300)
        o.update(steps=300)
        assert_bitwise(fl, o, f"{done + 300} steps")
        (vel, total), (smk, _) = advection_fallback_counts(o)
        counts.append((done + 300, vel, smk, total))
    for step, vel, smk, total in counts:
        print(f"after {step} steps: next step's advections leave the LDS stage on {vel} "
              f"(velocity) and {smk} (smoke) of {total} wavefronts", flush=True)
    # the FP64 sums of squares over 2 M cells run in another order on the
    # GPU: relative difference below n * 2^-53 (about 2.3e-10)
    g, r = fl.divergence_norm(), o.divergence_norm()
    assert g[0] == pytest.approx(r[0], rel=1e-9)
    assert g[1] == r[1]


@pytest.mark.parametrize("tile_width,tile_height,k", [(108, 44, 5), (96, 48, 8)])
def test_full_size_1000_steps_resident_bit_exact(tile_width, tile_height, k):
    """The same 1000 steps with the resident projection (all 50 sweeps of a
    step in one launch), in the shapes the autotuner picks at this size."""
    cfg = wl.plain_full()
    fl, o = pair(cfg, *wl.initial_state(cfg, 4))
    fl.set_tuning(projection_kernel=3, tile_width=tile_width, tile_height=tile_height,
                  sweeps_per_launch=k)
    for done in (500, 1000):
        fl.step(500)
        o.update(steps=500)
        assert_bitwise(fl, o, f"resident {tile_width}x{tile_height} k={k}, {done} steps")


def test_full_size_pressure_on_bit_exact():
    """The same grid with the pressure solve on (pressure sums in every
    sweep, the forces zeroing them): 20 steps bit-exact, pressure field and
    range included."""
    cfg = wl.plain_full(pressure=True)
    fl, o = pair(cfg, *wl.initial_state(cfg, 4))
    fl.step(20)
    o.update(steps=20)
    assert_bitwise(fl, o, "20 steps, pressure on")
    p = fl.get_field("pressure")
    assert np.array_equal(p.view(np.uint32), o.p.view(np.uint32))
    assert fl.pressure_range == o.pressure_range


FALLBACK_CASES = [
    # (cell size, velocity noise amplitude): |v| d_t / cs of up to 6.5, 10, 10, 7 cells
    (1, 130.0),
    (1, 200.0),
    (2, 400.0),
    (3, 400.0),
]


@pytest.mark.parametrize("cs,amp", FALLBACK_CASES)
def test_advection_hbm_fallback_bit_exact(cs, amp):
    """The samplers' HBM path: wavefronts of both advections sample outside
    their stage; 3 steps bit-exact against the oracle."""
    W, H = 256, 144
    cfg = make_config(W, H, cell_size=float(cs), wind_speed=2.0, g=-5.0, pipe_height=36,
                      smoke_height=36, obstacle_center_x=128, obstacle_center_y=72,
                      obstacle_radius=10.0, enable_pressure=1)
    u, v, s = wl.initial_state(cfg, 31)
    scale = np.float32(amp / 0.01)
    fl, o = pair(cfg, u * scale, v * scale, s)
    (vel, total), (smk, _) = advection_fallback_counts(o)
    print(f"cs={cs} amp={amp}: first step's advections leave the LDS stage on {vel} "
          f"(velocity) and {smk} (smoke) of {total} wavefronts", flush=True)
    assert vel > 0 and smk > 0
    fl.update()
    o.update()
    assert_bitwise(fl, o, f"cs={cs} amp={amp} 1 step")
    assert fl.pressure_range == o.pressure_range
    fl.step(2)
    o.update(steps=2)
    assert_bitwise(fl, o, f"cs={cs} amp={amp} 3 steps")
