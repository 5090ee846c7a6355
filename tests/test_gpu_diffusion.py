"""The diffusion pass on the GPU (fluid.viscosity != 0): bit-exact against
the oracle's Jacobi restatement through whole steps, for launch counts of
both parities (v moves with u on an odd count), every storage type, cell
size 2, pressure, an active source and the colour projection kernel; the
reference's default viscosity (0.001, a config without the key) runs."""
import json

import numpy as np
import pytest

from opensayal_amd import Batch, ConfigParser, Fluid, OpenSayalError, Source, workloads as wl
from oracle.oracle import OracleFluid
from tests.helpers import make_config

pytestmark = pytest.mark.gpu


def same(a, b):
    return np.array_equal(np.asarray(a, np.float32).view(np.uint32),
                          np.asarray(b, np.float32).view(np.uint32))


def pair(cfg, seed):
    fl = Fluid(cfg)
    o = OracleFluid(cfg)
    u, v, s = wl.initial_state(cfg, seed)
    fl.set_field("vel_x", u)
    fl.set_field("vel_y", v)
    fl.set_field("smoke", s)
    # the oracle starts from the state the GPU holds (16-bit storage rounds it)
    o.set_state(u=fl.get_field("vel_x"), v=fl.get_field("vel_y"), smoke=fl.get_field("smoke"))
    return fl, o


def check(fl, o, what):
    for name, attr in (("vel_x", "u"), ("vel_y", "v"), ("smoke", "smoke")):
        got, want = fl.get_field(name), getattr(o, attr)
        if not same(got, want):
            bad = np.argwhere(got.view(np.uint32) != np.asarray(want, np.float32).view(np.uint32))
            pytest.fail(f"{what} {name}: {len(bad)} cells differ, first at {tuple(bad[0])}")
    if fl.config.enable_pressure:
        assert same(fl.get_field("pressure"), o.p), what
        assert fl.pressure_range == o.pressure_range


CASES = [
    # (kwargs, steps) -- launch count ceil(n / 6): 1, 2, 3, 9
    (dict(projection_n=6, viscosity=0.5), 2),
    (dict(projection_n=12, viscosity=0.5), 2),
    (dict(projection_n=13, viscosity=0.02), 3),
    (dict(projection_n=50, viscosity=0.001), 2),
    (dict(projection_n=1, viscosity=3.0), 2),
    (dict(projection_n=20, viscosity=0.3, enable_pressure=1), 2),
    (dict(projection_n=17, viscosity=0.3, cell_size=2.0), 2),
    (dict(projection_n=15, viscosity=0.3, storage=1), 2),
    (dict(projection_n=15, viscosity=0.3, storage=2, enable_pressure=1), 2),
    (dict(projection_n=0, viscosity=0.3), 2),
]


@pytest.mark.parametrize("kw,steps", CASES)
def test_diffusion_steps_match_oracle(kw, steps):
    cfg = make_config(203, 117, wind_speed=2.0, g=-5.0, obstacle_center_x=70,
                      obstacle_center_y=58, obstacle_radius=11.0, **kw)
    fl, o = pair(cfg, 21)
    for s in range(steps):
        fl.update()
        o.update()
        check(fl, o, f"{kw} step {s + 1}")


def test_graph_steps_and_colour_kernel():
    cfg = make_config(150, 90, projection_n=9, viscosity=0.4, wind_speed=2.0, g=-5.0)
    fl, o = pair(cfg, 5)
    fl.step(3)
    o.update(steps=3)
    check(fl, o, "graph")
    fl.set_tuning(projection_kernel=0)
    fl.step(2)
    o.update(steps=2)
    check(fl, o, "colour kernel")


def test_active_source_with_diffusion():
    cfg = make_config(120, 80, projection_n=10, viscosity=0.2, wind_speed=1.0, g=-2.0)
    fl, o = pair(cfg, 8)
    src = Source()
    src.active, src.smoke, src.velocity = 1, 1.0, 0.5
    src.position_x, src.position_y = 60, 40
    fl.update(src)
    o.update(source=src)
    check(fl, o, "source")


def test_reference_default_viscosity_config_runs():
    """A config that leaves fluid.viscosity out gets the reference's 0.001
    and runs the diffusion pass."""
    text = json.dumps({"sim": {"width": 160, "height": 90, "projection": {"n": 14},
                               "wind_tunnel": {"speed": 2.0}, "physics": {"g": -5.0}}})
    cfg = ConfigParser.parse_string(text)
    assert cfg.viscosity == pytest.approx(0.001)
    fl, o = pair(cfg, 3)
    fl.step(2)
    o.update(steps=2)
    check(fl, o, "default viscosity")


def test_batch_refuses_diffusing_members():
    with Batch() as b:
        with pytest.raises(OpenSayalError) as e:
            b.add(Fluid(make_config(64, 36, viscosity=0.001)))
        assert "viscosity" in str(e.value)
