"""Serving on the GPU: requests of different sizes, storage types and step
counts arrive over time, are admitted under a budget small enough to make
some wait, share the device with whatever else is resident, and each result
is bit-identical to running that request alone."""
import numpy as np
import pytest

from opensayal_amd import device_bytes, workloads as wl
from opensayal_amd.serving import DeviceBackend, Server

pytestmark = pytest.mark.gpu

FIELDS = ("vel_x", "vel_y", "smoke")


def requests():
    out = []
    for t, spec in wl.serving_specs(count=10, seed=3, rate_hz=400.0):
        cfg = spec.cfg
        # keep the test short: cap the size and the step count
        if cfg.width > 400:
            cfg = wl.wind_tunnel(400, 225, pipe_height=56, cx=200, cy=112, radius=12.0,
                                 speed=float(cfg.wind_speed), n=int(cfg.projection_n),
                                 storage="f16" if cfg.storage == 1 else "f32")
        out.append((t, wl.SimSpec(cfg, spec.seed, 5 + spec.steps % 23)))
    return out


def test_results_identical_alone_and_in_the_mix():
    reqs = requests()
    sizes = sorted(device_bytes(s.cfg) for _, s in reqs)
    budget = sizes[-1] + sizes[-2] + sizes[0]  # forces some requests to wait
    backend = DeviceBackend(collect=FIELDS)
    try:
        rep = Server(backend, budget_bytes=budget, max_chunk=4).run(reqs)
    finally:
        backend.close()
    s = rep.summary()
    assert s["completed"] == len(reqs)
    assert rep.peak_bytes <= budget
    assert any(r.admitted - r.arrival > 1e-3 for r in rep.requests)  # someone queued
    assert s["step_ms_p99"] >= s["step_ms_p50"] > 0
    for r in rep.requests:
        alone = wl.start(r.spec)
        alone.step(r.spec.steps)
        for n in FIELDS:
            a, b = r.result[n], alone.get_field(n)
            assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), (r.rid, n)
        alone.close()
