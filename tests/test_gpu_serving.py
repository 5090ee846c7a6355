"""Serving on the GPU: requests of different sizes, storage types and step
counts arrive over time, are admitted under a budget (one small enough to
make some wait, and the 200 GB of the serving workload), share the device
with whatever else is resident, and each result is bit-identical to running
that request alone; one request is also checked against the oracle."""
import os

import numpy as np
import pytest

from opensayal_amd import device_bytes, workloads as wl
from opensayal_amd.serving import DeviceBackend, Server
from oracle.oracle import OracleFluid

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


def test_serving_workload_sizes():
    """The serving workload's own distribution: widths log-uniform in
    [64, 2048] (16:9), 50..2000 steps, n from {20, 50, 100}, FP32 or FP16
    storage, Poisson arrivals, 200 GB budget.  Every result is collected;
    a sample (the largest grid, the longest run, one of each storage type)
    is stepped alone and compared bit for bit, and the request with the
    least work is compared bit for bit with the oracle (in its storage
    mode)."""
    reqs = wl.serving_specs(count=12, seed=14, rate_hz=20.0)  # widths 71..2020
    backend = DeviceBackend(collect=FIELDS)
    try:
        rep = Server(backend, budget_bytes=200 * 10**9, max_chunk=32).run(reqs)
    finally:
        backend.close()
    s = rep.summary()
    assert s["completed"] == len(reqs)
    assert s["batch_steps"] == len(rep.step_ms) and s["step_ms_p99"] >= s["step_ms_p50"] > 0
    widths = [r.spec.cfg.width for r in rep.requests]
    assert max(widths) > 2000 and min(r.spec.steps for r in rep.requests) >= 50
    print(f"serving: {s['completed']} requests, widths {min(widths)}..{max(widths)}, "
          f"{s['batch_steps']} batch steps, step ms p50/p99 {s['step_ms_p50']:.3f}/"
          f"{s['step_ms_p99']:.3f}, completion s p50/p99 {s['completion_s_p50']:.2f}/"
          f"{s['completion_s_p99']:.2f}, {s['gcell_steps_s_wall']:.2f} Gcell-steps/s", flush=True)
    work = lambda r: r.spec.cfg.width * r.spec.cfg.height * r.spec.steps * r.spec.cfg.projection_n
    sample = {max(rep.requests, key=lambda r: r.spec.cfg.width).rid,
              max(rep.requests, key=lambda r: r.spec.steps).rid}
    for st in (0, 1):
        of = [r for r in rep.requests if r.spec.cfg.storage == st]
        if of:
            sample.add(min(of, key=work).rid)
    for r in rep.requests:
        if r.rid not in sample:
            continue
        alone = wl.start(r.spec)
        alone.step(r.spec.steps)
        for n in FIELDS:
            a, b = r.result[n], alone.get_field(n)
            assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), (r.rid, n)
        alone.close()
    r = min(rep.requests, key=work)
    o = OracleFluid(r.spec.cfg, nthreads=min(os.cpu_count() or 1, 16))
    u, v, sm = wl.initial_state(r.spec.cfg, r.spec.seed)
    o.set_state(u=u, v=v, smoke=sm)
    o.update(steps=r.spec.steps)
    for n, attr in zip(FIELDS, ("u", "v", "smoke")):
        assert np.array_equal(r.result[n].view(np.uint32), getattr(o, attr).view(np.uint32)), \
            (r.rid, n, r.spec.cfg.width, r.spec.steps)
