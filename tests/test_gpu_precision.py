"""The mixed-precision configuration: 3840x2160, obstacle r = 72 at the
centre, pipe 540, n = 100, o = 1.9, pressure on (FP32 pressure), u, v and
smoke stored in FP16 or BF16 with FP32 arithmetic.

* Bit-exact against the oracle's storage mode (which rounds every phase's
  outputs to the storage type, as the kernels do) after 2 steps, with the
  error against the FP64 oracle run from the same state bounded.
* The 200-step run of that configuration in FP32, FP16 and BF16 storage
  against the FP64 oracle: per field the relative L2 error, the max-abs
  error and the max-abs error over max |ref|, and the FP64 divergence
  residual (L2 and max over interior fluid cells, the formula of
  Fluid::get_divergence, src/fluid.cu:192-201) of the final state and of
  that state after one more projection of n sweeps.  The table is printed;
  the bounds asserted are stated below.
"""
import os
import time

import numpy as np
import pytest

from opensayal_amd import Fluid, workloads as wl
from oracle.oracle import OracleFluid

pytestmark = pytest.mark.gpu

NT = min(os.cpu_count() or 1, 16)
FIELDS = (("vel_x", "u"), ("vel_y", "v"), ("smoke", "smoke"), ("pressure", "p"))
SEED = 1


def errors(x: np.ndarray, ref: np.ndarray) -> dict:
    d = x.astype(np.float64) - ref
    nref = float(np.sqrt(np.sum(ref * ref)))
    mref = float(np.max(np.abs(ref)))
    linf = float(np.max(np.abs(d)))
    return {"l2_rel": float(np.sqrt(np.sum(d * d))) / nref if nref > 0 else 0.0,
            "linf_abs": linf, "linf_rel": linf / mref if mref > 0 else 0.0}


def gpu_run(storage, steps):
    cfg = wl.mixed_precision(storage)
    fl = Fluid(cfg)
    u, v, s = wl.initial_state(cfg, SEED)
    for name, arr in (("vel_x", u), ("vel_y", v), ("smoke", s)):
        fl.set_field(name, arr)
    fl.step(steps)
    return fl


def fp64_oracle(steps):
    cfg = wl.mixed_precision("f32")
    o = OracleFluid(cfg, np.float64, nthreads=NT)
    u, v, s = wl.initial_state(cfg, SEED)
    o.set_state(u=u.astype(np.float64), v=v.astype(np.float64), smoke=s.astype(np.float64))
    t0 = time.perf_counter()
    for k in range(steps):
        o.update()
        if (k + 1) % 50 == 0:
            print(f"fp64 oracle: step {k + 1}/{steps} ({time.perf_counter() - t0:.0f} s)",
                  flush=True)
    return o


# After 2 steps from the seeded state the 16-bit modes stay within these
# relative L2 errors of the FP64 run (u, v, smoke, pressure); the largest,
# u's, is 5.4e-3 in FP16 and 3.4e-2 in BF16.
SHORT_BOUND = {"f32": 1e-3, "f16": 2e-2, "bf16": 1e-1}


@pytest.mark.parametrize("storage", ["f32", "f16", "bf16"])
def test_mixed_precision_2_steps_bit_exact(storage):
    fl = gpu_run(storage, 2)
    cfg = wl.mixed_precision(storage)
    o = OracleFluid(cfg, nthreads=NT)
    u, v, s = wl.initial_state(cfg, SEED)
    o.set_state(u=u, v=v, smoke=s)
    o.update(steps=2)
    for name, attr in FIELDS:
        g = fl.get_field(name)
        ref = getattr(o, attr)
        bad = np.count_nonzero(g.view(np.uint32) != ref.view(np.uint32))
        assert bad == 0, f"{storage} {name}: {bad} cells differ from the oracle's storage mode"
    assert fl.pressure_range == o.pressure_range
    o64 = fp64_oracle(2)
    for name, attr in FIELDS:
        e = errors(fl.get_field(name), getattr(o64, attr))
        assert e["l2_rel"] < SHORT_BOUND[storage], (storage, name, e)
    fl.close()


# Bounds on the relative L2 error against the FP64 run after 200 steps, per
# storage type, and on the relative difference of the residual after one
# more projection from the FP64 run's.  The largest L2 errors, u's, are
# about 1.7e-3 (FP32), 2.1e-2 (FP16) and 0.26 (BF16): after 200 steps the
# wake behind the obstacle is chaotic, and BF16's 8-bit significand does not
# hold this flow.
LONG_BOUND = {"f32": (1e-2, 1e-3), "f16": (5e-2, 1e-2), "bf16": (0.5, 0.2)}


def test_mixed_precision_200_steps_against_fp64():
    steps = 200
    cfg = wl.mixed_precision("f32")
    n = cfg.projection_n
    runs = {}
    for storage in ("f32", "f16", "bf16"):
        fl = gpu_run(storage, steps)
        out = {name: fl.get_field(name) for name, _ in FIELDS}
        res_final = fl.divergence_norm()
        fl.project(sweeps=n)
        res_proj = fl.divergence_norm()
        fl.close()
        runs[storage] = (out, res_final, res_proj)
        print(f"gpu {storage}: {steps} steps done", flush=True)
    o = fp64_oracle(steps)
    ref = {name: getattr(o, attr).copy() for name, attr in FIELDS}
    ref_final = o.divergence_norm()
    o.projection()
    ref_proj = o.divergence_norm()
    print(f"3840x2160 n={n} pressure on, {steps} steps, errors against the FP64 oracle run")
    print(f"fp64 residual final L2/max {ref_final[0]:.6g} / {ref_final[1]:.6g}, "
          f"after projection {ref_proj[0]:.6g} / {ref_proj[1]:.6g}")
    failures = []
    for storage, (out, res_final, res_proj) in runs.items():
        l2_bound, res_bound = LONG_BOUND[storage]
        line = [f"{storage:5s}"]
        for name, _ in FIELDS:
            e = errors(out[name], ref[name])
            line.append(f"{name} L2 {e['l2_rel']:.3g} Linf {e['linf_abs']:.3g} "
                        f"({e['linf_rel']:.3g})")
            if not e["l2_rel"] < l2_bound:
                failures.append((storage, name, e["l2_rel"]))
        rel_res = abs(res_proj[0] - ref_proj[0]) / ref_proj[0]
        line.append(f"residual final {res_final[0]:.6g} / {res_final[1]:.6g}, after "
                    f"projection {res_proj[0]:.6g} / {res_proj[1]:.6g} (rel {rel_res:.2g})")
        if not rel_res < res_bound:
            failures.append((storage, "residual", rel_res))
        print(" | ".join(line), flush=True)
    assert not failures, failures
