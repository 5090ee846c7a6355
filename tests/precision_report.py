"""Precision-mode report (run on demand, not collected by pytest).

    python -m tests.precision_report [--steps 200] [--width 3840 --height 2160]
                                     [--out precision_report.json]

The 3840x2160 wind tunnel (obstacle r=72 at the centre, pipe 540, n=100,
o=1.9, pressure on) is stepped from one seeded state on the GPU with u, v and
smoke stored in FP16, in BF16 and in FP32 (pressure FP32 and FP32 arithmetic
in every case), and by the FP64 oracle.  For each storage it reports, per
field, the relative L2 error ||x - ref||_2 / ||ref||_2, the max-abs error and
the max-abs error relative to max |ref|, against the FP64 run; and the FP64
divergence residual (L2 and max over interior fluid cells, the formula of
Fluid::get_divergence, src/fluid.cu:192-201) of the final state and of that
state after one more projection of n sweeps.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

from opensayal_amd import Fluid, workloads as wl
from oracle.oracle import OracleFluid

FIELDS = (("vel_x", "u"), ("vel_y", "v"), ("smoke", "smoke"), ("pressure", "p"))


def errors(x: np.ndarray, ref: np.ndarray) -> dict:
    d = x.astype(np.float64) - ref
    nref = float(np.sqrt(np.sum(ref * ref)))
    mref = float(np.max(np.abs(ref)))
    linf = float(np.max(np.abs(d)))
    return {"l2_rel": float(np.sqrt(np.sum(d * d))) / nref if nref > 0 else None,
            "linf_abs": linf, "linf_rel": linf / mref if mref > 0 else None}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--width", type=int, default=3840)
    ap.add_argument("--height", type=int, default=2160)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--threads", type=int, default=min(16, os.cpu_count() or 1))
    ap.add_argument("--out", default="precision_report.json")
    args = ap.parse_args()

    W, H = args.width, args.height
    def config(storage):
        return wl.wind_tunnel(W, H, pipe_height=H // 4, cx=W // 2, cy=H // 2,
                              radius=72.0 * W / 3840, n=100, pressure=True, storage=storage)

    cfg64 = config("f32")
    u0, v0, s0 = wl.initial_state(cfg64, args.seed)
    n = cfg64.projection_n
    report = {"workload": f"{W}x{H} n={n} o=1.9 pressure on, seed {args.seed}",
              "steps": args.steps, "runs": {}}

    gpu = {}
    for storage in ("f16", "bf16", "f32"):
        t0 = time.perf_counter()
        fl = Fluid(config(storage))
        for name, arr in (("vel_x", u0), ("vel_y", v0), ("smoke", s0)):
            fl.set_field(name, arr)
        fl.step(args.steps)
        out = {name: fl.get_field(name) for name, _ in FIELDS}
        res_final = fl.divergence_norm()
        fl.project(sweeps=n)
        res_proj = fl.divergence_norm()
        fl.close()
        gpu[storage] = (out, res_final, res_proj)
        print(f"gpu {storage}: {args.steps} steps in {time.perf_counter() - t0:.1f} s",
              flush=True)

    o = OracleFluid(cfg64, np.float64, nthreads=args.threads)
    o.set_state(u=u0.astype(np.float64), v=v0.astype(np.float64), smoke=s0.astype(np.float64))
    t0 = time.perf_counter()
    for k in range(args.steps):
        o.update()
        if (k + 1) % 10 == 0:
            print(f"oracle fp64: step {k + 1}/{args.steps} "
                  f"({time.perf_counter() - t0:.0f} s)", flush=True)
    ref = {name: getattr(o, attr).copy() for name, attr in FIELDS}
    ref_final = o.divergence_norm()
    o.projection()
    ref_proj = o.divergence_norm()
    report["runs"]["fp64_oracle"] = {
        "residual_final": {"l2": ref_final[0], "max": ref_final[1]},
        "residual_after_projection": {"l2": ref_proj[0], "max": ref_proj[1]},
    }
    for storage, (out, res_final, res_proj) in gpu.items():
        report["runs"][storage] = {
            "fields": {name: errors(out[name], ref[name]) for name, _ in FIELDS},
            "residual_final": {"l2": res_final[0], "max": res_final[1]},
            "residual_after_projection": {"l2": res_proj[0], "max": res_proj[1]},
        }
    os.makedirs(os.path.dirname(args.out) or ".", exist_ok=True)
    with open(args.out, "w") as f:
        json.dump(report, f, indent=1)
    print(json.dumps(report), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
