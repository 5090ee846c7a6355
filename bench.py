"""Benchmark of the OpenSayal per-step solver on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload full|small]

A step is one Fluid::update of the wind-tunnel workload: 1920x1080 cells,
projection n = 50, o = 1.9, FP32, seeded +-0.01 velocity noise (synthetic
input, built on the host and uploaded before timing).  W untimed steps, then
exactly K steps between device synchronisations; prints one JSON line.

roofline: the dominant kernel is the projection.  Its algorithmic traffic is
17 B per cell per sweep (u and v read and written once, 4 B each, plus the
1-byte cell mask), so one launch that runs s sweeps over the grid moves
17 * s * W * H bytes; `achieved` divides that by the launch's average
duration, timed with HIP events on the stream the launches run on.
cpu_baseline: the oracle (a C restatement of the reference's update rule,
OpenMP over rows) timed on this host's cores on a bounded number of steps.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0

# HBM bytes per projection launch measured with rocprofv3 FETCH_SIZE and
# WRITE_SIZE passes (scaled by the same counters on a copy of known size),
# per (kernel, tile width, tile height, sweeps per launch, W, H); summaries
# and commands in profiles/step_1920x1080_n50_fp32.md and (kernel 3, one
# launch of all n = 50 sweeps) profiles/resident_1920x1080_n50_fp32.md.
MEASURED_LAUNCH_TRAFFIC = {
    (2, 108, 44, 5, 1920, 1080): 38_090_423,
    (2, 104, 40, 6, 1920, 1080): 38_554_377,
    (3, 108, 44, 5, 1920, 1080): 248_500_478,
    (3, 96, 48, 8, 1920, 1080): 256_970_333,
    (3, 104, 56, 6, 1920, 1080): 263_479_390,
    (3, 112, 48, 4, 1920, 1080): 256_964_802,
}
METRIC = "Gcell-steps/s at 1920x1080 n=50; % HBM roofline; p50/p99 batched step latency"


def launch_traffic(tuning: dict, W: int, H: int):
    """Measured HBM bytes of one projection launch of this shape, or None."""
    return MEASURED_LAUNCH_TRAFFIC.get((tuning["projection_kernel"], tuning["tile_width"],
                                        tuning["tile_height"], tuning["sweeps_per_launch"],
                                        W, H))


def step_bytes_per_cell(n: int) -> int:
    """Algorithmic bytes per cell-step, FP32, pressure off: forces 8 (v read
    and written), 17 per sweep, u/v advection 17, smoke advection 17."""
    return 42 + 17 * n


def alg_bytes_per_cell(cfg) -> int:
    """Algorithmic bytes per cell-step of one configuration: FP32 storage
    42 + 17n (46 + 25n with pressure), 16-bit storage 22 + 9n (26 + 17n with
    the FP32 pressure)."""
    n = cfg.projection_n
    if cfg.storage == 0:
        return 46 + 25 * n if cfg.enable_pressure else 42 + 17 * n
    return 26 + 17 * n if cfg.enable_pressure else 22 + 9 * n


def host_threads() -> int:
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(n, 16))


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(cfg, seed: int, budget_s: float = 15.0) -> dict:
    """Times the oracle (FP32, OpenMP over rows, threads bound to cores) on
    this host: three repetitions of a bounded number of steps, median."""
    # the OpenMP runtime reads these when the oracle library first runs
    os.environ.setdefault("OMP_PLACES", "cores")
    os.environ.setdefault("OMP_PROC_BIND", "close")
    from opensayal_amd import workloads as wl
    from oracle.oracle import OracleFluid

    threads = host_threads()
    o = OracleFluid(cfg, nthreads=threads)
    u, v, s = wl.initial_state(cfg, seed)
    o.set_state(u=u, v=v, smoke=s)
    t0 = time.perf_counter()
    o.update(steps=1)
    first = time.perf_counter() - t0
    steps = max(2, min(400, int(budget_s / 3 / max(first, 1e-6))))
    reps = []
    for _ in range(3):
        t0 = time.perf_counter()
        o.update(steps=steps)
        reps.append(time.perf_counter() - t0)
    dt = sorted(reps)[1]
    cells = cfg.width * cfg.height
    return {
        "value": cells * steps / dt / 1e9,
        "unit": "Gcell-steps/s",
        "cores": threads,
        "kind": "port",
        "sample": f"median of 3 runs of {steps} steps of the same {cfg.width}x{cfg.height} "
                  f"n={cfg.projection_n} workload after 1 untimed step "
                  f"({sum(reps):.1f} s in all), oracle FP32, OpenMP {threads} threads",
        "ms_per_step": dt / steps * 1e3,
        "cpu_model": cpu_model(),
        "host_cpus": os.cpu_count(),
    }


def pct(xs, q):
    import numpy as np

    return float(np.percentile(np.asarray(xs, dtype=np.float64), q))


KIND_NAMES = ("forces", "projection", "velocity_advection", "smoke_advection")


def item_timing(b) -> dict:
    """One batch step with per-item timing: per stage kind, the time items
    waited for their dependencies after being claimed and the time their
    body ran (microseconds), and the step's span from the first claim to
    the last completion."""
    import numpy as np

    b.set_timing(True)
    b.step(2)  # the first rebuilds the tables with the stamps
    r = b.item_times()
    b.set_timing(False)
    if len(r["kind"]) == 0:  # every member steps on its own per-grid graphs
        return {"items": 0}
    wait = (r["ready"] - r["claim"]) / 1e3
    body = (r["done"] - r["ready"]) / 1e3
    span = float(r["done"].max() - r["claim"].min()) / 1e3
    out = {"span_us": round(span, 2), "items": int(len(wait)),
           "cus": int(len(np.unique(r["hw_id"]))), "tick_ns": r["tick_ns"], "kinds": {}}
    for k, name in enumerate(KIND_NAMES):
        m = r["kind"] == k
        if not m.any():
            continue
        out["kinds"][name] = {
            "items": int(m.sum()),
            "wait_us_mean": round(float(wait[m].mean()), 3),
            "wait_us_p99": round(pct(list(wait[m]), 99), 3),
            "body_us_mean": round(float(body[m].mean()), 3),
            "body_us_p50": round(pct(list(body[m]), 50), 3),
            "body_us_p99": round(pct(list(body[m]), 99), 3),
            "wait_us_sum": round(float(wait[m].sum()), 1),
            "body_us_sum": round(float(body[m].sum()), 1),
        }
    return out


def batch_leg(specs, steps: int, warmup: int, plain_steps: int = 3, autotune: bool = True,
              timing: bool = False) -> dict:
    """Steps every simulation of `specs` together as one Batch (one persistent
    launch per step, replayed from a HIP graph), timing each step with HIP
    events on the stream the steps run on; then the same simulations one
    after another through the per-grid path (each its own captured graph)
    for comparison."""
    from opensayal_amd import Batch, Event, Stream, workloads as wl

    t0 = time.perf_counter()
    fluids = [wl.start(sp) for sp in specs]
    setup_s = time.perf_counter() - t0
    cells = sum(f.width * f.height for f in fluids)
    stream = Stream()
    b = Batch()
    for f in fluids:
        b.add(f)
    tuning, tune_ms = b.autotune() if autotune else (b.tuning, None)
    b.step_async(warmup, stream=stream)
    b.synchronize()
    evs = [Event() for _ in range(steps + 1)]
    evs[0].record(stream)
    for k in range(steps):
        b.step_async(1, stream=stream)
        evs[k + 1].record(stream)
    b.synchronize()
    per = [evs[k].elapsed_ms(evs[k + 1]) for k in range(steps)]
    total_ms = evs[0].elapsed_ms(evs[-1])
    info = b.info
    timed = item_timing(b) if timing else None
    b.close()

    # the per-grid path: each simulation alone, sequentially on one stream
    for f in fluids:
        f.step_async(1, stream=stream)  # capture its graphs (untimed)
    stream.synchronize()
    a, z = Event(), Event()
    a.record(stream)
    for _ in range(plain_steps):
        for f in fluids:
            f.step_async(1, stream=stream)
    z.record(stream)
    stream.synchronize()
    plain_ms = a.elapsed_ms(z) / plain_steps
    for f in fluids:
        f.close()
    value = cells * steps / (total_ms / 1e3) / 1e9
    plain = cells / (plain_ms / 1e3) / 1e9
    ns = sorted(sp.cfg.projection_n for sp in specs)
    alg = sum(sp.cfg.width * sp.cfg.height * alg_bytes_per_cell(sp.cfg) for sp in specs)
    achieved = alg * steps / (total_ms / 1e3) / 1e9
    return {
        "sims": len(specs), "cells_per_step": cells, "steps": steps, "warmup": warmup,
        "n_values": sorted(set(ns)),
        "value": round(value, 4), "unit": "Gcell-steps/s",
        "ms_per_step": round(total_ms / steps, 4),
        "step_ms_p50": round(pct(per, 50), 4), "step_ms_p99": round(pct(per, 99), 4),
        "tuning": tuning, "autotune_ms_per_step": None if tune_ms is None else round(tune_ms, 4),
        "work_items_per_step": info["work_items"], "groups": info["groups"],
        "plain_sequential": {"value": round(plain, 4), "ms_per_step": round(plain_ms, 4),
                             "steps": plain_steps},
        "speedup_vs_plain": round(value / plain, 3),
        "step_roofline": {"alg_bytes_per_step": int(alg), "achieved_GBs": round(achieved, 1),
                          "frac": round(achieved / HBM_PEAK_GBS, 4)},
        "setup_s": round(setup_s, 2),
        **({"item_timing": timed} if timed else {}),
    }


def diffusion_leg(seed: int, steps: int, warmup: int, autotune: bool = True) -> dict:
    """The same 1920x1080 step with the reference's default fluid.viscosity
    (0.001): the forces pass, n = 50 Jacobi diffusion rounds (9 launches) and
    then the plain step's projection and advections."""
    from opensayal_amd import Event, Fluid, Stream, workloads as wl

    cfg = wl.plain_full(viscosity=0.001)
    fl = Fluid(cfg)
    u, v, s = wl.initial_state(cfg, seed)
    fl.set_field("vel_x", u)
    fl.set_field("vel_y", v)
    fl.set_field("smoke", s)
    stream = Stream()
    if autotune:
        fl.autotune()
    fl.step_async(warmup, stream=stream)
    a, b = Event(), Event()
    a.record(stream)
    fl.step_async(steps, stream=stream)
    b.record(stream)
    stream.synchronize()
    ms = a.elapsed_ms(b) / steps
    fl.close()
    return {"viscosity": 0.001, "steps": steps, "ms_per_step": round(ms, 4),
            "value": round(cfg.width * cfg.height / (ms / 1e3) / 1e9, 4),
            "unit": "Gcell-steps/s"}


def precision_leg(seed: int, steps: int, warmup: int, autotune: bool = True) -> dict:
    """The mixed-precision configuration, 3840x2160, n = 100, pressure on
    (FP32 pressure), stepped with u, v, smoke stored as FP32, FP16 and BF16.
    Algorithmic bytes per cell-step: 46 + 25n in FP32, 26 + 17n with 16-bit
    storage (forces, the pressure zeroing, per sweep u and v read and written
    plus the mask and the pressure read and written, both advections)."""
    from opensayal_amd import Event, Fluid, Stream, workloads as wl

    out = {}
    for storage in ("f32", "f16", "bf16"):
        cfg = wl.mixed_precision(storage)
        n = cfg.projection_n
        fl = Fluid(cfg)
        u, v, s = wl.initial_state(cfg, seed)
        fl.set_field("vel_x", u)
        fl.set_field("vel_y", v)
        fl.set_field("smoke", s)
        stream = Stream()
        if autotune:
            fl.autotune()
        shape = fl.tuning
        fl.step_async(warmup, stream=stream)
        a, b = Event(), Event()
        a.record(stream)
        fl.step_async(steps, stream=stream)
        b.record(stream)
        stream.synchronize()
        ms = a.elapsed_ms(b) / steps
        fl.close()
        value = cfg.width * cfg.height / (ms / 1e3) / 1e9
        per_cell = 46 + 25 * n if storage == "f32" else 26 + 17 * n
        out[storage] = {"ms_per_step": round(ms, 4), "value": round(value, 4),
                        "alg_bytes_per_cell_step": per_cell,
                        "step_roofline_frac": round(value * 1e9 * per_cell / 8e12, 4),
                        "projection_shape": shape}
    out.update({"workload": "3840x2160 n=100 pressure on", "steps": steps,
                "unit": "Gcell-steps/s"})
    return out


def view_leg(fl, reps: int) -> dict:
    """What GraphicsHandler reads each frame, as the synchronous library
    calls a viewer makes: the RGBA frame (kernel plus the W*H*4-byte copy to
    host) and the path lines at the config defaults (distance 20, length 20).
    Wall-clock per call; the kernels' own times are in the rocprofv3 trace."""
    fl.render_rgba()
    fl.traces(20, 20)
    t0 = time.perf_counter()
    for _ in range(reps):
        fl.render_rgba()
    frame_ms = (time.perf_counter() - t0) * 1e3 / reps
    t0 = time.perf_counter()
    for _ in range(reps):
        fl.traces(20, 20)
    traces_ms = (time.perf_counter() - t0) * 1e3 / reps
    W, H = fl.width, fl.height
    return {"calls": reps, "ms_per_frame": round(frame_ms, 4),
            "frame_host_GBs": round(W * H * 4 / (frame_ms / 1e3) / 1e9, 2),
            "ms_per_traces": round(traces_ms, 4),
            "seeds": (W // 20) * (H // 20), "trace_length": 20}


def sweep_shapes(fl, stream, steps: int = 40) -> None:
    """Per-step time for a grid of projection launch shapes (one JSON line each)."""
    from opensayal_amd import Event

    shapes = [dict(projection_kernel=0)]
    for tw in (112, 176, 240):
        for th in (16, 24):
            for k in (3, 4):
                shapes.append(dict(projection_kernel=1, tile_width=tw, tile_height=th,
                                   sweeps_per_launch=k))
    for rows in (4, 8, 10, 12, 16):
        for k in (1, 2, 3, 4, 5, 6, 7, 8, 10, 12):
            if 8 * rows - 4 * k >= 1:
                shapes.append(dict(projection_kernel=2, tile_width=128 - 4 * k,
                                   tile_height=8 * rows - 4 * k, sweeps_per_launch=k))
    for sh in shapes:
        try:
            fl.set_tuning(**sh)
        except Exception:
            continue
        fl.step_async(3, stream=stream)
        a, b = Event(), Event()
        a.record(stream)
        fl.step_async(steps, stream=stream)
        b.record(stream)
        stream.synchronize()
        ms = a.elapsed_ms(b) / steps
        print(json.dumps({"shape": sh, "ms_per_step": round(ms, 4),
                          "gcell_steps_s": round(fl.width * fl.height / ms / 1e6, 3)}), flush=True)


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=300)
    ap.add_argument("--warmup", type=int, default=30)
    ap.add_argument("--workload", default="full",
                    choices=["full", "small", "batched", "mixed_small", "serving",
                             "lone_batch"])
    ap.add_argument("--item-timing", action="store_true",
                    help="batched workloads: add one step with per-item timing and report "
                         "wait and body time per stage kind")
    ap.add_argument("--batch-steps", type=int, default=20,
                    help="timed batch steps of the batched legs of the default run")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched legs")
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--kernel-reps", type=int, default=20)
    ap.add_argument("--no-view", action="store_true", help="skip the frame export leg")
    ap.add_argument("--no-precision", action="store_true",
                    help="skip the 3840x2160 FP32/FP16/BF16 storage leg")
    ap.add_argument("--no-diffusion", action="store_true",
                    help="skip the leg with the default viscosity")
    ap.add_argument("--view-reps", type=int, default=20)
    ap.add_argument("--requests", type=int, default=64, help="serving: number of requests")
    ap.add_argument("--serving-requests", type=int, default=32,
                    help="requests of the serving leg of the default run")
    ap.add_argument("--serving-seed", type=int, default=5,
                    help="seed of the serving leg's arrivals and requests")
    ap.add_argument("--no-serving", action="store_true", help="skip the serving leg")
    ap.add_argument("--rate", type=float, default=16.0, help="serving: arrivals per second")
    ap.add_argument("--budget-gb", type=float, default=200.0, help="serving: HBM state budget")
    ap.add_argument("--chunk", type=int, default=16, help="serving: steps between admissions")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="run only device-to-device copies of a known size (to scale "
                         "FETCH_SIZE / WRITE_SIZE counters)")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the default projection shape (for profiling runs)")
    ap.add_argument("--shape", default=None,
                    help="projection launch shape KERNEL,TILE_W,TILE_H,SWEEPS instead of "
                         "the autotuner's (for profiling one shape)")
    ap.add_argument("--sweep", action="store_true",
                    help="time a grid of projection launch shapes instead")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this build measures one GPU only; each simulation runs on a single "
              "device (multi-GPU sharding is designed, not built): run with --gpus 1",
              file=sys.stderr)
        return 2

    from opensayal_amd import Event, Fluid, Stream, lib as _lib, workloads as wl

    if args.calibrate_copy:
        nbytes = 512 * 1024 * 1024  # read and written per copy
        from opensayal_amd._abi import check

        check(_lib().osy_calibration_copy(nbytes, 5))
        print(json.dumps({"calibrate_copy": {"kernel": "k_copy", "bytes_read": nbytes,
                                             "bytes_written": nbytes, "copies": 5}}), flush=True)
        return 0

    if args.workload in ("batched", "mixed_small", "lone_batch"):
        specs = {"batched": wl.batched_specs, "mixed_small": wl.small_specs,
                 "lone_batch": lambda: [wl.SimSpec(wl.plain_full(), args.seed, 0)]}[args.workload]()
        out = batch_leg(specs, args.steps, args.warmup, autotune=not args.no_autotune,
                        timing=args.item_timing)
        out.update({"metric": METRIC, "n_gpus": 1, "higher_is_better": True,
                    "scaling": "weak", "vs_baseline": None, "dtype": "f32",
                    "data": "synthetic: seeded +-0.01 velocity noise, smoke 1 on the inflow",
                    "config": {"workload": args.workload}})
        print(json.dumps(out), flush=True)
        return 0
    if args.workload == "serving":
        return serving_main(args)

    cfg = wl.plain_full() if args.workload == "full" else wl.plain_small()
    W, H, n = cfg.width, cfg.height, cfg.projection_n
    fl = Fluid(cfg)
    u, v, s = wl.initial_state(cfg, args.seed)
    fl.set_field("vel_x", u)
    fl.set_field("vel_y", v)
    fl.set_field("smoke", s)

    stream = Stream()

    if args.sweep:
        sweep_shapes(fl, stream)
        return 0

    # the library's autotuner picks the projection launch shape (not timed)
    if args.shape:
        kern, tw, th, k = (int(x) for x in args.shape.split(","))
        fl.set_tuning(projection_kernel=kern, tile_width=tw, tile_height=th, sweeps_per_launch=k)
        tuning, tune_ms = fl.tuning, float("nan")
    elif args.no_autotune:
        tuning, tune_ms = fl.tuning, float("nan")
    else:
        tuning, tune_ms = fl.autotune()

    fl.step_async(args.warmup, stream=stream)
    stream.synchronize()

    # timed region: exactly K steps, HIP events on the stream they run on
    start, end = Event(), Event()
    t0 = time.perf_counter()
    start.record(stream)
    fl.step_async(args.steps, stream=stream)
    end.record(stream)
    stream.synchronize()
    wall = time.perf_counter() - t0
    gpu_ms = start.elapsed_ms(end)
    ms_per_step = gpu_ms / args.steps
    # four more repetitions of the same K steps, for the spread (not `value`)
    reps = [ms_per_step]
    for _ in range(4):
        start.record(stream)
        fl.step_async(args.steps, stream=stream)
        end.record(stream)
        stream.synchronize()
        reps.append(start.elapsed_ms(end) / args.steps)
    value = W * H * args.steps / (gpu_ms / 1e3) / 1e9

    # dominant kernel: kernel_reps projection launches of sweeps_per_launch
    # sweeps back to back (the resident kernel: kernel_reps launches of all n
    # sweeps), timed with events on the stream they run on
    resident = fl.projection_kernel == 3
    spl = n if resident else fl.sweeps_per_launch
    fl.project(sweeps=spl, stream=stream, sync=False)
    a, b = Event(), Event()
    a.record(stream)
    if resident:
        for _ in range(args.kernel_reps):
            fl.project(sweeps=spl, stream=stream, sync=False)
    else:
        fl.project(sweeps=spl * args.kernel_reps, stream=stream, sync=False)
    b.record(stream)
    stream.synchronize()
    launches = args.kernel_reps if resident else fl.launches_per_call(spl * args.kernel_reps)
    per_launch_ms = a.elapsed_ms(b) / launches
    alg_bytes_per_launch = 17 * spl * args.kernel_reps * W * H / launches
    achieved = alg_bytes_per_launch / (per_launch_ms / 1e3) / 1e9
    if launch_traffic(tuning, W, H) is None:
        # a shape without a FETCH_SIZE / WRITE_SIZE measurement: say so, and
        # carry the shapes that were measured rather than a silent null
        traffic_note = {"traffic_missing": "no rocprofv3 FETCH_SIZE/WRITE_SIZE measurement "
                                           "for this launch shape",
                        "traffic_measured": [{"shape": list(k[:4]), "grid": list(k[4:]),
                                              "bytes": b}
                                             for k, b in MEASURED_LAUNCH_TRAFFIC.items()]}
    else:
        traffic_note = {}
    roofline = {
        "bound": "hbm",
        "achieved": round(achieved, 1),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4),
        "traffic": launch_traffic(tuning, W, H),
        "kernel": fl.projection_kernel_name,
        "avg_launch_us": round(per_launch_ms * 1e3, 2),
        "alg_bytes_per_launch": int(alg_bytes_per_launch),
        **traffic_note,
    }
    step_alg = step_bytes_per_cell(n) * W * H * args.steps / (gpu_ms / 1e3) / 1e9
    out = {
        "metric": METRIC,
        "value": round(value, 4),
        "unit": "Gcell-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: seeded +-0.01 velocity noise, smoke 1 on the inflow",
        "config": {"workload": f"plain_{args.workload}_{W}x{H}_n{n}", "width": W, "height": H,
                   "projection_n": n, "projection_o": round(cfg.projection_o, 4),
                   "pipe_height": cfg.pipe_height, "obstacle_radius": cfg.obstacle_radius,
                   "seed": args.seed, "projection_shape": tuning,
                   "projection_ms_autotune": round(tune_ms, 4)},
        "roofline": roofline,
        "step_roofline": {"alg_bytes_per_cell_step": step_bytes_per_cell(n),
                          "achieved_GBs": round(step_alg, 1),
                          "frac": round(step_alg / HBM_PEAK_GBS, 4)},
        "wall_s": round(wall, 3),
        "reps_ms_per_step": [round(x, 4) for x in reps],
        "median_ms_per_step": round(sorted(reps)[2], 4),
    }
    if not args.no_view:
        out["frame_export"] = view_leg(fl, args.view_reps)
    fl.close()
    if not args.no_precision and args.workload == "full":
        out["mixed_precision"] = precision_leg(args.seed, max(4, args.steps // 15), 5,
                                               autotune=not args.no_autotune)
    if not args.no_diffusion and args.workload == "full":
        out["with_diffusion"] = diffusion_leg(args.seed, args.steps // 4, args.warmup,
                                              autotune=not args.no_autotune)
    if not args.no_batch and args.workload == "full":
        # 256 uneven grids (128..1024 wide, n from {20, 50, 100}) and 256
        # small ones (64..384 wide), each as one batch against the per-grid path
        out["batched"] = batch_leg(wl.batched_specs(), args.batch_steps, 3)
        out["batched_mixed_small"] = batch_leg(wl.small_specs(), args.batch_steps * 5, 5)
    if not args.no_serving and args.workload == "full":
        out["serving"] = serving_leg(args.serving_requests, args.rate, args.budget_gb,
                                     args.chunk, args.serving_seed)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(cfg, args.seed)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)
    return 0


def serving_leg(requests: int, rate_hz: float, budget_gb: float, chunk: int,
                seed: int) -> dict:
    """Poisson arrivals of requests from the serving workload's distribution
    (width log-uniform 64..2048 at 16:9, 50..2000 steps, n from {20, 50,
    100}, FP32 or FP16 storage), served by continuous batching under an HBM
    budget: the resident requests step together, one persistent launch per
    batch step, each batch step timed by an event pair on the batch's
    stream; requests arriving meanwhile are created and uploaded while a
    chunk runs and join at its end.  Throughput is all requests' cell-steps
    over the wall time from the first arrival to the last completion."""
    from opensayal_amd import workloads as wl
    from opensayal_amd.serving import DeviceBackend, Server

    arrivals = wl.serving_specs(count=requests, seed=seed, rate_hz=rate_hz)
    backend = DeviceBackend(max_chunk=chunk)
    try:
        rep = Server(backend, budget_bytes=int(budget_gb * 1e9), max_chunk=chunk).run(arrivals)
    finally:
        backend.close()
    s = rep.summary()
    out = {k: (round(v, 5) if isinstance(v, float) else v) for k, v in s.items()}
    out.update({"value": out["gcell_steps_s_wall"], "unit": "Gcell-steps/s",
                "gpu_busy_frac": round(s["gpu_busy_s"] / max(s["wall_s"], 1e-9), 4),
                "rate_hz": rate_hz, "budget_gb": budget_gb, "max_chunk": chunk, "seed": seed,
                "widths": [min(r.spec.cfg.width for r in rep.requests),
                           max(r.spec.cfg.width for r in rep.requests)]})
    return out


def serving_main(args) -> int:
    """Poisson arrivals of mixed requests served by continuous batching."""
    s = serving_leg(args.requests, args.rate, args.budget_gb, args.chunk, args.seed)
    out = {"metric": METRIC, "value": s["value"], "unit": "Gcell-steps/s",
           "n_gpus": 1, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "f32/f16 storage", "data": "synthetic requests: Poisson arrivals",
           "config": {"workload": "serving", "requests": args.requests, "rate_hz": args.rate,
                      "budget_gb": args.budget_gb, "max_chunk": args.chunk, "seed": args.seed},
           "serving": s}
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
