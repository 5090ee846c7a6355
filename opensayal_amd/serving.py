"""Serving loop: continuous batching of simulation requests on one GPU.

Requests (a configuration, a seed for the initial state, a step count)
arrive over time.  The loop admits waiting requests in arrival order while
their device memory fits the budget and a slot is free (a request that does
not fit waits; later, smaller ones may pass it), keeps every admitted
simulation's state resident in HBM as a member of one ``Batch``, and steps
all members together in chunks: one persistent kernel launch per step (per
storage group), so the work of every resident simulation overlaps on the
device.  A chunk never runs past the step count of any member, so each
request runs exactly its steps.  While a chunk runs on the GPU, the host
admits the requests that arrived meanwhile: it creates their handles and
uploads their initial state on their own streams, and they join the batch
at the end of the chunk, when finished requests are collected (fields
copied to the host if asked), leave the batch and free their memory.  Each
batch step is timed by an event pair on the batch's stream.

Each request's result is bit-identical to stepping it alone, because a batch
member's fields are (see ``Batch``).

The scheduler is separate from the device through a small backend interface,
so its admission and accounting rules run on the CPU in tests.
"""
from __future__ import annotations

import time
from collections import deque
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

from .workloads import SimSpec


@dataclass
class Request:
    rid: int
    spec: SimSpec
    arrival: float                 # seconds after the start of the run
    bytes: int = 0
    admitted: Optional[float] = None
    finished: Optional[float] = None
    steps_done: int = 0
    handle: object = None
    result: Optional[Dict[str, np.ndarray]] = None

    @property
    def remaining(self) -> int:
        return self.spec.steps - self.steps_done

    @property
    def cells(self) -> int:
        return self.spec.cfg.width * self.spec.cfg.height


@dataclass
class ServingReport:
    requests: List[Request]
    step_ms: List[float] = field(default_factory=list)   # one entry per batch step
    chunks: int = 0
    wall_s: float = 0.0
    gpu_s: float = 0.0
    peak_bytes: int = 0
    peak_members: int = 0

    def summary(self) -> dict:
        done = [r for r in self.requests if r.finished is not None]
        lat = np.array([r.finished - r.arrival for r in done]) if done else np.zeros(1)
        wait = np.array([r.admitted - r.arrival for r in done]) if done else np.zeros(1)
        steps = np.array(self.step_ms) if self.step_ms else np.zeros(1)
        cell_steps = float(sum(r.cells * r.spec.steps for r in done))
        return {
            "requests": len(self.requests),
            "completed": len(done),
            "cell_steps": cell_steps,
            "gcell_steps_s_wall": cell_steps / max(self.wall_s, 1e-9) / 1e9,
            "gcell_steps_s_gpu": cell_steps / max(self.gpu_s, 1e-9) / 1e9,
            "step_ms_p50": float(np.percentile(steps, 50)),
            "step_ms_p99": float(np.percentile(steps, 99)),
            "completion_s_p50": float(np.percentile(lat, 50)),
            "completion_s_p99": float(np.percentile(lat, 99)),
            "admission_wait_s_p50": float(np.percentile(wait, 50)),
            "admission_wait_s_p99": float(np.percentile(wait, 99)),
            "batch_steps": len(self.step_ms),
            "chunks": self.chunks,
            "wall_s": self.wall_s,
            "gpu_busy_s": self.gpu_s,
            "peak_device_bytes": self.peak_bytes,
            "peak_members": self.peak_members,
        }


class DeviceBackend:
    """The GPU side: one ``Batch`` whose members are the resident requests.

    ``prepare`` creates a request's handle and uploads its state (on the
    handle's own stream, so it may run while the batch steps); ``join`` adds
    it to the batch between chunks; ``launch`` enqueues steps with an event
    pair around each and returns at once; ``wait`` returns each step's GPU
    milliseconds."""

    def __init__(self, tuning: Optional[dict] = None, collect: Sequence[str] = (),
                 max_chunk: int = 64):
        from . import Batch, Event, Stream

        self.batch = Batch()
        if tuning:
            self.batch.set_tuning(**tuning)
        self.collect = tuple(collect)
        self.stream = Stream()
        self._events = [Event() for _ in range(max_chunk + 1)]
        self._launched = 0

    def bytes(self, spec: SimSpec) -> int:
        from . import device_bytes

        return device_bytes(spec.cfg)

    def prepare(self, spec: SimSpec):
        from .workloads import start

        return start(spec)

    def join(self, handle) -> None:
        self.batch.add(handle)

    def launch(self, steps: int) -> None:
        if steps > len(self._events) - 1:
            from . import Event

            self._events += [Event() for _ in range(steps + 1 - len(self._events))]
        ev = self._events
        ev[0].record(self.stream)
        for k in range(steps):
            self.batch.step_async(1, stream=self.stream)
            ev[k + 1].record(self.stream)
        self._launched = steps

    def wait(self) -> List[float]:
        self.batch.synchronize()
        ev = self._events
        return [ev[k].elapsed_ms(ev[k + 1]) for k in range(self._launched)]

    def retire(self, handle) -> Optional[Dict[str, np.ndarray]]:
        out = {n: handle.get_field(n) for n in self.collect} if self.collect else None
        self.batch.remove(handle)
        handle.close()
        return out

    def close(self) -> None:
        self.batch.close()
        self.stream.close()


class Server:
    """Continuous-batching scheduler over a backend (``DeviceBackend`` on the
    GPU).  ``budget_bytes`` bounds the resident requests' device memory,
    ``max_slots`` their number, ``max_chunk`` the steps run between
    admission rounds."""

    def __init__(self, backend, budget_bytes: int = 200 * 10**9, max_slots: int = 1024,
                 max_chunk: int = 32):
        self.backend = backend
        self.budget = int(budget_bytes)
        self.max_slots = int(max_slots)
        self.max_chunk = int(max_chunk)

    def run(self, arrivals: Sequence[Tuple[float, SimSpec]], realtime: bool = True,
            clock=time.perf_counter, sleep=time.sleep) -> ServingReport:
        reqs = [Request(i, spec, float(t)) for i, (t, spec) in enumerate(arrivals)]
        for r in reqs:
            r.bytes = int(self.backend.bytes(r.spec))
            if r.bytes > self.budget:
                raise ValueError(f"request {r.rid} needs {r.bytes} bytes, over the whole budget")
        pending = deque(sorted(reqs, key=lambda r: (r.arrival, r.rid)))
        waiting: List[Request] = []
        active: List[Request] = []
        rep = ServingReport(requests=reqs)
        used = 0  # device bytes of the resident and prepared requests
        t0 = clock()
        virtual = 0.0  # the clock when not running in real time: GPU time only

        def now() -> float:
            return clock() - t0 if realtime else virtual

        ready: List[Request] = []  # prepared while a chunk ran; join at its end

        def admit() -> None:
            nonlocal used
            for r in list(waiting):
                if len(active) + len(ready) >= self.max_slots:
                    break
                if used + r.bytes <= self.budget:
                    r.handle = self.backend.prepare(r.spec)
                    r.admitted = now()
                    used += r.bytes
                    ready.append(r)
                    waiting.remove(r)
            rep.peak_bytes = max(rep.peak_bytes, used)
            rep.peak_members = max(rep.peak_members, len(active) + len(ready))

        while pending or waiting or ready or active:
            t = now()
            while pending and pending[0].arrival <= t:
                waiting.append(pending.popleft())
            for r in ready:
                self.backend.join(r.handle)
                active.append(r)
            ready.clear()
            if not active:
                # nothing on the GPU: admit now (arrival order, skipping
                # requests that do not fit yet) and start them
                admit()
                if ready:
                    continue
                if pending:
                    if realtime:
                        sleep(max(0.0, pending[0].arrival - now()))
                    else:
                        virtual = max(virtual, pending[0].arrival)
                    continue
                raise RuntimeError("requests wait but none can be admitted")
            chunk = min(self.max_chunk, min(r.remaining for r in active))
            self.backend.launch(chunk)
            # while the chunk runs: requests that arrived are admitted
            t = now()
            while pending and pending[0].arrival <= t:
                waiting.append(pending.popleft())
            admit()
            step_ms = self.backend.wait()
            rep.chunks += 1
            rep.gpu_s += sum(step_ms) / 1e3
            rep.step_ms.extend(step_ms)
            if not realtime:
                virtual += sum(step_ms) / 1e3
            for r in active:
                r.steps_done += chunk
            for r in [r for r in active if r.remaining == 0]:
                r.result = self.backend.retire(r.handle)
                r.handle = None
                r.finished = now()
                used -= r.bytes
                active.remove(r)
        rep.wall_s = now()
        return rep
