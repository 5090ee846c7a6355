"""Host logic of the serving loop, on the CPU with a recording backend:
budget and slot limits hold at every moment, each request runs exactly its
step count, admission follows arrival order (skipping what does not fit),
and the latency accounting is consistent."""
import pytest

from opensayal_amd.serving import Server
from opensayal_amd.workloads import SimSpec
from tests.helpers import make_config


class FakeBackend:
    """Counts steps per resident handle; 1 ms of 'GPU' per step per member.
    Prepared handles hold memory but step only once they join."""

    def __init__(self):
        self.prepared = {}
        self.resident = {}
        self.log = []
        self.next = 0
        self.pending_steps = 0

    def bytes(self, spec):
        return spec.cfg.width * spec.cfg.height * 10

    def held(self):
        return list(self.prepared.values()) + [s for s, _ in self.resident.values()]

    def prepare(self, spec):
        h = self.next
        self.next += 1
        self.prepared[h] = spec
        held = self.held()
        self.log.append(("admit", h, sum(self.bytes(s) for s in held), len(held),
                         spec.cfg.width))
        return h

    def join(self, h):
        self.resident[h] = [self.prepared.pop(h), 0]

    def launch(self, steps):
        for v in self.resident.values():
            v[1] += steps
        self.log.append(("step", steps, [v[1] for v in self.resident.values()]))
        self.pending_steps = steps

    def wait(self):
        return [float(len(self.resident))] * self.pending_steps

    def retire(self, h):
        spec, done = self.resident.pop(h)
        self.log.append(("retire", h, done, spec.steps))
        return {"steps": done}


def spec(w, h, steps):
    return SimSpec(make_config(w, h), 0, steps)


def run(arrivals, **kw):
    be = FakeBackend()
    rep = Server(be, **kw).run(arrivals, realtime=False)
    return be, rep


def test_every_request_runs_exactly_its_steps():
    arrivals = [(0.001 * i, spec(16 + i, 9 + i, 5 + 7 * i)) for i in range(12)]
    be, rep = run(arrivals, max_chunk=8)
    for r in rep.requests:
        assert r.result == {"steps": r.spec.steps}
        assert r.finished >= r.admitted >= r.arrival
    for e in be.log:
        if e[0] == "retire":
            assert e[2] == e[3]


def test_budget_and_slots_are_never_exceeded():
    arrivals = [(0.0, spec(40, 30, 20 + i)) for i in range(10)]
    budget = 3 * 40 * 30 * 10
    be, rep = run(arrivals, budget_bytes=budget, max_slots=2)
    for e in be.log:
        if e[0] == "admit":
            assert e[2] <= budget and e[3] <= 2
    assert rep.peak_members == 2 and rep.peak_bytes <= budget
    assert all(r.finished is not None for r in rep.requests)


def test_admission_in_arrival_order_with_skip():
    big, small = spec(100, 50, 10), spec(10, 5, 10)
    budget = 100 * 50 * 10 + 10 * 5 * 10
    # the second big request does not fit while the first runs; the small
    # one behind it is admitted at once
    arrivals = [(0.0, big), (0.0, spec(100, 50, 10)), (0.0, small)]
    be, rep = run(arrivals, budget_bytes=budget)
    widths = [e[4] for e in be.log if e[0] == "admit"]
    assert widths == [100, 10, 100]
    r0, r1, r2 = rep.requests
    assert r2.admitted < r1.admitted


def test_chunks_do_not_run_past_any_member():
    arrivals = [(0.0, spec(20, 12, s)) for s in (3, 50, 17)]
    be, rep = run(arrivals, max_chunk=32)
    steps = [e[1] for e in be.log if e[0] == "step"]
    assert steps[0] == 3 and max(steps) <= 32
    assert sum(rep.step_ms) == pytest.approx(rep.gpu_s * 1e3)


def test_request_larger_than_budget_is_rejected():
    with pytest.raises(ValueError):
        run([(0.0, spec(100, 100, 5))], budget_bytes=1000)


def test_summary_fields():
    arrivals = [(0.01 * i, spec(16, 9, 10)) for i in range(5)]
    _, rep = run(arrivals)
    s = rep.summary()
    assert s["completed"] == 5 and s["cell_steps"] == 5 * 16 * 9 * 10
    assert s["step_ms_p99"] >= s["step_ms_p50"] > 0
    assert s["completion_s_p99"] >= s["completion_s_p50"] > 0


def test_admission_overlaps_the_running_chunk():
    """A request that arrives while others run is prepared during a chunk
    (after that chunk's launch, before its wait) and steps from the next."""
    arrivals = [(0.0, spec(30, 20, 40)), (0.0025, spec(20, 12, 8))]
    be, rep = run(arrivals, max_chunk=4)
    order = [e[0] for e in be.log]
    i_admit_b = [k for k, e in enumerate(be.log) if e[0] == "admit" and e[4] == 20][0]
    assert order[i_admit_b - 1] == "step"        # prepared while a chunk ran
    launched_before = [e for e in be.log[:i_admit_b] if e[0] == "step"]
    assert all(len(e[2]) == 1 for e in launched_before)  # B not stepped before it joined
    b = rep.requests[1]
    assert b.result == {"steps": 8} and b.admitted >= b.arrival
