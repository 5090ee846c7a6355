"""Host logic of the serving loop, on the CPU with a recording backend:
budget and slot limits hold at every moment, each request runs exactly its
step count, admission follows arrival order (skipping what does not fit),
and the latency accounting is consistent."""
import pytest

from opensayal_amd.serving import Server
from opensayal_amd.workloads import SimSpec
from tests.helpers import make_config


class FakeBackend:
    """Counts steps per resident handle; 1 ms of 'GPU' per step per member."""

    def __init__(self):
        self.resident = {}
        self.log = []
        self.next = 0

    def bytes(self, spec):
        return spec.cfg.width * spec.cfg.height * 10

    def admit(self, spec):
        h = self.next
        self.next += 1
        self.resident[h] = [spec, 0]
        self.log.append(("admit", h, sum(self.bytes(s) for s, _ in self.resident.values()),
                         len(self.resident), spec.cfg.width))
        return h

    def step(self, steps):
        for v in self.resident.values():
            v[1] += steps
        self.log.append(("step", steps, [v[1] for v in self.resident.values()]))
        return float(steps * len(self.resident))

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
