"""Host logic of the serving loop (no GPU): admission by free arena memory in arrival
order, chunked steps that stop exactly at each request's step count, retirement, the
deterministic step clock and the latency summary.  A stand-in batch records what the
loop asks of it."""
import numpy as np
import pytest

from pixelflow_amd import cases
from pixelflow_amd.serve import Request, Server, latency_stats


class FakeBatch:
    def __init__(self, max_cases, arena):
        self.max_cases, self.arena = max_cases, arena
        self.slots = {}
        self.calls = []
        self.diverge_at = set()

    def case_bytes(self, p):
        return p.cells * 100

    def free_bytes(self):
        return self.arena - sum(self.case_bytes(p) for p, _ in self.slots.values())

    def add(self, p, phi):
        assert self.case_bytes(p) <= self.free_bytes()
        s = min(set(range(self.max_cases)) - set(self.slots))
        self.slots[s] = (p, 0)
        self.calls.append(("add", s))
        return s

    def remove(self, s):
        del self.slots[s]
        self.calls.append(("remove", s))

    def case_steps(self, s):
        return self.slots[s][1] if s in self.slots else -1

    def step(self, n, raise_diverged=True):
        self.calls.append(("step", n, tuple(sorted(self.slots))))
        for s, (p, k) in list(self.slots.items()):
            self.slots[s] = (p, k + n)
        z = np.zeros((n, self.max_cases))
        e = z.copy()
        for s in self.diverge_at:
            if s in self.slots and self.slots[s][1] >= 2:
                e[-1, s] = np.nan
        return {"sor_error": e, "residual": z - 1, "sor_iters": z.astype(np.int32), "cycles": z.astype(np.int32)}

    def interior(self, s, f):
        return np.full(self.slots[s][0].interior_shape(), float(s))

    def close(self):
        pass


def _req(rid, m, steps, arrival):
    p = cases.CaseParams(dim=2, m=m, n=m // 2, width=1.0, height=0.5, time=1e-3, istep_max=steps, iter_max=5)
    return Request(rid, p, np.ones(p.cells), steps, arrival)


def test_admission_waits_for_memory_fifo():
    b = FakeBatch(max_cases=8, arena=100 * (32 * 16 + 32 * 16))       # room for two 32x16 cases
    srv = Server(None, 1, batch=b, clock="steps", chunk=3)
    reqs = [_req(0, 32, 5, 0), _req(1, 32, 7, 0), _req(2, 32, 4, 1), _req(3, 16, 2, 1)]
    for r in reqs:
        srv.submit(r)
    done = srv.run()
    assert [r.rid for r in sorted(done, key=lambda r: r.admitted)] == [0, 1, 2, 3]
    by = {r.rid: r for r in done}
    # request 2 waits for request 0 to retire (step 5); 3 fits only after 1 (FIFO, no overtaking)
    assert by[0].admitted == 0 and by[1].admitted == 0
    assert by[2].admitted == 5 and by[0].finished == 5
    assert by[3].admitted >= by[2].admitted
    for r in done:
        assert r.finished - r.admitted >= r.steps
    # every step chunk ends no later than the earliest remaining request's count
    steps = [c[1] for c in b.calls if c[0] == "step"]
    assert all(1 <= n <= 3 for n in steps)
    assert srv.steps_run == sum(steps)
    st = latency_stats(done)
    assert st["requests"] == 4 and st["p99_time_to_solution_s"] >= st["p50_time_to_solution_s"]


def test_idle_clock_jumps_to_next_arrival():
    b = FakeBatch(max_cases=2, arena=10**9)
    srv = Server(None, 1, batch=b, clock="steps", chunk=10)
    srv.submit(_req(0, 16, 3, 0))
    srv.submit(_req(1, 16, 2, 50))
    done = srv.run()
    assert [r.rid for r in done] == [0, 1]
    assert done[1].admitted == 50 and done[1].finished == 52


def test_case_larger_than_arena_raises():
    b = FakeBatch(max_cases=2, arena=10)
    srv = Server(None, 1, batch=b, clock="steps")
    srv.submit(_req(0, 16, 3, 0))
    with pytest.raises(RuntimeError, match="needs"):
        srv.run()


def test_results_downloaded_before_removal():
    b = FakeBatch(max_cases=4, arena=10**9)
    srv = Server(None, 1, batch=b, clock="steps", keep_fields=("u", "p"))
    srv.submit(_req(0, 16, 2, 0))
    done = srv.run()
    assert set(done[0].result) == {"u", "p"}
    assert done[0].result["u"].shape == (8, 16)
    assert ("remove", 0) in b.calls


def test_diverged_case_retires_without_result_and_others_finish():
    b = FakeBatch(max_cases=4, arena=10**9)
    b.diverge_at = {1}
    srv = Server(None, 1, batch=b, clock="steps", chunk=1)
    for r in (_req(0, 16, 5, 0), _req(1, 16, 5, 0)):
        srv.submit(r)
    done = {r.rid: r for r in srv.run()}
    assert done[1].error == "diverged" and done[1].result is None and done[1].finished == 2
    assert done[0].error is None and done[0].finished == 5
    st = latency_stats(list(done.values()))
    assert st["requests"] == 1 and st["failed"] == 1
