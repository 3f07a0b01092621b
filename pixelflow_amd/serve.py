"""Serving loop: continuous batching of independent cases on one GPU.

Requests (a case: parameters + porosity + number of steps) arrive over time.  The loop
keeps one device batch (pixelflow_amd.solver.Batch) whose fields stay resident between
steps; between step chunks it

1. admits arrived requests, oldest first, while the batch arena has room for the case's
   footprint (the arena was sized from the device's free memory) and a slot is free;
2. advances every resident case by one chunk of steps in shared launches;
3. retires cases that reached their step count (their results are downloaded first), and
   cases whose step became non-finite (marked ``error="diverged"``, without a result).

Admission happens at step boundaries, so requests that arrive while others run join them
and their work overlaps on the device.  A case's arithmetic does not depend on which
other cases share the batch: every result equals a run of that case alone.
"""
from __future__ import annotations

import dataclasses
import time
from collections import deque
from typing import Callable, Deque, Dict, List, Optional

import numpy as np

from .cases import CaseParams
from .solver import Batch, device_memory


@dataclasses.dataclass
class Request:
    rid: int
    params: CaseParams
    phi: np.ndarray
    steps: int
    arrival: float                      # seconds on the loop's clock
    admitted: Optional[float] = None
    finished: Optional[float] = None
    slot: int = -1
    result: Optional[Dict[str, np.ndarray]] = None
    last_stats: Optional[Dict[str, float]] = None
    error: Optional[str] = None         # "diverged": retired early, no result


class Server:
    """Continuous-batching loop over one resident batch of one kind of case."""

    def __init__(self, kind: CaseParams, precision: int, max_cases: int = 64, arena_bytes: Optional[int] = None,
                 memory_fraction: float = 0.5, chunk: int = 5, keep_fields=("u", "v", "w", "p"),
                 clock=None, batch_config: Optional[dict] = None,
                 batch=None, **solver_kw):
        """`batch`: an existing batch object to serve from (default: a new device Batch whose
        arena is `arena_bytes`, or `memory_fraction` of the free device memory)."""
        self.precision = precision
        self.solver_kw = solver_kw
        if batch is None:
            if arena_bytes is None:
                free, _total = device_memory()
                arena_bytes = int(free * memory_fraction)
            batch = Batch(kind, max_cases=max_cases, arena_bytes=arena_bytes, precision=precision, **solver_kw)
        self.batch = batch
        if batch_config:
            self.batch.set_config(**batch_config)
        self.chunk = int(chunk)
        self.keep_fields = keep_fields
        self.pending: Deque[Request] = deque()
        self.running: Dict[int, Request] = {}
        self.done: List[Request] = []
        self._t0 = time.perf_counter()
        self._offset = 0.0
        self._jump_idle = clock == "steps"
        if clock == "steps":
            # deterministic clock: time is the number of batch steps run; when the batch is
            # empty it jumps to the next arrival
            self.clock = lambda: self.steps_run + self._offset
        else:
            self.clock = clock or (lambda: time.perf_counter() - self._t0)
        self.steps_run = 0
        self.cell_steps = 0
        self.sor_cell_updates = 0
        self.busy_s = 0.0                # wall time inside batch steps (the device is busy)

    def footprint(self, r: Request) -> int:
        return self.batch.case_bytes(r.params)

    def submit(self, r: Request) -> None:
        self.pending.append(r)

    def _admit(self, now: float) -> None:
        while self.pending and self.pending[0].arrival <= now:
            r = self.pending[0]
            if len(self.running) >= self.batch.max_cases or self.footprint(r) > self.batch.free_bytes():
                break                    # wait for memory / a slot to free up (FIFO)
            self.pending.popleft()
            r.slot = self.batch.add(r.params, r.phi)
            r.admitted = now
            self.running[r.slot] = r

    def _retire(self, now: float) -> None:
        for slot, r in list(self.running.items()):
            if r.error or self.batch.case_steps(slot) >= r.steps:
                if not r.error:
                    r.result = {f: self.batch.interior(slot, f) for f in self.keep_fields
                                if not (f == "w" and r.params.dim == 2)}
                r.finished = now
                self.batch.remove(slot)
                del self.running[slot]
                self.done.append(r)

    def step_once(self) -> None:
        now = self.clock()
        self._admit(now)
        if not self.running:
            if self.pending and self.pending[0].arrival <= now:
                r = self.pending[0]
                raise RuntimeError(f"request {r.rid} needs {self.footprint(r)} bytes; the empty arena has "
                                   f"{self.batch.free_bytes()}")
            return
        todo = min(r.steps - self.batch.case_steps(s) for s, r in self.running.items())
        n = max(1, min(self.chunk, todo))
        t0 = time.perf_counter()
        st = self.batch.step(n, raise_diverged=False)
        self.busy_s += time.perf_counter() - t0
        for s, r in self.running.items():
            r.last_stats = {k: float(v[-1, s]) for k, v in st.items()}
            self.cell_steps += r.params.cells * n
            self.sor_cell_updates += r.params.cells * int(st["sor_iters"][:, s].sum())
            if not np.all(np.isfinite(st["sor_error"][:, s])):
                r.error = "diverged"
        self.steps_run += n
        self._retire(self.clock())

    def run(self, idle_sleep: float = 1e-3, max_wall: float = 3600.0) -> List[Request]:
        start = time.perf_counter()
        while (self.pending or self.running) and time.perf_counter() - start < max_wall:
            if not self.running and self.pending and self.pending[0].arrival > self.clock():
                if self._jump_idle:
                    self._offset += self.pending[0].arrival - self.clock()
                else:
                    time.sleep(min(idle_sleep, max(0.0, self.pending[0].arrival - self.clock())))
                continue
            self.step_once()
        return self.done

    def close(self) -> None:
        self.batch.close()


def latency_stats(done: List[Request]) -> Dict[str, float]:
    failed = sum(1 for r in done if r.error)
    done = [r for r in done if not r.error]
    if not done:
        return {"requests": 0, "failed": failed}
    tts = np.array([r.finished - r.arrival for r in done])
    wait = np.array([r.admitted - r.arrival for r in done])
    span = max(r.finished for r in done) - min(r.arrival for r in done)
    return {"requests": len(done), "failed": failed, "p50_time_to_solution_s": float(np.percentile(tts, 50)),
            "p99_time_to_solution_s": float(np.percentile(tts, 99)), "mean_queue_wait_s": float(wait.mean()),
            "cases_per_s": len(done) / span if span > 0 else 0.0, "span_s": span}
