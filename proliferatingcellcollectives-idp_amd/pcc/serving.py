"""Serving loop for colony-growth requests: continuous batching on one GPU.

A request is a colony's initial state, its parameters (hard or soft model,
growth rate, ...) and a number of timesteps.  The server keeps one batched
context per model (`engine.ColonyBatch`).  Between batch steps it

  * admits waiting requests, first come first served, while the estimated
    device footprint of the resident colonies plus the newcomer stays within
    an HBM budget: an admitted colony joins its model's batch and its state
    stays resident in HBM until its last step;
  * retires colonies that have run all their steps: the final state is
    downloaded and the colony leaves the batch, freeing its share of the
    budget for the next request.

The hard and the soft batch are stepped from two host threads, so their work
overlaps on the device (each context owns its own HIP stream).  Every colony's
result equals the result of a context that runs it alone: batched colonies
share launches, never data.

The reference runs one colony per MPI job (Domain::run, Domain.cpp:121-143);
this loop has no counterpart there.
"""
import collections
import threading
import time
from dataclasses import dataclass, field
from typing import Dict, List, Optional

import numpy as np

from . import abi_types as T
from .engine import ColonyBatch

# Device bytes per cell assumed for admission until a batch has been measured
# (body state and outputs, constraint set, solver-order and incidence arrays,
# detection scratch; dense 2D packings hold ~10 constraints per cell over the
# ReLCP phases of a step).
DEFAULT_BYTES_PER_CELL = 12_000


@dataclass
class Request:
    rid: int
    bodies: T.Bodies
    params: T.Params
    steps: int
    arrival: float = 0.0  # seconds after the start of the run


@dataclass
class Result:
    rid: int
    mode: int
    n_cells: int
    steps: int
    state: T.Bodies
    step_ms: List[float]
    statuses: List[int]
    arrival: float
    admitted: float
    finished: float


@dataclass
class _Member:
    req: Request
    admitted: float
    done: int = 0
    step_ms: List[float] = field(default_factory=list)
    statuses: List[int] = field(default_factory=list)


class ColonyServer:
    """Continuous-batching server over one batched context per model."""

    def __init__(self, hbm_budget_bytes: int, precision: int = T.PREC_MIXED, device: int = 0,
                 bytes_per_cell: Optional[float] = None, overlap: bool = True):
        self.budget = int(hbm_budget_bytes)
        self.precision = precision
        self.device = device
        self.overlap = overlap
        self.bytes_per_cell = float(bytes_per_cell or DEFAULT_BYTES_PER_CELL)
        self._fixed_rate = bytes_per_cell is not None
        self.batches: Dict[int, Optional[ColonyBatch]] = {T.MODE_HARD: None, T.MODE_SOFT: None}
        self.members: Dict[int, List[_Member]] = {T.MODE_HARD: [], T.MODE_SOFT: []}
        self.queue: collections.deque = collections.deque()
        self.results: Dict[int, Result] = {}
        self.peak_reserved = 0
        self.peak_device_bytes = 0
        self.batch_steps = 0
        self._peak_cells = 0

    # -- admission -----------------------------------------------------
    def footprint(self, n_cells: int) -> int:
        return int(self.bytes_per_cell * max(n_cells, 1))

    def reserved(self) -> int:
        return sum(self.footprint(m.req.bodies.n) for ms in self.members.values() for m in ms)

    def device_bytes(self) -> int:
        return sum(b.device_bytes() for b in self.batches.values() if b is not None)

    def submit(self, req: Request):
        if req.params.mode not in (T.MODE_HARD, T.MODE_SOFT):
            raise ValueError("Invalid mode")
        if req.params.precision != self.precision:
            req.params = req.params.replace(precision=self.precision)
        self.queue.append(req)

    def _admit(self, now: float) -> int:
        n = 0
        while self.queue:
            req = self.queue[0]
            need = self.footprint(req.bodies.n)
            if self.reserved() + need > self.budget and self.reserved() > 0:
                break  # first come, first served: the head waits for memory
            self.queue.popleft()
            mode = req.params.mode
            if self.batches[mode] is None:
                self.batches[mode] = ColonyBatch(req.params, device=self.device)
            self.batches[mode].add(req.bodies, req.params)
            self.members[mode].append(_Member(req, now))
            self.peak_reserved = max(self.peak_reserved, self.reserved())
            n += 1
        return n

    # -- stepping ------------------------------------------------------
    def _step_mode(self, mode: int, out: dict):
        b = self.batches[mode]
        t = time.perf_counter()
        r = b.step(self.batch_steps)
        out[mode] = (r, (time.perf_counter() - t) * 1e3)

    def _step_all(self) -> dict:
        modes = [m for m in (T.MODE_HARD, T.MODE_SOFT) if self.members[m]]
        out = {}
        if self.overlap and len(modes) > 1:
            th = [threading.Thread(target=self._step_mode, args=(m, out)) for m in modes]
            for x in th:
                x.start()
            for x in th:
                x.join()
        else:
            for m in modes:
                self._step_mode(m, out)
        missing = [m for m in modes if m not in out]
        if missing:
            raise RuntimeError(f"batch step failed for modes {missing}")
        self.batch_steps += 1
        return out

    def _retire(self, now: float):
        for mode, ms in self.members.items():
            b = self.batches[mode]
            for q in range(len(ms) - 1, -1, -1):
                m = ms[q]
                if m.done < m.req.steps:
                    continue
                state = b.get_state(q)
                b.remove(q)
                del ms[q]
                self.results[m.req.rid] = Result(
                    m.req.rid, mode, m.req.bodies.n, m.req.steps, state, m.step_ms, m.statuses,
                    m.req.arrival, m.admitted, now)

    def _learn_rate(self):
        # Device buffers keep their high-water size after colonies leave, so
        # the bytes-per-cell rate is only learned when the resident cell count
        # reaches a new peak (the buffers were sized for exactly that load).
        if self._fixed_rate:
            return
        cells = sum(m.req.bodies.n for ms in self.members.values() for m in ms)
        if cells > self._peak_cells:
            self._peak_cells = cells
            self.bytes_per_cell = max(self.bytes_per_cell, self.device_bytes() / cells)

    def run(self, requests: List[Request], time_limit: Optional[float] = None,
            progress=None, progress_every: float = 5.0) -> dict:
        """Serves `requests` (arrival times relative to the start) to
        completion, or until `time_limit` seconds have passed; returns the
        run's statistics.  Results are in self.results.  `progress(msg)`, if
        given, receives a status line every `progress_every` seconds of
        stepping."""
        pending = sorted(requests, key=lambda r: r.arrival)
        i = 0
        t0 = time.perf_counter()
        busy_cell_steps = 0
        first_step = None
        last_step = None
        next_report = progress_every
        while True:
            now = time.perf_counter() - t0
            while i < len(pending) and pending[i].arrival <= now:
                self.submit(pending[i])
                i += 1
            self._admit(now)
            active = sum(len(ms) for ms in self.members.values())
            if active == 0:
                if i >= len(pending) and not self.queue:
                    break
                if i < len(pending):
                    time.sleep(max(0.0, min(pending[i].arrival - now, 0.05)))
                continue
            if time_limit is not None and now > time_limit:
                break
            if first_step is None:
                first_step = time.perf_counter()
            out = self._step_all()
            last_step = time.perf_counter()
            for mode, (res, ms_) in out.items():
                for q, m in enumerate(self.members[mode]):
                    m.done += 1
                    m.step_ms.append(ms_)
                    m.statuses.append(res[q]["status"])
                    busy_cell_steps += m.req.bodies.n
            self.peak_device_bytes = max(self.peak_device_bytes, self.device_bytes())
            self._learn_rate()
            self._retire(time.perf_counter() - t0)
            if progress is not None and time.perf_counter() - t0 >= next_report:
                next_report += progress_every
                progress(f"t={time.perf_counter() - t0:.1f}s batch_steps={self.batch_steps} "
                         f"resident={active} queued={len(self.queue)} "
                         f"completed={len(self.results)}/{len(requests)} "
                         f"cell_steps={busy_cell_steps}")
        wall = (last_step - first_step) if first_step is not None else 0.0
        lat = np.array([x for r in self.results.values() for x in r.step_ms]
                       + [x for ms in self.members.values() for m in ms for x in m.step_ms])
        return {
            "requests": len(requests),
            "completed": len(self.results),
            "cell_steps": int(busy_cell_steps),
            "wall_s": wall,
            "cell_steps_per_s": busy_cell_steps / wall if wall > 0 else 0.0,
            "p50_step_ms": float(np.percentile(lat, 50)) if len(lat) else 0.0,
            "p99_step_ms": float(np.percentile(lat, 99)) if len(lat) else 0.0,
            "batch_steps": self.batch_steps,
            "peak_reserved_bytes": int(self.peak_reserved),
            "peak_device_bytes": int(self.peak_device_bytes),
            "budget_bytes": self.budget,
        }

    def close(self):
        for b in self.batches.values():
            if b is not None:
                b.close()
        self.batches = {T.MODE_HARD: None, T.MODE_SOFT: None}
