"""Serving loop for colony-growth requests: continuous batching on one GPU.

A request is a colony's initial state, its parameters (hard or soft model,
growth rate, ...) and a number of timesteps.  A timestep is the reference's
time-loop body (Domain::run, Domain.cpp:121-143): divide the cells that have
reached twice their rest length and renumber by x, on the device
(pcc_batch_divide, seeded by the request id and the colony's own step count,
so a colony's run does not depend on what it is batched with), then step.
The server keeps one batched context per model (`engine.ColonyBatch`).
Between batch steps it

  * admits waiting requests, first come first served, while the estimated
    device footprint of the resident colonies plus the newcomer stays within
    an HBM budget (sized for the most cells the colony can reach by division
    in its steps) and, per model (hard: per model and size bucket), the
    resident cells stay within a compute budget (the cells a batch step may
    carry: a step's time grows with its cells, so this bounds the per-step
    latency of every resident colony; the device is shared by the batches of
    both models): an admitted colony joins its model's batch and its state
    stays resident in HBM until its last step;
  * retires colonies that have run all their steps: the final state is
    downloaded and the colony leaves the batch, freeing its share of the
    budget for the next request.

Colonies are also split by size into buckets, one batched context per
(model, bucket): a batched step runs its colonies' ReLCP phases and BBPGD
solves in lockstep, so the slowest colony of a batch sets that batch's step
time, and bucketing keeps small colonies from waiting on large ones.  For the
same reason a newly admitted colony first runs `warm_steps` steps in a warm-up
batch of its (model, bucket): the first steps of a fresh colony relax its
initial overlaps and can take orders of magnitude more BBPGD iterations than
the steps after them, and only colonies in the same situation wait for them.
It then moves, state unchanged, to the steady batch.  All batches are stepped
from their own host threads, so their work overlaps on the device (each
context owns its own HIP stream).  Every colony's result equals the result of
a context that runs it alone: batched colonies share launches, never data,
and a move between batches carries the full state.

The reference runs one colony per MPI job (Domain::run, Domain.cpp:121-143);
this loop has no counterpart there.
"""
import collections
import threading
import time
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

from . import abi_types as T
from .engine import ColonyBatch

# Device bytes per cell charged against the budget at admission (body state
# and outputs, constraint set, solver-order and incidence arrays, detection
# scratch; dense 2D packings hold ~10 constraints per cell over the ReLCP
# phases of a step).  The run reports the measured peak beside it.
DEFAULT_BYTES_PER_CELL = 12_000

# Resident cells that admission allows (the compute budget), per model, or
# per model and size bucket (a tuple, one cap per bucket).  A hard step of a
# relaxed dense packing runs at ~1M cell-steps/s on one MI355X (a few hundred
# BBPGD iterations per ReLCP phase), a soft step at ~100M.  The hard caps keep
# a large-colony batch step near a quarter of a second and the small-colony
# batches near 20-100 ms (charged at the most cells a colony can reach), each bucket with a budget of its own, so a large
# colony waiting for room never holds back small ones.
DEFAULT_MAX_CELLS = {T.MODE_HARD: (40_000, 100_000, 250_000), T.MODE_SOFT: 4_000_000}


@dataclass
class Request:
    rid: int
    bodies: T.Bodies
    params: T.Params
    steps: int
    arrival: float = 0.0  # seconds after the start of the run


def division_seed(rid: int, step: int) -> int:
    """Seed of the divisions before a request's step `step` (the device
    generator is keyed by seed, step argument 0 and parent index)."""
    return ((int(rid) & 0xFFFFFFFF) << 32) | (int(step) & 0xFFFFFFFF)


def max_cells(req: "Request", length_noise: float = 0.0) -> int:
    """Bound on a colony's cell count over `steps` steps of growth with
    division.  Each step multiplies a length by at most 1 + dt/tau (the
    impedance exp(-lambda sigma) is <= 1 since stresses are >= 0) and
    division conserves length, so the total length stays below
    L0 (1 + dt/tau)^steps; every cell is at least as long as the shortest
    initial cell or the shortest child (half of >= 2 l0, less the length
    noise), so N <= total length / that minimum."""
    b = req.bodies
    if b.n == 0:
        return 0
    total = float(np.sum(b.length)) * (1.0 + req.params.dt / req.params.tau) ** req.steps
    shortest = min(float(np.min(b.length)),
                   float(np.min(b.l0)) * max(1.0 - 2.0 * length_noise, 0.05))
    return max(b.n, int(np.ceil(total / shortest * (1.0 + 1e-9))) + 1)


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
    state: Optional[T.Bodies] = None  # current state when moved between batches


class ColonyServer:
    """Continuous-batching server: one worker thread per (model, size bucket,
    stage) batch steps its colonies back to back; the caller's thread admits
    arrivals first come first served against the HBM budget and hands them to
    the warm-up batch of their (model, bucket) (or straight to the steady
    batch when warm_steps is 0), whose worker adds them between two of its
    steps."""

    def __init__(self, hbm_budget_bytes: int, precision: int = T.PREC_MIXED, device: int = 0,
                 bytes_per_cell: Optional[float] = None, overlap: bool = True,
                 bucket_edges: Sequence[int] = (4_000, 20_000), warm_steps: int = 2,
                 divide: bool = True, length_noise: float = 0.0, angle_noise: float = 0.01,
                 max_cells_resident: Optional[Dict[int, object]] = None):
        self.budget = int(hbm_budget_bytes)
        # compute budget: resident cells per model (None: HBM alone decides)
        self.max_resident = dict(DEFAULT_MAX_CELLS if max_cells_resident is None
                                 else max_cells_resident)
        self.precision = precision
        self.device = device
        self.overlap = overlap  # False: one worker steps every batch in turn
        self.bytes_per_cell = float(bytes_per_cell or DEFAULT_BYTES_PER_CELL)
        self.bucket_edges = tuple(sorted(bucket_edges))
        self.warm_steps = int(warm_steps)
        self.divide = bool(divide)  # False: cells grow but never divide
        self.length_noise = float(length_noise)
        self.angle_noise = float(angle_noise)
        stages = (0, 1) if self.warm_steps > 0 else (1,)  # 0: warm-up, 1: steady
        self.keys = [(m, b, g) for m in (T.MODE_HARD, T.MODE_SOFT)
                     for b in range(len(self.bucket_edges) + 1) for g in stages]
        self.batches: Dict[Tuple[int, int, int], Optional[ColonyBatch]] = {k: None for k in self.keys}
        self.members: Dict[Tuple[int, int, int], List[_Member]] = {k: [] for k in self.keys}
        self.handoff: Dict[Tuple[int, int, int], List[_Member]] = {k: [] for k in self.keys}
        self.queue: collections.deque = collections.deque()
        self.results: Dict[int, Result] = {}
        self.peak_reserved = 0
        self.peak_device_bytes = 0
        self.batch_steps = 0
        self.cell_steps = 0
        # per batch step: (key, colonies, cells, divide ms, step ms, most BBPGD
        # iterations of a colony, most ReLCP phases of a colony)
        self.step_log: List[tuple] = []
        self._reserved = 0
        self._cells = {T.MODE_HARD: 0, T.MODE_SOFT: 0}  # resident cells per model
        self.peak_cells = {T.MODE_HARD: 0, T.MODE_SOFT: 0}
        self._bucket_cells: Dict[Tuple[int, int], int] = collections.defaultdict(int)
        self._lock = threading.Condition()
        self._stop = False
        self._errors: List[str] = []
        self._t0 = 0.0
        self._first_step = None
        self._last_step = None

    # -- admission (caller's thread) -------------------------------------
    def footprint(self, n_cells: int) -> int:
        return int(self.bytes_per_cell * max(n_cells, 1))

    def reservation(self, req: Request) -> int:
        """Device bytes reserved for a request from admission to retirement."""
        return self.footprint(max_cells(req, self.length_noise) if self.divide
                              else req.bodies.n)

    def reserved(self) -> int:
        return self._reserved

    def _bucket(self, n_cells: int) -> int:
        return sum(n_cells >= e for e in self.bucket_edges)

    def submit(self, req: Request):
        if req.params.mode not in (T.MODE_HARD, T.MODE_SOFT):
            raise ValueError("Invalid mode")
        if req.params.precision != self.precision:
            req.params = req.params.replace(precision=self.precision)
        self.queue.append(req)

    def cells_charge(self, req: Request) -> int:
        """Cells a request counts against its model's compute budget: the
        most it can reach in its steps (the HBM bound)."""
        return max_cells(req, self.length_noise) if self.divide else req.bodies.n

    def _budget_key(self, req: Request):
        """The compute budget a request counts against: (model, bucket) when
        the model's cap is per bucket, else (model, None)."""
        mode = req.params.mode
        cap = self.max_resident.get(mode)
        if isinstance(cap, (tuple, list)):
            b = self._bucket(req.bodies.n)
            return (mode, b), cap[min(b, len(cap) - 1)]
        return (mode, None), cap

    def _admit(self, now: float) -> int:
        """First come, first served: a request that does not fit the HBM
        budget stops admission; one that does not fit its compute budget
        (its model's, or its model's and size bucket's) stops admission for
        that budget only: later requests of another budget may pass it, never
        those of its own."""
        n = 0
        with self._lock:
            blocked = set()
            i = 0
            while i < len(self.queue):
                req = self.queue[i]
                mode = req.params.mode
                bkey, cap = self._budget_key(req)
                if bkey in blocked:
                    i += 1
                    continue
                need = self.reservation(req)
                if self._reserved + need > self.budget and self._reserved > 0:
                    break  # the head waits for memory
                cells = self.cells_charge(req)
                held = self._bucket_cells[bkey]
                if cap is not None and held > 0 and held + cells > cap:
                    blocked.add(bkey)  # waits for compute
                    i += 1
                    continue
                del self.queue[i]
                self._reserved += need
                self._bucket_cells[bkey] += cells
                self._cells[mode] += cells
                self.peak_cells[mode] = max(self.peak_cells[mode], self._cells[mode])
                self.peak_reserved = max(self.peak_reserved, self._reserved)
                key = (req.params.mode, self._bucket(req.bodies.n),
                       0 if self.warm_steps > 0 else 1)
                self.handoff[key].append(_Member(req, now))
                n += 1
            if n:
                self._lock.notify_all()
        return n

    def _resident(self) -> int:
        return sum(len(v) for v in self.members.values()) + \
            sum(len(v) for v in self.handoff.values())

    # -- one batch (its worker thread) -----------------------------------
    def _serve_once(self, key) -> bool:
        """Adds handed-off colonies, steps the batch once, retires finished
        colonies.  Returns False when the batch had nothing to do."""
        ms = self.members[key]
        with self._lock:
            # moved in one critical section so that run() never sees a colony
            # in neither list
            new = self.handoff[key]
            self.handoff[key] = []
            ms.extend(new)
        for m in new:
            if self.batches[key] is None:
                self.batches[key] = ColonyBatch(m.req.params, device=self.device)
            q = self.batches[key].add(m.state if m.state is not None else m.req.bodies,
                                      m.req.params)
            # a colony moved from the warm-up batch keeps its step count (the
            # key of its Brownian draws), as a context running it alone does
            self.batches[key].set_step_count(m.done, q)
            m.state = None
        if not ms:
            return False
        b = self.batches[key]
        t = time.perf_counter()
        if self.divide:
            b.divide([division_seed(m.req.rid, m.done) for m in ms], 0, self.length_noise,
                     self.angle_noise)
        td = time.perf_counter()
        res = b.step(self.batch_steps)
        t1 = time.perf_counter()
        dt_ms = (t1 - t) * 1e3
        cells = 0
        sizes = b.sizes
        for q, m in enumerate(ms):
            m.done += 1
            m.step_ms.append(dt_ms)
            m.statuses.append(res[q]["status"])
            cells += sizes[q]
        dev = b.device_bytes()
        self.step_log.append((key, len(ms), cells, (td - t) * 1e3, (t1 - td) * 1e3,
                              max(r["bbpgd_iterations"] for r in res),
                              max(r["constraint_iterations"] for r in res)))
        gone = []  # (q, member, final state or None if moved), q descending
        for q in range(len(ms) - 1, -1, -1):
            m = ms[q]
            if m.done < m.req.steps:
                if key[2] == 0 and m.done >= self.warm_steps:
                    m.state = b.get_state(q)  # warm-up over: on to the steady batch
                    b.remove(q)
                    gone.append((q, m, None))
                continue
            state = b.get_state(q)
            b.remove(q)
            gone.append((q, m, state))
        now = time.perf_counter() - self._t0
        with self._lock:
            self.batch_steps += 1
            self.cell_steps += cells
            self.peak_device_bytes = max(self.peak_device_bytes, dev)
            if self._first_step is None or t < self._first_step:
                self._first_step = t
            self._last_step = t1 if self._last_step is None else max(self._last_step, t1)
            for q, m, state in gone:  # descending q
                del ms[q]
                if state is None:
                    self.handoff[(key[0], key[1], 1)].append(m)
                    continue
                self._reserved -= self.reservation(m.req)
                self._cells[m.req.params.mode] -= self.cells_charge(m.req)
                self._bucket_cells[self._budget_key(m.req)[0]] -= self.cells_charge(m.req)
                self.results[m.req.rid] = Result(
                    m.req.rid, key[0], m.req.bodies.n, m.req.steps, state, m.step_ms,
                    m.statuses, m.req.arrival, m.admitted, now)
            self._lock.notify_all()
        return True

    def _worker(self, keys):
        try:
            while True:
                busy = False
                for key in keys:
                    if self._stop:
                        return
                    busy = self._serve_once(key) or busy
                if not busy:
                    with self._lock:
                        if self._stop:
                            return
                        if not any(self.handoff[k] for k in keys):
                            self._lock.wait(0.02)
        except Exception as e:  # surfaced by run()
            with self._lock:
                self._errors.append(f"{keys}: {e!r}")
                self._lock.notify_all()

    def run(self, requests: List[Request], time_limit: Optional[float] = None,
            progress=None, progress_every: float = 5.0) -> dict:
        """Serves `requests` (arrival times relative to the start) to
        completion, or until `time_limit` seconds have passed; returns the
        run's statistics.  Results are in self.results.  `progress(msg)`, if
        given, receives a status line every `progress_every` seconds."""
        pending = sorted(requests, key=lambda r: r.arrival)
        i = 0
        self._t0 = time.perf_counter()
        groups = [[k] for k in self.keys] if self.overlap else [list(self.keys)]
        workers = [threading.Thread(target=self._worker, args=(g,)) for g in groups]
        for w in workers:
            w.start()
        next_report = progress_every
        try:
            while True:
                now = time.perf_counter() - self._t0
                while i < len(pending) and pending[i].arrival <= now:
                    self.submit(pending[i])
                    i += 1
                self._admit(now)
                with self._lock:
                    if self._errors:
                        raise RuntimeError("batch worker failed: " + "; ".join(self._errors))
                    idle = i >= len(pending) and not self.queue and self._resident() == 0
                    if idle:
                        break
                    if time_limit is not None and now > time_limit:
                        break
                    if progress is not None and now >= next_report:
                        next_report += progress_every
                        progress(f"t={now:.1f}s batch_steps={self.batch_steps} "
                                 f"resident={self._resident()} queued={len(self.queue)} "
                                 f"completed={len(self.results)}/{len(requests)} "
                                 f"cell_steps={self.cell_steps}")
                    wait = 0.05
                    if i < len(pending):
                        wait = max(0.0, min(wait, pending[i].arrival - now))
                    self._lock.wait(wait)
        finally:
            with self._lock:
                self._stop = True
                self._lock.notify_all()
            for w in workers:
                w.join()
        if self._errors:
            raise RuntimeError("batch worker failed: " + "; ".join(self._errors))
        wall = (self._last_step - self._first_step) if self._first_step is not None else 0.0
        lat = np.array([x for r in self.results.values() for x in r.step_ms]
                       + [x for ms in self.members.values() for m in ms for x in m.step_ms])
        return {
            "requests": len(requests),
            "completed": len(self.results),
            "cell_steps": int(self.cell_steps),
            "wall_s": wall,
            "cell_steps_per_s": self.cell_steps / wall if wall > 0 else 0.0,
            "p50_step_ms": float(np.percentile(lat, 50)) if len(lat) else 0.0,
            "p99_step_ms": float(np.percentile(lat, 99)) if len(lat) else 0.0,
            "batch_steps": self.batch_steps,
            "peak_reserved_bytes": int(self.peak_reserved),
            "peak_resident_cells": {"hard": int(self.peak_cells[T.MODE_HARD]),
                                    "soft": int(self.peak_cells[T.MODE_SOFT])},
            "max_resident_cells": {"hard": self.max_resident.get(T.MODE_HARD),
                                   "soft": self.max_resident.get(T.MODE_SOFT)},
            "peak_device_bytes": int(self.peak_device_bytes),
            "budget_bytes": self.budget,
        }

    def step_summary(self) -> dict:
        """Per batch (model-bucket-stage): its batch steps with the median
        colonies and cells, the divide and step times and the most BBPGD
        iterations of a colony per step (p50 / p99)."""
        out = {}
        for key in self.keys:
            rows = [r for r in self.step_log if r[0] == key]
            if not rows:
                continue
            a = np.array([r[1:] for r in rows], dtype=float)
            out["-".join(str(x) for x in key)] = {
                "batch_steps": len(rows),
                "colonies_p50": float(np.percentile(a[:, 0], 50)),
                "cells_p50": float(np.percentile(a[:, 1], 50)),
                "divide_ms_p50": float(np.percentile(a[:, 2], 50)),
                "step_ms_p50": float(np.percentile(a[:, 3], 50)),
                "step_ms_p99": float(np.percentile(a[:, 3], 99)),
                "max_iterations_p50": float(np.percentile(a[:, 4], 50)),
                "max_iterations_p99": float(np.percentile(a[:, 4], 99)),
                "max_phases_p99": float(np.percentile(a[:, 5], 99)),
            }
        return out

    def close(self):
        for k, b in self.batches.items():
            if b is not None:
                b.close()
            self.batches[k] = None
