"""Host logic of the continuous-batching server (no GPU): admission against
the HBM budget, first come first served, joining running batches, retirement
after exactly the requested steps.  The batched context is replaced by an
in-memory stand-in that advances each colony's lengths by one per step."""
import time

import numpy as np
import pytest

from pcc import abi_types as T
from pcc import generators as G
from pcc import serving


class FakeBatch:
    instances = []

    def __init__(self, params, device=0):
        self.mode = params.mode
        self.cols = []
        self.events = []
        self.divides = []
        FakeBatch.instances.append(self)

    def add(self, bodies, params):
        assert params.mode == self.mode
        self.events.append(("add", len(self.cols), tuple(c[1] for c in self.cols)))
        self.cols.append([bodies.copy(), 0])
        return len(self.cols) - 1

    def remove(self, q):
        del self.cols[q]

    def set_step_count(self, count, q):
        assert self.cols[q][1] == 0
        self.cols[q][1] = count

    @property
    def sizes(self):
        return [c[0].n for c in self.cols]

    def divide(self, seeds, step, length_noise=0.0, angle_noise=0.01):
        # no cell reaches twice its rest length here; record the call
        self.divides.append((list(seeds), step))
        return self.sizes

    def step(self, i):
        # a device step takes time; a colony whose first length is exactly
        # 100 is fresh and unrelaxed: its first step is slow
        slow = any(c[0].length[0] == 100.0 for c in self.cols)
        time.sleep(0.3 if slow else 0.002)
        for c in self.cols:
            c[0].length = c[0].length + 1.0
            c[1] += 1
        return [{"status": 0, "n_constraints": 0, "constraint_iterations": 0,
                 "bbpgd_iterations": 0, "residual": 0.0, "max_overlap": 0.0}
                for _ in self.cols]

    def get_state(self, q):
        return self.cols[q][0]

    def device_bytes(self):
        return 1000 * sum(c[0].n for c in self.cols)

    def close(self):
        pass


@pytest.fixture(autouse=True)
def fake_batch(monkeypatch):
    FakeBatch.instances = []
    monkeypatch.setattr(serving, "ColonyBatch", FakeBatch)


def make_requests(spec):
    reqs = []
    for rid, (n, mode, steps, arrival) in enumerate(spec):
        p = T.reference_params(mode=mode, precision=T.PREC_MIXED)
        reqs.append(serving.Request(rid, G.packed_colony(n, seed=rid), p, steps, arrival))
    return reqs


def test_every_request_runs_exactly_its_steps():
    reqs = make_requests([(100, 0, 3, 0.0), (200, 1, 5, 0.0), (150, 0, 2, 0.0), (80, 1, 4, 0.0)])
    srv = serving.ColonyServer(10**12, bytes_per_cell=1000, warm_steps=0, divide=False)
    st = srv.run(reqs)
    assert st["completed"] == 4
    for r in reqs:
        res = srv.results[r.rid]
        assert len(res.step_ms) == r.steps and res.statuses == [0] * r.steps
        assert np.array_equal(res.state.length, r.bodies.length + r.steps)
        assert res.mode == r.params.mode
    # one batch per (model, size bucket) in use: all sizes here fall in the
    # smallest bucket
    assert sorted(b.mode for b in FakeBatch.instances) == [0, 1]


def test_size_buckets_get_their_own_batches():
    reqs = make_requests([(100, 0, 2, 0.0), (5000, 0, 2, 0.0), (25000, 0, 2, 0.0),
                          (150, 0, 2, 0.0)])
    srv = serving.ColonyServer(10**12, bytes_per_cell=1000, bucket_edges=(4000, 20000),
                               warm_steps=0, divide=False)
    st = srv.run(reqs)
    assert st["completed"] == 4
    assert len(FakeBatch.instances) == 3
    assert sorted(len(b.events) for b in FakeBatch.instances) == [1, 1, 2]


def test_admission_respects_budget_and_order():
    # budget for 300 cells at 1000 B/cell: the requests must take turns
    reqs = make_requests([(200, 0, 3, 0.0), (100, 0, 5, 0.0), (250, 0, 2, 0.0), (50, 0, 1, 0.0)])
    srv = serving.ColonyServer(300 * 1000, bytes_per_cell=1000, divide=False)
    st = srv.run(reqs)
    assert st["completed"] == 4
    assert st["peak_reserved_bytes"] <= 300 * 1000
    r = srv.results
    # first come, first served: request 2 (250 cells) waits for 0 and 1 to
    # leave although 3 (50 cells) would fit earlier
    assert r[2].admitted >= r[0].finished and r[2].admitted >= r[1].finished
    assert r[3].admitted >= r[2].admitted
    for q in reqs:
        assert np.array_equal(r[q.rid].state.length, q.bodies.length + q.steps)


def test_compute_budget_bounds_resident_cells_per_model():
    # HBM is plentiful; the hard model may hold 300 resident cells, the soft
    # one is unbounded: hard requests take turns, the soft one runs at once
    reqs = make_requests([(200, 0, 3, 0.0), (200, 0, 3, 0.0), (150, 0, 2, 0.0),
                          (1000, 1, 2, 0.0)])
    srv = serving.ColonyServer(10**12, bytes_per_cell=1000, divide=False,
                               max_cells_resident={T.MODE_HARD: 300})
    st = srv.run(reqs)
    assert st["completed"] == 4
    assert st["peak_resident_cells"]["hard"] <= 300
    assert st["peak_resident_cells"]["soft"] == 1000
    r = srv.results
    assert r[1].admitted >= r[0].finished  # 200 + 200 > 300: waited for compute
    assert r[2].admitted >= r[1].admitted  # first come, first served
    assert r[3].admitted < r[0].finished   # the other model's budget is its own
    for q in reqs:
        assert np.array_equal(r[q.rid].state.length, q.bodies.length + q.steps)


def test_compute_budget_per_bucket_lets_small_colonies_pass():
    # hard budgets per size bucket (edges 150 and 10^6 cells): a large colony
    # waiting for room in its bucket does not hold back a small colony that
    # arrived after it; within a bucket, first come first served still holds
    reqs = make_requests([(400, 0, 3, 0.0), (400, 0, 3, 0.0), (100, 0, 2, 0.0),
                          (120, 0, 2, 0.0)])
    srv = serving.ColonyServer(10**12, bytes_per_cell=1000, divide=False,
                               bucket_edges=(150, 10**6),
                               max_cells_resident={T.MODE_HARD: (130, 500)})
    st = srv.run(reqs)
    assert st["completed"] == 4
    r = srv.results
    assert r[1].admitted >= r[0].finished  # 400 + 400 > 500: waits for its bucket
    assert r[2].admitted < r[0].finished   # the small bucket's budget is its own
    assert r[3].admitted >= r[2].finished  # 100 + 120 > 130 in the small bucket
    for q in reqs:
        assert np.array_equal(r[q.rid].state.length, q.bodies.length + q.steps)


def test_late_arrival_joins_a_running_batch():
    # the first colony stays resident for at least 500 x 2 ms; the second
    # arrives well after its first step, also on a loaded machine
    reqs = make_requests([(100, 0, 500, 0.0), (100, 0, 3, 0.25)])
    srv = serving.ColonyServer(10**12, bytes_per_cell=1000, warm_steps=0, divide=False)
    srv.run(reqs)
    adds = FakeBatch.instances[0].events  # both colonies fall in the same bucket
    # the second colony was added while the first had already stepped
    assert len(adds) == 2 and adds[1][2][0] > 0
    assert np.array_equal(srv.results[1].state.length, reqs[1].bodies.length + 3)


def test_oversized_request_runs_alone():
    reqs = make_requests([(500, 0, 2, 0.0), (100, 0, 2, 0.0)])
    srv = serving.ColonyServer(100 * 1000, bytes_per_cell=1000, divide=False)
    st = srv.run(reqs)
    assert st["completed"] == 2
    assert srv.results[1].admitted >= srv.results[0].finished


def test_warm_up_moves_keep_state_and_step_count():
    """Colonies run warm_steps steps in the warm-up batch, then move to the
    steady batch of their (model, bucket) with their state; each still runs
    exactly its steps (requests shorter than the warm-up never move)."""
    spec = [(100 + 10 * i, i % 2, 1 + i, 0.001 * i) for i in range(7)]
    reqs = make_requests(spec)
    srv = serving.ColonyServer(10**12, bytes_per_cell=1000, warm_steps=3)
    st = srv.run(reqs)
    assert st["completed"] == len(reqs)
    for r in reqs:
        res = srv.results[r.rid]
        assert len(res.step_ms) == r.steps
        want = r.bodies.length.copy()
        for _ in range(r.steps):  # the stand-in adds 1 per step
            want = want + 1.0
        assert np.array_equal(res.state.length, want)
    # warm-up batches saw every colony, steady batches only those with > 3 steps
    added = {}
    for b in FakeBatch.instances:
        added.setdefault(b.mode, []).append(len(b.events))
    assert sum(sum(v) for v in added.values()) == len(reqs) + sum(r.steps > 3 for r in reqs)


def slow_pair(warm_steps):
    # the steady colony has left the warm-up batch long before the fresh one
    # arrives (also on a loaded machine) and is still running when it does
    a = make_requests([(100, 0, 500, 0.0)])[0]
    b = make_requests([(777, 0, 5, 0.0)])[0]
    b.rid, b.arrival = 1, 0.25
    b.bodies.length[0] = 100.0  # fresh: slow first step
    # compute budgets off: both colonies resident together
    srv = serving.ColonyServer(10**12, bytes_per_cell=1000, warm_steps=warm_steps,
                               max_cells_resident={})
    srv.run([a, b])
    return srv.results


def test_warm_up_batch_shields_steady_colonies():
    """A fresh colony's slow first step runs in the warm-up batch: the steady
    colony beside it in the same (model, bucket) keeps its step time; without
    the warm-up stage it waits for the slow step."""
    r = slow_pair(2)
    assert r[1].step_ms[0] >= 300
    # a 2 ms stand-in step; the margin absorbs a loaded host, and stays below
    # the 300 ms step the steady colony would have waited for
    assert max(r[0].step_ms) < 250
    r0 = slow_pair(0)
    assert max(r0[0].step_ms) >= 300


def test_division_before_each_step_seeded_per_colony():
    """Every batch step divides first, each colony with the seed of its own
    request id and step count (so a colony's run is the same alone and in
    any mix); admission reserves memory for the cells division can add."""
    reqs = make_requests([(100, 0, 6, 0.0), (120, 0, 4, 0.004), (90, 1, 5, 0.0)])
    srv = serving.ColonyServer(10**12, bytes_per_cell=1000, warm_steps=2)
    st = srv.run(reqs)
    assert st["completed"] == 3
    seen = {}
    for b in FakeBatch.instances:
        for seeds, step in b.divides:
            assert step == 0
            for sd in seeds:
                rid, k = sd >> 32, sd & 0xFFFFFFFF
                seen.setdefault(rid, []).append(k)
    for r in reqs:
        assert sorted(seen[r.rid]) == list(range(r.steps))
        assert serving.max_cells(r) > r.bodies.n
    assert st["peak_reserved_bytes"] >= 1000 * max(serving.max_cells(r) for r in reqs)


def test_step_log_and_batch_summary():
    """Every batch step is logged with its colonies, cells and the most
    BBPGD iterations of a colony; the summary groups them per batch."""
    reqs = make_requests([(100, 0, 3, 0.0), (120, 0, 3, 0.0), (90, 1, 2, 0.0)])
    srv = serving.ColonyServer(10**12, bytes_per_cell=1000, warm_steps=0, divide=False)
    srv.run(reqs)
    hard = [r for r in srv.step_log if r[0][0] == 0]
    soft = [r for r in srv.step_log if r[0][0] == 1]
    assert sum(r[1] for r in hard) == 6 and sum(r[1] for r in soft) == 2
    n0, n1 = reqs[0].bodies.n, reqs[1].bodies.n
    assert all(r[2] in ((n0, n1) if r[1] == 1 else (n0 + n1,)) for r in hard)
    summ = srv.step_summary()
    assert set(summ) == {"0-0-1", "1-0-1"}
    assert summ["0-0-1"]["batch_steps"] == len(hard)
    assert summ["0-0-1"]["max_iterations_p99"] == 0.0  # the fake batch reports none
