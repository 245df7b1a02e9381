"""The continuous-batching server on the GPU: colonies join and leave the
hard and the soft batches between steps while others are mid-run, divide
before every step, move from the warm-up to the steady batch, and every
result is bit-identical to a context that runs the same request alone
(divide with the request's seed, then step)."""
import math

import pytest

from pcc import abi_types as T
from pcc import generators as G

from test_gpu_parity import assert_same_state

pytestmark = pytest.mark.gpu
LN2 = math.log(2.0)


def request_params(mode, rate):
    if mode == T.MODE_SOFT:
        return T.reference_params(mode=T.MODE_SOFT, dt=1e-3 / 30, tolerance=1e-6, tau=1 / rate,
                                  kcc=4.0 / 3.0 * 1e4, precision=T.PREC_MIXED)
    return T.reference_params(dt=1e-3, tolerance=1e-6, tau=1 / rate, cutoff_phase0=0.05,
                              cutoff_phaseK=0.05, precision=T.PREC_MIXED)


def test_served_results_equal_alone():
    from pcc.engine import ParticleManager
    from pcc.serving import ColonyServer, Request, division_seed
    spec = [(600, 0, 4, 0.0), (300, 1, 3, 0.0), (900, 0, 2, 0.0), (150, 1, 5, 0.0),
            (400, 0, 3, 0.0), (1200, 1, 2, 0.0), (250, 0, 6, 0.0)]
    reqs = []
    for rid, (n, mode, steps, arr) in enumerate(spec):
        rate = (0.5 + 0.15 * rid) * LN2
        b = G.packed_colony(n, seed=90 + rid)
        b.length[b.length > 1.9] = 2.0  # these divide before the first step
        reqs.append(Request(rid, b, request_params(mode, rate), steps, arr))
    # room for ~2500 reserved cells at a time (a reservation covers the cells
    # division can add): later requests wait and then join
    # batches whose other colonies are mid-run
    srv = ColonyServer(2500 * 10_000, precision=T.PREC_MIXED, bytes_per_cell=10_000)
    st = srv.run(reqs)
    srv.close()
    print(f"[serving] {st}", flush=True)
    assert st["completed"] == len(reqs)
    assert st["peak_reserved_bytes"] <= 2500 * 10_000
    waited = [r for r in srv.results.values() if r.admitted > 0.0]
    assert waited, "no request had to wait for memory"
    for r in reqs:
        res = srv.results[r.rid]
        assert res.statuses == [T.PCC_OK] * r.steps
        pm = ParticleManager(r.params)
        pm.set_state(r.bodies)
        for k in range(r.steps):
            pm.divide(division_seed(r.rid, k), 0, 0.0, 0.01)
            pm.step(k)
        assert res.state.n > r.bodies.n  # cells divided
        assert_same_state(res.state, pm.get_state())
        pm.close()


def test_concurrent_contexts_allocate_during_graph_capture():
    """One thread steps a hard batch (every BBPGD phase captures a graph)
    while another keeps growing soft contexts (device allocations, frees and
    synchronisations): both must run to the end and give their solo results."""
    import threading
    from pcc.engine import ColonyBatch, ParticleManager
    hp = request_params(T.MODE_HARD, LN2)
    sp = request_params(T.MODE_SOFT, LN2)
    hard_cols = [G.dense_colony(n, seed=120 + i) for i, n in enumerate([3000, 1200, 5000])]
    soft_cols = [G.dense_colony(n, seed=130 + i) for i, n in enumerate([500, 1500, 4000, 9000])]
    out, errors = {}, []

    def hard_worker():
        try:
            cb = ColonyBatch(hp)
            cb.set_state(hard_cols, [hp] * 3)
            for k in range(6):
                cb.step(k)
            out["hard"] = [cb.get_state(q) for q in range(3)]
            cb.close()
        except Exception as e:  # reported below
            errors.append(repr(e))

    def soft_worker():
        try:
            res = []
            for b in soft_cols:  # growing sizes: every context allocates afresh
                pm = ParticleManager(sp)
                pm.set_state(b)
                for k in range(3):
                    pm.step(k)
                res.append(pm.get_state())
                pm.close()
            out["soft"] = res
        except Exception as e:
            errors.append(repr(e))

    th = [threading.Thread(target=hard_worker), threading.Thread(target=soft_worker)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    assert not errors, errors
    cb = ColonyBatch(hp)
    cb.set_state(hard_cols, [hp] * 3)
    for k in range(6):
        cb.step(k)
    for q in range(3):
        assert_same_state(out["hard"][q], cb.get_state(q))
    cb.close()
    for b, got in zip(soft_cols, out["soft"]):
        pm = ParticleManager(sp)
        pm.set_state(b)
        for k in range(3):
            pm.step(k)
        assert_same_state(got, pm.get_state())
        pm.close()


def test_config5_request_mix_under_a_binding_budget():
    """Config 5's request mix over a short horizon: Poisson arrivals at 20/s,
    colonies of 1k-100k cells (log-uniform, the largest hard and soft ones
    near 100k), hard or soft 50/50, growth rate U[0.5, 1.5] ln 2, 3 steps
    each.  The HBM budget holds the largest colony's reservation but not the
    two largest together, so requests wait for memory.  The budget is never exceeded, every request
    completes, and the largest hard and the largest soft request equal their
    alone runs bit for bit."""
    import numpy as np
    from pcc.engine import ParticleManager
    from pcc.serving import ColonyServer, Request, division_seed
    rng = np.random.default_rng(5)
    sizes = [int(np.exp(rng.uniform(np.log(1e3), np.log(1e5)))) for _ in range(8)]
    sizes[0], sizes[1] = 99_000, 97_000  # the largest hard and the largest soft
    reqs, t = [], 0.0
    for rid, n in enumerate(sizes):
        t += rng.exponential(1 / 20.0)
        mode = T.MODE_HARD if rid % 2 == 0 else T.MODE_SOFT
        rate = rng.uniform(0.5, 1.5) * LN2
        p = request_params(mode, rate).replace(temperature=0.0)
        reqs.append(Request(rid, G.packed_colony(n, seed=300 + rid), p, 3, t))
    bpc = 12_000
    srv = ColonyServer(1, precision=T.PREC_MIXED, bytes_per_cell=bpc)
    budget = int(1.2 * max(srv.reservation(r) for r in reqs))
    srv = ColonyServer(budget, precision=T.PREC_MIXED, bytes_per_cell=bpc)
    print(f"[config5] sizes {sizes}, budget {budget / 2**30:.2f} GiB", flush=True)
    st = srv.run(reqs, time_limit=900)
    srv.close()
    print(f"[config5] {st}", flush=True)
    assert st["completed"] == len(reqs)
    assert st["peak_reserved_bytes"] <= budget
    assert st["peak_device_bytes"] <= budget
    waits = [srv.results[r.rid].admitted - r.arrival for r in reqs]
    assert max(waits) > 0.05, "no request waited for memory"
    for rid in (0, 1):
        r = reqs[rid]
        res = srv.results[rid]
        assert res.statuses == [T.PCC_OK] * r.steps
        pm = ParticleManager(r.params)
        pm.set_state(r.bodies)
        for k in range(r.steps):
            pm.divide(division_seed(r.rid, k), 0, 0.0, 0.01)
            pm.step(k)
        print(f"[config5] request {rid} ({'hard' if rid == 0 else 'soft'}, {r.bodies.n} cells): "
              f"served == alone", flush=True)
        assert_same_state(res.state, pm.get_state())
        pm.close()
