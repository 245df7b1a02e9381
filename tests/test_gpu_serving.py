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
