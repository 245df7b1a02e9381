"""Config 4 at its configured scale: 128 independent hard-model colonies with
sizes log-uniform in [500, 50,000] cells (~1.58M cells), mixed precision, in
one batched context with captured BBPGD graphs and per-colony convergence.

Four time-loop iterations (Domain::run, Domain.cpp:121-143: device division
seeded per colony and step, then the step; the first relaxes the packings'
initial overlaps).  Every colony's every step must succeed with its BBPGD
residual at or below the tolerance and its overlap at or below relcp_tol,
and the smallest, the median and the largest colony must be bit-identical
to a context stepping that colony alone through the same iterations.
"""
import math
import time

import numpy as np
import pytest

from pcc import generators as G
from pcc.abi_types import PCC_OK, PREC_MIXED, reference_params

from test_gpu_parity import assert_same_constraints, assert_same_state

pytestmark = pytest.mark.gpu
STEPS = 4


def params():
    return reference_params(dt=1e-3, tolerance=1e-6, tau=1 / math.log(2), lam=1e-2,
                            cutoff_phase0=0.05, cutoff_phaseK=0.05, relcp_tol=1e-6,
                            precision=PREC_MIXED, max_bbpgd_iterations=2_000_000,
                            temperature=0.0)


def test_ensemble_128_colonies_at_scale():
    from pcc.engine import ColonyBatch, ParticleManager
    from pcc.serving import division_seed
    t0 = time.time()
    cols = G.ensemble(128, seed=0)
    n = [c.n for c in cols]
    print(f"[ensemble] 128 colonies, {sum(n)} cells ({min(n)}..{max(n)}), "
          f"generated in {time.time() - t0:.1f} s", flush=True)
    assert len(cols) == 128 and 1_400_000 < sum(n) < 1_800_000
    assert min(n) >= 500 and max(n) <= 50_000
    p = params()
    cb = ColonyBatch(p)
    cb.set_state(cols, [p] * len(cols))
    for i in range(STEPS):
        t = time.time()
        cb.divide([division_seed(q, i) for q in range(len(cols))], 0, 0.0, 0.01)
        res = cb.step(i)
        its = [r["bbpgd_iterations"] for r in res]
        print(f"[ensemble] step {i}: {time.time() - t:.2f} s, {cb.n} cells, BBPGD iterations "
              f"{min(its)}..{max(its)}", flush=True)
        for q, r in enumerate(res):
            assert r["status"] == PCC_OK, (i, q, r)
            assert r["residual"] <= p.bbpgd_tol, (i, q, r)
            assert r["max_overlap"] <= p.relcp_tol, (i, q, r)
    assert cb.tuning()["tuned"]
    order = np.argsort(n, kind="stable")
    sample = [int(order[0]), int(order[len(order) // 2]), int(order[-1])]
    for q in sample:
        pm = ParticleManager(p)
        pm.set_state(cols[q])
        for i in range(STEPS):
            pm.divide(division_seed(q, i), 0, 0.0, 0.01)
            pm.step(i)
        print(f"[ensemble] colony {q}: {cols[q].n} -> {pm.n} cells, alone == batched",
              flush=True)
        assert_same_state(cb.get_state(q), pm.get_state())
        assert_same_constraints(cb.constraints(q), pm.constraints())
        pm.close()
    cb.close()
