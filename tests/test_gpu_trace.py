"""BBPGD trace on the device (pcc_set_trace / pcc_get_trace) against the
oracle's traced step: FP64 records (iteration, phase, constraint count,
residual, step size, the three energy terms and their sum) bit-exact on
planar and 3D colonies, the traced step's state bit-exact too (tracing
checks the residual every iteration, in both); truncation at the capacity;
untraced steps after it; per-colony traces in a batch equal to alone; mixed
precision within tolerance of the FP64 oracle."""
import math

import numpy as np
import pytest

import oracle
from pcc import generators as G
from pcc.abi_types import PREC_MIXED, reference_params

pytestmark = pytest.mark.gpu
LN2 = math.log(2.0)
FIELDS = ("pos", "quat", "length", "ldot", "velocity", "omega", "stress")


def random_3d(n, seed):
    rng = np.random.default_rng(seed)
    pos = rng.uniform(-4, 4, (n, 3))
    q = rng.normal(size=(n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    o = np.argsort(pos[:, 0], kind="stable")
    return G.Bodies(pos[o], q[o], rng.uniform(1, 2, n)[o])


def same_records(a, b):
    assert len(a) == len(b)
    for f in a.dtype.names:
        assert np.array_equal(a[f], b[f]), f


CASES = [
    ("colony3000", lambda: G.dense_colony(3000, seed=220), 1e-5),
    ("random3d", lambda: random_3d(300, 7), 1e-3),
]


@pytest.mark.parametrize("name,make,tol", CASES, ids=[c[0] for c in CASES])
def test_trace_fp64_bitexact_against_oracle(name, make, tol):
    from pcc.engine import ParticleManager
    p = reference_params(dt=1e-2, tolerance=tol, tau=1 / LN2, cutoff_phase0=0.05,
                         cutoff_phaseK=0.05)
    b = make()
    print(f"[trace] {name} n={b.n}", flush=True)
    st, res, want, _, _, err, rec, n = oracle.traced_step(p, b)
    assert st == 0, err
    pm = ParticleManager(p)
    pm.set_state(b)
    pm.set_trace(True, 100000)
    r = pm.step(0)
    got, cnt = pm.trace()
    assert cnt == n and r["bbpgd_iterations"] == res["bbpgd_iterations"]
    same_records(got, rec)
    s = pm.get_state()
    for f in FIELDS:
        assert np.array_equal(getattr(s, f), getattr(want, f)), f
    # off again: the next step is untraced and equals the oracle's step
    pm.set_trace(False)
    r2 = pm.step(1)
    assert pm.trace()[1] == 0
    st2, res2, want2 = oracle.step(p, want, 1)[:3]
    assert r2["bbpgd_iterations"] == res2["bbpgd_iterations"]
    assert np.array_equal(pm.get_state().pos, want2.pos)
    pm.close()


def test_trace_capacity_truncates_but_counts():
    from pcc.engine import ParticleManager
    p = reference_params(dt=1e-2, tolerance=1e-5, tau=1 / LN2)
    b = G.dense_colony(1500, seed=222)
    pm = ParticleManager(p)
    pm.set_state(b)
    pm.set_trace(True, 100000)
    pm.step(0)
    full, n = pm.trace()
    pm.set_state(b)
    pm.set_trace(True, 5)
    pm.step(0)
    got, n5 = pm.trace()
    assert n5 == n > 5 and len(got) == 5
    same_records(got, full[:5])
    pm.close()


@pytest.mark.parametrize("compact", [False, True], ids=["full_grids", "compacted"])
def test_batch_traces_equal_alone(compact, monkeypatch):
    """Per-colony traces of a batch equal the alone traces; with compaction
    forced (from one body block on, as soon as one colony has finished), the
    grids are rebuilt mid-solve, and the energy partials and block partials
    the trace reads follow the layout."""
    if compact:
        monkeypatch.setenv("PCC_COMPACT_MIN_BLOCKS", "1")
        monkeypatch.setenv("PCC_COMPACT_RATIO", "0.95")
    from pcc.engine import ColonyBatch, ParticleManager
    cols = [G.dense_colony(n, seed=230 + i) for i, n in enumerate([900, 2000, 150])]
    ps = [reference_params(dt=1e-2 * (1 + 0.3 * i), tolerance=1e-5, tau=1 / LN2)
          for i in range(3)]
    cb = ColonyBatch(ps[0])
    cb.set_state(cols, ps)
    cb.set_trace(True, 50000)
    cb.step(0)
    if compact:
        assert cb.solve_stats()["compactions"] >= 1
    for q in range(3):
        pm = ParticleManager(ps[q])
        pm.set_state(cols[q])
        pm.set_trace(True, 50000)
        pm.step(0)
        a, na = pm.trace()
        b, nb = cb.trace(q)
        assert na == nb
        same_records(b, a)
        pm.close()
    cb.close()


def test_trace_mixed_close_to_fp64_oracle():
    from pcc.engine import ParticleManager
    p = reference_params(dt=1e-2, tolerance=1e-6, tau=1 / LN2, precision=PREC_MIXED)
    b = G.dense_colony(3000, seed=240)
    rec = oracle.traced_step(p.replace(precision=0), b)[6]
    pm = ParticleManager(p)
    pm.set_state(b)
    pm.set_trace(True, 100000)
    pm.step(0)
    got, n = pm.trace()
    # phase 0 starts from the same warm start: the first records agree to
    # the FP32 storage of the operator
    k = min(5, len(got), len(rec))
    for f in ("linear", "quadratic", "growth", "total"):
        scale = np.max(np.abs(rec[f][:k])) + 1e-300
        err = np.max(np.abs(got[f][:k] - rec[f][:k])) / scale
        print(f"[trace mixed] {f} rel err {err:.2e}", flush=True)
        assert err < 1e-5
    assert got["residual"][-1] <= p.bbpgd_tol
    pm.close()
