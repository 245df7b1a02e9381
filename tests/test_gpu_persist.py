"""The one-launch BBPGD for small colonies (a workgroup runs a colony's whole
solve) against the multi-kernel path (PCC_PERSIST=0) and the oracle: every
step bit-identical, FP64 and mixed, planar and 3D, single colonies and
batches of small colonies, converged and failing solves, and a growth run
from one cell with device division."""
import math

import numpy as np
import pytest

import oracle
from pcc import generators as G
from pcc.abi_types import PREC_MIXED, reference_params

from test_gpu_parity import assert_same_constraints, assert_same_state, random_3d

pytestmark = pytest.mark.gpu
LN2 = math.log(2.0)


def hard(**kw):
    base = dict(dt=1e-3, tolerance=1e-6, tau=1 / LN2, cutoff_phase0=0.05, cutoff_phaseK=0.05)
    base.update(kw)
    return reference_params(**base)


def force_persist(monkeypatch, on):
    """PCC_PERSIST on with no size limit (every case takes the one-launch
    path), or off."""
    monkeypatch.setenv("PCC_PERSIST", "1" if on else "0")
    monkeypatch.setenv("PCC_PERSIST_MAX_C", "1000000000")
    monkeypatch.setenv("PCC_PERSIST_MAX_N", "1000000000")


def run(monkeypatch, persist, params, bodies, steps=2):
    from pcc.engine import ParticleManager
    force_persist(monkeypatch, persist)
    pm = ParticleManager(params)
    pm.set_state(bodies)
    out = [pm.step(k, raise_on_error=False) for k in range(steps)]
    res = (out, pm.get_state(), pm.constraints(), pm.phase_stats())
    pm.close()
    return res


CASES = [
    ("colony1500", lambda: G.dense_colony(1500, seed=300), {}),
    ("packed700", lambda: G.packed_colony(700, seed=301), {}),
    ("two_rods", lambda: G.two_rods(0.9), {}),
    ("single", lambda: G.single_cell(), {}),
    # 3D and a solve that hits the iteration cap: failures must agree too
    ("random3d_capped", lambda: random_3d(300, 41),
     dict(max_bbpgd_iterations=3000, max_recursive_iterations=3)),
]


@pytest.mark.parametrize("precision", [0, 1], ids=["fp64", "mixed"])
@pytest.mark.parametrize("name,make,kw", CASES, ids=[c[0] for c in CASES])
def test_persistent_equals_multikernel(name, make, kw, precision, monkeypatch):
    b = make()
    p = hard(precision=precision, **kw)
    print(f"[persist] {name} n={b.n} precision={precision}", flush=True)
    a = run(monkeypatch, True, p, b)
    m = run(monkeypatch, False, p, b)
    assert a[0] == m[0]
    assert_same_state(a[1], m[1])
    assert_same_constraints(a[2], m[2])
    assert a[3] == m[3]


def test_persistent_fp64_matches_oracle(monkeypatch):
    b = G.dense_colony(2000, seed=302)
    p = hard()
    st, res, want, cons, ph, err = oracle.step(p, b)
    got = run(monkeypatch, True, p, b, steps=1)
    assert got[0][0]["bbpgd_iterations"] == res["bbpgd_iterations"]
    for f in ("pos", "quat", "length", "ldot", "velocity", "omega", "stress"):
        assert np.array_equal(getattr(got[1], f), getattr(want, f)), f


@pytest.mark.parametrize("precision", [0, 1], ids=["fp64", "mixed"])
def test_persistent_batch_equals_alone(precision, monkeypatch):
    """A batch of small colonies (one persistent launch, a block per colony,
    each with its own convergence) equals each colony stepped alone."""
    from pcc.engine import ColonyBatch
    cols = [G.dense_colony(n, seed=310 + i) for i, n in enumerate([40, 900, 1, 2500, 300])]
    ps = [hard(precision=precision, tau=1 / (LN2 * (0.6 + 0.2 * i))) for i in range(5)]
    want = [run(monkeypatch, False, ps[q], cols[q]) for q in range(5)]
    force_persist(monkeypatch, True)
    cb = ColonyBatch(ps[0])
    cb.set_state(cols, ps)
    got = [cb.step(k) for k in range(2)]
    for q in range(5):
        for k in range(2):
            assert got[k][q] == want[q][0][k], (q, k)
        assert_same_state(cb.get_state(q), want[q][1])
        assert_same_constraints(cb.constraints(q), want[q][2])
    cb.close()


def test_persistent_growth_from_one_cell_matches_oracle(monkeypatch):
    """Config-1 parameters (dt 1e-3, tol 1e-6, growth at ln2, exact halves
    turned by N(0, 0.01)): divide and step on the device to 64 cells, equal
    to the oracle at every step."""
    from pcc.engine import ParticleManager
    force_persist(monkeypatch, True)
    p = hard(cutoff_phase0=0.25, cutoff_phaseK=1e-6)
    want = G.single_cell()
    pm = ParticleManager(p)
    pm.set_state(want)
    k = 0
    while want.n < 64:
        st, res, want, _, _, err = oracle.step(p, want, k)
        assert st == 0, err
        r = pm.step(k)
        assert r["bbpgd_iterations"] == res["bbpgd_iterations"], k
        want = oracle.divide(want, 0, k, 0.0, 0.01)
        assert pm.divide(0, k, 0.0, 0.01) == want.n
        k += 1
    print(f"[persist growth] {k} steps -> {want.n} cells", flush=True)
    got = pm.get_state()
    for f in ("pos", "quat", "length", "l0", "diameter", "ldot"):
        assert np.array_equal(getattr(got, f), getattr(want, f)), f
    pm.close()
