"""Batched ensembles on the GPU: many colonies in one context, one launch
sequence per BBPGD iteration with per-colony step sizes and convergence flags.

The bar: every colony's results in the batch are the ones a context holding it
alone produces, bit for bit (state, constraint set, SolverSolution summary and
solver status), in FP64 and in mixed precision, for the hard and the soft
model.  The single-colony path itself is pinned bit-exact to the FP64 oracle
in test_gpu_parity.py.
"""
import math

import numpy as np
import pytest

import oracle
from pcc import generators as G
from pcc.abi_types import (MODE_SOFT, PCC_E_BBPGD_NOT_CONVERGED, PCC_E_RELCP_MAX_ITERS, PCC_OK,
                           PREC_MIXED, reference_params)

from test_gpu_parity import assert_same_constraints, assert_same_state, random_3d

pytestmark = pytest.mark.gpu
LN2 = math.log(2.0)


def hard(**kw):
    base = dict(dt=1e-3, tolerance=1e-6, tau=1 / LN2, cutoff_phase0=0.05, cutoff_phaseK=0.05)
    base.update(kw)
    return reference_params(**base)


def alone(params, bodies, steps=1):
    from pcc.engine import ParticleManager
    pm = ParticleManager(params)
    pm.set_state(bodies)
    out = []
    for k in range(steps):
        out.append(pm.step(k, raise_on_error=False))
    res = (out, pm.get_state(), pm.constraints(), pm.phase_stats())
    pm.close()
    return res


def check_batch_equals_alone(ctx_params, colonies, params, steps=1):
    from pcc.engine import ColonyBatch
    want = [alone(p, b, steps) for p, b in zip(params, colonies)]
    cb = ColonyBatch(ctx_params)
    cb.set_state(colonies, params)
    got_r = [cb.step(k) for k in range(steps)]
    for q in range(len(colonies)):
        print(f"  colony {q}: n={colonies[q].n} status={got_r[-1][q]['status']} "
              f"C={got_r[-1][q]['n_constraints']} its={got_r[-1][q]['bbpgd_iterations']}",
              flush=True)
        w_r, w_state, w_cons, w_ph = want[q]
        for k in range(steps):
            assert got_r[k][q] == w_r[k], (q, k, got_r[k][q], w_r[k])
        assert_same_state(cb.get_state(q), w_state)
        assert_same_constraints(cb.constraints(q), w_cons)
        assert cb.phase_stats(q) == w_ph
    cb.close()


SIZES = [37, 400, 1500, 90, 2600, 1]


@pytest.mark.parametrize("precision", [0, 1], ids=["fp64", "mixed"])
def test_batch_hard_equals_alone(precision):
    """Colonies of uneven sizes with their own growth rates and tolerances."""
    print(f"[batch-hard] precision={precision}", flush=True)
    colonies = [G.packed_colony(n, seed=20 + i) for i, n in enumerate(SIZES)]
    params = [hard(precision=precision, tau=1 / (LN2 * (0.5 + 0.2 * i)),
                   tolerance=1e-6 if i % 2 == 0 else 1e-5) for i in range(len(SIZES))]
    check_batch_equals_alone(params[0], colonies, params, steps=2)


@pytest.mark.parametrize("precision", [0, 1], ids=["fp64", "mixed"])
def test_solo_colonies_equal_alone(precision, monkeypatch):
    """With a solo threshold of 1000 cells the colonies of 1500 and 2600
    cells run their ReLCP loops alone, one after the other, after the
    lockstep loop of the small ones (whose constraints stay in the set while
    the solo loops append and regroup theirs); every colony still equals its
    alone run bit for bit, a colony with rotational diffusion included."""
    monkeypatch.setenv("PCC_SOLO_CELLS", "1000")
    print(f"[batch-solo] precision={precision}", flush=True)
    colonies = [G.packed_colony(n, seed=40 + i) for i, n in enumerate(SIZES)]
    params = [hard(precision=precision, tau=1 / (LN2 * (0.5 + 0.2 * i)),
                   temperature=0.05 if i == 4 else 0.0) for i in range(len(SIZES))]
    check_batch_equals_alone(params[0], colonies, params, steps=2)


@pytest.mark.parametrize("precision", [0, 1], ids=["fp64", "mixed"])
def test_compacted_grids_equal_alone(precision, monkeypatch):
    """Eight 3,000-cell colonies, half of them with a looser tolerance: once
    the loose ones have converged, the solve rebuilds its grids over the
    colonies still iterating (compaction from 8 body blocks on here).  Every
    colony still equals its alone run bit for bit, and the batch did compact."""
    monkeypatch.setenv("PCC_COMPACT_MIN_BLOCKS", "8")
    print(f"[batch-compact] precision={precision}", flush=True)
    from pcc.engine import ColonyBatch
    colonies = [G.packed_colony(3000, seed=60 + i) for i in range(8)]
    params = [hard(precision=precision, tolerance=1e-3 if i % 2 else 1e-6) for i in range(8)]
    check_batch_equals_alone(params[0], colonies, params, steps=2)
    cb = ColonyBatch(params[0])
    cb.set_state(colonies, params)
    cb.step(0)
    n = cb.solve_stats()["compactions"]
    print(f"  compactions in one step: {n}", flush=True)
    assert n >= 1
    cb.close()


def test_batch_solver_failures_stay_per_colony():
    """One colony exhausts BBPGD, one exhausts ReLCP phases, one is empty; the
    others are unaffected and every status is the one it has alone."""
    print("[batch-fail]", flush=True)
    colonies = [G.packed_colony(800, seed=30), G.two_rods(0.9), G.two_rods(0.9),
                G.Bodies(np.zeros((0, 3)), np.zeros((0, 4)), np.zeros(0)),
                G.packed_colony(700, seed=31)]
    params = [hard(), hard(max_bbpgd_iterations=5), hard(max_recursive_iterations=1, relcp_tol=1e-6),
              hard(), hard(relcp_tol=1e-4)]
    from pcc.engine import ColonyBatch
    cb = ColonyBatch(params[0])
    cb.set_state(colonies, params)
    r = cb.step()
    assert r[1]["status"] == PCC_E_BBPGD_NOT_CONVERGED
    assert r[2]["status"] == PCC_E_RELCP_MAX_ITERS
    assert r[0]["status"] == PCC_OK and r[3]["status"] == PCC_OK and r[4]["status"] == PCC_OK
    cb.close()
    check_batch_equals_alone(params[0], colonies, params)


def test_batch_3d_and_planar_colonies():
    """A 3D colony in the batch switches every colony to the 3D kernels, which
    give the planar kernels' values."""
    print("[batch-3d]", flush=True)
    colonies = [G.dense_colony(3000, seed=40), random_3d(300, 41), G.packed_colony(500, seed=42)]
    params = [hard(), hard(relcp_tol=1e-3, max_recursive_iterations=8), hard()]
    check_batch_equals_alone(params[0], colonies, params)


@pytest.mark.parametrize("precision", [0, 1], ids=["fp64", "mixed"])
def test_batch_soft_equals_alone(precision):
    print(f"[batch-soft] precision={precision}", flush=True)
    colonies = [G.packed_colony(n, seed=50 + i) for i, n in enumerate([300, 2000, 40])]
    params = [reference_params(mode=MODE_SOFT, precision=precision, dt=1e-3 / 30, tolerance=1e-6,
                               tau=1 / (LN2 * (0.6 + 0.3 * i)), kcc=4.0 / 3.0 * 1e4)
              for i in range(3)]
    check_batch_equals_alone(params[0], colonies, params, steps=3)


def test_batch_matches_oracle_fp64():
    """Straight to the oracle as well: a batched FP64 colony is bit-identical
    to the FP64 restatement."""
    print("[batch-oracle]", flush=True)
    colonies = [G.packed_colony(600, seed=60), G.random_disc(500, seed=61)]
    params = [hard(), hard(tau=1.0)]
    from pcc.engine import ColonyBatch
    cb = ColonyBatch(params[0])
    cb.set_state(colonies, params)
    r = cb.step()
    for q in range(2):
        st, res, want_b, want_c, ph, err = oracle.step(params[q], colonies[q])
        assert r[q]["status"] == st
        for k in res:
            assert r[q][k] == res[k], (q, k)
        assert_same_state(cb.get_state(q), want_b)
        assert_same_constraints(cb.constraints(q), want_c)
    cb.close()


def test_batch_add_remove_keeps_resident_state():
    """Continuous batching: colonies join and leave between steps; the others'
    device-resident state carries on as if they ran alone."""
    print("[batch-add-remove]", flush=True)
    from pcc.engine import ColonyBatch, ParticleManager
    p = hard(precision=PREC_MIXED)
    a, b, c = G.packed_colony(700, seed=70), G.packed_colony(300, seed=71), G.packed_colony(1200, seed=72)
    ref = {}
    for name, bod, steps in (("a", a, 3), ("b", b, 1), ("c", c, 2)):
        pm = ParticleManager(p)
        pm.set_state(bod)
        for k in range(steps):
            pm.step(k)
        ref[name] = pm.get_state()
        pm.close()
    cb = ColonyBatch(p)
    cb.set_state([a, b], [p, p])
    cb.step(0)  # a: 1, b: 1
    assert_same_state(cb.get_state(1), ref["b"])
    cb.remove(1)  # b leaves
    qc = cb.add(c, p)  # c joins
    assert qc == 1
    cb.step(1)  # a: 2, c: 1
    cb.step(2)  # a: 3, c: 2
    assert_same_state(cb.get_state(0), ref["a"])
    assert_same_state(cb.get_state(1), ref["c"])
    cb.close()


@pytest.mark.parametrize("precision", [0, 1], ids=["fp64", "mixed"])
def test_graph_replays_equal_plain_launches(precision, monkeypatch):
    """BBPGD chunks run as captured HIP graphs by default; every replay must
    give what plain launches give (PCC_GRAPH=0), step after step, with the
    host's eager copies and synchronisations between the replays."""
    print(f"[graph] precision={precision}", flush=True)
    from pcc.engine import ColonyBatch
    colonies = [G.packed_colony(n, seed=80 + i) for i, n in enumerate([900, 150, 2200])]
    params = [hard(precision=precision, tau=1 / (LN2 * (0.7 + 0.3 * i))) for i in range(3)]
    runs = []
    for env in ("1", "0"):
        monkeypatch.setenv("PCC_GRAPH", env)
        cb = ColonyBatch(params[0])
        cb.set_state(colonies, params)
        steps = []
        for k in range(4):
            r = cb.step(k)
            steps.append((r, [cb.get_state(q) for q in range(3)]))
        cb.close()
        runs.append(steps)
    for (ra, sa), (rb, sb) in zip(runs[0], runs[1]):
        assert ra == rb
        for q in range(3):
            assert_same_state(sa[q], sb[q])


def test_autotuned_body_kernel_gives_the_same_bits(monkeypatch):
    """The autotuner picks the body kernel's staging size on the first solve
    with >= 1e5 constraints; every candidate sums in the same order, so a
    tuned step equals an untuned one bit for bit (FP64)."""
    from pcc.engine import ParticleManager
    print("[autotune]", flush=True)
    b = G.dense_colony(45000, seed=160)
    p = hard()
    runs = []
    for env in ("1", "0"):
        monkeypatch.setenv("PCC_AUTOTUNE", env)
        pm = ParticleManager(p)
        pm.set_state(b)
        r = pm.step()
        runs.append((r, pm.get_state(), pm.tuning()))
        pm.close()
    print(f"  tuning: {runs[0][2]}", flush=True)
    assert runs[0][2]["tuned"] and not runs[1][2]["tuned"]
    assert runs[0][0] == runs[1][0]
    assert_same_state(runs[0][1], runs[1][1])


@pytest.mark.parametrize("precision", [0, 1], ids=["fp64", "mixed"])
def test_every_body_kernel_variant_gives_the_same_bits(precision, monkeypatch):
    """Each staging size and threads-per-block variant (forced with
    PCC_BODY_VARIANT) sums every body in the same order: planar and 3D steps
    equal the default variant's bit for bit.  PCC_PERSIST=0 keeps the small
    3D case on the multi-kernel path, which the variants belong to."""
    from pcc.engine import BODY_VARIANTS, ParticleManager
    monkeypatch.setenv("PCC_PERSIST", "0")
    for name, b, p in (
            ("planar", G.dense_colony(6000, seed=161), hard(precision=precision)),
            # a hard 3D packing, capped: failing solves must agree bit for bit too
            ("3d", random_3d(300, 41), hard(precision=precision, max_bbpgd_iterations=3000,
                                            max_recursive_iterations=4))):
        runs = []
        for v in range(BODY_VARIANTS):
            monkeypatch.setenv("PCC_BODY_VARIANT", str(v))
            pm = ParticleManager(p)
            pm.set_state(b)
            r = pm.step(raise_on_error=False)
            t = pm.tuning()
            runs.append((r, pm.get_state(), pm.constraints()))
            pm.close()
            assert t["variant"] == v
            print(f"  [{name} variant {v}] chunk={t['body_chunk']} threads={t['body_threads']} "
                  f"prefetch={t['prefetch']} "
                  f"its={r['bbpgd_iterations']}", flush=True)
        for r, st, cons in runs[1:]:
            assert r == runs[0][0]
            assert_same_state(st, runs[0][1])
            assert_same_constraints(cons, runs[0][2])

