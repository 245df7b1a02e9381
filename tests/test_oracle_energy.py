"""Pins the oracle's BBPGD energy and trace (HardModelGradient::energy,
HardModelGradient.cpp:160-203; BBPGD.cpp:57-80).

  * two-body head-on closed form of the three terms,
      E = gamma phi + dt/2 K gamma^2 + dt/(lambda tau) (l_a + l_b) e^{-lambda gamma/2},
    with K = 1/(xi l_a) + 1/(xi l_b) (collinear rods: no torque, s = 1/2);
  * the gradient of E by central differences equals the BBPGD gradient when
    gamma_old = 0 and ldot_prev = 0 (the energy is then the solver's
    objective), on a dense colony with rotations;
  * trace semantics: the residual is checked every iteration, so each
    phase's records run 0..its with only the last one at or below the
    tolerance, step 0 has step size 1/residual, and a traced phase 0 stops no
    later than the untraced one; the energy drops over a traced phase 0.
"""
import math

import numpy as np
import pytest

import oracle
from pcc import generators as G
from pcc.abi_types import reference_params

LN2 = math.log(2.0)


def test_two_body_energy_closed_form():
    la, lb, sep = 1.0, 1.4, 1.1
    p = reference_params(dt=1e-3, tolerance=1e-9, tau=1 / LN2, lam=0.37, xi=1.3,
                         cutoff_phase0=0.05, warm_start_factor=0.0)
    b = G.two_rods(sep, 0.0, la, lb)
    C, _, _ = oracle.phase0_energy(p, b)
    assert C == 1
    phi = sep - (la + lb) / 2
    K = 1 / (p.xi * la) + 1 / (p.xi * lb)
    for gam in (0.0, 0.25, 3.0, 40.0):
        _, E, _ = oracle.phase0_energy(p, b, np.array([gam]))
        lin = gam * phi
        quad = p.dt / 2 * K * gam * gam
        grow = p.dt / (p.lam * p.tau) * (la + lb) * math.exp(-p.lam * gam / 2)
        assert E[0] == pytest.approx(lin, rel=1e-14, abs=1e-300)
        assert E[1] == pytest.approx(quad, rel=1e-13, abs=1e-300)
        assert E[2] == pytest.approx(grow, rel=1e-13)
        assert E[3] == pytest.approx(lin + quad + grow, rel=1e-13)


@pytest.mark.parametrize("seed", [3, 4])
def test_energy_gradient_is_the_bbpgd_gradient(seed):
    p = reference_params(dt=1e-2, tolerance=1e-4, tau=1 / LN2, lam=0.05,
                         warm_start_factor=0.0)
    b = G.dense_colony(150, seed=seed)
    C, _, _ = oracle.phase0_energy(p, b)
    assert C > 50
    rng = np.random.default_rng(seed)
    gam = rng.random(C) * 2.0
    _, E, g = oracle.phase0_energy(p, b, gam)
    for c in rng.choice(C, 12, replace=False):
        h = 1e-4
        gp, gm = gam.copy(), gam.copy()
        gp[c] += h
        gm[c] -= h
        Ep = oracle.phase0_energy(p, b, gp)[1][3]
        Em = oracle.phase0_energy(p, b, gm)[1][3]
        fd = (Ep - Em) / (2 * h)
        assert fd == pytest.approx(g[c], rel=1e-6, abs=1e-9 * max(1.0, abs(E[3])))


def test_trace_records_every_iteration_of_every_phase():
    p = reference_params(dt=1e-2, tolerance=1e-5, tau=1 / LN2)
    b = G.dense_colony(400, seed=7)
    st, res, nb, cons, phases, err, rec, n = oracle.traced_step(p, b)
    assert st == 0, err
    phases = phases[:res["constraint_iterations"]]  # solved phases
    assert len(phases) >= 2
    assert n == len(rec) == sum(int(x) + 1 for x in phases)
    o = 0
    for k, its in enumerate(phases):
        r = rec[o:o + its + 1]
        o += its + 1
        assert np.array_equal(r["step"], np.arange(its + 1))
        assert np.all(r["recursive_iteration"] == k)
        assert len(set(r["total_constraints"])) == 1
        assert r["step_size"][0] == 1.0 / r["residual"][0]
        assert r["residual"][-1] <= p.bbpgd_tol
        assert np.all(r["residual"][:-1] > p.bbpgd_tol)
        assert np.array_equal(r["total"], r["linear"] + r["quadratic"] + r["growth"])
    assert res["bbpgd_iterations"] == sum(int(x) for x in phases)


def test_traced_phase0_stops_no_later_and_descends():
    p = reference_params(dt=1e-2, tolerance=1e-6, tau=1 / LN2, warm_start_factor=0.0)
    b = G.dense_colony(300, seed=8)
    untraced = oracle.step(p, b)[4]
    traced, rec = oracle.traced_step(p, b)[4], oracle.traced_step(p, b)[6]
    assert traced[0] <= untraced[0]
    assert untraced[0] - traced[0] < 10
    r0 = rec[rec["recursive_iteration"] == 0]
    assert r0["total"][-1] < r0["total"][0]
