"""Pins the oracle's hard (ReLCP/BBPGD) and soft (Hertz) steps.

Independent checks, each chosen so that a plausible slip (a sign, a dropped
mobility factor, a wrong torque lever, the lagging-iterate quirk) fails:
  * two-body head-on LCP in closed form (lambda = 0);
  * two rods in a T, where the rotational mobility 12/(xi l^3) enters K;
  * an independent dense QP solve (scipy L-BFGS-B) of the linear
    complementarity problem of one phase: body velocities agree;
  * Newton's third law and angular-momentum balance of D^T gamma and Hertz;
  * exponential growth of an isolated cell, l_k = l_0 (1 + dt/tau)^k;
  * BBPGD quirks: iterations are multiples of 10, >= 10 with no contacts;
  * failure modes and their messages.
"""
import math

import numpy as np
import pytest
from scipy.optimize import minimize

import oracle
from pcc import generators as G
from pcc.abi_types import (MODE_SOFT, PCC_E_BBPGD_NOT_CONVERGED, PCC_E_INVALID_ARGUMENT,
                           PCC_E_INVALID_MODE, PCC_E_RELCP_MAX_ITERS, reference_params)

LN2 = math.log(2.0)


def hard_params(**kw):
    base = dict(dt=1e-3, tolerance=1e-9, tau=1 / LN2, lam=0.0, cutoff_phase0=0.05,
                cutoff_phaseK=0.05, relcp_tol=math.inf)
    base.update(kw)
    return reference_params(**base)


def test_two_body_head_on_closed_form():
    la, lb, sep = 1.0, 1.4, 1.1
    b = G.two_rods(sep, 0.0, la, lb)
    p = hard_params()
    st, res, nb, cons, ph, err = oracle.step(p, b)
    assert st == 0, err
    phi = sep - (la + lb) / 2
    g0 = max(0.0, -500 * phi)
    K = 1 / la + 1 / lb
    growth = 0.5 * (la + lb) / p.tau
    gstar = max(0.0, g0 + (p.dt * growth - phi) / (p.dt * K))
    assert cons["gamma"][0] == pytest.approx(gstar, rel=1e-9)
    # the bodies move apart along x by dt * (gamma - gamma0)/(xi l)
    imp = gstar - g0
    assert nb.pos[0, 0] == pytest.approx(-p.dt * imp / la, rel=1e-7)
    assert nb.pos[1, 0] == pytest.approx(sep + p.dt * imp / lb, rel=1e-7)
    assert res["bbpgd_iterations"] % 10 == 0


def test_t_configuration_uses_rotational_mobility():
    # B perpendicular to A, its end cap pressing on A's side off-centre
    la, lb = 2.0, 1.0
    xb, yb = 0.4, 0.7  # B's lower cap centre at y = 0.7 - 0.25 = 0.45 -> overlap 0.05
    b = G.Bodies(np.array([[0, 0, 0], [xb, yb, 0.0]]),
                 np.array([[1, 0, 0, 0], [math.cos(math.pi / 4), 0, 0, math.sin(math.pi / 4)]]),
                 np.array([la, lb]))
    p = hard_params()
    st, res, nb, cons, ph, err = oracle.step(p, b)
    assert st == 0, err
    r = oracle.detect(b, 0.05)
    phi = r["signed_distance"][0]
    assert phi == pytest.approx(-0.05, abs=1e-12)
    n = r["normal"][0]
    rI, rJ = r["rPosI"][0], r["rPosJ"][0]
    tI = np.cross(rI, n)
    tJ = np.cross(rJ, -n)
    K = (n @ n) / la + (tI @ tI) * 12 / la ** 3 + (n @ n) / lb + (tJ @ tJ) * 12 / lb ** 3
    g0 = -500 * phi
    growth = (r["stressI"][0] * la + r["stressJ"][0] * lb) / p.tau
    gstar = g0 + (p.dt * growth - phi) / (p.dt * K)
    assert cons["gamma"][0] == pytest.approx(gstar, rel=1e-8)
    # A rotates (torque about z), orientation stays unit and in-plane
    assert abs(nb.omega[0, 2]) > 0
    assert np.allclose(np.linalg.norm(nb.quat, axis=1), 1.0, atol=1e-15)
    assert np.all(nb.quat[:, 1:3] == 0) and np.all(nb.pos[:, 2] == 0)


def _phase0_qp(b, p):
    """Independent solve of phase 0 as the convex QP
    min_{gamma>=0} 1/2 (gamma-g0)^T A (gamma-g0) + c^T gamma,  A = dt D M D^T,
    c = phi - dt L ldot_free (lambda = 0).  Returns body velocities."""
    r = oracle.detect(b, p.cutoff_phase0)
    C, N = len(r["gidI"]), b.n
    D = np.zeros((C, 6 * N))
    L = np.zeros((C, N))
    for c in range(C):
        i, j, n = r["gidI"][c], r["gidJ"][c], r["normal"][c]
        D[c, 6 * i:6 * i + 3] = n
        D[c, 6 * i + 3:6 * i + 6] = np.cross(r["rPosI"][c], n)
        D[c, 6 * j:6 * j + 3] = -n
        D[c, 6 * j + 3:6 * j + 6] = np.cross(r["rPosJ"][c], -n)
        L[c, i] = r["stressI"][c]
        L[c, j] = r["stressJ"][c]
    m = np.zeros(6 * N)
    for i in range(N):
        m[6 * i:6 * i + 3] = 1 / (p.xi * b.length[i])
        m[6 * i + 3:6 * i + 6] = 12 / (p.xi * b.length[i] ** 3)
    A = p.dt * D @ (m[:, None] * D.T)
    g0 = r["gamma"]
    c = r["signed_distance"] - p.dt * L @ (b.length / p.tau)

    def f(g):
        d = g - g0
        return 0.5 * d @ A @ d + c @ g, A @ d + c

    sol = minimize(f, np.maximum(g0, 0), jac=True, method="L-BFGS-B",
                   bounds=[(0, None)] * C, options=dict(maxiter=20000, ftol=1e-15, gtol=1e-12))
    U = m * (D.T @ (sol.x - g0))
    return U.reshape(N, 6), C


@pytest.mark.parametrize("n,seed", [(12, 0), (25, 1)])
def test_phase0_against_independent_qp(n, seed):
    b = G.random_disc(n, seed=seed, area_fraction=0.6)
    p = hard_params(tolerance=1e-8)
    st, res, nb, cons, ph, err = oracle.step(p, b)
    assert st == 0, err
    U, C = _phase0_qp(b, p)
    assert C > 3
    Uo = np.concatenate([nb.velocity, nb.omega], axis=1)
    scale = np.abs(U).max()
    assert np.abs(Uo - U).max() <= 2e-5 * scale


def test_momentum_balance_of_constraint_forces():
    b = G.random_disc(300, seed=3, area_fraction=0.8)
    p = hard_params(tolerance=1e-6, lam=1e-2)
    st, res, nb, cons, ph, err = oracle.step(p, b)
    assert st == 0, err
    F, T = nb.force, nb.torque
    scale = np.abs(F).max()
    assert np.abs(F.sum(0)).max() <= 1e-10 * scale * b.n
    # total angular momentum about the origin: sum x_i x F_i + tau_i = 0, with
    # x the pre-step centres (the forces act at the configuration they were
    # assembled on)
    L = np.cross(b.pos, F).sum(0) + T.sum(0)
    assert np.abs(L).max() <= 1e-9 * scale * b.n


def test_isolated_cell_grows_exponentially_and_bbpgd_quirks():
    p = reference_params(dt=1e-3, tolerance=1e-6, tau=1 / LN2)
    b = G.single_cell()
    lens = [1.0]
    for k in range(50):
        st, res, b, cons, ph, err = oracle.step(p, b, k)
        assert st == 0
        assert res["bbpgd_iterations"] == 10      # Q4: residual only checked at it % 10 == 0
        assert res["constraint_iterations"] == 1  # phase 1 finds no overlap and stops
        assert res["n_constraints"] == 0
        lens.append(b.length[0])
    want = [(1 + p.dt / p.tau) ** k for k in range(51)]
    assert np.allclose(lens, want, rtol=1e-13)
    assert np.all(b.pos == 0)


def test_length_clamped_at_twice_l0():
    p = reference_params(dt=0.5, tolerance=1e-6, tau=1.0)
    b = G.single_cell(length=1.9)
    st, res, b, cons, ph, err = oracle.step(p, b)
    assert b.length[0] == 2.0


def test_failure_modes():
    b = G.two_rods(0.9)
    st, *_, err = oracle.step(hard_params(max_bbpgd_iterations=5), b)
    assert st == PCC_E_BBPGD_NOT_CONVERGED and err == "BBPGD did not converge"
    st, res, nb, *_, err = oracle.step(hard_params(max_recursive_iterations=1, relcp_tol=1e-6), b)
    assert st == PCC_E_RELCP_MAX_ITERS
    assert err == "Maximum number of constraint iterations reached"
    assert nb.length[0] > b.length[0]  # grown before the throw (Hard.cpp:185-190)
    st, *_, err = oracle.step(hard_params(mode=7), b)
    assert st == PCC_E_INVALID_MODE and err == "Invalid mode"
    st, *_, err = oracle.step(hard_params(temperature=-1.0), b)
    assert st == PCC_E_INVALID_ARGUMENT and err == "invalid temperature"


def test_relcp_grown_colony_converges_in_few_phases():
    p = reference_params(dt=1e-3, tolerance=1e-6, tau=1 / LN2)
    b = G.packed_colony(200, seed=1)
    st, res, nb, cons, ph, err = oracle.step(p, b)
    assert st == 0, err
    assert 1 <= res["constraint_iterations"] < 10
    assert res["max_overlap"] <= 1e-6
    assert np.all(cons["gamma"] >= 0)


# ----------------------------------------------------------------- soft model

def test_soft_two_rods_hertz_closed_form():
    p = reference_params(mode=MODE_SOFT, dt=1e-3 / 30, tolerance=1e-6, tau=1 / LN2,
                         kcc=4.0 / 3.0 * 1e4, lam=0.0)
    b = G.two_rods(0.95)
    st, res, nb, cons, ph, err = oracle.step(p, b)
    assert st == 0
    delta = 0.05
    F = math.sqrt(0.25 * 0.25 / 0.5) * p.kcc * delta ** 1.5
    assert nb.force[0] == pytest.approx([-F, 0, 0], rel=1e-12)
    assert nb.force[1] == pytest.approx([F, 0, 0], rel=1e-12)
    assert nb.stress == pytest.approx([0.5 * F, 0.5 * F], rel=1e-12)
    assert nb.pos[0, 0] == pytest.approx(-p.dt * F / 1.0, rel=1e-12)
    assert res["max_overlap"] == pytest.approx(delta, abs=1e-14)
    assert res["bbpgd_iterations"] == 0 and res["constraint_iterations"] == 0


def test_soft_momentum_balance():
    p = reference_params(mode=MODE_SOFT, dt=1e-3 / 30, tolerance=1e-6, tau=1 / LN2,
                         kcc=4.0 / 3.0 * 1e4)
    b = G.packed_colony(1000, seed=2)
    st, res, nb, cons, ph, err = oracle.step(p, b)
    assert st == 0 and res["n_constraints"] > 100
    F, T = nb.force, nb.torque
    scale = np.abs(F).max()
    assert scale > 0
    assert np.abs(F.sum(0)).max() <= 1e-9 * scale * b.n
    L = np.cross(b.pos, F).sum(0) + T.sum(0)
    assert np.abs(L).max() <= 1e-8 * scale * b.n
