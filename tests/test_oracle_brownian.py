"""Pins the oracle's Brownian input (temperature T > 0):
calculate_external_velocities (code/cpp/src/dynamics/Physics.cpp:59-81) with
Particle::calculateBrownianVelocity (code/cpp/src/simulation/Particle.cpp:174-190).

The reference draws from an unseeded global std::mt19937 (Physics.cpp:19), so
no value of it can be matched.  The checks are what the formula
omega_k = ((T u1) u_k) iota iota iota, u ~ U[-1, 1), fixes for any generator,
each chosen so that a plausible slip fails one of them:
  * |omega| <= T iota^3, and omega = 0 exactly for iota = 0 (a cell that was
    never stepped, Particle.cpp:25) and for T = 0;
  * monolayer (the reference default, CLIParser.h:183): omega_x = omega_y = 0;
  * moments over many isolated cells: mean 0, E[omega^2] = T^2 iota^6 / 9 and
    E[omega^4] = T^4 iota^12 / 25 (a product of two uniforms; a missing u1, an
    iota^2 or a T^2 moves them);
  * off the monolayer the three components share u1, so
    Cov(omega_x^2, omega_y^2) = T^4 iota^12 (1/5 - 1/9) / 9 > 0 (one factor
    drawn per component gives 0);
  * the velocity is U_ext of phase 0 (Hard.cpp:140-144): an isolated cell
    turns by exactly dt omega / 2 before renormalisation (Particle.cpp:51-57),
    its centre does not move, and a later ReLCP phase adds nothing;
  * the soft model adds it on every step (Soft.cpp:94);
  * draws are a function of (seed, step index, body index): repeatable,
    different for another seed or step.
"""
import math

import numpy as np
import pytest

import oracle
from pcc import generators as G
from pcc.abi_types import MODE_SOFT, PCC_E_RELCP_MAX_ITERS, Bodies, reference_params

LN2 = math.log(2.0)
T = 0.7
IOTA = 0.9


def isolated_cells(n, iota=IOTA, spacing=5.0):
    """n cells on a square lattice far apart (no contacts), impedance iota."""
    side = int(math.ceil(math.sqrt(n)))
    k = np.arange(n)
    pos = np.zeros((n, 3))
    pos[:, 0] = (k % side) * spacing
    pos[:, 1] = (k // side) * spacing
    order = np.argsort(pos[:, 0], kind="stable")
    quat = np.zeros((n, 4))
    quat[:, 0] = 1.0
    return Bodies(pos[order], quat, np.ones(n), impedance=np.full(n, iota))


def hard(**kw):
    base = dict(dt=1e-3, tolerance=1e-6, tau=1 / LN2, temperature=T, cutoff_phase0=0.05,
                cutoff_phaseK=0.05)
    base.update(kw)
    return reference_params(**base)


def omega_after_step(params, b, iter_=0):
    st, res, nb, cons, ph, err = oracle.step(params, b, iter_)
    assert st == 0, err
    assert res["n_constraints"] == 0
    return nb, nb.omega


def test_bounds_and_zero_cases():
    b = isolated_cells(400)
    nb, w = omega_after_step(hard(), b)
    bound = T * IOTA ** 3
    assert np.all(np.abs(w[:, 2]) <= bound)
    assert np.abs(w[:, 2]).max() > 0.9 * bound  # the bound is reached, not a smaller one
    assert np.all(w[:, :2] == 0.0)  # monolayer
    # a cell never stepped has impedance 0: no Brownian velocity at all
    nb0, w0 = omega_after_step(hard(), isolated_cells(50, iota=0.0))
    assert np.all(w0 == 0.0)
    assert np.all(nb0.quat == b.quat[:50])
    # T = 0: none either
    nbt, wt = omega_after_step(hard(temperature=0.0), b)
    assert np.all(wt == 0.0)
    # no translation in any case
    assert np.all(nb.pos == b.pos)


def test_moments_monolayer():
    n = 20000
    _, w = omega_after_step(hard(), isolated_cells(n))
    wz = w[:, 2]
    s2 = T ** 2 * IOTA ** 6
    assert abs(wz.mean()) < 5 * math.sqrt(s2 / 9 / n)
    assert np.mean(wz ** 2) == pytest.approx(s2 / 9, rel=0.05)
    assert np.mean(wz ** 4) == pytest.approx(s2 ** 2 / 25, rel=0.12)


def test_shared_factor_off_monolayer():
    n = 20000
    _, w = omega_after_step(hard(monolayer=0), isolated_cells(n))
    s2 = T ** 2 * IOTA ** 6
    for k in range(3):
        assert np.mean(w[:, k] ** 2) == pytest.approx(s2 / 9, rel=0.05)
    cov = np.mean(w[:, 0] ** 2 * w[:, 1] ** 2) - np.mean(w[:, 0] ** 2) * np.mean(w[:, 1] ** 2)
    want = s2 ** 2 * (1 / 5 - 1 / 9) / 9
    assert cov == pytest.approx(want, rel=0.25)
    # the components are uncorrelated (independent u_x, u_y)
    assert abs(np.mean(w[:, 0] * w[:, 1])) < 5 * (s2 / 9) / math.sqrt(n)


def test_turn_of_an_isolated_cell():
    p = hard(monolayer=0)
    b = isolated_cells(64)
    nb, w = omega_after_step(p, b)
    # q = (1, 0, 0, 0): dC = G U gives (0, wx/2, wy/2, wz/2) (Physics.cpp:220-265)
    q = np.stack([np.ones(b.n), p.dt * w[:, 0] / 2, p.dt * w[:, 1] / 2, p.dt * w[:, 2] / 2], 1)
    q /= np.linalg.norm(q, axis=1)[:, None]
    assert np.allclose(nb.quat, q, rtol=0, atol=1e-15)
    assert np.any(nb.quat[:, 1:] != 0.0)


def test_phase_zero_only():
    """With contacts present the step runs a second ReLCP phase; the final
    velocity is that phase's, which carries no U_ext: isolated cells of the
    same run then have omega = 0."""
    pair = G.packed_colony(100, seed=1)  # overlapping cells: phases 0 and 1 solve
    cells = isolated_cells(16)
    cells.pos[:, 0] += pair.pos[:, 0].max() + 50.0
    pos = np.concatenate([pair.pos, cells.pos])
    quat = np.concatenate([pair.quat, cells.quat])
    b = Bodies(pos, quat, np.concatenate([pair.length, cells.length]),
               impedance=np.full(pair.n + cells.n, IOTA))
    # an overlap test stricter than the BBPGD tolerance forces a second phase;
    # two phases are allowed, so the step ends with the reference's
    # max-iterations error after both phases have moved the bodies
    p = hard(relcp_tol=1e-15, max_recursive_iterations=2)
    st, res, nb, cons, ph, err = oracle.step(p, b)
    assert st == PCC_E_RELCP_MAX_ITERS, err
    assert res["constraint_iterations"] == 2
    m = pair.n
    assert np.all(nb.omega[m:] == 0.0)  # last phase: no Brownian term
    # but phase 0 turned them
    assert np.any(nb.quat[m:, 3] != 0.0)


def test_soft_every_step_and_keys():
    p = reference_params(mode=MODE_SOFT, dt=1e-3 / 30, tolerance=1e-6, tau=1 / LN2,
                         temperature=T, kcc=4.0 / 3.0 * 1e4)
    b = isolated_cells(256)
    st, res, nb, *_ = oracle.step(p, b, 3)
    assert st == 0
    w3 = nb.omega.copy()
    assert np.all(np.abs(w3[:, 2]) <= T * IOTA ** 3) and np.any(w3[:, 2] != 0)
    # the next step draws again, with the impedance the step wrote (exp(0) = 1)
    assert np.all(nb.impedance == 1.0)
    st, res, nb2, *_ = oracle.step(p, nb, 4)
    w4 = nb2.omega
    assert np.all(np.abs(w4[:, 2]) <= T) and np.any(np.abs(w4[:, 2]) > T * IOTA ** 3)
    # repeatable; another step index or seed gives other draws
    _, _, again, *_ = oracle.step(p, b, 3)
    assert np.array_equal(again.omega, w3)
    _, _, other, *_ = oracle.step(p, b, 5)
    assert not np.array_equal(other.omega, w3)
    _, _, seeded, *_ = oracle.step(p.replace(brownian_seed=99), b, 3)
    assert not np.array_equal(seeded.omega, w3)


def test_division_keeps_parent_impedance():
    b = G.single_cell(length=2.0)
    b.impedance[:] = 0.5
    d = oracle.divide(b, seed=1, step=0)
    assert d.n == 2
    assert sorted(d.impedance.tolist()) == [0.0, 0.5]
