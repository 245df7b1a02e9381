"""Mixed precision (FP32 operator storage, FP64 iterates/accumulation/
reductions/integration) against the FP64 oracle.

Tolerances (normwise, ||x - x_ref||_inf / ||x_ref||_inf):
  impulses gamma             <= 1e-5   (the target for the solve)
  body velocities D^T gamma  <= 1e-5   (unique even where gamma is not)
  residual overlap           <= the BBPGD tolerance (1e-6) in the linearised gap
  position drift, 100 steps  <= 1e-6 * colony radius (hard), 1e-6 (soft)
  lengths after 100 steps    <= 1e-7 relative
The contact list of ReLCP phase 0 (same input state) stays bit-exact.
"""
import math

import numpy as np
import pytest

import oracle
from pcc import generators as G
from pcc.abi_types import MODE_SOFT, PREC_MIXED, reference_params

pytestmark = pytest.mark.gpu
LN2 = math.log(2.0)


def rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def one_phase(**kw):
    base = dict(dt=1e-3, tolerance=1e-6, tau=1 / LN2, cutoff_phase0=0.05, cutoff_phaseK=0.05,
                relcp_tol=math.inf)
    base.update(kw)
    return reference_params(**base)


@pytest.mark.parametrize("gen,n,seed", [("random_disc", 3000, 10), ("dense_colony", 20000, 11)])
def test_mixed_single_phase_against_fp64_oracle(gen, n, seed):
    from pcc.engine import ParticleManager
    b = getattr(G, gen)(n, seed=seed)
    p64 = one_phase()
    st, res, want, cons, ph, err = oracle.step(p64, b)
    pm = ParticleManager(p64.replace(precision=PREC_MIXED))
    pm.set_state(b)
    r = pm.step()
    got = pm.get_state()
    gc = pm.constraints()
    # phase 0 ran on identical inputs: its contacts are bit-exact; the phase-1
    # detection that ends the ReLCP loop sees the moved bodies
    m0 = cons["iteration"] == 0
    g0 = gc["iteration"] == 0
    for k in ("gidI", "gidJ", "normal", "rPosI", "rPosJ", "contact_point", "stressI", "stressJ",
              "gid"):
        assert np.array_equal(gc[k][g0], cons[k][m0]), k
    eg = rel(gc["gamma"][g0], cons["gamma"][m0])
    eu = rel(np.concatenate([got.velocity, got.omega], 1),
             np.concatenate([want.velocity, want.omega], 1))
    print(f"[mixed n={n}] iters fp64={res['bbpgd_iterations']} mixed={r['bbpgd_iterations']} "
          f"gamma_err={eg:.3e} U_err={eu:.3e} max|gamma|={np.abs(cons['gamma']).max():.3e}")
    assert r["residual"] <= p64.bbpgd_tol
    assert eg <= 1e-5
    assert eu <= 1e-5
    # residual overlap after the step: the linearised gaps stored back satisfy
    # g >= -tol (projected-gradient residual <= tol)
    assert (-gc["signed_distance"][g0]).max() <= p64.bbpgd_tol
    pm.close()


def test_mixed_relcp_colony_positions_and_contacts():
    from pcc.engine import ParticleManager
    p = reference_params(dt=1e-3, tolerance=1e-6, tau=1 / LN2)
    b = G.packed_colony(3000, seed=12)
    st, res, want, cons, ph, err = oracle.step(p, b)
    pm = ParticleManager(p.replace(precision=PREC_MIXED))
    pm.set_state(b)
    r = pm.step()
    got = pm.get_state()
    gc = pm.constraints()
    ph0 = cons["iteration"] == 0
    gph0 = gc["iteration"] == 0
    assert np.array_equal(gc["gidI"][gph0], cons["gidI"][ph0])
    assert np.array_equal(gc["gidJ"][gph0], cons["gidJ"][ph0])
    assert r["max_overlap"] <= p.relcp_tol
    R = G.disc_radius(b)
    drift = np.abs(got.pos - want.pos).max()
    print(f"[mixed relcp] phases fp64={res['constraint_iterations']} "
          f"mixed={r['constraint_iterations']} drift={drift:.3e}")
    assert drift <= 1e-6 * R
    pm.close()


def test_mixed_hard_growth_100_steps_drift():
    from pcc.engine import ParticleManager
    p = reference_params(dt=1e-3, tolerance=1e-6, tau=1 / LN2)
    b0 = G.packed_colony(400, seed=13)
    pm = ParticleManager(p.replace(precision=PREC_MIXED))
    pm.set_state(b0)
    bo = b0
    for k in range(100):
        st, res, bo, *_ = oracle.step(p, bo, k)
        pm.step(k)
    got = pm.get_state()
    R = G.disc_radius(b0)
    drift = np.abs(got.pos - bo.pos).max()
    dq = np.abs(got.quat - bo.quat).max()
    print(f"[mixed hard 100 steps] drift={drift:.3e} quat={dq:.3e} R={R:.1f}")
    assert drift <= 1e-6 * R
    # growth couples to gamma through sigma = L^T gamma: lengths drift ~1e-8
    assert np.allclose(got.length, bo.length, rtol=1e-7)
    pm.close()


def test_mixed_soft_100_steps_drift():
    from pcc.engine import ParticleManager
    p = reference_params(mode=MODE_SOFT, dt=1e-3 / 30, tolerance=1e-6, tau=1 / LN2,
                         kcc=4.0 / 3.0 * 1e4)
    b0 = G.packed_colony(3000, seed=14)
    pm = ParticleManager(p.replace(precision=PREC_MIXED))
    pm.set_state(b0)
    bo = b0
    for k in range(100):
        st, res, bo, *_ = oracle.step(p, bo, k)
        pm.step(k)
    got = pm.get_state()
    drift = np.abs(got.pos - bo.pos).max()
    print(f"[mixed soft 100 steps] drift={drift:.3e}")
    assert drift <= 1e-6
    pm.close()
