"""Parity at the benchmark's full size (300,000 cells).

* Contact detection against the oracle, bit-exact (records, grid bucket and
  body permutation): the correlated colony of the headline benchmark and the
  i.i.d. disc at area fraction 0.9 of the configured 300k workload, at the
  phase-0 cutoffs 0.05 and 0.25.
* One ReLCP phase of the hard step against the oracle (relcp_tol = inf):
  - the dense colony at cutoff 0.05 (phase 0 converges in ~20 BBPGD
    iterations): the FP64 GPU step is bit-exact (state, impulses, gaps,
    constraints, iteration count, residual); the mixed step's impulses are
    within 1e-5 of the oracle's normwise, its body velocities within 1e-5;
  - config 2's i.i.d. disc at area fraction 0.9, cutoff 0.05: its phase-0
    LCP converges slowly (crossing rods; residual ~0.7 after 3,000
    iterations), so the test caps BBPGD at 100 iterations: both sides fail
    at the same iteration with the same residual, bit for bit, which pins
    the first 100 gradient evaluations and BB steps.
* Mixed precision against FP64 on the full 300k ReLCP step (the oracle needs
  minutes per phase-2 solve at this size): the same number of phases,
  overlap at or below the tolerance, positions within 1e-6 of the colony
  radius.
"""
import math

import numpy as np
import pytest

import oracle
from pcc import generators as G
from pcc.abi_types import PCC_E_BBPGD_NOT_CONVERGED, PREC_MIXED, reference_params

pytestmark = pytest.mark.gpu
LN2 = math.log(2.0)
N = 300_000


def rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


CASES = [("colony300k_c0.05", lambda: G.dense_colony(N, seed=0), 0.05),
         ("disc300k_c0.25", lambda: G.random_disc(N, seed=5), 0.25)]


@pytest.mark.parametrize("name,make,cutoff", CASES, ids=[c[0] for c in CASES])
def test_detect_bitexact_at_full_size(name, make, cutoff):
    from pcc.engine import ParticleManager
    b = make()
    print(f"[detect full] {name} n={b.n} cutoff={cutoff}", flush=True)
    pm = ParticleManager(reference_params())
    pm.set_state(b)
    got, bucket, perm = pm.detect(cutoff)
    want = oracle.detect(b, cutoff)
    g = oracle.grid(b)
    assert np.array_equal(bucket, g["bucket"])
    assert np.array_equal(perm, g["perm"])
    print(f"  contacts {len(want['gidI'])}", flush=True)
    assert len(got["gidI"]) == len(want["gidI"]) > N
    for k in want:
        assert np.array_equal(got[k], want[k]), k
    pm.close()


def step_both(params, b):
    from pcc.engine import ParticleManager
    out = []
    for prec in (0, PREC_MIXED):
        pm = ParticleManager(params.replace(precision=prec))
        pm.set_state(b)
        r = pm.step()
        out.append((r, pm.get_state(), pm.constraints(), pm.phase_stats()))
        pm.close()
    return out


STATE = ("pos", "quat", "length", "ldot", "force", "torque", "velocity", "omega", "impedance",
         "stress", "num_constraints")


def one_phase_params(**kw):
    base = dict(dt=1e-3, tolerance=1e-6, tau=1 / LN2, cutoff_phase0=0.05, cutoff_phaseK=0.05,
                relcp_tol=math.inf, temperature=0.0)
    base.update(kw)
    return reference_params(**base)


def test_fp64_and_mixed_single_phase_vs_oracle_dense_colony():
    import time
    from pcc.engine import ParticleManager
    b = G.dense_colony(N, seed=0)
    p = one_phase_params()
    t0 = time.time()
    st, res, want, wc, ph, err = oracle.step(p, b)
    print(f"[oracle 300k one phase] {res} in {time.time() - t0:.1f} s", flush=True)
    assert st == 0, err
    pm = ParticleManager(p)
    pm.set_state(b)
    r = pm.step()
    for k in res:
        assert r[k] == res[k], (k, r[k], res[k])
    got = pm.get_state()
    for f in STATE:
        assert np.array_equal(getattr(got, f), getattr(want, f)), f
    gc = pm.constraints()
    for k in wc:
        assert np.array_equal(gc[k], wc[k]), k
    pm.close()
    pm = ParticleManager(p.replace(precision=PREC_MIXED))
    pm.set_state(b)
    rm = pm.step()
    cm = pm.constraints()
    sm = pm.get_state()
    # phase 0's contacts (the set also holds the phase-1 detection batch,
    # whose geometry follows each path's own positions after phase 0)
    m0, g0 = wc["iteration"] == 0, cm["iteration"] == 0
    for k in ("gidI", "gidJ", "normal", "rPosI", "rPosJ", "contact_point", "gid"):
        assert np.array_equal(cm[k][g0], wc[k][m0]), k
    gm, gw = cm["gamma"][g0], wc["gamma"][m0]
    eg = float(np.linalg.norm(gm - gw) / np.linalg.norm(gw))
    egmax = rel(gm, gw)
    eu = rel(np.concatenate([sm.velocity, sm.omega], 1),
             np.concatenate([want.velocity, want.omega], 1))
    print(f"[mixed vs oracle, 300k one phase] C0={int(m0.sum())} its "
          f"oracle={res['bbpgd_iterations']} mixed={rm['bbpgd_iterations']} "
          f"gamma_err normwise={eg:.3e} max={egmax:.3e} U_err={eu:.3e}", flush=True)
    assert rm["residual"] <= p.bbpgd_tol
    assert eg <= 1e-5 and egmax <= 1e-5 and eu <= 1e-5
    pm.close()


def test_fp64_capped_phase_vs_oracle_iid_disc():
    import time
    from pcc.engine import ParticleManager
    b = G.random_disc(N, seed=5)
    p = one_phase_params(max_bbpgd_iterations=100)
    t0 = time.time()
    st, res, want, wc, ph, err = oracle.step(p, b)
    print(f"[oracle 300k disc, 100 its] status {st} {res} in {time.time() - t0:.1f} s",
          flush=True)
    assert st == PCC_E_BBPGD_NOT_CONVERGED
    pm = ParticleManager(p)
    pm.set_state(b)
    r = pm.step(raise_on_error=False)
    assert r["status"] == st
    for k in res:
        assert r[k] == res[k], (k, r[k], res[k])
    gc = pm.constraints()
    assert len(gc["gidI"]) == res["n_constraints"] > N
    for k in ("gidI", "gidJ", "signed_distance", "normal", "rPosI", "rPosJ", "contact_point",
              "stressI", "stressJ", "gamma", "gid", "iteration"):
        assert np.array_equal(gc[k], wc[k]), k
    pm.close()


def test_mixed_single_phase_at_full_size():
    b = G.dense_colony(N, seed=0)
    p = reference_params(dt=1e-3, tolerance=1e-6, tau=1 / LN2, cutoff_phase0=0.05,
                         cutoff_phaseK=0.05, relcp_tol=math.inf)
    (r64, s64, c64, _), (rm, sm, cm, _) = step_both(p, b)
    m0, g0 = c64["iteration"] == 0, cm["iteration"] == 0
    for k in ("gidI", "gidJ", "normal", "rPosI", "rPosJ", "contact_point", "gid"):
        assert np.array_equal(cm[k][g0], c64[k][m0]), k
    eg = rel(cm["gamma"][g0], c64["gamma"][m0])
    eu = rel(np.concatenate([sm.velocity, sm.omega], 1),
             np.concatenate([s64.velocity, s64.omega], 1))
    print(f"[mixed full, one phase] C={int(m0.sum())} its fp64={r64['bbpgd_iterations']} "
          f"mixed={rm['bbpgd_iterations']} gamma_err={eg:.3e} U_err={eu:.3e}", flush=True)
    assert rm["residual"] <= p.bbpgd_tol
    assert eg <= 1e-5 and eu <= 1e-5


def test_mixed_full_relcp_step_at_full_size():
    b = G.dense_colony(N, seed=0)
    p = reference_params(dt=1e-3, tolerance=1e-6, tau=1 / LN2, cutoff_phase0=0.05,
                         cutoff_phaseK=0.05)
    (r64, s64, c64, ph64), (rm, sm, cm, phm) = step_both(p, b)
    R = float(np.abs(b.pos[:, :2]).max())
    drift = float(np.abs(sm.pos - s64.pos).max())
    print(f"[mixed full, ReLCP] phases fp64={len(ph64)} mixed={len(phm)} "
          f"its fp64={r64['bbpgd_iterations']} mixed={rm['bbpgd_iterations']} "
          f"overlap fp64={r64['max_overlap']:.2e} mixed={rm['max_overlap']:.2e} "
          f"drift={drift:.3e} R={R:.1f}", flush=True)
    assert r64["constraint_iterations"] == rm["constraint_iterations"]
    assert rm["max_overlap"] <= p.relcp_tol and r64["max_overlap"] <= p.relcp_tol
    assert drift <= 1e-6 * R
    assert rel(sm.length, s64.length) <= 1e-7
