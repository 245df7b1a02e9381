"""Parity at the benchmark's full size (300,000 cells).

* Contact detection against the oracle, bit-exact (records, grid bucket and
  body permutation): the correlated colony of the headline benchmark and the
  i.i.d. disc at area fraction 0.9 of the configured 300k workload, at the
  phase-0 cutoffs 0.05 and 0.25.
* Mixed precision against FP64 on the same 300k hard step.  The FP64 GPU
  path is the reference here: it is bit-identical to the oracle on every case
  the oracle finishes in seconds, and the oracle needs minutes per step at
  this size.  One ReLCP phase: impulses within 1e-5 normwise, body velocities
  within 1e-5, phase-0 contacts bit-exact.  The full ReLCP step: the same
  number of phases, overlap at or below the tolerance, positions within
  1e-6 of the colony radius.
"""
import math

import numpy as np
import pytest

import oracle
from pcc import generators as G
from pcc.abi_types import PREC_MIXED, reference_params

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
