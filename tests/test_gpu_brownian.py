"""GPU parity of the Brownian input (temperature > 0) against the FP64 oracle.

calculate_external_velocities (Physics.cpp:59-81) with
Particle::calculateBrownianVelocity (Particle.cpp:174-190): both sides draw
from the same counter-based generator, so the FP64 step is bit-exact, in the
planar monolayer kernels, in the 3D kernels (monolayer off turns bodies out
of the plane), through device-resident steps with division, in a batch, and
in the soft model (rtol 1e-13: its pow()).
"""
import math

import numpy as np
import pytest

import oracle
from pcc import generators as G
from pcc.abi_types import MODE_SOFT, reference_params

pytestmark = pytest.mark.gpu
LN2 = math.log(2.0)
FIELDS = ("pos", "quat", "length", "ldot", "force", "torque", "velocity", "omega",
          "impedance", "stress", "num_constraints", "age")


def same(got, want, rtol=0.0):
    for f in FIELDS:
        a, b = getattr(got, f), getattr(want, f)
        if rtol == 0.0:
            assert np.array_equal(a, b), f
        else:
            assert np.allclose(a, b, rtol=rtol, atol=1e-300), f


def hard(**kw):
    base = dict(dt=1e-3, tolerance=1e-6, tau=1 / LN2, temperature=0.05, cutoff_phase0=0.05,
                cutoff_phaseK=0.05, brownian_seed=11)
    base.update(kw)
    return reference_params(**base)


def colony(n, seed):
    b = G.packed_colony(n, seed=seed)
    b.impedance[:] = 0.5 + 0.5 * G.uniform(seed, 77, n)  # as if stepped before
    return b


@pytest.mark.parametrize("mono", [1, 0], ids=["monolayer", "3d"])
def test_hard_brownian_steps_bitexact(mono):
    from pcc.engine import ParticleManager
    p = hard(monolayer=mono)
    b = colony(700, 3)
    print(f"[brownian-hard] monolayer={mono} n={b.n}", flush=True)
    pm = ParticleManager(p)
    pm.set_state(b)
    want = b
    for k in range(4):
        st, res, want, wc, ph, err = oracle.step(p, want, k)
        r = pm.step(k)
        assert r["status"] == st == 0, err
        assert r["bbpgd_iterations"] == res["bbpgd_iterations"]
        same(pm.get_state(), want)
    if not mono:
        assert np.any(want.pos[:, 2] != 0.0)  # left the plane: the 3D kernels ran
    pm.close()


def test_hard_brownian_with_device_division():
    from pcc.engine import ParticleManager
    p = hard(tau=0.02)  # fast growth: cells divide within the first steps
    b = colony(300, 5)
    pm = ParticleManager(p)
    pm.set_state(b)
    want = b
    for k in range(4):
        n = pm.divide(seed=21, step=k)
        want = oracle.divide(want, seed=21, step=k)
        assert n == want.n
        st, res, want, *_ = oracle.step(p, want, k)
        r = pm.step(k)
        assert r["status"] == st == 0
        got = pm.get_state()
        same(got, want)
    assert want.n > 300  # cells divided; new cells started from impedance 0
    pm.close()


def test_batch_brownian_equals_alone():
    from pcc.engine import ColonyBatch, ParticleManager
    cols = [colony(n, 40 + i) for i, n in enumerate([250, 900, 60])]
    ps = [hard(brownian_seed=100 + i, temperature=0.02 * (i + 1)) for i in range(3)]
    cb = ColonyBatch(ps[0])
    cb.set_state(cols, ps)
    for k in range(2):
        cb.step(k)
    for q in range(3):
        pm = ParticleManager(ps[q])
        pm.set_state(cols[q])
        for k in range(2):
            pm.step(k)
        same(cb.get_state(q), pm.get_state())
        pm.close()
    cb.close()


def test_soft_brownian_matches():
    from pcc.engine import ParticleManager
    p = reference_params(mode=MODE_SOFT, dt=1e-3 / 30, tolerance=1e-6, tau=1 / LN2,
                         kcc=4.0 / 3.0 * 1e4, temperature=0.3, monolayer=0, brownian_seed=5)
    b = colony(1200, 9)
    pm = ParticleManager(p)
    pm.set_state(b)
    want = b
    for k in range(5):
        st, res, want, *_ = oracle.step(p, want, k)  # k: the colony's step count
        pm.step(k)
        same(pm.get_state(), want, rtol=1e-11)  # pow()'s last bit may differ
    assert pm.step_count() == 5
    pm.close()


def test_step_count_keys_and_snapshot():
    """The draws follow the colony's own step count: a colony added to a
    batch that has already stepped gets the draws it gets alone; the device
    snapshot keeps the count, and setting it resumes a run."""
    from pcc.engine import ColonyBatch, ParticleManager
    p = hard()
    a, c = colony(400, 61), colony(350, 62)
    alone = ParticleManager(p)
    alone.set_state(c)
    alone.step(0)
    alone.step(1)
    want = alone.get_state()
    assert alone.step_count() == 2
    alone.close()
    cb = ColonyBatch(p)
    cb.set_state([a], [p])
    cb.step(0)
    q = cb.add(c, p)
    cb.step(7)  # the call's step index does not key the draws
    cb.step(8)
    same(cb.get_state(q), want)
    assert cb.step_count(0) == 3 and cb.step_count(q) == 2
    cb.close()
    # snapshot / restore and set_step_count
    pm = ParticleManager(p)
    pm.set_state(c)
    pm.step(0)
    pm.snapshot()
    pm.step(1)
    first = pm.get_state()
    pm.restore()
    assert pm.step_count() == 1
    pm.step(1)
    same(pm.get_state(), first)
    pm.set_state(c)
    pm.step(0)
    mid = pm.get_state()
    pm.close()
    pm = ParticleManager(p)
    pm.set_state(mid)
    pm.set_step_count(1)
    pm.step(1)
    same(pm.get_state(), want)
    pm.close()
