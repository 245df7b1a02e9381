"""GPU parity: libpcc (HIP, through the C ABI) against the FP64 oracle.

FP64 bar: bit-exact.  Contact lists and grid ordering are integer/index work;
the FP64 solver evaluates the same expressions in the same order with the same
reduction trees, so positions, quaternions, lengths, forces, velocities,
impulses and gaps must match bit for bit (the soft model's pow() is the one
libm call whose last bit may differ: 1e-14 relative there).
"""
import math

import numpy as np
import pytest

import oracle
from pcc import generators as G
from pcc.abi_types import (MODE_SOFT, PCC_E_BBPGD_NOT_CONVERGED, PCC_E_INVALID_MODE,
                           PCC_E_RELCP_MAX_ITERS, PREC_FP64, reference_params)

pytestmark = pytest.mark.gpu
LN2 = math.log(2.0)
STATE_FIELDS = ("pos", "quat", "length", "l0", "diameter", "ldot", "force", "torque",
                "velocity", "omega", "impedance", "stress", "num_constraints", "age")


def pm_for(params, bodies):
    from pcc.engine import ParticleManager
    pm = ParticleManager(params)
    pm.set_state(bodies)
    return pm


def random_3d(n, seed):
    rng = np.random.default_rng(seed)
    pos = rng.uniform(-4, 4, (n, 3))
    q = rng.normal(size=(n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    o = np.argsort(pos[:, 0], kind="stable")
    return G.Bodies(pos[o], q[o], rng.uniform(1, 2, n)[o])


def equal_x_lattice(cols=30, rows=40, seed=5):
    """Horizontal rods of length 2 stacked in columns (spacing 0.48 < d): the
    contacts of a column share one contact x, so the contact order rests on
    its (I, J) tie-break; gids are shuffled so that I order is not x order."""
    rng = np.random.default_rng(seed)
    x, y = np.meshgrid(3.0 * np.arange(cols), 0.48 * np.arange(rows), indexing="ij")
    n = cols * rows
    pos = np.zeros((n, 3))
    pos[:, 0] = x.ravel()
    pos[:, 1] = y.ravel()
    o = rng.permutation(n)
    q = np.tile([1.0, 0.0, 0.0, 0.0], (n, 1))
    return G.Bodies(pos[o], q, np.full(n, 2.0))


DETECT_CASES = [
    ("disc2000_c0.05", lambda: G.random_disc(2000, seed=0), 0.05),
    ("equal_x_lattice", lambda: equal_x_lattice(), 0.05),
    ("disc2000_c0.25", lambda: G.random_disc(2000, seed=1), 0.25),
    ("disc1500_c0", lambda: G.random_disc(1500, seed=2), 0.0),
    ("colony5000", lambda: G.packed_colony(5000, seed=3), 0.05),
    ("random3d", lambda: random_3d(600, 4), 0.05),
    ("two_rods", lambda: G.two_rods(0.9), 0.05),
    ("single", lambda: G.single_cell(), 0.05),
]


@pytest.mark.parametrize("name,make,cutoff", DETECT_CASES, ids=[c[0] for c in DETECT_CASES])
def test_detect_bitexact(name, make, cutoff):
    b = make()
    print(f"[detect] {name} n={b.n} cutoff={cutoff}", flush=True)
    pm = pm_for(reference_params(), b)
    got, bucket, perm = pm.detect(cutoff)
    want = oracle.detect(b, cutoff)
    g = oracle.grid(b)
    assert np.array_equal(bucket, g["bucket"])
    assert np.array_equal(perm, g["perm"])
    assert len(got["gidI"]) == len(want["gidI"])
    for k in want:
        assert np.array_equal(got[k], want[k]), k
    pm.close()


def test_detect_bitexact_with_crowded_cells():
    """Grid cells of size 12 hold up to ~200 bodies of a dense packing: the
    writing pass walks buckets of more than 64 bodies again and uses the
    counting pass's mask of accepted positions for the others.  The list is
    bit-exact against the oracle on the same grid, and equal to the list of
    the default grid (the contacts do not depend on the cell size)."""
    b = G.dense_colony(3000, seed=17)
    h = 12.0
    g = oracle.grid(b, h=h)
    counts = np.bincount(g["bucket"])
    assert counts.max() > 64 and ((counts > 0) & (counts <= 64)).any()
    print(f"[detect crowded] n={b.n} h={h} largest bucket {counts.max()}", flush=True)
    pm = pm_for(reference_params(linked_cell_size=h), b)
    got, bucket, perm = pm.detect(0.05)
    want = oracle.detect(b, 0.05, h=h)
    assert np.array_equal(bucket, g["bucket"]) and np.array_equal(perm, g["perm"])
    default = oracle.detect(b, 0.05)
    assert len(got["gidI"]) == len(want["gidI"]) > b.n
    for k in want:
        assert np.array_equal(got[k], want[k]), k
        assert np.array_equal(got[k], default[k]), k
    pm.close()


def assert_same_state(a, b, rtol=0.0):
    for f in STATE_FIELDS:
        x, y = getattr(a, f), getattr(b, f)
        if rtol == 0.0 or f == "num_constraints":
            assert np.array_equal(x, y), f
        else:
            assert np.allclose(x, y, rtol=rtol, atol=rtol * max(1.0, np.abs(y).max())), f


def assert_same_constraints(a, b, rtol=0.0):
    assert set(a) == set(b)
    for k in b:
        if rtol == 0.0 or a[k].dtype.kind == "i":
            assert np.array_equal(a[k], b[k]), k
        else:
            assert np.allclose(a[k], b[k], rtol=rtol, atol=rtol * max(1.0, np.abs(b[k]).max())), k


def hard(**kw):
    base = dict(dt=1e-3, tolerance=1e-6, tau=1 / LN2, cutoff_phase0=0.05, cutoff_phaseK=0.05,
                relcp_tol=math.inf)
    base.update(kw)
    return reference_params(**base)


STEP_CASES = [
    ("two_rods", lambda: G.two_rods(0.9), lambda: hard()),
    ("t_config", lambda: G.Bodies(np.array([[0, 0, 0], [0.4, 0.7, 0.0]]),
                                  np.array([[1, 0, 0, 0], [math.cos(math.pi / 4), 0, 0,
                                                           math.sin(math.pi / 4)]]),
                                  np.array([2.0, 1.0])), lambda: hard()),
    ("disc600_one_phase", lambda: G.random_disc(600, seed=5), lambda: hard()),
    ("colony1500_relcp", lambda: G.packed_colony(1500, seed=6),
     lambda: reference_params(dt=1e-3, tolerance=1e-6, tau=1 / LN2)),
    ("random3d_relcp", lambda: random_3d(300, 7),
     lambda: hard(relcp_tol=1e-3, max_recursive_iterations=8)),
    ("empty", lambda: G.Bodies(np.zeros((0, 3)), np.zeros((0, 4)), np.zeros(0)), lambda: hard()),
    ("single_cell", lambda: G.single_cell(), lambda: reference_params(dt=1e-3, tolerance=1e-6)),
]


@pytest.mark.parametrize("name,make,params", STEP_CASES, ids=[c[0] for c in STEP_CASES])
def test_hard_step_fp64_bitexact(name, make, params):
    b = make()
    p = params()
    print(f"[step-fp64] {name} n={b.n}", flush=True)
    st, res, want_b, want_c, ph, err = oracle.step(p, b)
    pm = pm_for(p, b)
    r = pm.step(0, raise_on_error=False)
    assert r["status"] == st, (r, err)
    for k in res:
        assert r[k] == res[k], (k, r[k], res[k])
    got_b = pm.get_state()
    assert_same_state(got_b, want_b)
    assert_same_constraints(pm.constraints(), want_c)
    pm.close()


def test_growth_sequence_fp64_bitexact():
    """Config-1 style growth of a 40-cell colony: 60 steps with host division
    between steps, the state uploaded before each step; compared every step."""
    p = reference_params(dt=1e-3, tolerance=1e-6, tau=1 / LN2)
    b = G.packed_colony(40, seed=8)
    from pcc.engine import ParticleManager
    pm = ParticleManager(p)
    for k in range(60):
        st, res, want_b, want_c, ph, err = oracle.step(p, b, k)
        pm.set_state(b)
        pm.set_step_count(k)  # the Brownian key the oracle got (set_state resets it)
        r = pm.step(k)
        assert r["bbpgd_iterations"] == res["bbpgd_iterations"]
        got_b = pm.get_state()
        assert_same_state(got_b, want_b)
        b = G.divide(want_b, 0, k)
    pm.close()


def test_soft_step_fp64_matches():
    p = reference_params(mode=MODE_SOFT, dt=1e-3 / 30, tolerance=1e-6, tau=1 / LN2,
                         kcc=4.0 / 3.0 * 1e4)
    b = G.packed_colony(3000, seed=9)
    from pcc.engine import ParticleManager
    pm = ParticleManager(p)
    pm.set_state(b)
    for k in range(10):
        st, res, want_b, want_c, ph, err = oracle.step(p, b, k)
        r = pm.step(k)
        assert r["n_constraints"] == res["n_constraints"]
        got_b = pm.get_state()
        assert_same_state(got_b, want_b, rtol=1e-13)
        assert_same_constraints(pm.constraints(), want_c, rtol=1e-13)
        b = want_b
        pm.set_state(b)
        pm.set_step_count(k + 1)
    pm.close()


def test_failure_modes_on_gpu():
    from pcc.engine import ParticleManager, PccError
    b = G.two_rods(0.9)
    pm = pm_for(hard(max_bbpgd_iterations=5), b)
    with pytest.raises(PccError) as e:
        pm.step()
    assert e.value.status == PCC_E_BBPGD_NOT_CONVERGED and str(e.value) == "BBPGD did not converge"
    pm.close()
    p = hard(max_recursive_iterations=1, relcp_tol=1e-6)
    st, res, want_b, *_ = oracle.step(p, b)
    pm = pm_for(p, b)
    with pytest.raises(PccError) as e:
        pm.step()
    assert e.value.status == PCC_E_RELCP_MAX_ITERS
    assert str(e.value) == "Maximum number of constraint iterations reached"
    assert_same_state(pm.get_state(), want_b)  # advanced and grown, as the reference
    pm.close()
    with pytest.raises(PccError) as e:
        ParticleManager(hard(mode=5))
    assert e.value.status == PCC_E_INVALID_MODE


def test_repeat_runs_bitwise_identical():
    """Deterministic reductions, no float atomics: two runs give the same bits
    (the race detector of the build)."""
    p = hard(precision=1)
    b = G.random_disc(3000, seed=10)
    pm = pm_for(p, b)
    pm.snapshot()
    r1 = pm.step()
    s1 = pm.get_state()
    c1 = pm.constraints()
    pm.restore()
    r2 = pm.step()
    assert r1 == r2
    assert_same_state(pm.get_state(), s1)
    assert_same_constraints(pm.constraints(), c1)
    pm.close()


@pytest.mark.parametrize("precision", [0, 1], ids=["fp64", "mixed"])
def test_planar_kernels_equal_3d_kernels(precision, monkeypatch):
    """The 2D monolayer kernels skip terms that are exact zeros: the results
    equal the 3D kernels' (forced with PCC_PLANAR=0) value for value."""
    from pcc.engine import ParticleManager
    p = reference_params(dt=1e-3, tolerance=1e-6, tau=1 / LN2, cutoff_phase0=0.05,
                         cutoff_phaseK=0.05, precision=precision)
    b = G.dense_colony(4000, seed=15)
    pm2 = ParticleManager(p)
    pm2.set_state(b)
    r2 = pm2.step()
    monkeypatch.setenv("PCC_PLANAR", "0")
    pm3 = ParticleManager(p)
    pm3.set_state(b)
    r3 = pm3.step()
    assert r2 == r3
    assert_same_state(pm2.get_state(), pm3.get_state())
    assert_same_constraints(pm2.constraints(), pm3.constraints())
    pm2.close()
    pm3.close()
