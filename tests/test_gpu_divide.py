"""Device-side division and gid reassignment (pcc_divide / pcc_batch_divide)
against the oracle's restatement: bit-exact, alone and batched, and a growth
run that stays on the device from one cell, stepped and divided in turn,
compared with the oracle at every step."""
import math

import numpy as np
import pytest

import oracle
from pcc import generators as G
from pcc.abi_types import reference_params

pytestmark = pytest.mark.gpu
LN2 = math.log(2.0)
FIELDS = ("pos", "quat", "length", "l0", "diameter", "ldot", "age")


def same_bodies(a, b):
    assert a.n == b.n
    for f in FIELDS:
        assert np.array_equal(getattr(a, f), getattr(b, f)), f


def with_dividers(b, every=3):
    """Every `every`-th cell at its division length; generations 1..5 so
    that both the carried and the incremented ages are checked."""
    b = b.copy()
    b.length[::every] = 2 * b.l0[::every]
    b.age = (1 + np.arange(b.n) % 5).astype(np.int32)
    return b


@pytest.mark.parametrize("length_noise,angle_noise", [(0.0, 0.01), (0.02, 0.01), (0.02, 0.0)])
def test_divide_matches_oracle(length_noise, angle_noise):
    from pcc.engine import ParticleManager
    b = with_dividers(G.dense_colony(3000, seed=140))
    print(f"[divide] n={b.n} noise=({length_noise}, {angle_noise})", flush=True)
    pm = ParticleManager(reference_params())
    pm.set_state(b)
    n = pm.divide(seed=77, step=12, length_noise=length_noise, angle_noise=angle_noise)
    want = oracle.divide(b, 77, 12, length_noise, angle_noise)
    assert n == want.n == b.n + len(b.length[::3])
    same_bodies(pm.get_state(), want)
    pm.close()


def test_divide_ties_keep_staging_order():
    """A column of vertical rods at one x, all at their division length: both
    children of every cell keep the parent's x, so the new gids of the whole
    column rest on the tie-break by staging order."""
    from pcc.engine import ParticleManager
    n = 200
    pos = np.zeros((n, 3))
    pos[:, 0] = 1.25
    pos[:, 1] = 2.1 * np.arange(n)
    h = math.sqrt(0.5)
    q = np.tile([h, 0.0, 0.0, h], (n, 1))  # turned by pi/2: along y
    b = G.Bodies(pos, q, np.full(n, 2.0))
    b.age = (1 + np.arange(n) % 3).astype(np.int32)
    pm = ParticleManager(reference_params())
    pm.set_state(b)
    got_n = pm.divide(seed=3, step=1, length_noise=0.0, angle_noise=0.0)
    want = oracle.divide(b, 3, 1, 0.0, 0.0)
    assert got_n == want.n == 2 * n
    assert np.unique(want.pos[:, 0]).size < 4  # ties everywhere
    same_bodies(pm.get_state(), want)
    pm.close()


def test_batch_divide_equals_alone():
    from pcc.engine import ColonyBatch
    cols = [with_dividers(G.dense_colony(n, seed=150 + i), every=2 + i)
            for i, n in enumerate([800, 2500, 60])]
    seeds = [5, 6, 7]
    p = reference_params()
    cb = ColonyBatch(p)
    cb.set_state(cols, [p] * 3)
    sizes = cb.divide(seeds, step=4, length_noise=0.02, angle_noise=0.01)
    for q in range(3):
        want = oracle.divide(cols[q], seeds[q], 4, 0.02, 0.01)
        assert sizes[q] == want.n
        same_bodies(cb.get_state(q), want)
    cb.close()


def test_device_resident_growth_from_one_cell():
    """Config-1 style: one cell grows at rate ln2, divides at 2 l0 (exact
    halves, N(0, 0.01) turns), FP64 hard model; step and divide alternate on
    the device and the state equals the oracle's after every step."""
    from pcc.engine import ParticleManager
    p = reference_params(dt=1e-2, tolerance=1e-6, tau=1 / LN2, cutoff_phase0=0.05,
                         cutoff_phaseK=0.05)
    b = G.single_cell()
    pm = ParticleManager(p)
    pm.set_state(b)
    want = b
    k = 0
    while want.n < 40 and k < 800:
        st, res, want, _, _, err = oracle.step(p, want, k)
        r = pm.step(k)
        assert r["bbpgd_iterations"] == res["bbpgd_iterations"], k
        want = oracle.divide(want, 0, k, 0.0, 0.01)
        n = pm.divide(0, k, 0.0, 0.01)
        assert n == want.n, k
        k += 1
    print(f"[growth] {k} steps -> {want.n} cells", flush=True)
    assert want.n >= 40
    same_bodies(pm.get_state(), want)
    pm.close()


def expected_dt(state, dt, cfl_factor, tol, alpha):
    """Domain::adaptDt (Domain.cpp:69-118) restated with numpy: median of
    |v| + ldot (mean of the two middle values for an even count), smoothed
    CFL step, clamped to [0.9 dt, 1.1 dt]."""
    u = state.velocity
    v = np.sqrt(u[:, 0] * u[:, 0] + u[:, 1] * u[:, 1] + u[:, 2] * u[:, 2]) + state.ldot
    s = np.sort(v)
    n = len(s)
    med = float(s[n // 2]) if n % 2 else 0.5 * (float(s[n // 2 - 1]) + float(s[n // 2]))
    if med == 0.0:
        return dt, med
    nd = dt * (1.0 - alpha) + alpha * ((cfl_factor * tol) / med)
    lo, hi = 0.9 * dt, 1.1 * dt
    return (lo if nd < lo else hi if hi < nd else nd), med


@pytest.mark.parametrize("n,tol", [(2000, 1e-3), (2001, 1e-6), (1, 1e-2)])
def test_adapt_dt_matches_reference_rule(n, tol):
    from pcc.engine import ParticleManager
    p = reference_params(dt=1e-3, tolerance=tol, tau=1 / LN2, cutoff_phase0=0.05,
                         cutoff_phaseK=0.05)
    b = G.dense_colony(n, seed=170) if n > 1 else G.single_cell()
    pm = ParticleManager(p)
    pm.set_state(b)
    pm.step(0)
    want_dt, want_med = expected_dt(pm.get_state(), p.dt, 0.5, tol, 0.01)
    dt, med = pm.adapt_dt(0.5, 0.01)
    print(f"[adapt_dt] n={n} median={med:.6g} dt={dt:.6g}", flush=True)
    assert med == want_med and dt == want_dt
    pm.step(1)  # the next step runs with the new dt
    pm.close()


def test_batch_adapt_dt_per_colony():
    from pcc.engine import ColonyBatch, ParticleManager
    cols = [G.dense_colony(n, seed=180 + i) for i, n in enumerate([1500, 301, 2400])]
    ps = [reference_params(dt=1e-3 * (1 + i), tolerance=1e-3, tau=1 / LN2, cutoff_phase0=0.05,
                           cutoff_phaseK=0.05) for i in range(3)]
    cb = ColonyBatch(ps[0])
    cb.set_state(cols, ps)
    cb.step(0)
    dts, meds = cb.adapt_dt(0.5, 0.01)
    for q in range(3):
        want_dt, want_med = expected_dt(cb.get_state(q), ps[q].dt, 0.5, ps[q].relcp_tol, 0.01)
        assert meds[q] == want_med and dts[q] == want_dt
    cb.close()


def test_adapt_dt_one_colony_batch_equals_single():
    """Both entry points take the tolerance from the colony's relcp_tol: a
    batch of one colony and the single-colony context adapt to the same dt."""
    from pcc.engine import ColonyBatch, ParticleManager
    p = reference_params(dt=1e-3, tolerance=3e-4, tau=1 / LN2, cutoff_phase0=0.05,
                         cutoff_phaseK=0.05)
    b = G.dense_colony(900, seed=185)
    pm = ParticleManager(p)
    pm.set_state(b)
    pm.step(0)
    cb = ColonyBatch(p)
    cb.set_state([b], [p])
    cb.step(0)
    dt1, med1 = pm.adapt_dt(0.5, 0.01)
    dts, meds = cb.adapt_dt(0.5, 0.01)
    print(f"[adapt_dt] single {dt1!r} batch {dts[0]!r}", flush=True)
    assert dt1 == dts[0] and med1 == meds[0]
    pm.close()
    cb.close()
