"""Seeded division + gid reassignment in the oracle (Particle.cpp:122-150,
Domain.cpp:492-508), pinned by closed forms and invariants:

* no noise: exact halves at x -/+ (L/4) t, parent orientation kept;
* the split conserves length; l1 stays in [1-a, 1+a] L/2;
* each child's in-plane turn equals the rotation by the drawn angle (checked
  against numpy's cos/sin), and the angles have mean 0, sd angle_noise;
* cells below 2 l0 are untouched; the output is sorted by x;
* draws depend on (seed, step, parent) only;
* age (the generation, Particle.cpp:26,144,148): k rounds of division of a
  cell of age 1 give 2^k cells of age k + 1; over any sequence of divisions
  sum 2^-(age-1) is conserved; a cell that does not divide keeps its age.
"""
import math

import numpy as np

import oracle
from pcc import generators as G
from pcc.abi_types import Bodies


def rods(theta, L, x=None):
    n = len(L)
    pos = np.zeros((n, 3))
    pos[:, 0] = np.arange(n) * 10.0 if x is None else x
    q = np.zeros((n, 4))
    q[:, 0] = np.cos(np.asarray(theta) / 2)
    q[:, 3] = np.sin(np.asarray(theta) / 2)
    return Bodies(pos, q, np.asarray(L, float))


def angle(q):
    return 2 * np.arctan2(q[:, 3], q[:, 0])


def test_noiseless_division_is_exact_halves():
    b = rods([0.3, 1.1, 2.0], [2.0, 1.5, 2.0])
    out = oracle.divide(b, seed=1, step=0, length_noise=0.0, angle_noise=0.0)
    assert out.n == 5
    for i in (0, 2):
        t = np.array([math.cos([0.3, 1.1, 2.0][i]), math.sin([0.3, 1.1, 2.0][i]), 0.0])
        c = b.pos[i]
        want = sorted([tuple(c - 0.5 * t), tuple(c + 0.5 * t)])
        got = sorted(tuple(p) for p, ln in zip(out.pos, out.length)
                     if abs(p[0] - c[0]) < 2 and ln == 1.0)
        assert np.allclose(got, want, rtol=0, atol=1e-15)
    assert sorted(out.length.tolist()) == [1.0, 1.0, 1.0, 1.0, 1.5]
    # orientation kept (up to the renormalisation's last bit)
    for p, q in zip(out.pos, out.quat):
        src = int(np.argmin(np.abs(b.pos[:, 0] - p[0])))
        assert np.allclose(q, b.quat[src], rtol=0, atol=1e-15)


def test_length_noise_conserves_length_and_stays_in_range():
    rng = np.random.default_rng(0)
    n = 400
    b = rods(rng.uniform(0, np.pi, n), np.full(n, 2.0))
    out = oracle.divide(b, seed=7, step=3, length_noise=0.02, angle_noise=0.0)
    assert out.n == 2 * n
    assert np.all(out.length >= 0.98 - 1e-15) and np.all(out.length <= 1.02 + 1e-15)
    assert abs(out.length.sum() - 2.0 * n) < 1e-9
    assert out.length.std() > 0.005  # the noise is there


def test_turn_matches_rotation_and_angle_statistics():
    rng = np.random.default_rng(1)
    n = 3000
    th = rng.uniform(0, np.pi, n)
    b = rods(th, np.full(n, 2.0), x=np.arange(n) * 10.0)
    sd = 0.01
    out = oracle.divide(b, seed=11, step=5, length_noise=0.0, angle_noise=sd)
    assert out.n == 2 * n
    # children of parent i lie at x_i -/+ 0.5 t_i; pair them by x
    parent = np.rint(out.pos[:, 0] / 10.0).astype(int)
    d = angle(out.quat) - th[parent]
    d = (d + np.pi) % (2 * np.pi) - np.pi
    assert abs(d.mean()) < 3 * sd / math.sqrt(2 * n)
    assert abs(d.std() / sd - 1) < 0.05
    assert np.abs(d).max() <= 6 * sd + 1e-12  # Irwin-Hall bound
    # the quaternion is exactly a z-rotation of the parent's (numpy cos/sin)
    q0 = b.quat[parent]
    c, s = np.cos(d / 2), np.sin(d / 2)
    r = np.stack([c * q0[:, 0] - s * q0[:, 3], c * q0[:, 1] - s * q0[:, 2],
                  c * q0[:, 2] + s * q0[:, 1], c * q0[:, 3] + s * q0[:, 0]], 1)
    assert np.allclose(out.quat, r, rtol=0, atol=2e-15)
    assert np.allclose(np.linalg.norm(out.quat, axis=1), 1.0, rtol=0, atol=2e-16)


def test_untouched_cells_sorted_output_and_planarity():
    b = G.dense_colony(3000, seed=4)
    b.length[::3] = 2 * b.l0[::3]  # every third cell divides
    out = oracle.divide(b, seed=2, step=9, length_noise=0.02, angle_noise=0.01)
    assert out.n == b.n + len(b.length[::3])
    assert np.all(np.diff(out.pos[:, 0]) >= 0)
    keep = b.length < 2 * b.l0
    kept = {tuple(p) for p in b.pos[keep]}
    got = {tuple(p) for p in out.pos}
    assert kept <= got
    assert np.all(out.pos[:, 2] == 0.0)
    assert np.all(out.quat[:, 1] == 0.0) and np.all(out.quat[:, 2] == 0.0)


def test_draws_depend_on_seed_and_step_only():
    b = G.dense_colony(500, seed=5)
    b.length[:] = 2 * b.l0
    a1 = oracle.divide(b, seed=3, step=1, length_noise=0.02, angle_noise=0.01)
    a2 = oracle.divide(b.copy(), seed=3, step=1, length_noise=0.02, angle_noise=0.01)
    a3 = oracle.divide(b.copy(), seed=3, step=2, length_noise=0.02, angle_noise=0.01)
    assert np.array_equal(a1.pos, a2.pos) and np.array_equal(a1.quat, a2.quat)
    assert not np.array_equal(a1.length, a3.length)


def test_synchronous_generations_have_closed_form_ages():
    b = rods([0.2, 0.9], [2.0, 2.0], x=[0.0, 50.0])
    assert list(b.age) == [1, 1]
    for k in range(1, 5):
        b = oracle.divide(b, seed=3, step=k, length_noise=0.0, angle_noise=0.0)
        assert b.n == 2 * 2**k
        assert np.all(b.age == k + 1), k
        b.length[:] = 2.0  # every cell at its division length again


def test_age_weight_is_conserved_and_non_dividers_keep_their_age():
    rng = np.random.default_rng(4)
    n = 300
    b = rods(rng.uniform(0, np.pi, n), rng.uniform(1.0, 2.4, n))
    b.age = rng.integers(1, 7, n).astype(np.int32)
    weight = np.sum(2.0 ** -(b.age - 1.0))
    for k in range(6):
        divides = b.length >= 2 * b.l0
        keep = sorted(zip(b.pos[~divides, 0], b.age[~divides]))
        b = oracle.divide(b, seed=11, step=k, length_noise=0.02, angle_noise=0.01)
        assert np.sum(2.0 ** -(b.age - 1.0)) == weight, k
        # cells that did not divide are where they were, with their age
        by_x = dict(zip(b.pos[:, 0], b.age))
        assert all(by_x[x] == a for x, a in keep)
        b.length = b.length * rng.uniform(1.0, 2.2, b.n)
