"""Pins the oracle's broad and narrow phase.

* the grid ordering is a stable counting sort of bodies by hashed cell;
* the grid candidate set equals brute force O(N^2) whenever
  (l_i+l_j)/2 + cutoff <= h (true for cutoff 0.05, h = 2.2, l <= 2);
* every record satisfies the closed forms of tryCreateConstraint
  (CollisionDetector.cpp:201-311): phi = dist - d, |n| = 1, r = p - x,
  s = |n.t|/2, gamma0 = 500*overlap; the batch is sorted by (p.x, I, J).
"""
import math

import numpy as np
import pytest

import oracle
from pcc import generators as G


def test_grid_is_stable_counting_sort():
    b = G.random_disc(3000, seed=1)
    g = oracle.grid(b)
    cell = np.floor((b.pos + 1.0) / 2.2).astype(np.int64)
    assert np.array_equal(g["cell"], cell)
    perm = g["perm"]
    assert sorted(perm.tolist()) == list(range(b.n))
    keys = g["bucket"][perm]
    assert np.all(np.diff(keys.astype(np.int64)) >= 0)
    # stable: equal buckets keep ascending body index
    same = np.diff(keys.astype(np.int64)) == 0
    assert np.all(np.diff(perm)[same] > 0)
    start = g["bucket_start"]
    assert start[-1] == b.n and np.all(np.diff(start) >= 0)


@pytest.mark.parametrize("n,seed,cutoff", [(400, 0, 0.05), (1500, 2, 0.05), (800, 3, 0.0)])
def test_grid_detection_equals_brute_force(n, seed, cutoff):
    b = G.random_disc(n, seed=seed)
    a = oracle.detect(b, cutoff)
    c = oracle.detect(b, cutoff, brute=True)
    for k in a:
        assert np.array_equal(a[k], c[k]), k


def test_grid_detection_equals_brute_force_3d():
    rng = np.random.default_rng(9)
    n = 300
    pos = rng.uniform(-4, 4, (n, 3))
    q = rng.normal(size=(n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    b = G.Bodies(pos[np.argsort(pos[:, 0])], q, rng.uniform(1, 2, n))
    a = oracle.detect(b, 0.05)
    c = oracle.detect(b, 0.05, brute=True)
    assert len(a["gidI"]) > 50
    for k in a:
        assert np.array_equal(a[k], c[k]), k


@pytest.mark.parametrize("h", [2.2, 5.0, 12.0])
def test_grid_detection_does_not_depend_on_cell_size(h):
    """Any cell size at least the largest contact reach gives the same list
    (records and order) as brute force: crowded cells (h = 12 puts ~200
    bodies in a cell) only add candidates that the filters reject."""
    b = G.dense_colony(800, seed=21)
    a = oracle.detect(b, 0.05, h=h)
    c = oracle.detect(b, 0.05, brute=True)
    assert len(a["gidI"]) > b.n
    for k in a:
        assert np.array_equal(a[k], c[k]), (h, k)


def test_records_closed_forms_and_order():
    b = G.random_disc(1000, seed=4)
    r = oracle.detect(b, 0.05)
    m = len(r["gidI"])
    assert m > 1000
    I, J = r["gidI"], r["gidJ"]
    assert np.all(I < J)
    # order: (p.x, I, J)
    key = list(zip(r["contact_point"][:, 0], I, J))
    assert key == sorted(key)
    n = r["normal"]
    assert np.allclose(np.linalg.norm(n, axis=1), 1.0, atol=1e-14)
    p = r["contact_point"]
    assert np.allclose(r["rPosI"], p - b.pos[I], atol=1e-15)
    assert np.allclose(r["rPosJ"], p - b.pos[J], atol=1e-15)
    th = 2 * np.arctan2(b.quat[:, 3], b.quat[:, 0])
    tI = np.stack([np.cos(th[I]), np.sin(th[I])], 1)
    tJ = np.stack([np.cos(th[J]), np.sin(th[J])], 1)
    assert np.allclose(r["stressI"], 0.5 * np.abs((n[:, :2] * tI).sum(1)), atol=1e-14)
    assert np.allclose(r["stressJ"], 0.5 * np.abs((n[:, :2] * tJ).sum(1)), atol=1e-14)
    phi = r["signed_distance"]
    assert np.all(phi <= 0.05)
    assert np.allclose(r["gamma"], np.where(phi >= 0, 0.0, -500.0 * phi))
    assert np.array_equal(r["gid"], np.arange(m))
    # phi = core distance - d, recomputed independently by brute force
    for c in range(0, m, max(1, m // 50)):
        P = [b.pos[I[c], :] + s * (b.length[I[c]] / 2 - 0.25) * np.r_[tI[c], 0] for s in (-1, 1)]
        Q = [b.pos[J[c], :] + s * (b.length[J[c]] / 2 - 0.25) * np.r_[tJ[c], 0] for s in (-1, 1)]
        d = oracle.dcp(P[0], P[1], Q[0], Q[1])[9]
        assert phi[c] == pytest.approx(d - 0.5, abs=1e-12)


def test_two_rods_head_on_record():
    b = G.two_rods(0.9)  # overlap 0.1 along x
    r = oracle.detect(b, 0.05)
    assert len(r["gidI"]) == 1
    assert r["signed_distance"][0] == pytest.approx(-0.1, abs=1e-15)
    assert np.allclose(r["normal"][0], [-1, 0, 0])
    assert r["stressI"][0] == pytest.approx(0.5) and r["stressJ"][0] == pytest.approx(0.5)
    assert r["gamma"][0] == pytest.approx(50.0)
    # centre-distance reject (CollisionDetector.cpp:210-213)
    assert len(oracle.detect(G.two_rods(1.06), 0.05)["gidI"]) == 0
    assert len(oracle.detect(G.two_rods(1.04), 0.05)["gidI"]) == 1
    assert math.isclose(oracle.detect(G.two_rods(1.04), 0.05)["signed_distance"][0], 0.04,
                        abs_tol=1e-14)
