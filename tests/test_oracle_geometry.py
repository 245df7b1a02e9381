"""Pins the oracle's geometry against the reference's own code and closed forms.

* golden vectors written by tests/golden/gen_golden.py from the reference's
  DCPQuery / getDirectionVector / normalize compiled in place (bit-exact);
* a larger random sweep against oracle/_ref/libref.so when it is built;
* closed forms of segment-segment distance and the deterministic exp.
"""
import ctypes as C
import math
import os

import numpy as np
import pytest

import oracle

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden", "geometry_golden.npz")
REF = os.path.join(os.path.dirname(HERE), "oracle", "_ref", "libref.so")


def test_dcp_matches_reference_golden_bitwise():
    z = np.load(GOLD)
    for a, want in zip(z["dcp_in"], z["dcp_out"]):
        got = oracle.dcp(a[0:3], a[3:6], a[6:9], a[9:12])
        assert np.array_equal(got, want), (a, got, want)


def test_direction_and_normalize_match_reference_golden_bitwise():
    z = np.load(GOLD)
    for q, t in zip(z["quat"], z["direction"]):
        assert np.array_equal(oracle.direction(q), t)
    for v, t in zip(z["vec"], z["normalized"]):
        assert np.array_equal(oracle.normalize(v), t)


@pytest.mark.skipif(not os.path.exists(REF), reason="oracle/_ref/libref.so not built")
def test_dcp_random_sweep_against_compiled_reference():
    L = C.CDLL(REF)
    dp = C.POINTER(C.c_double)
    L.ref_dcp.argtypes = [dp, dp, dp, dp, dp]
    rng = np.random.default_rng(7)
    for _ in range(3000):
        pts = [np.ascontiguousarray(rng.uniform(-2, 2, 3)) for _ in range(4)]
        if rng.random() < 0.5:
            for p in pts:
                p[2] = 0.0
        if rng.random() < 0.2:  # parallel pairs
            d = pts[1] - pts[0]
            pts[3] = pts[2] + d * rng.uniform(-2, 2)
        want = np.zeros(10)
        L.ref_dcp(*[p.ctypes.data_as(dp) for p in pts], want.ctypes.data_as(dp))
        got = oracle.dcp(*pts)
        assert np.array_equal(got, want)


def _seg_dist_brute(P0, P1, Q0, Q1, n=2001):
    s = np.linspace(0, 1, n)
    a = P0[None, :] + s[:, None] * (P1 - P0)[None, :]
    b = Q0[None, :] + s[:, None] * (Q1 - Q0)[None, :]
    d = np.linalg.norm(a[:, None, :] - b[None, :, :], axis=2)
    return d.min()


@pytest.mark.parametrize("case,dist", [
    (([-1, 0, 0], [1, 0, 0], [0, -1, 0], [0, 1, 0]), 0.0),            # crossing
    (([0, 0, 0], [2, 0, 0], [1, 0.3, 0], [3, 0.3, 0]), 0.3),           # parallel overlap
    (([0, 0, 0], [1, 0, 0], [2, 0, 0], [3, 0, 0]), 1.0),               # collinear gap
    (([0, 0, 0], [2, 0, 0], [1, 0.5, 0], [1, 1.5, 0]), 0.5),           # T / point-segment
    (([0.5, 0.5, 0], [0.5, 0.5, 0], [0, 0, 0], [1, 0, 0]), 0.5),       # degenerate core
    (([0, 0, 0], [1, 0, 0], [0.5, -1, 1], [0.5, 1, 1]), 1.0),          # skew 3D
    (([0, 0, 0], [1, 0, 0], [4, 3, 0], [5, 3, 0]), math.hypot(3, 3)),  # endpoint-endpoint
])
def test_dcp_closed_forms(case, dist):
    P0, P1, Q0, Q1 = (np.array(v, float) for v in case)
    o = oracle.dcp(P0, P1, Q0, Q1)
    assert o[9] == pytest.approx(dist, abs=1e-14)
    # closest points lie on the segments and realise the distance
    assert np.linalg.norm(o[2:5] - o[5:8]) == pytest.approx(dist, abs=1e-14)
    # symmetry
    o2 = oracle.dcp(Q0, Q1, P0, P1)
    assert o2[9] == pytest.approx(dist, abs=1e-14)


def test_dcp_against_brute_force_random():
    rng = np.random.default_rng(3)
    for _ in range(40):
        P0, P1, Q0, Q1 = (rng.uniform(-1, 1, 3) for _ in range(4))
        o = oracle.dcp(P0, P1, Q0, Q1)
        assert o[9] <= _seg_dist_brute(P0, P1, Q0, Q1) + 1e-12
        assert o[9] >= _seg_dist_brute(P0, P1, Q0, Q1) - 2e-3


def test_direction_planar_closed_form():
    for th in np.linspace(0, math.pi, 17):
        t = oracle.direction([math.cos(th / 2), 0, 0, math.sin(th / 2)])
        assert t == pytest.approx([math.cos(th), math.sin(th), 0.0], abs=1e-15)


def test_deterministic_exp_accuracy():
    rng = np.random.default_rng(11)
    xs = np.concatenate([rng.uniform(-700, 700, 4000), rng.uniform(-1, 1, 4000),
                         -np.logspace(-12, 2.5, 500), [0.0, -0.0, 1.0, -1.0, 709.0, -744.0]])
    for x in xs:
        got = oracle.exp(x)
        want = math.exp(x)
        if want == 0.0 or not np.isfinite(want):
            continue
        if want < 2.3e-308:  # subnormal results: compare absolutely
            assert abs(got - want) <= 2 * 5e-324
            continue
        ulp = np.spacing(want)
        assert abs(got - want) <= 2 * ulp, (x, got, want)
    assert oracle.exp(0.0) == 1.0
    assert oracle.exp(-800.0) == 0.0
    assert math.isinf(oracle.exp(800.0))


def test_block_sum_is_the_fixed_tree():
    rng = np.random.default_rng(5)
    for n in (0, 1, 63, 64, 255, 256, 257, 1000, 70000):
        v = rng.normal(size=n)
        s = oracle.block_sum(v)
        assert s == pytest.approx(math.fsum(v), abs=1e-12 * max(1.0, np.abs(v).sum()))
        # restate the tree in numpy and require bitwise equality
        nb = (n + 255) // 256
        pad = np.zeros(nb * 256)
        pad[:n] = v

        def block(x):
            w = []
            for k in range(4):
                y = x[64 * k:64 * (k + 1)].copy()
                off = 32
                while off >= 1:
                    y = y + y[np.arange(64) ^ off]
                    off //= 2
                w.append(y[0])
            return ((w[0] + w[1]) + w[2]) + w[3]

        parts = [block(pad[256 * b:256 * (b + 1)]) for b in range(nb)]
        acc = []
        for t in range(256):
            a = 0.0
            for i in range(t, nb, 256):
                a += parts[i]
            acc.append(a)
        assert s == block(np.array(acc))
