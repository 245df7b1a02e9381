"""The evidence behind the mixed-precision design, rerun on the CPU: the
phase-0 LCP of an i.i.d. random disc packing (N = 3000, area fraction 0.7,
seed 7, cutoff 0.05, tol 1e-6, impulses near 1e6) solved by BBPGD with the
applies in different arithmetic (oracle/precision_study.py).  FP32
arithmetic in the applies cannot reach the tolerance; FP32 storage of the
operator, Phi and the evaluated gradient with FP64 accumulation (the
library's mixed path) converges to within 1e-5 of the FP64 impulses."""
import numpy as np
import pytest

import oracle
import precision_study as ps
from pcc import generators as G
from pcc.abi_types import reference_params


@pytest.fixture(scope="module")
def problem():
    p = reference_params(dt=1e-3, tolerance=1e-6)
    b = G.random_disc(3000, seed=7, area_fraction=0.7)
    cons = oracle.detect(b, 0.05)
    D, m = ps.operator(b, cons, p.xi)
    phi = cons["signed_distance"]
    gamma0 = np.where(phi < 0.0, -phi * 500.0, 0.0)
    return p, D, m, phi, gamma0


def solve(problem, variant, max_its=20000):
    p, D, m, phi, gamma0 = problem
    return ps.bbpgd(ps.make_grad(D, m, phi, p.dt, variant), gamma0, p.bbpgd_tol, max_its)


def test_fp32_arithmetic_stalls_fp32_storage_converges(problem):
    g64, its64, res64 = solve(problem, "fp64")
    assert res64 <= 1e-6 and g64.max() > 5e5  # impulses approach 1e6
    _, _, res_acc = solve(problem, "fp32_acc")
    assert res_acc > 1e-5  # FP32 sums: stalls two orders above the tolerance
    _, _, res_u = solve(problem, "fp32_u")
    assert res_u > 1e-6  # FP64 sums, U rounded to FP32: stalls just above it
    gm, _, res_m = solve(problem, "fp32_store_g")
    assert res_m <= 1e-6
    assert np.linalg.norm(gm - g64) / np.linalg.norm(g64) <= 1e-5
