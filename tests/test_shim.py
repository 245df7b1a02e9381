"""The header-only C++ shim (include/pcc_particle_manager.hpp) that keeps the
reference's ParticleManager::step signature compiles and links against
libpcc.so; on a GPU it runs and matches the oracle."""
import math
import os
import subprocess

import numpy as np
import pytest

import oracle
from pcc import engine
from pcc import generators as G
from pcc.abi_types import reference_params

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(REPO, "tests", "cpp", "shim_demo.cpp")
PKG = os.path.dirname(engine.LIB_PATH)


def build(tmp_path):
    exe = str(tmp_path / "shim_demo")
    subprocess.run(["g++", "-std=c++17", "-O1", "-I", os.path.join(REPO, "include"), SRC, "-L",
                    PKG, "-lpcc", f"-Wl,-rpath,{PKG}", "-o", exe], check=True)
    return exe


@pytest.mark.skipif(not os.path.exists(engine.LIB_PATH), reason="libpcc.so not built")
def test_shim_compiles_and_links(tmp_path):
    assert os.path.exists(build(tmp_path))


@pytest.mark.gpu
def test_shim_runs_and_matches_oracle(tmp_path):
    exe = build(tmp_path)
    out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout.splitlines()
    p = reference_params(dt=1e-3, tolerance=1e-6, tau=1 / math.log(2))
    st, res, nb, cons, ph, err = oracle.step(p, G.two_rods(0.9))
    head = out[0].split()
    assert int(head[1]) == res["n_constraints"]
    assert int(head[3]) == res["constraint_iterations"]
    assert int(head[5]) == res["bbpgd_iterations"]
    assert int(head[7]) == 1
    gam = [float(l.split()[1]) for l in out if l.startswith("gamma")]
    assert np.array_equal(gam, cons["gamma"])
    xs = [float(l.split()[1]) for l in out if l.startswith("x ")]
    assert np.array_equal(xs, nb.pos[:, 0])
    assert out[-1] == "error: BBPGD did not converge"
