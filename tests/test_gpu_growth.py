"""The configured CPU-path workload on the device: one cell grown to 1,000
cells, state resident on the GPU, bit-for-bit against the oracle's growth.

The fixture (tests/golden/config1_growth.npz, written by gen_config1.py with
the FP64 oracle) holds the BBPGD iterations and ReLCP phases of all 10,548
time-loop iterations, the cell count after each, and the full state at the
first step reaching 125, 250, 500 and 1,000 cells.  Each iteration here is
the same: the device division (seed 0, the step index keying the draws,
angle noise 0.01) and renumbering by x (Domain.cpp:121-143), then the hard
step with dt = 1e-3, the reference's tolerances at 1e-6 and rotational
diffusion off (FP64).  The state never leaves the device between
iterations; the counts are read from the step summaries.  Growth to 250
cells (8,138 iterations) runs in the default GPU suite; the full 1,000-cell
run (49.5M BBPGD iterations) is in the same test behind the
PCC_GROWTH_FULL=1 environment switch and in bench.py --workload growth.
"""
import math
import os
import time

import numpy as np
import pytest

from pcc import generators as G
from pcc.abi_types import reference_params

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
FIELDS = ("pos", "quat", "length", "l0", "diameter", "ldot", "impedance", "age", "force",
          "torque", "velocity", "omega", "stress", "num_constraints")


def fixture():
    return np.load(os.path.join(HERE, "golden", "config1_growth.npz"))


def params():
    return reference_params(dt=1e-3, tolerance=1e-6, tau=1 / math.log(2), temperature=0.0)


def test_fixture_shape():
    f = fixture()
    n = f["n_cells"]
    assert len(f["bbpgd_iterations"]) == len(n) == len(f["relcp_phases"]) == 10548
    assert n[-1] >= 1000 > n[-2]
    assert np.all(np.diff(n) >= 0)
    for m, st in zip(f["marks"], f["mark_steps"]):
        assert n[st] >= m and (st == 0 or n[st - 1] < m)
        assert len(f[f"n{m}_pos"]) == n[st]


def grow_and_compare(last_mark):
    from pcc.engine import ParticleManager
    f = fixture()
    marks = {int(m): int(s) for m, s in zip(f["marks"], f["mark_steps"])}
    steps = marks[last_mark] + 1
    pm = ParticleManager(params())
    pm.set_state(G.single_cell())
    its = np.zeros(steps, np.int64)
    phases = np.zeros(steps, np.int32)
    counts = np.zeros(steps, np.int32)
    t0 = last = time.time()
    for k in range(steps):
        pm.divide(0, k, 0.0, 0.01)
        r = pm.step(k)
        its[k], phases[k], counts[k] = r["bbpgd_iterations"], r["constraint_iterations"], pm.n
        if counts[k] != f["n_cells"][k] or its[k] != f["bbpgd_iterations"][k]:
            break
        if k % 500 == 0 or time.time() - last > 30.0:
            last = time.time()
            print(f"[growth] iteration {k}: {pm.n} cells, {its[k]} BBPGD iterations, "
                  f"{last - t0:.1f} s", flush=True)
        for m, st in marks.items():
            if st == k:
                got = pm.get_state()
                for fld in FIELDS:
                    assert np.array_equal(getattr(got, fld), f[f"n{m}_{fld}"]), (m, fld)
                print(f"[growth] {m} cells at iteration {k}: state bit-identical, "
                      f"{time.time() - t0:.1f} s", flush=True)
    pm.close()
    assert np.array_equal(counts, f["n_cells"][:steps])
    assert np.array_equal(its, f["bbpgd_iterations"][:steps])
    assert np.array_equal(phases, f["relcp_phases"][:steps])
    return time.time() - t0


def test_config1_growth_bitexact():
    full = os.environ.get("PCC_GROWTH_FULL") == "1"
    wall = grow_and_compare(1000 if full else 250)
    print(f"[growth] device {wall:.1f} s", flush=True)
