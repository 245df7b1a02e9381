"""Why the benchmark does not use i.i.d. centres at area fraction 0.9
(checker-side, CPU only): the phase-0 LCP of such a packing, solved by the
oracle's BBPGD (BBPGD.cpp:32-127) at tolerance 1e-6 with the iteration cap
raised step by step, keeps a residual orders of magnitude above the
tolerance.  Crossing rods, whose normal is the arbitrary (1, 0, 0) of
ArrayMath::normalize (ArrayMath.h:224-229), make the contact set
practically infeasible; the correlated nematic-domain packing at the same
size converges.

Run: make -C oracle && python oracle/iid_disc_study.py [--n 5000] [--caps 1000,10000,100000]
"""
import argparse
import math
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path[:0] = [HERE, os.path.join(os.path.dirname(HERE), "proliferatingcellcollectives-idp_amd")]

import oracle  # noqa: E402
from pcc import generators as G  # noqa: E402
from pcc.abi_types import reference_params  # noqa: E402


def crossing_fraction(b, cutoff):
    """Share of phase-0 contacts whose normal is the (1, 0, 0) fallback."""
    c = oracle.detect(b, cutoff)
    n = c["normal"]
    fb = (n[:, 0] == 1.0) & (n[:, 1] == 0.0) & (n[:, 2] == 0.0)
    return len(n), float(fb.mean()) if len(n) else 0.0


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=5000)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--caps", default="1000,10000,100000")
    a = ap.parse_args()
    p = reference_params(dt=1e-3, tolerance=1e-6, tau=1 / math.log(2), cutoff_phase0=0.05,
                         cutoff_phaseK=0.05, relcp_tol=math.inf)
    for name, b in (("i.i.d. disc, area fraction 0.9", G.random_disc(a.n, seed=a.seed)),
                    ("nematic-domain packing", G.dense_colony(a.n, seed=a.seed))):
        m, frac = crossing_fraction(b, 0.05)
        print(f"{name}: n={b.n}, {m} phase-0 contacts, {100 * frac:.1f}% with the fallback normal",
              flush=True)
        for cap in [int(x) for x in a.caps.split(",")]:
            t = time.time()
            st, res, *_ = oracle.step(p.replace(max_bbpgd_iterations=cap), b, fast=True)
            print(f"  cap {cap:>8d}: status {st}, iterations {res['bbpgd_iterations']}, "
                  f"residual {res['residual']:.3e} ({time.time() - t:.1f} s)", flush=True)
            if st == 0:
                break


if __name__ == "__main__":
    main()
