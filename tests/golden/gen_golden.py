"""Writes tests/golden/geometry_golden.npz: input/output vectors of the
reference's own geometry code, compiled in place from the reference checkout
into oracle/_ref/libref.so (oracle/Makefile target `ref`):

  dcp        DCPQuery<double,3>::operator()   util/SpherocylinderCell.h:64-261
  direction  Quaternion::getDirectionVector   util/Quaternion.cpp:36-38
  normalize  ArrayMath::normalize             util/ArrayMath.h:218-231

Run: make -C oracle ref && python tests/golden/gen_golden.py
The fixture is data only (inputs and the reference's outputs).
"""
import ctypes as C
import math
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
LIB = os.path.join(REPO, "oracle", "_ref", "libref.so")


def special_segments():
    """Hand-made cases that reach every region of the DCP case analysis."""
    cases = []
    z = 0.0
    # crossing segments (distance 0)
    cases.append(([-1, 0, z], [1, 0, z], [0, -1, z], [0, 1, z]))
    # parallel, overlapping, offset in y
    cases.append(([0, 0, z], [2, 0, z], [1, 0.3, z], [3, 0.3, z]))
    # parallel, disjoint along the axis (endpoint-endpoint)
    cases.append(([0, 0, z], [1, 0, z], [2, 0.2, z], [3, 0.2, z]))
    # anti-parallel
    cases.append(([0, 0, z], [1, 0, z], [3, 0.2, z], [2, 0.2, z]))
    # collinear, touching end to end
    cases.append(([0, 0, z], [1, 0, z], [1, 0, z], [2, 0, z]))
    # collinear overlapping
    cases.append(([0, 0, z], [2, 0, z], [1, 0, z], [3, 0, z]))
    # T configuration (point-segment)
    cases.append(([0, 0, z], [2, 0, z], [1, 0.5, z], [1, 1.5, z]))
    # degenerate P (zero-length core: length == diameter)
    cases.append(([0.5, 0.5, z], [0.5, 0.5, z], [0, 0, z], [1, 0, z]))
    # degenerate Q
    cases.append(([0, 0, z], [1, 0, z], [0.3, -0.4, z], [0.3, -0.4, z]))
    # both degenerate
    cases.append(([0, 0, z], [0, 0, z], [1, 1, z], [1, 1, z]))
    # skew 3D
    cases.append(([0, 0, 0], [1, 0, 0], [0.5, -1, 1], [0.5, 1, 1]))
    # nearly parallel (det tiny but positive)
    cases.append(([0, 0, z], [1, 0, z], [0, 0.1, z], [1, 0.1 + 1e-9, z]))
    # endpoint regions of the s>0 branch
    cases.append(([0, 0, z], [1, 0, z], [1.5, 0.5, z], [2.5, 1.5, z]))
    cases.append(([0, 0, z], [1, 0, z], [-1.5, 0.5, z], [-0.5, 1.5, z]))
    cases.append(([0, 0, z], [1, 0, z], [0.5, 2.0, z], [0.7, 0.5, z]))
    cases.append(([0, 0, z], [1, 0, z], [0.5, -2.0, z], [0.7, -0.5, z]))
    return [tuple(np.array(v, dtype=np.float64) for v in c) for c in cases]


def random_segments(rng, n, planar):
    out = []
    for _ in range(n):
        c1 = rng.uniform(-2, 2, 3)
        c2 = c1 + rng.uniform(-1.5, 1.5, 3)
        if planar:
            c1[2] = 0.0
            c2[2] = 0.0
        h1 = rng.uniform(0, 0.75)
        h2 = rng.uniform(0, 0.75)
        if planar:
            a1, a2 = rng.uniform(0, math.pi, 2)
            t1 = np.array([math.cos(a1), math.sin(a1), 0.0])
            t2 = np.array([math.cos(a2), math.sin(a2), 0.0])
        else:
            t1 = rng.normal(size=3)
            t1 /= np.linalg.norm(t1)
            t2 = rng.normal(size=3)
            t2 /= np.linalg.norm(t2)
        out.append((c1 - h1 * t1, c1 + h1 * t1, c2 - h2 * t2, c2 + h2 * t2))
    return out


def main():
    if not os.path.exists(LIB):
        raise SystemExit(f"{LIB} missing: run `make -C oracle ref` first")
    L = C.CDLL(LIB)
    dp = C.POINTER(C.c_double)
    L.ref_dcp.argtypes = [dp, dp, dp, dp, dp]
    L.ref_direction.argtypes = [dp, dp]
    L.ref_normalize.argtypes = [dp, dp]

    def d(a):
        return a.ctypes.data_as(dp)

    rng = np.random.default_rng(20251203)
    segs = special_segments() + random_segments(rng, 300, True) + random_segments(rng, 200, False)
    dcp_in = np.zeros((len(segs), 12))
    dcp_out = np.zeros((len(segs), 10))
    for i, (P0, P1, Q0, Q1) in enumerate(segs):
        o = np.zeros(10)
        args = [np.ascontiguousarray(v) for v in (P0, P1, Q0, Q1)]
        L.ref_dcp(*[d(a) for a in args], d(o))
        dcp_in[i] = np.concatenate(args)
        dcp_out[i] = o

    qs = rng.normal(size=(200, 4))
    qs /= np.linalg.norm(qs, axis=1, keepdims=True)
    th = rng.uniform(0, math.pi, 100)
    qp = np.zeros((100, 4))
    qp[:, 0] = np.cos(th / 2)
    qp[:, 3] = np.sin(th / 2)
    quats = np.concatenate([qs, qp, np.array([[1.0, 0, 0, 0]])])
    dirs = np.zeros((len(quats), 3))
    for i, q in enumerate(quats):
        q = np.ascontiguousarray(q)
        t = np.zeros(3)
        L.ref_direction(d(q), d(t))
        dirs[i] = t

    vecs = np.concatenate([rng.normal(size=(100, 3)), np.array([[0.0, 0, 0], [1e-17, 0, 0],
                                                                 [3e-16, 0, 0]])])
    nrm = np.zeros_like(vecs)
    for i, v in enumerate(vecs):
        v = np.ascontiguousarray(v)
        o = np.zeros(3)
        L.ref_normalize(d(v), d(o))
        nrm[i] = o

    out = os.path.join(HERE, "geometry_golden.npz")
    np.savez_compressed(out, dcp_in=dcp_in, dcp_out=dcp_out, quat=quats, direction=dirs,
                        vec=vecs, normalized=nrm)
    print(f"wrote {out}: {len(segs)} DCP cases, {len(quats)} quaternions, {len(vecs)} vectors")


if __name__ == "__main__":
    main()
