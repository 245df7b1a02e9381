"""pcc_detect_noalloc (contact detection that allocates nothing) against
pcc_detect, through the C ABI.

The no-allocation variant runs the same detection pipeline on the library's
fixed per-device scratch array instead of a context's buffers, for one colony
held in caller-owned device memory.  Its records must equal pcc_detect's bit
for bit (integer and index work plus the same FP64 expressions).  Three calls
in a row share the scratch, each on a smaller input than the one before
(2D, then 2D, then 3D), so stale contents of a larger call would show.  The
test allocates the caller's device buffers itself with hipMalloc on the HIP
runtime libpcc links (plumbing; the entry point allocates nothing).  Also:
inputs beyond the documented maximum are rejected with a message and write
nothing; a capacity smaller than the contact count writes the first records
and reports the full count.
"""
import ctypes as C

import numpy as np
import pytest

from pcc import generators as G
from pcc.abi_types import PCC_E_INVALID_ARGUMENT, ConstraintsC, reference_params

pytestmark = pytest.mark.gpu

MAX_BODIES = 32768  # PCC_NOALLOC_MAX_BODIES
FIELDS = ("gidI", "gidJ", "signed_distance", "normal", "rPosI", "rPosJ", "contact_point",
          "stressI", "stressJ", "gamma", "gid", "iteration")


class DeviceBodies(C.Structure):
    _fields_ = [("n", C.c_int64), ("pos", C.c_void_p), ("quat", C.c_void_p),
                ("length", C.c_void_p), ("diameter", C.c_void_p)]


def lib():
    from pcc import engine
    L = engine.lib()
    from pcc.abi_types import Params
    L.pcc_detect_noalloc.argtypes = [C.c_int, C.c_void_p, C.POINTER(Params), C.c_double,
                                     C.POINTER(DeviceBodies), C.POINTER(ConstraintsC),
                                     C.POINTER(C.c_int64)]
    L.pcc_detect_noalloc.restype = C.c_int
    L.pcc_last_error.argtypes = [C.c_void_p]
    L.pcc_last_error.restype = C.c_char_p
    return L


class Hip:
    """hipMalloc / hipMemcpy / hipFree of the HIP runtime libpcc.so links."""

    H2D, D2H = 1, 2

    def __init__(self):
        lib()  # libpcc.so first: its libamdhip64 is the one loaded below
        self.L = C.CDLL("libamdhip64.so.7")
        self.L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.L.hipFree.argtypes = [C.c_void_p]
        self.L.hipSetDevice.argtypes = [C.c_int]
        for f in ("hipMalloc", "hipMemcpy", "hipFree", "hipSetDevice", "hipDeviceSynchronize"):
            getattr(self.L, f).restype = C.c_int
        assert self.L.hipSetDevice(0) == 0
        self.held = []

    def put(self, a):
        a = np.ascontiguousarray(a)
        p = C.c_void_p()
        assert self.L.hipMalloc(C.byref(p), max(a.nbytes, 8)) == 0
        self.held.append(p)
        if a.nbytes:
            assert self.L.hipMemcpy(p, a.ctypes.data_as(C.c_void_p), a.nbytes, self.H2D) == 0
        return p.value

    def get(self, ptr, like):
        out = np.empty_like(like)
        assert self.L.hipMemcpy(out.ctypes.data_as(C.c_void_p), C.c_void_p(ptr), out.nbytes,
                                self.D2H) == 0
        return out

    def free(self):
        assert self.L.hipDeviceSynchronize() == 0
        for p in self.held:
            assert self.L.hipFree(p) == 0
        self.held = []


def noalloc_detect(bodies, cutoff, params, capacity):
    hip = Hip()
    db = DeviceBodies(bodies.n, hip.put(bodies.pos), hip.put(bodies.quat),
                      hip.put(bodies.length), hip.put(bodies.diameter))
    cap = max(int(capacity), 1)
    init = {
        "gidI": np.full(cap, -7, np.int32), "gidJ": np.full(cap, -7, np.int32),
        "signed_distance": np.full(cap, np.nan), "normal": np.full((cap, 3), np.nan),
        "rPosI": np.full((cap, 3), np.nan), "rPosJ": np.full((cap, 3), np.nan),
        "contact_point": np.full((cap, 3), np.nan), "stressI": np.full(cap, np.nan),
        "stressJ": np.full(cap, np.nan), "gamma": np.full(cap, np.nan),
        "gid": np.full(cap, -7, np.int32), "iteration": np.full(cap, -7, np.int32),
    }
    dev = {f: hip.put(init[f]) for f in FIELDS}
    oc = ConstraintsC()
    oc.capacity = int(capacity)
    for f in FIELDS:
        ct = C.c_int32 if init[f].dtype == np.int32 else C.c_double
        setattr(oc, f, C.cast(C.c_void_p(dev[f]), C.POINTER(ct)))
    cnt = C.c_int64(-1)
    st = lib().pcc_detect_noalloc(0, None, C.byref(params), float(cutoff), C.byref(db),
                                  C.byref(oc), C.byref(cnt))
    host = {f: hip.get(dev[f], init[f]) for f in FIELDS}
    hip.free()
    return st, int(cnt.value), host


def reference_detect(bodies, cutoff, params):
    from pcc.engine import ParticleManager
    pm = ParticleManager(params)
    pm.set_state(bodies)
    got, _, _ = pm.detect(cutoff)
    pm.close()
    return got


def random_3d(n, seed):
    rng = np.random.default_rng(seed)
    pos = rng.uniform(-4, 4, (n, 3))
    q = rng.normal(size=(n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    o = np.argsort(pos[:, 0], kind="stable")
    return G.Bodies(pos[o], q[o], rng.uniform(1, 2, n)[o])


CALLS = [("colony6000_c0.05", lambda: G.packed_colony(6000, seed=31), 0.05),
         ("disc1500_c0.25", lambda: G.random_disc(1500, seed=32), 0.25),
         ("random3d_500_c0.05", lambda: random_3d(500, 33), 0.05)]


def test_noalloc_equals_detect_on_consecutive_calls():
    p = reference_params()
    sizes = []
    for name, make, cutoff in CALLS:
        b = make()
        sizes.append(b.n)
        want = reference_detect(b, cutoff, p)
        m = len(want["gidI"])
        print(f"[noalloc] {name} n={b.n} contacts={m}", flush=True)
        st, cnt, got = noalloc_detect(b, cutoff, p, m + 5)
        assert st == 0, lib().pcc_last_error(None).decode()
        assert cnt == m > 0
        for f in FIELDS:
            assert np.array_equal(got[f][:m], want[f]), f
        assert np.all(got["gidI"][m:] == -7)  # nothing past the count
    assert sizes[0] > sizes[1] > sizes[2]


def test_noalloc_capacity_and_limits():
    p = reference_params()
    b = G.packed_colony(2000, seed=34)
    want = reference_detect(b, 0.05, p)
    m = len(want["gidI"])
    st, cnt, got = noalloc_detect(b, 0.05, p, 100)
    assert st == 0 and cnt == m > 100
    for f in FIELDS:
        assert np.array_equal(got[f][:100], want[f][:100]), f
    big = G.random_disc(MAX_BODIES + 1, seed=35)
    st, cnt, got = noalloc_detect(big, 0.05, p, 10)
    assert st == PCC_E_INVALID_ARGUMENT
    assert "PCC_NOALLOC_MAX_BODIES" in lib().pcc_last_error(None).decode()
    assert np.all(got["gidI"] == -7)
    # a valid call after the rejected one still matches
    st, cnt, got = noalloc_detect(b, 0.05, p, m)
    assert st == 0 and cnt == m
    for f in FIELDS:
        assert np.array_equal(got[f], want[f]), f
