"""Python mirror of the reference's step interface over libpcc's C ABI.

`ParticleManager.step()` has the meaning of
`ParticleManager::SolverSolution ParticleManager::step(int, std::function<void()>)`
(code/cpp/src/simulation/ParticleManager.h:50, ParticleManager.cpp:150-166): it
advances the colony by one timestep (hard ReLCP/BBPGD or soft Hertz model) and
returns the SolverSolution fields.  The three solver failures raise
`PccError` with the reference's messages ("BBPGD did not converge",
"Maximum number of constraint iterations reached", "Invalid mode").

There is no CPU fallback: if libpcc.so is missing or no GPU is visible, the
constructor raises.
"""
import ctypes as C
import os

import numpy as np

from . import abi_types as T

_PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(_PKG_DIR, "libpcc.so")
_lib = None
BODY_VARIANTS = 9  # PCC_BODY_VARIANTS


def lib():
    """Loads libpcc.so (the HIP library); raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built; run __graft_entry__.build()")
        L = C.CDLL(LIB_PATH)
        vp = C.c_void_p
        L.pcc_default_params.argtypes = [C.POINTER(T.Params)]
        L.pcc_create.argtypes = [C.POINTER(T.Params), C.c_int, C.POINTER(vp)]
        L.pcc_destroy.argtypes = [vp]
        L.pcc_set_params.argtypes = [vp, C.POINTER(T.Params)]
        L.pcc_last_error.argtypes = [vp]
        L.pcc_last_error.restype = C.c_char_p
        L.pcc_status_string.argtypes = [C.c_int]
        L.pcc_status_string.restype = C.c_char_p
        L.pcc_set_state.argtypes = [vp, C.POINTER(T.BodiesC)]
        L.pcc_get_state.argtypes = [vp, C.POINTER(T.BodiesC)]
        L.pcc_snapshot.argtypes = [vp]
        L.pcc_restore.argtypes = [vp]
        L.pcc_step.argtypes = [vp, C.c_int32, C.POINTER(T.StepResult)]
        L.pcc_get_constraints.argtypes = [vp, C.POINTER(T.ConstraintsC), C.POINTER(C.c_int64)]
        L.pcc_detect.argtypes = [vp, C.c_double, C.POINTER(T.ConstraintsC),
                                 C.POINTER(C.c_int64), C.POINTER(C.c_uint32),
                                 C.POINTER(C.c_int32)]
        L.pcc_set_timing.argtypes = [vp, C.c_int]
        L.pcc_get_timing.argtypes = [vp, C.c_int, C.POINTER(C.c_int64), C.POINTER(C.c_double),
                                     C.POINTER(C.c_double)]
        L.pcc_kernel_name.argtypes = [C.c_int]
        L.pcc_kernel_name.restype = C.c_char_p
        L.pcc_get_phase_stats.argtypes = [vp, C.POINTER(C.c_int64), C.POINTER(C.c_int64),
                                          C.c_int64, C.POINTER(C.c_int64)]
        i64 = C.c_int64
        L.pcc_batch_set_state.argtypes = [vp, i64, C.POINTER(T.BodiesC), C.POINTER(T.Params)]
        L.pcc_batch_add.argtypes = [vp, C.POINTER(T.BodiesC), C.POINTER(T.Params),
                                    C.POINTER(i64)]
        L.pcc_batch_remove.argtypes = [vp, i64]
        L.pcc_batch_set_params.argtypes = [vp, i64, C.POINTER(T.Params)]
        L.pcc_batch_size.argtypes = [vp, C.POINTER(i64), C.POINTER(i64)]
        L.pcc_batch_step.argtypes = [vp, C.c_int32, C.POINTER(T.StepResult),
                                     C.POINTER(C.c_int)]
        L.pcc_batch_get_state.argtypes = [vp, i64, C.POINTER(T.BodiesC)]
        L.pcc_batch_get_constraints.argtypes = [vp, i64, C.POINTER(T.ConstraintsC),
                                                C.POINTER(i64)]
        L.pcc_batch_get_phase_stats.argtypes = [vp, i64, C.POINTER(i64), C.POINTER(i64), i64,
                                                C.POINTER(i64)]
        L.pcc_device_bytes.argtypes = [vp, C.POINTER(i64)]
        L.pcc_batch_get_step_count.argtypes = [vp, i64, C.POINTER(i64)]
        L.pcc_batch_set_step_count.argtypes = [vp, i64, i64]
        L.pcc_get_tuning.argtypes = [vp, C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                                     C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                                     C.POINTER(C.c_double)]
        dpt = C.POINTER(C.c_double)
        L.pcc_adapt_dt.argtypes = [vp, C.c_double, C.c_double, dpt, dpt]
        L.pcc_batch_adapt_dt.argtypes = [vp, C.c_double, C.c_double, dpt, dpt]
        L.pcc_divide.argtypes = [vp, C.c_uint64, C.c_int32, C.c_double, C.c_double,
                                 C.POINTER(i64)]
        L.pcc_batch_divide.argtypes = [vp, C.POINTER(C.c_uint64), C.c_int32, C.c_double,
                                       C.c_double, C.POINTER(i64)]
        L.pcc_measure_copy_bandwidth.argtypes = [C.c_int, i64, C.c_int, C.POINTER(C.c_double)]
        L.pcc_measure_read_bandwidth.argtypes = [C.c_int, i64, C.c_int, C.POINTER(C.c_double)]
        L.pcc_get_solve_stats.argtypes = [vp, C.POINTER(i64)]
        L.pcc_set_trace.argtypes = [vp, C.c_int, i64]
        L.pcc_get_trace.argtypes = [vp, i64, vp, i64, C.POINTER(i64)]
        for f in ("pcc_create", "pcc_set_params", "pcc_set_state", "pcc_get_state",
                  "pcc_snapshot", "pcc_restore", "pcc_step", "pcc_get_constraints",
                  "pcc_detect", "pcc_set_timing", "pcc_get_timing", "pcc_get_phase_stats",
                  "pcc_batch_set_state", "pcc_batch_add", "pcc_batch_remove",
                  "pcc_batch_set_params", "pcc_batch_size", "pcc_batch_step",
                  "pcc_batch_get_state", "pcc_batch_get_constraints",
                  "pcc_batch_get_phase_stats", "pcc_device_bytes",
                  "pcc_measure_copy_bandwidth", "pcc_measure_read_bandwidth", "pcc_divide",
                  "pcc_batch_divide", "pcc_get_tuning", "pcc_get_solve_stats", "pcc_adapt_dt",
                  "pcc_batch_adapt_dt", "pcc_set_trace", "pcc_get_trace",
                  "pcc_batch_get_step_count", "pcc_batch_set_step_count"):
            getattr(L, f).restype = C.c_int
        _lib = L
    return _lib


EXPORTED = ("pcc_default_params", "pcc_create", "pcc_destroy", "pcc_set_params",
            "pcc_last_error", "pcc_status_string", "pcc_set_state", "pcc_get_state",
            "pcc_snapshot", "pcc_restore", "pcc_step", "pcc_get_constraints", "pcc_detect",
            "pcc_set_timing", "pcc_get_timing", "pcc_kernel_name", "pcc_get_phase_stats",
            "pcc_batch_set_state", "pcc_batch_add", "pcc_batch_remove", "pcc_batch_set_params",
            "pcc_batch_size", "pcc_batch_step", "pcc_batch_get_state",
            "pcc_batch_get_constraints", "pcc_batch_get_phase_stats", "pcc_device_bytes",
            "pcc_measure_copy_bandwidth", "pcc_measure_read_bandwidth", "pcc_divide",
            "pcc_batch_divide", "pcc_get_tuning", "pcc_get_solve_stats",
            "pcc_adapt_dt", "pcc_batch_adapt_dt", "pcc_set_trace", "pcc_get_trace",
            "pcc_batch_get_step_count", "pcc_batch_set_step_count")


def copy_bandwidth(nbytes: int = 1 << 30, reps: int = 10, device: int = 0) -> float:
    """Achievable device copy bandwidth in GB/s (bytes read + written)."""
    g = C.c_double(0)
    st = lib().pcc_measure_copy_bandwidth(int(device), int(nbytes), int(reps), C.byref(g))
    if st != T.PCC_OK:
        raise PccError(st, "pcc_measure_copy_bandwidth failed")
    return float(g.value)


def read_bandwidth(nbytes: int = 1 << 30, reps: int = 10, device: int = 0) -> float:
    """Achievable device read bandwidth in GB/s (bytes read)."""
    g = C.c_double(0)
    st = lib().pcc_measure_read_bandwidth(int(device), int(nbytes), int(reps), C.byref(g))
    if st != T.PCC_OK:
        raise PccError(st, "pcc_measure_read_bandwidth failed")
    return float(g.value)


class PccError(RuntimeError):
    def __init__(self, status, message):
        super().__init__(message)
        self.status = status


def default_params() -> T.Params:
    p = T.Params()
    lib().pcc_default_params(C.byref(p))
    return p


class ParticleManager:
    """Owns one libpcc context (one HIP stream on one device) and the colony's
    device-resident state."""

    def __init__(self, params: T.Params, device: int = 0):
        self._ctx = C.c_void_p()
        self.params = params.copy()
        st = lib().pcc_create(C.byref(self.params), int(device), C.byref(self._ctx))
        if st != T.PCC_OK:
            raise PccError(st, "pcc_create failed: " +
                           lib().pcc_status_string(st).decode())
        self.n = 0

    def close(self):
        if self._ctx:
            lib().pcc_destroy(self._ctx)
            self._ctx = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, st):
        if st != T.PCC_OK:
            msg = lib().pcc_last_error(self._ctx).decode() or lib().pcc_status_string(st).decode()
            raise PccError(st, msg)

    def set_params(self, params: T.Params):
        self.params = params.copy()
        self._check(lib().pcc_set_params(self._ctx, C.byref(self.params)))

    def set_state(self, bodies: T.Bodies):
        s = bodies.c_struct()
        self._check(lib().pcc_set_state(self._ctx, C.byref(s)))
        self.n = bodies.n

    def get_state(self) -> T.Bodies:
        n = self.n
        b = T.Bodies(np.zeros((n, 3)), np.zeros((n, 4)), np.zeros(n), np.zeros(n), np.zeros(n),
                     np.zeros(n))
        s = b.c_struct()
        self._check(lib().pcc_get_state(self._ctx, C.byref(s)))
        return b

    def adapt_dt(self, cfl_factor: float = 0.5, alpha: float = 0.01):
        """Domain::adaptDt on the device, with relcp_tol as the reference's
        tolerance; returns (new dt, median speed)."""
        dt, med = C.c_double(0), C.c_double(0)
        self._check(lib().pcc_adapt_dt(self._ctx, float(cfl_factor), float(alpha),
                                       C.byref(dt), C.byref(med)))
        self.params.dt = float(dt.value)
        return float(dt.value), float(med.value)

    def tuning(self) -> dict:
        """Autotuner state: the body-kernel variant in use (its index,
        staging size, threads per block, whether it prefetches the next
        pass), whether it was tuned, and the measured time of each of the
        BODY_VARIANTS candidates (us)."""
        va, ch, th, tu = C.c_int32(0), C.c_int32(0), C.c_int32(0), C.c_int32(0)
        us = (C.c_double * BODY_VARIANTS)()
        self._check(lib().pcc_get_tuning(self._ctx, C.byref(va), C.byref(ch), C.byref(th),
                                         C.byref(tu), us))
        return {"variant": int(va.value), "body_chunk": int(ch.value),
                "body_threads": int(th.value), "prefetch": int(va.value) >= 6,
                "tuned": bool(tu.value), "from_cache": int(tu.value) == 2,
                "variant_us": [float(x) for x in us]}

    def solve_stats(self) -> dict:
        """Counters of this context's BBPGD solves: grid compactions."""
        n = C.c_int64(0)
        self._check(lib().pcc_get_solve_stats(self._ctx, C.byref(n)))
        return {"compactions": int(n.value)}

    def set_trace(self, enable: bool = True, capacity: int = 100000):
        """BBPGD trace of the following steps (the reference's bbpgd_logger):
        one record per iteration of every phase, residual checked every
        iteration while on."""
        self._check(lib().pcc_set_trace(self._ctx, int(bool(enable)), int(capacity)))
        self._trace_cap = int(capacity) if enable else 0

    def trace(self, q: int = 0):
        """(records, iterations traced) of the last step for colony q:
        records is a numpy array of dtype abi_types.BBPGD_RECORD."""
        cnt = C.c_int64(0)
        self._check(lib().pcc_get_trace(self._ctx, int(q), None, 0, C.byref(cnt)))
        n = min(int(cnt.value), getattr(self, "_trace_cap", 0))  # records stored
        buf = np.zeros(max(n, 1), T.BBPGD_RECORD)
        self._check(lib().pcc_get_trace(self._ctx, int(q), buf.ctypes.data_as(C.c_void_p), n,
                                        C.byref(cnt)))
        return buf[:n].copy(), int(cnt.value)

    def step_count(self, q: int = 0) -> int:
        """Steps colony q has taken since it was set or added (the key of its
        Brownian draws)."""
        n = C.c_int64(0)
        self._check(lib().pcc_batch_get_step_count(self._ctx, int(q), C.byref(n)))
        return int(n.value)

    def set_step_count(self, count: int, q: int = 0):
        self._check(lib().pcc_batch_set_step_count(self._ctx, int(q), int(count)))

    def device_bytes(self) -> int:
        """Device memory held by this context."""
        b = C.c_int64(0)
        self._check(lib().pcc_device_bytes(self._ctx, C.byref(b)))
        return int(b.value)

    def divide(self, seed: int, step: int, length_noise: float = 0.0,
               angle_noise: float = 0.01) -> int:
        """Divides every cell with length >= 2 l0 on the device and renumbers
        the cells by x; returns the new cell count."""
        n = C.c_int64(0)
        self._check(lib().pcc_divide(self._ctx, C.c_uint64(seed & 0xFFFFFFFFFFFFFFFF), int(step),
                                     float(length_noise), float(angle_noise), C.byref(n)))
        self.n = int(n.value)
        return self.n

    def snapshot(self):
        self._check(lib().pcc_snapshot(self._ctx))

    def restore(self):
        self._check(lib().pcc_restore(self._ctx))

    def step(self, i: int = 0, raise_on_error: bool = True):
        """One timestep; returns the SolverSolution summary as a dict."""
        r = T.StepResult()
        st = lib().pcc_step(self._ctx, int(i), C.byref(r))
        out = r.as_dict()
        out["status"] = st
        if st != T.PCC_OK and raise_on_error:
            self._check(st)
        return out

    def constraints(self, count=None) -> dict:
        cnt = C.c_int64(0)
        probe = T.Constraints(1)
        ps = probe.c_struct()
        ps.capacity = 0
        self._check(lib().pcc_get_constraints(self._ctx, C.byref(ps), C.byref(cnt)))
        n = int(cnt.value)
        cons = T.Constraints(max(n, 1))
        cs = cons.c_struct()
        self._check(lib().pcc_get_constraints(self._ctx, C.byref(cs), C.byref(cnt)))
        cons.count = n
        return cons.trimmed()

    def detect(self, cutoff: float, capacity=None):
        """One detection batch at the current state: (constraints dict,
        bucket per body, stable body permutation by bucket)."""
        cnt = C.c_int64(0)
        cap = capacity if capacity is not None else 16 * max(self.n, 1) + 64
        cons = T.Constraints(cap)
        cs = cons.c_struct()
        bucket = np.zeros(max(self.n, 1), np.uint32)
        perm = np.zeros(max(self.n, 1), np.int32)
        self._check(lib().pcc_detect(self._ctx, float(cutoff), C.byref(cs), C.byref(cnt),
                                     bucket.ctypes.data_as(C.POINTER(C.c_uint32)),
                                     perm.ctypes.data_as(C.POINTER(C.c_int32))))
        if cnt.value > cap:
            return self.detect(cutoff, capacity=int(cnt.value))
        cons.count = int(cnt.value)
        return cons.trimmed(), bucket[: self.n], perm[: self.n]

    def phase_stats(self):
        """[(bbpgd_iterations, n_constraints)] per solved ReLCP phase of the
        last step."""
        it = np.zeros(256, np.int64)
        nc = np.zeros(256, np.int64)
        cnt = C.c_int64(0)
        self._check(lib().pcc_get_phase_stats(self._ctx, it.ctypes.data_as(C.POINTER(C.c_int64)),
                                              nc.ctypes.data_as(C.POINTER(C.c_int64)), 256,
                                              C.byref(cnt)))
        n = min(int(cnt.value), 256)
        return list(zip(it[:n].tolist(), nc[:n].tolist()))

    def set_timing(self, enable: bool):
        self._check(lib().pcc_set_timing(self._ctx, int(bool(enable))))

    def timing(self) -> dict:
        out = {}
        for k in range(T.K_COUNT):
            n = C.c_int64(0)
            ms = C.c_double(0)
            by = C.c_double(0)
            self._check(lib().pcc_get_timing(self._ctx, k, C.byref(n), C.byref(ms), C.byref(by)))
            out[lib().pcc_kernel_name(k).decode()] = {
                "launches": int(n.value), "total_ms": float(ms.value),
                "bytes_per_launch": float(by.value)}
        return out


class ColonyBatch(ParticleManager):
    """One libpcc context holding many independent colonies (same mode and
    precision, every other parameter per colony).  `step()` advances all of
    them with shared launches and returns one SolverSolution summary per
    colony; each colony's results equal those of a ParticleManager holding it
    alone.  Colony indices follow the order of set/add; `remove(q)` shifts the
    later colonies down."""

    def __init__(self, params: T.Params, device: int = 0):
        super().__init__(params, device)
        self.sizes = []
        self.colony_params = []  # host copy of every colony's parameters

    def set_state(self, colonies, params=None):
        cs = [b.c_struct() for b in colonies]
        arr = (T.BodiesC * len(cs))(*cs)
        parr = None
        if params is not None:
            parr = (T.Params * len(params))(*[p.copy() for p in params])
        self._check(lib().pcc_batch_set_state(self._ctx, len(cs), arr, parr))
        self.sizes = [b.n for b in colonies]
        self.colony_params = [(params[q] if params is not None else self.params).copy()
                              for q in range(len(cs))]
        self.n = sum(self.sizes)

    def add(self, bodies: T.Bodies, params: T.Params = None) -> int:
        s = bodies.c_struct()
        idx = C.c_int64(-1)
        pp = C.byref(params.copy()) if params is not None else None
        self._check(lib().pcc_batch_add(self._ctx, C.byref(s), pp, C.byref(idx)))
        self.sizes.append(bodies.n)
        self.colony_params.append((params if params is not None else self.params).copy())
        self.n = sum(self.sizes)
        return int(idx.value)

    def remove(self, q: int):
        self._check(lib().pcc_batch_remove(self._ctx, int(q)))
        del self.sizes[q]
        del self.colony_params[q]
        self.n = sum(self.sizes)

    def set_colony_params(self, q: int, params: T.Params):
        self._check(lib().pcc_batch_set_params(self._ctx, int(q), C.byref(params.copy())))
        self.colony_params[q] = params.copy()

    def __len__(self):
        return len(self.sizes)

    def step(self, i: int = 0):
        """One timestep of every colony: a list of SolverSolution dicts, each
        with the colony's solver status under "status"."""
        B = len(self.sizes)
        res = (T.StepResult * B)()
        sts = (C.c_int * B)()
        self._check(lib().pcc_batch_step(self._ctx, int(i), res, sts))
        out = []
        for q in range(B):
            d = res[q].as_dict()
            d["status"] = int(sts[q])
            out.append(d)
        return out

    def get_state(self, q: int) -> T.Bodies:
        n = self.sizes[q]
        b = T.Bodies(np.zeros((n, 3)), np.zeros((n, 4)), np.zeros(n), np.zeros(n), np.zeros(n),
                     np.zeros(n))
        s = b.c_struct()
        self._check(lib().pcc_batch_get_state(self._ctx, int(q), C.byref(s)))
        return b

    def constraints(self, q: int) -> dict:
        cnt = C.c_int64(0)
        probe = T.Constraints(1)
        ps = probe.c_struct()
        ps.capacity = 0
        self._check(lib().pcc_batch_get_constraints(self._ctx, int(q), C.byref(ps),
                                                    C.byref(cnt)))
        n = int(cnt.value)
        cons = T.Constraints(max(n, 1))
        cs = cons.c_struct()
        self._check(lib().pcc_batch_get_constraints(self._ctx, int(q), C.byref(cs),
                                                    C.byref(cnt)))
        cons.count = n
        return cons.trimmed()

    def phase_stats(self, q: int = 0):
        it = np.zeros(256, np.int64)
        nc = np.zeros(256, np.int64)
        cnt = C.c_int64(0)
        self._check(lib().pcc_batch_get_phase_stats(
            self._ctx, int(q), it.ctypes.data_as(C.POINTER(C.c_int64)),
            nc.ctypes.data_as(C.POINTER(C.c_int64)), 256, C.byref(cnt)))
        n = min(int(cnt.value), 256)
        return list(zip(it[:n].tolist(), nc[:n].tolist()))

    def detect(self, cutoff, capacity=None):
        raise PccError(T.PCC_E_INVALID_ARGUMENT, "detect() needs a single-colony context")

    def adapt_dt(self, cfl_factor: float = 0.5, alpha: float = 0.01):
        """Domain::adaptDt for every colony (its relcp_tol as the tolerance);
        returns (dts, medians)."""
        B = len(self.sizes)
        dts, med = (C.c_double * B)(), (C.c_double * B)()
        self._check(lib().pcc_batch_adapt_dt(self._ctx, float(cfl_factor), float(alpha), dts,
                                             med))
        for q in range(B):
            self.colony_params[q].dt = float(dts[q])
        return [float(x) for x in dts], [float(x) for x in med]

    def divide(self, seeds, step: int, length_noise: float = 0.0, angle_noise: float = 0.01):
        """Divides the cells of every colony (seeds[q] per colony) and
        renumbers each colony by x; returns the new colony sizes."""
        B = len(self.sizes)
        sd = (C.c_uint64 * B)(*[int(x) & 0xFFFFFFFFFFFFFFFF for x in seeds])
        out = (C.c_int64 * B)()
        self._check(lib().pcc_batch_divide(self._ctx, sd, int(step), float(length_noise),
                                           float(angle_noise), out))
        self.sizes = [int(x) for x in out]
        self.n = sum(self.sizes)
        return list(self.sizes)
