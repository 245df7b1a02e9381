"""ctypes mirrors of the structs declared in include/pcc.h.

The C ABI is the drop-in boundary of the collision pipeline; these classes only
describe its memory layout (no library is loaded here).
"""
import ctypes as C

import numpy as np

PCC_OK = 0
PCC_E_BBPGD_NOT_CONVERGED = 1
PCC_E_RELCP_MAX_ITERS = 2
PCC_E_INVALID_MODE = 3
PCC_E_INVALID_ARGUMENT = 4
PCC_E_DEVICE = 5
PCC_E_OUT_OF_MEMORY = 6
PCC_E_UNSUPPORTED = 7
PCC_E_NO_STATE = 8

MODE_HARD = 0
MODE_SOFT = 1
PREC_FP64 = 0
PREC_MIXED = 1

K_BODY_APPLY = 0
K_CONSTR_APPLY = 1
K_DETECT = 2
K_ASSEMBLE = 3
K_INTEGRATE = 4
K_HERTZ = 5
K_COUNT = 6

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)


class Params(C.Structure):
    """pcc_params: SimulationParameters (code/cpp/src/util/Config.h:9-56) with
    the single `tolerance` split into bbpgd_tol / relcp_tol / cutoffs."""

    _fields_ = [
        ("mode", C.c_int32),
        ("precision", C.c_int32),
        ("dt", C.c_double),
        ("xi", C.c_double),
        ("tau", C.c_double),
        ("lam", C.c_double),
        ("temperature", C.c_double),
        ("kcc", C.c_double),
        ("bbpgd_tol", C.c_double),
        ("relcp_tol", C.c_double),
        ("cutoff_phase0", C.c_double),
        ("cutoff_phaseK", C.c_double),
        ("max_bbpgd_iterations", C.c_int64),
        ("max_recursive_iterations", C.c_int64),
        ("linked_cell_size", C.c_double),
        ("grid_origin", C.c_double),
        ("warm_start_factor", C.c_double),
        ("monolayer", C.c_int32),
        ("reserved0", C.c_int32),
        ("brownian_seed", C.c_uint64),
    ]

    def copy(self):
        p = Params()
        C.pointer(p)[0] = self
        return p

    def replace(self, **kw):
        p = self.copy()
        for k, v in kw.items():
            setattr(p, k, v)
        return p


def reference_params(**kw) -> Params:
    """Reference defaults (code/cpp/src/util/CLIParser.h:167-194,232) with the
    single-tolerance coupling (Hard.cpp:121,132,153): temperature 1e-28 and
    monolayer on (CLIParser.h:181,183), Brownian seed 0.  Workloads with
    rotational diffusion off pass temperature=0."""
    tol = kw.pop("tolerance", 1e-3)
    p = Params(
        mode=MODE_HARD,
        precision=PREC_FP64,
        dt=1e-2,
        xi=1.0,
        tau=1.0,
        lam=1e-2,
        temperature=1e-28,
        kcc=20000.0,
        bbpgd_tol=tol,
        relcp_tol=tol,
        cutoff_phase0=0.25,
        cutoff_phaseK=tol,
        max_bbpgd_iterations=500000,
        max_recursive_iterations=50,
        linked_cell_size=2.2,
        grid_origin=-1.0,
        warm_start_factor=500.0,
        monolayer=1,
        reserved0=0,
        brownian_seed=0,
    )
    return p.replace(**kw) if kw else p


class BodiesC(C.Structure):
    _fields_ = [
        ("n", C.c_int64),
        ("pos", _dp),
        ("quat", _dp),
        ("length", _dp),
        ("l0", _dp),
        ("diameter", _dp),
        ("ldot", _dp),
        ("force", _dp),
        ("torque", _dp),
        ("velocity", _dp),
        ("omega", _dp),
        ("impedance", _dp),
        ("stress", _dp),
        ("num_constraints", _ip),
        ("age", _ip),
    ]


class ConstraintsC(C.Structure):
    _fields_ = [
        ("capacity", C.c_int64),
        ("gidI", _ip),
        ("gidJ", _ip),
        ("signed_distance", _dp),
        ("normal", _dp),
        ("rPosI", _dp),
        ("rPosJ", _dp),
        ("contact_point", _dp),
        ("stressI", _dp),
        ("stressJ", _dp),
        ("gamma", _dp),
        ("gid", _ip),
        ("iteration", _ip),
    ]


class StepResult(C.Structure):
    """SolverSolution minus the constraint vector (ParticleManager.h:32-38)."""

    _fields_ = [
        ("n_constraints", C.c_int64),
        ("constraint_iterations", C.c_int32),
        ("reserved", C.c_int32),
        ("bbpgd_iterations", C.c_int64),
        ("residual", C.c_double),
        ("max_overlap", C.c_double),
    ]

    def as_dict(self):
        return {
            "n_constraints": int(self.n_constraints),
            "constraint_iterations": int(self.constraint_iterations),
            "bbpgd_iterations": int(self.bbpgd_iterations),
            "residual": float(self.residual),
            "max_overlap": float(self.max_overlap),
        }


# pcc_bbpgd_record: BBPGDStep (logger/BBPGDLogger.h:7-17) with the energy
# terms of HardModelGradient::energy at the iterate
BBPGD_RECORD = np.dtype([
    ("step", "<i8"), ("recursive_iteration", "<i8"), ("total_constraints", "<i8"),
    ("residual", "<f8"), ("step_size", "<f8"), ("linear", "<f8"), ("quadratic", "<f8"),
    ("growth", "<f8"), ("total", "<f8"),
])


def _ptr(a, ct):
    if a is None:
        return C.cast(None, C.POINTER(ct))
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(C.POINTER(ct))


class Bodies:
    """Host-side body state in gid order (numpy, FP64; age, the generation of
    ParticleData.h:16, int32: 1 for a new cell; impedance, that of the last
    step, 0 for a new cell, is the input of the Brownian velocity)."""

    FIELDS_IN = ("pos", "quat", "length", "l0", "diameter", "ldot", "age", "impedance")

    def __init__(self, pos, quat, length, l0=None, diameter=None, ldot=None, age=None,
                 impedance=None):
        n = len(length)
        self.pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(n, 3)
        self.quat = np.ascontiguousarray(quat, dtype=np.float64).reshape(n, 4)
        self.length = np.ascontiguousarray(length, dtype=np.float64).reshape(n)
        self.l0 = np.ascontiguousarray(np.ones(n) if l0 is None else l0, dtype=np.float64)
        self.diameter = np.ascontiguousarray(
            np.full(n, 0.5) if diameter is None else diameter, dtype=np.float64)
        self.ldot = np.ascontiguousarray(np.zeros(n) if ldot is None else ldot, dtype=np.float64)
        self.force = np.zeros((n, 3))
        self.torque = np.zeros((n, 3))
        self.velocity = np.zeros((n, 3))
        self.omega = np.zeros((n, 3))
        self.impedance = np.ascontiguousarray(
            np.zeros(n) if impedance is None else impedance, dtype=np.float64).reshape(n)
        self.stress = np.zeros(n)
        self.num_constraints = np.zeros(n, dtype=np.int32)
        self.age = np.ascontiguousarray(np.ones(n) if age is None else age,
                                        dtype=np.int32).reshape(n)

    @property
    def n(self):
        return len(self.length)

    def copy(self):
        b = Bodies(self.pos.copy(), self.quat.copy(), self.length.copy(), self.l0.copy(),
                   self.diameter.copy(), self.ldot.copy(), self.age.copy())
        for f in ("force", "torque", "velocity", "omega", "impedance", "stress",
                  "num_constraints"):
            setattr(b, f, getattr(self, f).copy())
        return b

    def c_struct(self) -> BodiesC:
        s = BodiesC()
        s.n = self.n
        for f in ("pos", "quat", "length", "l0", "diameter", "ldot", "force", "torque",
                  "velocity", "omega", "impedance", "stress"):
            setattr(s, f, _ptr(getattr(self, f), C.c_double))
        s.num_constraints = _ptr(self.num_constraints, C.c_int32)
        s.age = _ptr(self.age, C.c_int32)
        return s

    def subset(self, idx):
        idx = np.asarray(idx)
        return Bodies(self.pos[idx], self.quat[idx], self.length[idx], self.l0[idx],
                      self.diameter[idx], self.ldot[idx], self.age[idx], self.impedance[idx])


class Constraints:
    """Host-side constraint set (SoA image of std::vector<Constraint>,
    code/cpp/src/dynamics/Constraint.h:6-49)."""

    def __init__(self, capacity):
        c = int(max(capacity, 1))
        self.capacity = c
        self.gidI = np.zeros(c, np.int32)
        self.gidJ = np.zeros(c, np.int32)
        self.signed_distance = np.zeros(c)
        self.normal = np.zeros((c, 3))
        self.rPosI = np.zeros((c, 3))
        self.rPosJ = np.zeros((c, 3))
        self.contact_point = np.zeros((c, 3))
        self.stressI = np.zeros(c)
        self.stressJ = np.zeros(c)
        self.gamma = np.zeros(c)
        self.gid = np.zeros(c, np.int32)
        self.iteration = np.zeros(c, np.int32)
        self.count = 0

    def c_struct(self) -> ConstraintsC:
        s = ConstraintsC()
        s.capacity = self.capacity
        for f in ("gidI", "gidJ", "gid", "iteration"):
            setattr(s, f, _ptr(getattr(self, f), C.c_int32))
        for f in ("signed_distance", "normal", "rPosI", "rPosJ", "contact_point", "stressI",
                  "stressJ", "gamma"):
            setattr(s, f, _ptr(getattr(self, f), C.c_double))
        return s

    def trimmed(self, count=None):
        """Dict of arrays cut to `count` (default: self.count)."""
        n = self.count if count is None else count
        n = min(n, self.capacity)
        return {f: getattr(self, f)[:n].copy() for f in
                ("gidI", "gidJ", "signed_distance", "normal", "rPosI", "rPosJ",
                 "contact_point", "stressI", "stressJ", "gamma", "gid", "iteration")}
