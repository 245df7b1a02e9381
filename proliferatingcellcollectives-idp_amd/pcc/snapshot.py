"""Snapshots of colony state: a lossless binary checkpoint and the reference's
VTU/PVTU output.

Checkpoint (`save_checkpoint` / `load_checkpoint`)
    One file per snapshot: the 8-byte magic ``PCCSNAP1``, a little-endian
    uint64 header length, a UTF-8 JSON header, then one section per body field
    in structure-of-arrays form, every section starting on a 256-byte
    boundary.  The sections are the FP64 (and int32) arrays the C ABI's
    pcc_bodies struct points at, colony after colony, so a loader can
    memory-map the file and hand the arrays to pcc_set_state /
    pcc_batch_set_state without a copy.  Every float is stored bit for bit, so
    a run resumed from a checkpoint steps exactly as the uninterrupted run
    (the reference's VTU files keep 6 significant digits).  The header holds
    the step index, the simulation time, each colony's size and its
    parameters.

VTU / PVTU (`VTKDataLogger`, `ParticleLogger`, `ConstraintLogger`,
`SimulationLogger`, `BBPGDLogger`, `VTKStateLoader`)
    Files laid out as the reference writes them (code/cpp/src/logger/VTK.h:
    180-313 for the directory layout and the XML, ParticleLogger.cpp:61-74,
    ConstraintLogger.cpp, SimulationLogger.cpp:14-29 for the array names), so
    ParaView, the reference's analysis notebooks (analysis/parse_vtu.py) and
    its resume path (code/cpp/src/loader/VTKStateLoader.cpp) read them:

        <dir>/<prefix>_<step:07d>.pvtu
        <dir>/data/<prefix>_<step:07d>_rank_<rank:04d>.vtu

    ``precision="reference"`` prints numbers as the reference's ostreams do
    (6 significant digits, positions as Float32); ``precision="lossless"``
    prints every float with the shortest text that reads back to the same
    double and declares positions Float64, so a VTU round trip is exact.
    Field data is written without exponents either way: the reference's
    field-data reader (VTKStateLoader.cpp:318-319) takes digits, '.' and '-'
    only.
"""
from __future__ import annotations

import json
import os
import re
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence

import numpy as np

from . import abi_types as T

# --------------------------------------------------------------------------
# binary checkpoint
# --------------------------------------------------------------------------

MAGIC = b"PCCSNAP1"
VERSION = 1
ALIGN = 256

# (field, dtype, components); gid is the index inside the colony
CHECKPOINT_FIELDS = (
    ("pos", "<f8", 3),
    ("quat", "<f8", 4),
    ("length", "<f8", 1),
    ("l0", "<f8", 1),
    ("diameter", "<f8", 1),
    ("ldot", "<f8", 1),
    ("force", "<f8", 3),
    ("torque", "<f8", 3),
    ("velocity", "<f8", 3),
    ("omega", "<f8", 3),
    ("impedance", "<f8", 1),
    ("stress", "<f8", 1),
    ("num_constraints", "<i4", 1),
    ("age", "<i4", 1),
)
# sections a checkpoint may lack (written before they existed): their default
OPTIONAL_FIELDS = {"age": 1}


def params_to_dict(p: T.Params) -> dict:
    return {name: getattr(p, name) for name, _ in T.Params._fields_}


def params_from_dict(d: dict) -> T.Params:
    known = {name for name, _ in T.Params._fields_}
    unknown = set(d) - known
    if unknown:
        raise ValueError(f"checkpoint parameters not in pcc_params: {sorted(unknown)}")
    missing = known - set(d)
    if missing:
        raise ValueError(f"checkpoint parameters missing: {sorted(missing)}")
    p = T.Params()
    for k, v in d.items():
        setattr(p, k, v)
    return p


@dataclass
class Checkpoint:
    colonies: List[T.Bodies]
    params: Optional[List[T.Params]]
    step: int = 0
    simulation_time_s: float = 0.0
    user: dict = field(default_factory=dict)
    step_counts: Optional[List[int]] = None  # each colony's step count (Brownian key)


def _pad(n: int) -> int:
    return (n + ALIGN - 1) // ALIGN * ALIGN


def save_checkpoint(path: str, colonies, params=None, step: int = 0,
                    simulation_time_s: float = 0.0, user: Optional[dict] = None,
                    step_counts=None) -> int:
    """Writes `colonies` (one Bodies or a list of them) with their `params`
    (one Params per colony, or one for all, or None) and, if given, each
    colony's step count; returns the file size."""
    if isinstance(colonies, T.Bodies):
        colonies = [colonies]
    colonies = list(colonies)
    if params is not None and isinstance(params, T.Params):
        params = [params] * len(colonies)
    if params is not None and len(params) != len(colonies):
        raise ValueError("one Params per colony")
    sizes = [b.n for b in colonies]
    n = sum(sizes)
    header = {
        "format": "pcc-snapshot",
        "version": VERSION,
        "n_bodies": n,
        "colony_sizes": sizes,
        "step": int(step),
        "simulation_time_s": float(simulation_time_s),
        "params": None if params is None else [params_to_dict(p) for p in params],
        "step_counts": None if step_counts is None else [int(x) for x in step_counts],
        "user": user or {},
        "fields": {},
    }
    # section offsets depend on the header length, which depends on the
    # offsets' digits: lay out against a generous header bound
    layout = []
    for name, dt, comp in CHECKPOINT_FIELDS:
        layout.append((name, dt, comp, n * comp * np.dtype(dt).itemsize))
    probe = dict(header)
    probe["fields"] = {name: {"dtype": dt, "shape": [n, comp] if comp > 1 else [n],
                              "offset": 10**15} for name, dt, comp, _ in layout}
    start = _pad(16 + len(json.dumps(probe).encode()) + 64)
    off = start
    for name, dt, comp, nbytes in layout:
        header["fields"][name] = {"dtype": dt, "shape": [n, comp] if comp > 1 else [n],
                                  "offset": off}
        off = _pad(off + nbytes)
    hb = json.dumps(header).encode()
    assert 16 + len(hb) <= start
    with open(path, "wb") as f:
        f.write(MAGIC)
        f.write(np.uint64(len(hb)).tobytes())
        f.write(hb)
        for name, dt, comp, nbytes in layout:
            f.seek(header["fields"][name]["offset"])
            for b in colonies:
                a = np.ascontiguousarray(getattr(b, name), dtype=dt)
                f.write(a.tobytes())
        f.truncate(off)
    return off


def read_checkpoint_header(path: str) -> dict:
    with open(path, "rb") as f:
        magic = f.read(8)
        if magic != MAGIC:
            raise ValueError(f"{path}: not a pcc checkpoint (magic {magic!r})")
        hl = int(np.frombuffer(f.read(8), "<u8")[0])
        size = os.fstat(f.fileno()).st_size
        if hl > size - 16:
            raise ValueError(f"{path}: truncated header")
        header = json.loads(f.read(hl).decode())
    if header.get("format") != "pcc-snapshot" or header.get("version") != VERSION:
        raise ValueError(f"{path}: unsupported checkpoint version {header.get('version')}")
    n = header["n_bodies"]
    if sum(header["colony_sizes"]) != n:
        raise ValueError(f"{path}: colony sizes do not add up to n_bodies")
    for name, dt, comp in CHECKPOINT_FIELDS:
        s = header["fields"].get(name)
        if s is None and name in OPTIONAL_FIELDS:
            continue
        if s is None:
            raise ValueError(f"{path}: missing section {name}")
        if s["dtype"] != dt or s["shape"] != ([n, comp] if comp > 1 else [n]):
            raise ValueError(f"{path}: section {name} has the wrong type or shape")
        if s["offset"] % ALIGN or s["offset"] + n * comp * np.dtype(dt).itemsize > size:
            raise ValueError(f"{path}: section {name} out of bounds")
    return header


def load_checkpoint(path: str, mmap: bool = False) -> Checkpoint:
    """Reads a checkpoint.  With mmap=True every array is a read-only view of
    the mapped file (the colonies share it)."""
    h = read_checkpoint_header(path)
    n = h["n_bodies"]
    arrays = {}
    for name, dt, comp in CHECKPOINT_FIELDS:
        s = h["fields"].get(name)
        if s is None:  # an optional section this file predates
            arrays[name] = np.full(n, OPTIONAL_FIELDS[name], dtype=dt)
            continue
        shape = tuple(s["shape"])
        if mmap:
            arrays[name] = np.memmap(path, dtype=dt, mode="r", offset=s["offset"], shape=shape)
        else:
            with open(path, "rb") as f:
                f.seek(s["offset"])
                arrays[name] = np.fromfile(f, dtype=dt, count=n * comp).reshape(shape)
    colonies = []
    o = 0
    for sz in h["colony_sizes"]:
        sl = slice(o, o + sz)
        b = T.Bodies(arrays["pos"][sl], arrays["quat"][sl], arrays["length"][sl],
                     arrays["l0"][sl], arrays["diameter"][sl], arrays["ldot"][sl],
                     arrays["age"][sl])
        for f_ in ("force", "torque", "velocity", "omega", "impedance", "stress",
                   "num_constraints"):
            setattr(b, f_, np.ascontiguousarray(arrays[f_][sl]))
        colonies.append(b)
        o += sz
    params = None if h["params"] is None else [params_from_dict(d) for d in h["params"]]
    return Checkpoint(colonies, params, h["step"], h["simulation_time_s"], h["user"],
                      h.get("step_counts"))


def checkpoint_context(pm, path: str, step: int = 0, simulation_time_s: float = 0.0,
                       user: Optional[dict] = None) -> int:
    """Writes the state of a ParticleManager (one colony) or a ColonyBatch
    (every colony, with its own parameters) after a step."""
    from .engine import ColonyBatch
    if isinstance(pm, ColonyBatch):
        cols = [pm.get_state(q) for q in range(len(pm))]
        counts = [pm.step_count(q) for q in range(len(pm))]
        return save_checkpoint(path, cols, list(pm.colony_params), step, simulation_time_s, user,
                               counts)
    return save_checkpoint(path, [pm.get_state()], [pm.params], step, simulation_time_s, user,
                           [pm.step_count()])


def restore_context(path: str, device: int = 0, batch: Optional[bool] = None):
    """Opens a context on `device` holding the checkpoint's colonies: a
    ParticleManager for a single colony (unless batch=True), else a
    ColonyBatch.  Returns (context, Checkpoint)."""
    from .engine import ColonyBatch, ParticleManager
    ck = load_checkpoint(path)
    if ck.params is None:
        raise ValueError(f"{path}: checkpoint holds no parameters; load_checkpoint() it and "
                         "pass your own")
    if batch is None:
        batch = len(ck.colonies) != 1
    if not batch:
        if len(ck.colonies) != 1:
            raise ValueError("a ParticleManager holds one colony")
        pm = ParticleManager(ck.params[0], device=device)
        pm.set_state(ck.colonies[0])
        if ck.step_counts:
            pm.set_step_count(ck.step_counts[0])
        return pm, ck
    cb = ColonyBatch(ck.params[0], device=device)
    cb.set_state(ck.colonies, ck.params)
    for q, n in enumerate(ck.step_counts or []):
        cb.set_step_count(n, q)
    return cb, ck


# --------------------------------------------------------------------------
# VTU / PVTU
# --------------------------------------------------------------------------

_VTK_TYPES = {
    np.dtype(np.int32): "Int32",
    np.dtype(np.float32): "Float32",
    np.dtype(np.float64): "Float64",
    np.dtype(np.uint8): "UInt8",
    np.dtype(np.int64): "Int64",
    np.dtype(np.uint64): "UInt64",
}
_NP_TYPES = {v: k for k, v in _VTK_TYPES.items()}
PRECISIONS = ("reference", "lossless")


def _fmt_values(a: np.ndarray, precision: str) -> List[str]:
    """Text of every value of `a` (flattened): %g with 6 digits as an
    ostream prints by default, or the shortest round-trip text."""
    flat = a.reshape(-1)
    if flat.dtype.kind in "iu":
        return [str(x) for x in flat.tolist()]
    if precision == "reference":
        return ["%g" % x for x in flat.tolist()]
    return [repr(x) for x in flat.astype(np.float64).tolist()]


def _fmt_scalar_positional(x, precision: str) -> str:
    if isinstance(x, (int, np.integer)):
        return str(int(x))
    x = float(x)
    if precision == "reference":
        s = "%g" % x
        if "e" not in s:
            return s
        # the reference prints an exponent here; its own reader stops at the
        # 'e', so keep the digits without one
        return np.format_float_positional(float(s), trim="-")
    return np.format_float_positional(x, unique=True, trim="-")


class _Array:
    def __init__(self, name: str, data: np.ndarray, components: int):
        self.name = name
        self.data = data
        self.components = components

    @property
    def vtk_type(self) -> str:
        t = _VTK_TYPES.get(self.data.dtype)
        if t is None:
            raise TypeError(f"VTK array {self.name}: unsupported dtype {self.data.dtype}")
        return t

    @property
    def tuples(self) -> int:
        return self.data.size // self.components


class VTKDataLogger:
    """Writes one VTU piece per rank and the PVTU index per step, with the
    reference's VTKDataLogger layout and life cycle (VTK.h:180-313): the
    constructor creates <dir>/data and, unless preserve_existing, removes this
    prefix's old .pvtu/.vtu files; write() emits the current step and clears
    the grid."""

    def __init__(self, output_directory: str, file_prefix: str, only_rank_zero: bool = False,
                 preserve_existing: bool = False, step: int = 0, rank: int = 0, size: int = 1,
                 precision: str = "reference"):
        if precision not in PRECISIONS:
            raise ValueError(f"precision must be one of {PRECISIONS}")
        self.output_directory = output_directory
        self.file_prefix = file_prefix
        self.only_rank_zero = only_rank_zero
        self.step = int(step)
        self.rank = int(rank)
        self.size = int(size)
        self.precision = precision
        self._reset()
        if self.rank == 0:
            data_dir = os.path.join(output_directory, "data")
            os.makedirs(data_dir, exist_ok=True)
            if not preserve_existing:
                for d, ext in ((output_directory, ".pvtu"), (data_dir, ".vtu")):
                    for f in os.listdir(d):
                        p = os.path.join(d, f)
                        if f.startswith(file_prefix) and f.endswith(ext) and os.path.isfile(p):
                            os.remove(p)

    def _reset(self):
        self.points = np.zeros((0, 3))
        self.cells: List[Sequence[int]] = []
        self.point_data: List[_Array] = []
        self.field_data: List[_Array] = []

    def add_point_data(self, name: str, data, num_components: Optional[int] = None):
        a = np.asarray(data)
        if num_components is None:
            num_components = 1 if a.ndim == 1 else int(a.shape[-1])
        self.point_data.append(_Array(name, np.ascontiguousarray(a), num_components))

    def add_field_data(self, name: str, value):
        if isinstance(value, (bool, np.bool_)):
            raise TypeError(f"field {name}: booleans have no VTK type here")
        a = np.asarray([value])
        if a.dtype.kind == "f":
            a = a.astype(np.float64)
        elif a.dtype.kind in "iu":
            a = a.astype(np.int64 if a.dtype.kind == "i" else np.uint64)
        self.field_data.append(_Array(name, a, 1))

    def add_points(self, positions):
        self.points = np.asarray(positions, dtype=np.float64).reshape(-1, 3)

    def add_cells(self, connectivity):
        self.cells = [list(c) for c in connectivity]

    def _basename(self) -> str:
        return f"{self.file_prefix}_{self.step:07d}"

    def vtu_path(self, rank: Optional[int] = None) -> str:
        r = self.rank if rank is None else rank
        return os.path.join(self.output_directory, "data", f"{self._basename()}_rank_{r:04d}.vtu")

    def pvtu_path(self) -> str:
        return os.path.join(self.output_directory, self._basename() + ".pvtu")

    def write(self):
        if not (self.only_rank_zero and self.rank != 0):
            self._write_vtu(self.vtu_path())
        if self.rank == 0:
            self._write_pvtu(self.pvtu_path())
        self.step += 1
        self._reset()

    # -- XML --------------------------------------------------------------
    def _points_type(self) -> str:
        return "Float32" if self.precision == "reference" else "Float64"

    def _write_vtu(self, path: str):
        pr = self.precision
        pts = self.points.astype(np.float32) if pr == "reference" else self.points
        out = ['<?xml version="1.0"?>\n',
               '<VTKFile type="UnstructuredGrid" version="0.1" byte_order="LittleEndian">\n',
               "  <UnstructuredGrid>\n"]
        if self.field_data:
            out.append("    <FieldData>\n")
            for a in self.field_data:
                out.append(f'      <DataArray type="{a.vtk_type}" Name="{a.name}" '
                           f'NumberOfTuples="{a.tuples}" format="ascii">\n')
                out.append("        " + " ".join(_fmt_scalar_positional(x, pr)
                                                 for x in a.data.tolist()) + "\n")
                out.append("      </DataArray>\n")
            out.append("    </FieldData>\n")
        out.append(f'    <Piece NumberOfPoints="{len(pts)}" NumberOfCells="{len(self.cells)}">\n')
        out.append("      <Points>\n")
        out.append(f'        <DataArray type="{self._points_type()}" Name="Points" '
                   'NumberOfComponents="3" format="ascii">\n')
        v = _fmt_values(pts, pr)
        out.extend(f"          {v[3 * i]} {v[3 * i + 1]} {v[3 * i + 2]}\n"
                   for i in range(len(pts)))
        out.append("        </DataArray>\n")
        out.append("      </Points>\n")
        out.append("      <Cells>\n")
        conn, offs, o = [], [], 0
        for c in self.cells:
            conn.append("          " + "".join(f"{i} " for i in c) + "\n")
            o += len(c)
            offs.append(f"          {o}\n")
        out.append('        <DataArray type="Int32" Name="connectivity" format="ascii">\n')
        out.extend(conn)
        out.append("        </DataArray>\n")
        out.append('        <DataArray type="Int32" Name="offsets" format="ascii">\n')
        out.extend(offs)
        out.append("        </DataArray>\n")
        out.append('        <DataArray type="UInt8" Name="types" format="ascii">\n')
        out.extend("          1\n" for _ in self.cells)  # VTK_VERTEX
        out.append("        </DataArray>\n")
        out.append("      </Cells>\n")
        if self.point_data:
            out.append("      <PointData>\n")
            for a in self.point_data:
                out.append(f'        <DataArray type="{a.vtk_type}" Name="{a.name}" '
                           f'NumberOfComponents="{a.components}" format="ascii">\n')
                v = _fmt_values(a.data, pr)
                k = a.components
                if k == 1:
                    body = " ".join(v)
                else:
                    body = "\n          ".join(" ".join(v[i * k:(i + 1) * k])
                                               for i in range(a.tuples))
                out.append("          " + body + "\n")
                out.append("        </DataArray>\n")
            out.append("      </PointData>\n")
        out.append("    </Piece>\n")
        out.append("  </UnstructuredGrid>\n")
        out.append("</VTKFile>\n")
        with open(path, "w") as f:
            f.write("".join(out))

    def _write_pvtu(self, path: str):
        out = ['<?xml version="1.0"?>\n',
               '<VTKFile type="PUnstructuredGrid" version="0.1" byte_order="LittleEndian">\n',
               '  <PUnstructuredGrid GhostLevel="0">\n']
        for tag, arrays in (("PFieldData", self.field_data), ("PPointData", self.point_data)):
            if arrays:
                out.append(f"    <{tag}>\n")
                for a in arrays:
                    out.append(f'      <PDataArray type="{a.vtk_type}" Name="{a.name}" '
                               f'NumberOfComponents="{a.components}"/>\n')
                out.append(f"    </{tag}>\n")
        out.append("    <PPoints>\n")
        out.append(f'      <PDataArray type="{self._points_type()}" Name="Points" '
                   'NumberOfComponents="3"/>\n')
        out.append("    </PPoints>\n")
        for r in range(1 if self.only_rank_zero else self.size):
            out.append(f'    <Piece Source="data/{self._basename()}_rank_{r:04d}.vtu"/>\n')
        out.append("  </PUnstructuredGrid>\n")
        out.append("</VTKFile>\n")
        with open(path, "w") as f:
            f.write("".join(out))


def direction_vectors(quat: np.ndarray) -> np.ndarray:
    """rotate(q, e_x) for every row of `quat` ((w, x, y, z)), in the
    expression order of Quaternion.cpp:7-38."""
    q = np.asarray(quat, dtype=np.float64).reshape(-1, 4)
    w, x, y, z = q[:, 0], q[:, 1], q[:, 2], q[:, 3]
    vx, vy, vz = 1.0, 0.0, 0.0
    c1x = y * vz - z * vy
    c1y = z * vx - x * vz
    c1z = x * vy - y * vx
    c2x = y * c1z - z * c1y
    c2y = z * c1x - x * c1z
    c2z = x * c1y - y * c1x
    return np.stack([vx + 2.0 * w * c1x + 2.0 * c2x,
                     vy + 2.0 * w * c1y + 2.0 * c2y,
                     vz + 2.0 * w * c1z + 2.0 * c2z], axis=1)


class ParticleLogger:
    """ParticleLogger::log (ParticleLogger.cpp:20-77): one VTU point per cell
    with the reference's fourteen arrays.  `age` is the generation
    (Particle::getAge, ParticleLogger.cpp:51): bodies.age unless given."""

    def __init__(self, output_directory: str, base_filename: str = "particles",
                 preserve_existing: bool = False, step: int = 0, precision: str = "reference",
                 rank: int = 0, size: int = 1):
        self.logger = VTKDataLogger(output_directory, base_filename, False, preserve_existing,
                                    step, rank, size, precision)

    def log(self, bodies: T.Bodies, age=None, gid_offset: int = 0):
        n = bodies.n
        t = direction_vectors(bodies.quat)
        ang = np.arctan2(t[:, 1], t[:, 0])
        ang = np.where(ang < 0, ang + np.pi, ang)
        L = self.logger
        L.add_points(bodies.pos)
        L.add_point_data("gid", np.arange(gid_offset, gid_offset + n, dtype=np.int32))
        L.add_point_data("quaternion", bodies.quat, 4)
        L.add_point_data("orientation", t, 3)
        L.add_point_data("orientation_angle", ang)
        L.add_point_data("lengths", np.stack([bodies.length, bodies.diameter, bodies.diameter],
                                             axis=1), 3)
        L.add_point_data("impedance", bodies.impedance)
        L.add_point_data("ldot", bodies.ldot)
        L.add_point_data("stress", bodies.stress)
        L.add_point_data("number_of_constraints", bodies.num_constraints.astype(np.int32))
        L.add_point_data("age", np.asarray(bodies.age if age is None else age, dtype=np.int32))
        L.add_point_data("forces", bodies.force, 3)
        L.add_point_data("velocity_linear", bodies.velocity, 3)
        L.add_point_data("velocity_angular", bodies.omega, 3)
        L.add_point_data("rank", np.full(n, L.rank, np.int32))
        L.write()


class ConstraintLogger:
    """ConstraintLogger::log (ConstraintLogger.cpp): one VTU point per
    contact at its contact point, from a constraints dict of the engine."""

    def __init__(self, output_directory: str, base_filename: str = "constraints",
                 preserve_existing: bool = False, step: int = 0, precision: str = "reference",
                 rank: int = 0, size: int = 1):
        self.logger = VTKDataLogger(output_directory, base_filename, False, preserve_existing,
                                    step, rank, size, precision)

    def log(self, cons: Dict[str, np.ndarray]):
        n = len(cons["gidI"])
        L = self.logger
        L.add_points(cons["contact_point"])
        L.add_point_data("gidI", cons["gidI"].astype(np.int32))
        L.add_point_data("gidJ", cons["gidJ"].astype(np.int32))
        L.add_point_data("gids", cons["gid"].astype(np.int32))
        L.add_point_data("signed_distance", cons["signed_distance"].astype(np.float64))
        L.add_point_data("normals", cons["normal"], 3)
        # the reference stores the phase index in a double array
        L.add_point_data("iteration", cons["iteration"].astype(np.float64))
        L.add_point_data("gamma", cons["gamma"].astype(np.float64))
        L.add_point_data("rank", np.full(n, L.rank, np.int32))
        L.write()


class BBPGDLogger:
    """BBPGDLogger::log (BBPGDLogger.cpp): one rank-0 VTU per traced BBPGD
    iteration, its record as field data (step, recursive_iteration,
    total_constraints, residual, step_size, linear, quadratic, growth, total),
    under the reference's "logs/bbpgdtrace_<k>" names (Hard.cpp:110).
    `records` is what ParticleManager.trace() returns."""

    FIELDS = ("step", "recursive_iteration", "total_constraints", "residual", "step_size",
              "linear", "quadratic", "growth", "total")

    def __init__(self, output_directory: str = "logs", base_filename: str = "bbpgdtrace",
                 step: int = 0, precision: str = "reference"):
        self.logger = VTKDataLogger(output_directory, base_filename, True, False, step, 0, 1,
                                    precision)

    def log(self, records):
        for r in records:
            for f in self.FIELDS:
                v = r[f]
                self.logger.add_field_data(f, int(v) if np.issubdtype(type(v), np.integer)
                                           else float(v))
            self.logger.write()


# SimulationStep (SimulationLogger.h:7-30): name -> integer-valued
SIMULATION_FIELDS = (
    ("simulation_time_s", False), ("time_since_last_log_s", False), ("step", True),
    ("num_particles", True), ("num_constraints", True), ("colony_radius", False),
    ("recursive_iterations", True), ("bbpgd_iterations", True), ("max_overlap", False),
    ("residual", False), ("memory_usage_mb", False), ("peak_memory_mb", False),
    ("cpu_time_s", False), ("mpi_comm_time_s", False), ("load_imbalance", False),
    ("dt_s", False),
)


class SimulationLogger:
    """SimulationLogger::log (SimulationLogger.cpp:8-31): the step's metrics
    as field data of a rank-0-only VTU.  Missing metrics are written as 0."""

    def __init__(self, output_directory: str, base_filename: str = "simulation",
                 preserve_existing: bool = False, step: int = 0, precision: str = "reference"):
        self.logger = VTKDataLogger(output_directory, base_filename, True, preserve_existing,
                                    step, 0, 1, precision)

    def log(self, metrics: Dict[str, float]):
        unknown = set(metrics) - {n for n, _ in SIMULATION_FIELDS}
        if unknown:
            raise ValueError(f"not a SimulationStep field: {sorted(unknown)}")
        for name, is_int in SIMULATION_FIELDS:
            v = metrics.get(name, 0)
            self.logger.add_field_data(name, int(v) if is_int else float(v))
        self.logger.write()


# -- loader -----------------------------------------------------------------

_ARRAY_RE = re.compile(r"<DataArray\b([^>]*?)/?>(?:(.*?)</DataArray>)?", re.S)
_ATTR_RE = re.compile(r'(\w+)="([^"]*)"')
_SECTION_RE = {s: re.compile(rf"<{s}>(.*?)</{s}>", re.S)
               for s in ("FieldData", "PointData", "Points")}
_STEP_RE = re.compile(r"simulation_(\d{7})\.pvtu")


@dataclass
class VTUArray:
    name: str
    type: str
    num_components: int
    data: np.ndarray


@dataclass
class LoadedState:
    """LoadedState (VTKStateLoader.h:14-19) with the cells as Bodies in gid
    order; `gids` holds the file's gids in that order."""
    bodies: T.Bodies
    gids: np.ndarray
    simulation_time_s: float
    dt_s: float
    step: int


def _parse_arrays(text: str) -> Dict[str, VTUArray]:
    out = {}
    for m in _ARRAY_RE.finditer(text):
        attrs = dict(_ATTR_RE.findall(m.group(1)))
        name = attrs.get("Name")
        if name is None:
            continue
        typ = attrs.get("type", "Float64")
        comp = int(attrs.get("NumberOfComponents", "1"))
        body = m.group(2) or ""
        dt = _NP_TYPES.get(typ, np.dtype(np.float64))
        vals = body.split()
        data = np.array(vals, dtype=dt if dt.kind in "iu" else np.float64)
        if comp > 1:
            if data.size % comp:
                raise ValueError(f"array {name}: {data.size} values for {comp} components")
            data = data.reshape(-1, comp)
        out[name] = VTUArray(name, typ, comp, data)
    return out


def _strip_debug(text: str) -> str:
    # VTKStateLoader skips lines that start with "line:"
    return re.sub(r"(?m)^line:.*(?:\n|$)", "", text)


class VTKStateLoader:
    """Resume from VTU output (VTKStateLoader.cpp:26-162): the latest (or a
    given) step's simulation_<step>.pvtu field data and particles_<step>.pvtu
    pieces.  As in the reference, l0 is 1 and the diameter 0.5 for every
    loaded cell (VTKStateLoader.cpp:131-136), torque and angular parts of the
    force are 0, and ldot is not read."""

    def __init__(self, vtk_directory: str):
        self.vtk_directory = vtk_directory
        self.data_dir = os.path.join(vtk_directory, "data")
        if not os.path.isdir(vtk_directory) or not os.path.isdir(self.data_dir):
            raise FileNotFoundError("VTK directory or data subdirectory does not exist")

    def find_latest_timestep(self) -> Optional[int]:
        best = None
        for f in os.listdir(self.vtk_directory):
            m = _STEP_RE.fullmatch(f)
            if m and os.path.isfile(os.path.join(self.vtk_directory, f)):
                k = int(m.group(1))
                best = k if best is None or k > best else best
        return best

    def get_latest_timestep(self) -> int:
        k = self.find_latest_timestep()
        if k is None:
            raise FileNotFoundError("No timesteps found in VTK directory")
        return k

    def load_latest_state(self) -> LoadedState:
        return self.load_state(self.get_latest_timestep())

    def load_state(self, timestep: int) -> LoadedState:
        sim = os.path.join(self.vtk_directory, f"simulation_{timestep:07d}.pvtu")
        fd = self.read_field_data(sim)
        bodies, gids = self.load_particles(timestep)
        return LoadedState(bodies, gids, float(fd.get("simulation_time_s", 0.0)),
                           float(fd.get("dt_s", 0.0)), int(fd.get("step", 0)))

    def get_piece_files(self, pvtu_file: str) -> List[str]:
        with open(pvtu_file) as f:
            return re.findall(r'<Piece\s+Source="([^"]+)"', f.read())

    def read_vtu_file(self, vtu_file: str) -> Dict[str, VTUArray]:
        with open(vtu_file) as f:
            return _parse_arrays(_strip_debug(f.read()))

    def read_field_data(self, pvtu_file: str) -> Dict[str, float]:
        pieces = self.get_piece_files(pvtu_file)
        if not pieces:
            raise ValueError(f"No piece file found in PVTU: {pvtu_file}")
        with open(os.path.join(os.path.dirname(pvtu_file), pieces[0])) as f:
            text = _strip_debug(f.read())
        m = _SECTION_RE["FieldData"].search(text)
        if m is None:
            return {}
        out = {}
        for name, a in _parse_arrays(m.group(1)).items():
            if a.data.size:
                out[name] = a.data.reshape(-1)[0].item()
        return out

    def load_particles(self, timestep: int):
        """(Bodies in gid order, their gids) from every piece of
        particles_<step>.pvtu."""
        pv = os.path.join(self.vtk_directory, f"particles_{timestep:07d}.pvtu")
        parts = {k: [] for k in ("pos", "gid", "quat", "len", "force", "vel", "omega")}
        for piece in self.get_piece_files(pv):
            arr = self.read_vtu_file(os.path.join(self.vtk_directory, piece))
            pts = arr["Points"].data.reshape(-1, 3)
            n = len(pts)

            def get(name, comp):
                a = arr[name].data.reshape(-1, comp) if comp > 1 else arr[name].data.reshape(-1)
                if len(a) != n:
                    raise ValueError(f"{piece}: {name} has {len(a)} tuples for {n} points")
                return a
            parts["pos"].append(pts.astype(np.float64))
            parts["gid"].append(get("gid", 1).astype(np.int64))
            parts["quat"].append(get("quaternion", 4).astype(np.float64))
            parts["len"].append(get("lengths", 3)[:, 0].astype(np.float64))
            parts["force"].append(get("forces", 3).astype(np.float64))
            parts["vel"].append(get("velocity_linear", 3).astype(np.float64))
            parts["omega"].append(get("velocity_angular", 3).astype(np.float64))
        cat = {k: (np.concatenate(v) if v else np.zeros((0,))) for k, v in parts.items()}
        gid = cat["gid"].reshape(-1)
        if len(np.unique(gid)) != len(gid):
            raise ValueError(f"{pv}: duplicate gids")
        order = np.argsort(gid, kind="stable")
        n = len(gid)
        b = T.Bodies(cat["pos"].reshape(-1, 3)[order], cat["quat"].reshape(-1, 4)[order],
                     cat["len"].reshape(-1)[order], np.ones(n), np.full(n, 0.5), np.zeros(n))
        b.force = np.ascontiguousarray(cat["force"].reshape(-1, 3)[order])
        b.velocity = np.ascontiguousarray(cat["vel"].reshape(-1, 3)[order])
        b.omega = np.ascontiguousarray(cat["omega"].reshape(-1, 3)[order])
        return b, gid[order]
