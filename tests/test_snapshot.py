"""Snapshot formats on the CPU: the binary checkpoint (bit-exact round trip,
resume that steps as the uninterrupted run under the oracle, rejection of
damaged files) and the VTU/PVTU loggers and loader (the reference's file
layout and array names, its 6-digit printing, exact round trip in lossless
mode, field data its own reader can take, multi-piece loads)."""
import os
import re
import xml.etree.ElementTree as ET

import numpy as np
import pytest

import oracle
from pcc import abi_types as T
from pcc import generators as G
from pcc import snapshot as S

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden", "geometry_golden.npz")

# ParticleLogger.cpp:61-74, in order, with their component counts and types
PARTICLE_ARRAYS = [
    ("gid", 1, "Int32"), ("quaternion", 4, "Float64"), ("orientation", 3, "Float64"),
    ("orientation_angle", 1, "Float64"), ("lengths", 3, "Float64"),
    ("impedance", 1, "Float64"), ("ldot", 1, "Float64"), ("stress", 1, "Float64"),
    ("number_of_constraints", 1, "Int32"), ("age", 1, "Int32"), ("forces", 3, "Float64"),
    ("velocity_linear", 3, "Float64"), ("velocity_angular", 3, "Float64"),
    ("rank", 1, "Int32"),
]


def random_bodies(n, seed):
    rng = np.random.default_rng(seed)
    q = rng.normal(size=(n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    b = T.Bodies(rng.normal(size=(n, 3)) * 7, q, 1 + rng.random(n), 1 + 0.1 * rng.random(n),
                 0.5 + 0.01 * rng.random(n), rng.random(n) * 1e-3)
    for f in ("force", "torque", "velocity", "omega"):
        setattr(b, f, rng.normal(size=(n, 3)) * 10.0 ** rng.integers(-12, 12))
    b.impedance = rng.random(n)
    b.stress = rng.normal(size=n) * 1e-7
    b.num_constraints = rng.integers(0, 9, n).astype(np.int32)
    b.age = rng.integers(1, 40, n).astype(np.int32)
    return b


def same(a, b, fields=tuple(f for f, _, _ in S.CHECKPOINT_FIELDS)):
    assert a.n == b.n
    for f in fields:
        x, y = getattr(a, f), getattr(b, f)
        assert x.dtype == y.dtype and np.array_equal(x.view(np.uint8), y.view(np.uint8)), f


# -- checkpoint -----------------------------------------------------------------

@pytest.mark.parametrize("mmap", [False, True])
def test_checkpoint_round_trip_is_bitwise(tmp_path, mmap):
    cols = [random_bodies(n, 10 + n) for n in (37, 0, 1000)]
    cols[2].pos[5, 1] = -0.0
    cols[2].ldot[7] = 5e-324  # subnormal
    ps = [T.reference_params(dt=1e-3 * (q + 1), mode=q % 2, precision=T.PREC_MIXED)
          for q in range(3)]
    path = str(tmp_path / "c.pccsnap")
    size = S.save_checkpoint(path, cols, ps, step=1234, simulation_time_s=56.25,
                             user={"seed": 9})
    assert os.path.getsize(path) == size
    h = S.read_checkpoint_header(path)
    assert all(s["offset"] % 256 == 0 for s in h["fields"].values())
    ck = S.load_checkpoint(path, mmap=mmap)
    assert (ck.step, ck.simulation_time_s, ck.user) == (1234, 56.25, {"seed": 9})
    assert [b.n for b in ck.colonies] == [37, 0, 1000]
    for a, b in zip(ck.colonies, cols):
        same(a, b)
    for p, q in zip(ck.params, ps):
        assert S.params_to_dict(p) == S.params_to_dict(q)


def test_checkpoint_sections_are_the_abi_arrays(tmp_path):
    """Each section is the concatenation over colonies of the array the
    pcc_bodies struct points at: read one with a plain offset."""
    cols = [random_bodies(20, 1), random_bodies(30, 2)]
    path = str(tmp_path / "c.pccsnap")
    S.save_checkpoint(path, cols)
    h = S.read_checkpoint_header(path)
    s = h["fields"]["quat"]
    raw = np.fromfile(path, dtype="<f8", count=50 * 4, offset=s["offset"]).reshape(50, 4)
    assert np.array_equal(raw, np.concatenate([cols[0].quat, cols[1].quat]))
    assert S.load_checkpoint(path).params is None


def test_checkpoint_without_age_section_loads_with_age_one(tmp_path, monkeypatch):
    """A checkpoint written before the age section existed still loads; its
    cells are of age 1, as after the reference's restart."""
    cols = [random_bodies(25, 3)]
    path = str(tmp_path / "old.pccsnap")
    monkeypatch.setattr(S, "CHECKPOINT_FIELDS",
                        tuple(f for f in S.CHECKPOINT_FIELDS if f[0] != "age"))
    S.save_checkpoint(path, cols)
    monkeypatch.undo()
    assert "age" not in S.read_checkpoint_header(path)["fields"]
    got = S.load_checkpoint(path).colonies[0]
    assert got.age.dtype == np.int32 and np.all(got.age == 1)
    same(got, cols[0], tuple(f for f, _, _ in S.CHECKPOINT_FIELDS if f != "age"))


def test_checkpoint_rejects_damaged_files(tmp_path):
    path = str(tmp_path / "c.pccsnap")
    S.save_checkpoint(path, [random_bodies(100, 3)], T.reference_params())
    data = open(path, "rb").read()
    bad = str(tmp_path / "bad")
    open(bad, "wb").write(b"NOTSNAP1" + data[8:])
    with pytest.raises(ValueError, match="magic"):
        S.load_checkpoint(bad)
    open(bad, "wb").write(data[:-300])
    with pytest.raises(ValueError, match="out of bounds"):
        S.load_checkpoint(bad)
    open(bad, "wb").write(data[:20])
    with pytest.raises(ValueError, match="truncated"):
        S.load_checkpoint(bad)
    with pytest.raises(ValueError):
        S.params_from_dict({"dt": 1.0})


def test_resume_from_checkpoint_steps_as_the_uninterrupted_run(tmp_path):
    """Oracle hard steps 0..5 without a stop equal steps 0..2, a checkpoint,
    a load and steps 3..5: the checkpoint holds every input the step reads
    (lengths, l0 and diameters differ from their defaults here)."""
    p = T.reference_params(dt=1e-2, tolerance=1e-4, tau=1 / np.log(2.0))
    b = G.dense_colony(300, seed=41)
    b.l0 = b.l0 * 1.05
    b.diameter = b.diameter * 0.98
    want = b
    for k in range(6):
        want = oracle.step(p, want, k)[2]
    cur = b
    for k in range(3):
        cur = oracle.step(p, cur, k)[2]
    path = str(tmp_path / "r.pccsnap")
    S.save_checkpoint(path, cur, p, step=3)
    ck = S.load_checkpoint(path)
    cur, p2 = ck.colonies[0], ck.params[0]
    for k in range(ck.step, 6):
        cur = oracle.step(p2, cur, k)[2]
    same(cur, want, T.Bodies.FIELDS_IN)


# -- VTU / PVTU ------------------------------------------------------------------

def test_direction_vectors_match_reference_golden_bitwise():
    z = np.load(GOLD)
    assert np.array_equal(S.direction_vectors(z["quat"]), z["direction"])


def write_step(d, b, k, precision, **sim):
    S.ParticleLogger(str(d), precision=precision, step=k, preserve_existing=True).log(b)
    S.SimulationLogger(str(d), precision=precision, step=k, preserve_existing=True).log(sim)


def test_vtu_layout_and_names_follow_the_reference(tmp_path):
    b = random_bodies(50, 4)
    write_step(tmp_path, b, 17, "reference", step=17, dt_s=1e-3, simulation_time_s=0.17,
               num_particles=50)
    assert sorted(os.listdir(tmp_path)) == ["data", "particles_0000017.pvtu",
                                            "simulation_0000017.pvtu"]
    assert sorted(os.listdir(tmp_path / "data")) == ["particles_0000017_rank_0000.vtu",
                                                     "simulation_0000017_rank_0000.vtu"]
    root = ET.parse(tmp_path / "data" / "particles_0000017_rank_0000.vtu").getroot()
    assert root.attrib == {"type": "UnstructuredGrid", "version": "0.1",
                           "byte_order": "LittleEndian"}
    piece = root.find("UnstructuredGrid/Piece")
    assert piece.attrib == {"NumberOfPoints": "50", "NumberOfCells": "0"}
    pts = piece.find("Points/DataArray")
    assert pts.attrib["type"] == "Float32" and pts.attrib["NumberOfComponents"] == "3"
    arrays = piece.findall("PointData/DataArray")
    got = [(a.attrib["Name"], int(a.attrib["NumberOfComponents"]), a.attrib["type"])
           for a in arrays]
    assert got == PARTICLE_ARRAYS
    for a, (_, k, _) in zip(arrays, PARTICLE_ARRAYS):
        assert len(a.text.split()) == 50 * k
    ages = next(a for a in arrays if a.attrib["Name"] == "age")
    assert [int(v) for v in ages.text.split()] == b.age.tolist()
    proot = ET.parse(tmp_path / "particles_0000017.pvtu").getroot()
    pp = [(a.attrib["Name"], int(a.attrib["NumberOfComponents"]), a.attrib["type"])
          for a in proot.findall("PUnstructuredGrid/PPointData/PDataArray")]
    assert pp == PARTICLE_ARRAYS
    assert [e.attrib["Source"] for e in proot.findall("PUnstructuredGrid/Piece")] == \
        ["data/particles_0000017_rank_0000.vtu"]
    sroot = ET.parse(tmp_path / "data" / "simulation_0000017_rank_0000.vtu").getroot()
    names = [a.attrib["Name"] for a in sroot.findall("UnstructuredGrid/FieldData/DataArray")]
    assert names == [n for n, _ in S.SIMULATION_FIELDS]


def test_reference_precision_prints_six_significant_digits(tmp_path):
    b = random_bodies(40, 5)
    b.pos[0] = (123456.789, -1.5e-7, 0.0)
    write_step(tmp_path, b, 0, "reference", step=0)
    st = S.VTKStateLoader(str(tmp_path)).load_state(0)
    want = np.array([float("%g" % float(np.float32(x))) for x in b.pos.reshape(-1)])
    assert np.array_equal(st.bodies.pos.reshape(-1), want)
    assert st.bodies.pos[0, 0] == 123457.0
    want_q = np.array([float("%g" % x) for x in b.quat.reshape(-1)])
    assert np.array_equal(st.bodies.quat.reshape(-1), want_q)
    # orientation_angle in [0, pi)
    arr = S.VTKStateLoader(str(tmp_path)).read_vtu_file(
        str(tmp_path / "data" / "particles_0000000_rank_0000.vtu"))
    ang = arr["orientation_angle"].data
    assert ang.min() >= 0 and ang.max() < np.pi + 1e-5


def test_lossless_round_trip_and_reference_loader_defaults(tmp_path):
    b = random_bodies(500, 6)
    write_step(tmp_path, b, 3, "lossless", step=3, dt_s=1e-5, simulation_time_s=12.345678901234)
    st = S.VTKStateLoader(str(tmp_path)).load_latest_state()
    assert (st.step, st.dt_s, st.simulation_time_s) == (3, 1e-5, 12.345678901234)
    got = st.bodies
    for f in ("pos", "quat", "length", "force", "velocity", "omega"):
        assert np.array_equal(getattr(got, f), getattr(b, f)), f
    # VTKStateLoader.cpp:131-136 and :139-160: l0 = 1, d = 0.5, no torque, and
    # the Particle constructor's age 1 (Particle.cpp:26)
    assert np.all(got.l0 == 1.0) and np.all(got.diameter == 0.5) and np.all(got.ldot == 0.0)
    assert np.all(got.age == 1) and np.any(b.age != 1)
    assert np.all(got.torque == 0.0)
    assert np.array_equal(st.gids, np.arange(500))


def test_field_data_has_no_exponent_for_the_reference_reader(tmp_path):
    write_step(tmp_path, random_bodies(3, 7), 0, "reference", step=0, dt_s=1e-5,
               simulation_time_s=3e-9, colony_radius=2.5e11)
    text = open(tmp_path / "data" / "simulation_0000000_rank_0000.vtu").read()
    # the reference's reader: first run of [digits . -] on the line after the tag
    lines = text.splitlines()
    got = {}
    for i, ln in enumerate(lines):
        if "</FieldData>" in ln:
            break
        m = re.search(r'<DataArray.*Name="([^"]+)".*>', ln)
        if m:
            got[m.group(1)] = float(re.search(r"\s*([\d.-]+)\s*", lines[i + 1]).group(1))
    assert got["dt_s"] == 1e-5 and got["simulation_time_s"] == 3e-9
    assert got["colony_radius"] == 2.5e11


def test_multi_piece_load_orders_by_gid(tmp_path):
    a, b = random_bodies(30, 8), random_bodies(20, 9)
    # rank 0 holds gids 20..49, rank 1 gids 0..19
    S.ParticleLogger(str(tmp_path), rank=0, size=2, precision="lossless").log(a, gid_offset=20)
    S.ParticleLogger(str(tmp_path), rank=1, size=2, precision="lossless",
                     preserve_existing=True).log(b, gid_offset=0)
    S.SimulationLogger(str(tmp_path), preserve_existing=True).log({"step": 0})
    L = S.VTKStateLoader(str(tmp_path))
    assert L.get_piece_files(str(tmp_path / "particles_0000000.pvtu")) == [
        "data/particles_0000000_rank_0000.vtu", "data/particles_0000000_rank_0001.vtu"]
    st = L.load_state(0)
    assert np.array_equal(st.gids, np.arange(50))
    assert np.array_equal(st.bodies.pos, np.concatenate([b.pos, a.pos]))


def test_logger_clears_old_output_unless_preserved(tmp_path):
    b = random_bodies(5, 10)
    lg = S.ParticleLogger(str(tmp_path))
    lg.log(b)
    lg.log(b)
    assert sorted(os.listdir(tmp_path / "data")) == ["particles_0000000_rank_0000.vtu",
                                                     "particles_0000001_rank_0000.vtu"]
    S.ParticleLogger(str(tmp_path), preserve_existing=True, step=2).log(b)
    assert len(os.listdir(tmp_path / "data")) == 3
    S.ParticleLogger(str(tmp_path))
    assert os.listdir(tmp_path / "data") == [] and os.listdir(tmp_path) == ["data"]


def test_loader_latest_step_and_debug_lines(tmp_path):
    b = random_bodies(10, 11)
    for k in (4, 12, 9):
        write_step(tmp_path, b, k, "lossless", step=k)
    L = S.VTKStateLoader(str(tmp_path))
    assert L.get_latest_timestep() == 12
    p = tmp_path / "data" / "particles_0000012_rank_0000.vtu"
    text = open(p).read().replace("<PointData>", "<PointData>\nline: 42 debug")
    open(p, "w").write(text)
    st = L.load_latest_state()
    assert st.step == 12 and np.array_equal(st.bodies.pos, b.pos)
    with pytest.raises(FileNotFoundError):
        S.VTKStateLoader(str(tmp_path / "nowhere"))


def test_constraint_logger_arrays(tmp_path):
    p = T.reference_params()
    b = G.dense_colony(200, seed=12)
    res = oracle.step(p, b, 0)
    cons = res[3]
    S.ConstraintLogger(str(tmp_path)).log(cons)
    root = ET.parse(tmp_path / "data" / "constraints_0000000_rank_0000.vtu").getroot()
    names = [a.attrib["Name"] for a in root.findall("UnstructuredGrid/Piece/PointData/DataArray")]
    assert names == ["gidI", "gidJ", "gids", "signed_distance", "normals", "iteration", "gamma",
                     "rank"]
    n = int(root.find("UnstructuredGrid/Piece").attrib["NumberOfPoints"])
    assert n == len(cons["gidI"]) > 0


def test_bbpgd_logger_writes_one_file_per_iteration(tmp_path):
    p = T.reference_params(dt=1e-2, tolerance=1e-4)
    rec = oracle.traced_step(p, G.dense_colony(200, seed=13))[6]
    S.BBPGDLogger(str(tmp_path)).log(rec[:7])
    assert sorted(os.listdir(tmp_path)) == ["bbpgdtrace_%07d.pvtu" % k for k in range(7)] + \
        ["data"]
    L = S.VTKStateLoader(str(tmp_path))
    fd = L.read_field_data(str(tmp_path / "bbpgdtrace_0000003.pvtu"))
    assert list(fd) == list(S.BBPGDLogger.FIELDS)
    assert fd["step"] == 3 and fd["total_constraints"] == rec["total_constraints"][3]
    assert fd["residual"] == float("%g" % rec["residual"][3])
