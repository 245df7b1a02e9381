"""Checkpoint and resume on the device: a context stepped, checkpointed,
stepped on equals a fresh context restored from the checkpoint and stepped
the same way, bit for bit (single colony and a batch with per-colony
parameters, FP64 and mixed precision); the VTU output of a device state loads
back exactly in lossless mode."""
import math

import numpy as np
import pytest

from pcc import generators as G
from pcc import snapshot as S
from pcc.abi_types import PREC_FP64, PREC_MIXED, reference_params

pytestmark = pytest.mark.gpu
LN2 = math.log(2.0)
FIELDS = ("pos", "quat", "length", "l0", "diameter", "ldot", "velocity", "omega", "force",
          "torque", "stress", "impedance", "num_constraints")


def same(a, b):
    assert a.n == b.n
    for f in FIELDS:
        assert np.array_equal(getattr(a, f), getattr(b, f)), f


@pytest.mark.parametrize("precision", [PREC_FP64, PREC_MIXED])
def test_resume_single_colony(tmp_path, precision):
    from pcc.engine import ParticleManager
    p = reference_params(dt=1e-2, tolerance=1e-4, tau=1 / LN2, precision=precision)
    b = G.dense_colony(4000, seed=190)
    print(f"[resume] n={b.n} precision={precision}", flush=True)
    pm = ParticleManager(p)
    pm.set_state(b)
    for k in range(3):
        pm.step(k)
    dt, _ = pm.adapt_dt(0.5)
    path = str(tmp_path / "a.pccsnap")
    S.checkpoint_context(pm, path, step=3, simulation_time_s=0.03)
    for k in range(3, 6):
        pm.step(k)
    want = pm.get_state()
    pm.close()
    pm2, ck = S.restore_context(path)
    assert ck.step == 3 and ck.params[0].dt == dt
    for k in range(ck.step, 6):
        pm2.step(k)
    same(pm2.get_state(), want)
    pm2.close()


def test_resume_batch_with_per_colony_params(tmp_path):
    from pcc.engine import ColonyBatch
    cols = [G.dense_colony(n, seed=200 + i) for i, n in enumerate([1200, 300, 2500])]
    ps = [reference_params(dt=1e-2 * (1 + 0.5 * i), tolerance=1e-4, tau=1 / LN2,
                           precision=PREC_MIXED) for i in range(3)]
    print("[resume batch] sizes=1200,300,2500", flush=True)
    cb = ColonyBatch(ps[0])
    cb.set_state(cols, ps)
    for k in range(2):
        cb.step(k)
    cb.divide([1, 2, 3], step=2, length_noise=0.02)
    cb.adapt_dt(0.5)
    path = str(tmp_path / "b.pccsnap")
    S.checkpoint_context(cb, path, step=2)
    for k in range(2, 5):
        cb.step(k)
    want = [cb.get_state(q) for q in range(3)]
    cb.close()
    cb2, ck = S.restore_context(path)
    assert [p.dt for p in ck.params] != [p.dt for p in ps]  # adapted dt travelled
    for k in range(2, 5):
        cb2.step(k)
    for q in range(3):
        same(cb2.get_state(q), want[q])
    cb2.close()


def test_vtu_export_of_device_state_round_trips(tmp_path):
    from pcc.engine import ParticleManager
    p = reference_params(dt=1e-2, tolerance=1e-4, tau=1 / LN2)
    pm = ParticleManager(p)
    pm.set_state(G.dense_colony(3000, seed=210))
    r = pm.step(0)
    st = pm.get_state()
    S.ParticleLogger(str(tmp_path), precision="lossless").log(st)
    S.ConstraintLogger(str(tmp_path), precision="lossless").log(pm.constraints())
    S.SimulationLogger(str(tmp_path), precision="lossless").log(
        {"step": 0, "dt_s": p.dt, "num_particles": st.n, "num_constraints": r["n_constraints"],
         "bbpgd_iterations": r["bbpgd_iterations"], "residual": r["residual"]})
    got = S.VTKStateLoader(str(tmp_path)).load_latest_state()
    for f in ("pos", "quat", "length", "force", "velocity", "omega"):
        assert np.array_equal(getattr(got.bodies, f), getattr(st, f)), f
    assert got.dt_s == p.dt
    pm.close()
