"""Humanoid (free joint + 21 hinges, nv 27, condim 6 elliptic cones) on the
GPU against the FP64 oracle, through the C ABI."""
from __future__ import annotations

import numpy as np
import pytest

import ref64
import states
from mjmlx_hip import capi, scenes

pytestmark = pytest.mark.gpu


def _states(model, B, seed):
    orc = ref64.Oracle(ref64.fields_from(model.field))
    rng = np.random.default_rng(seed)
    q = np.tile(model.field("qpos0"), (B, 1))
    qv = rng.normal(0, 0.3, (B, model.nv))
    qa = np.zeros((B, model.nv))
    lens = rng.integers(5, 80, size=B)
    for k in range(int(lens.max())):
        idx = np.nonzero(lens > k)[0]
        ctrl = rng.uniform(-0.4, 0.4, (len(idx), model.nu))
        qs, vs, as_ = q[idx].copy(), qv[idx].copy(), qa[idx].copy()
        orc.batch_step(qs, vs, as_, ctrl, 4)
        q[idx], qv[idx], qa[idx] = qs, vs, as_
    f32 = lambda a: np.ascontiguousarray(a, dtype=np.float32)
    return f32(q), f32(qv), f32(qa), f32(rng.uniform(-0.4, 0.4, (B, model.nu))), orc


def test_humanoid_single_step_parity():
    model = capi.Model(scenes.humanoid_xml())
    B = 64
    q, qv, qa, ctrl, orc = _states(model, B, 2)
    sim = capi.BatchedSim(model, B)
    sim.set_state(q, qv)
    sim.set_qacc_warmstart(qa)
    dump = sim.debug_step(ctrl)
    st = sim.stats()
    assert not (st["status"] & ~3).any(), st["status"]
    worst = worstc = worstf = worsts = 0.0
    ncon = 0
    for e in range(B):
        states.oracle_step(orc, q[e], qv[e], qa[e], ctrl[e])
        con = dump.get(1, e, "contact").reshape(-1, 16)[: st["ncon"][e]]
        ok, msg = states.compare_contacts(con[:, 13].view(np.int32), con[:, 12], orc.contacts(), margin_tol=1e-4, ncon=st["ncon"][e])
        assert ok, f"env {e}: {msg}"
        ncon += len(con)
        b = orc.get("qacc")
        qk = dump.get(2, e, "qacc")
        eq, ef = states.qacc_force_errors(orc, qk, dump.get(2, e, "efc_force"), st["nefc"][e])
        assert st["nefc"][e] == orc.nefc(), (e, st["nefc"][e], orc.nefc())
        worst, worstf = max(worst, eq), max(worstf, ef or 0.0)
        if orc.nefc():
            c_opt = orc.cost(b)[0]
            worstc = max(worstc, (orc.cost(qk.astype(np.float64))[0] - c_opt) / states.cost_scale(orc, c_opt))
    for e in range(0, B, 8):
        worsts = max(worsts, states.input_sensitivity(orc, q[e], qv[e], qa[e], ctrl[e], seed=e))
    print(f"humanoid: contacts {ncon}, worst qacc rel err {worst:.2e}, worst efc_force rel err {worstf:.2e}, "
          f"cost gap {worstc:.2e}, oracle's own FP32-input sensitivity {worsts:.2e}, "
          f"mean iterations {st['niter'].mean():.2f}")
    assert ncon > B // 2
    assert worst < 1e-4 and worstf < 1e-4 and worstc < 1e-8


def test_humanoid_rollout_stays_finite():
    model = capi.Model(scenes.humanoid_xml())
    B = 256
    sim = capi.BatchedSim(model, B)
    rng = np.random.default_rng(0)
    for k in range(100):
        sim.step(rng.uniform(-0.4, 0.4, (B, model.nu)).astype(np.float32), 4)
    q, qv = sim.state()
    st = sim.stats()
    assert np.isfinite(q).all() and np.isfinite(qv).all()
    assert not (st["status"] & ~3).any()
    print(f"after 100 control steps: mean root height {q[:, 2].mean():.3f}, mean iterations {st['niter'].mean():.2f}")
