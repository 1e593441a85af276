"""Domain randomisation on the GPU through the C ABI
(mjmlx_batched_set_env_params / set_free_drag / num_free_joints, reference
src/batched.cpp:1508-1545) against the FP64 oracle with the same per-env
parameters."""
from __future__ import annotations

import numpy as np
import pytest

import ref64
import states
from mjmlx_hip import capi, scenes

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float(np.max(np.abs(np.asarray(a) - np.asarray(b))) / max(1.0, float(np.max(np.abs(b)))))


def test_randomised_single_step_parity():
    model = capi.Model(scenes.quadruped_xml())
    B = 64
    q, qv, qa, ctrl, orc = states.quadruped_states(ref64.fields_from(model.field), B, seed=41)
    sim = capi.BatchedSim(model, B, domain_rand=True)
    nb, nu = orc.nbody, orc.nu
    assert sim.num_free_joints() == orc.nfree() == 1
    rng = np.random.default_rng(41)
    ms = rng.uniform(0.7, 1.3, (B, nb))
    cd = rng.uniform(-0.02, 0.02, (B, nb, 3))
    fs = rng.uniform(0.5, 1.5, B)
    gs = rng.uniform(0.7, 1.3, (B, nu))
    dk = rng.uniform(0.0, 0.5, (B, 1))
    # the first half of the envs set in one call, the rest through env_idx
    # in two calls with different fields
    h = B // 2
    assert sim.set_env_params(None, ms, cd, fs, gs) == 0
    idx = np.arange(h, B)
    assert sim.set_env_params(idx, None, None, fs[h:] * 1.0, None) == 0
    assert sim.set_free_drag(dk) == 0
    sim.set_state(q, qv)
    sim.set_qacc_warmstart(qa)
    dump = sim.debug_step(ctrl)
    st = sim.stats()
    assert not (st["status"] & ~3).any(), st["status"]
    worst_q = worst_f = 0.0
    for e in range(B):
        orc.set_env_params(ms[e], cd[e], [fs[e]], gs[e], dk[e])
        states.oracle_step(orc, q[e], qv[e], qa[e], ctrl[e])
        con = dump.get(1, e, "contact").reshape(-1, 16)[: st["ncon"][e]]
        ok, msg = states.compare_contacts(con[:, 13].view(np.int32), con[:, 12], orc.contacts(), ncon=st["ncon"][e])
        assert ok, f"env {e}: {msg}"
        worst_q = max(worst_q, _rel(dump.get(2, e, "qacc"), orc.get("qacc")))
        n = orc.nefc()
        assert st["nefc"][e] == n
        if n:
            worst_f = max(worst_f, _rel(dump.get(2, e, "efc_force")[:n], orc.get("efc_force")))
    print(f"randomised quadruped: worst qacc rel {worst_q:.2e}, efc_force rel {worst_f:.2e}")
    assert worst_q < 1e-4 and worst_f < 1e-4


def test_identity_parameters_and_error_codes():
    """A domain_rand sim at its default (identity) parameters steps like a
    plain sim (to FP32 rounding: subtree masses are summed per env); the
    setters reject bad handles and indices with the reference's codes."""
    model = capi.Model(scenes.quadruped_xml())
    B = 32
    q0 = scenes.quadruped_initial_qpos(B, 4).astype(np.float32)
    rng = np.random.default_rng(4)
    acts = [scenes.quadruped_actions(B, rng).astype(np.float32) for _ in range(5)]
    out = []
    for rand in (False, True):
        sim = capi.BatchedSim(model, B, domain_rand=rand)
        sim.set_state(q0, np.zeros((B, model.nv), np.float32))
        for a in acts:
            sim.step(a, 4)
        out.append(sim.state())
        if rand:
            assert sim.set_env_params(np.array([B], np.int32), friction_scale=np.ones(1)) == -2
            assert sim.set_env_params(np.array([-1], np.int32), friction_scale=np.ones(1)) == -2
            assert sim.set_free_drag(np.zeros((B, 1)), np.array([0, B], np.int32)) == -2
        else:
            assert sim.set_env_params(None, friction_scale=np.ones(B)) == -1
            assert sim.set_free_drag(np.zeros((B, 1))) == -1
            assert sim.num_free_joints() == 0
        del sim
    for a, b in zip(out[0], out[1]):
        assert np.max(np.abs(a - b)) < 1e-3


def test_randomised_boxes_device_memory_workspace(monkeypatch):
    """Randomised masses and friction with the workspace in device memory
    (box scene, Newton) against the oracle with the same parameters."""
    monkeypatch.setenv("MJMLX_HIP_GLOBAL_WORKSPACE", "1")
    model = capi.Model(scenes.boxes_xml(6, 3, solver="Newton", iterations=50))
    B = 24
    orc = ref64.Oracle(ref64.fields_from(model.field)).set_caps()
    q = np.tile(model.field("qpos0"), (B, 1))
    qv = np.zeros((B, model.nv))
    qa = np.zeros((B, model.nv))
    rng = np.random.default_rng(5)
    lens = rng.integers(20, 80, size=B)
    for k in range(int(lens.max())):
        idx = np.nonzero(lens > k)[0]
        qs_, vs_, as_ = q[idx].copy(), qv[idx].copy(), qa[idx].copy()
        orc.batch_step(qs_, vs_, as_, np.zeros((len(idx), 0)), 4)
        q[idx], qv[idx], qa[idx] = qs_, vs_, as_
    f32 = lambda a: np.ascontiguousarray(a, dtype=np.float32)
    q, qv, qa = f32(q), f32(qv), f32(qa)
    sim = capi.BatchedSim(model, B, domain_rand=True)
    nb = orc.nbody
    ms = rng.uniform(0.6, 1.4, (B, nb)).astype(np.float32)
    fs = rng.uniform(0.4, 1.6, B).astype(np.float32)
    assert sim.set_env_params(None, ms, None, fs, None) == 0
    assert sim.num_free_joints() == 6
    sim.set_state(q, qv)
    sim.set_qacc_warmstart(qa)
    dump = sim.debug_step(None)
    st = sim.stats()
    assert not (st["status"] & ~3).any()
    worst = worstf = 0.0
    for e in range(B):
        orc.set_env_params(mass_scale=ms[e], friction_scale=[fs[e]])
        states.oracle_step(orc, q[e], qv[e], qa[e], np.zeros(0, np.float32))
        assert st["nefc"][e] == orc.nefc()
        eq, ef = states.qacc_force_errors(orc, dump.get(2, e, "qacc"), dump.get(2, e, "efc_force"), st["nefc"][e])
        worst, worstf = max(worst, eq), max(worstf, ef or 0.0)
    # qacc relative to the balanced accelerations (resting boxes: qacc ~ 0, g balanced)
    print(f"randomised boxes (device memory): worst qacc rel {worst:.2e}, efc_force rel {worstf:.2e}")
    assert worst < 1e-4 and worstf < 1e-4
