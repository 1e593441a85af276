"""The kernel source, run by its serial interpreter on the CPU, against the FP64
oracle (same bars as the GPU parity tests: identical contact pair lists,
qacc / efc_force within 1e-4 relative). The GPU tests then show that the
gfx950 build of the same source matches the interpreter.
"""
from __future__ import annotations

import numpy as np
import pytest

import emu
import ref64
import states
from mjmlx_hip import scenes


@pytest.mark.parametrize("self_collision", [True, False])
def test_quadruped_single_step(self_collision):
    xml = scenes.quadruped_xml(self_collision=self_collision)
    B = 48
    sim = emu.EmuSim(xml, B, self_collision=int(self_collision), maxcon=32, maxefc=96)
    q, qv, qa, ctrl, orc = states.quadruped_states(ref64.fields_from(sim.field), B, seed=21)
    sim.set_state(q, qv, qa)
    sim.step(ctrl, nsub=1, dump=True)
    st = sim.stats()
    assert (st["status"] & ~3 == 0).all()
    worst_q = worst_f = 0.0
    same_iters = 0
    for e in range(B):
        states.oracle_step(orc, q[e], qv[e], qa[e], ctrl[e])
        con = sim.dump(1, e, "contact").reshape(-1, 16)[: st["ncon"][e]]
        ok, msg = states.compare_contacts(con[:, 13].view(np.int32), con[:, 12], orc.contacts(), ncon=st["ncon"][e])
        assert ok, f"env {e}: {msg}"
        b = orc.get("qacc")
        worst_q = max(worst_q, np.max(np.abs(sim.dump(2, e, "qacc") - b)) / max(1.0, np.max(np.abs(b))))
        n = orc.nefc()
        assert st["nefc"][e] == n
        if n:
            fb = orc.get("efc_force")
            worst_f = max(worst_f, np.max(np.abs(sim.dump(2, e, "efc_force")[:n] - fb)) / max(1.0, np.max(np.abs(fb))))
        same_iters += int(st["niter"][e] == orc.niter())
    assert worst_q < 1e-4 and worst_f < 1e-4, (worst_q, worst_f)
    assert same_iters >= 0.9 * B  # FP32 Newton takes the FP64 iteration path


def test_smooth_quantities_match():
    xml = scenes.quadruped_xml()
    B = 8
    sim = emu.EmuSim(xml, B, maxcon=32, maxefc=96)
    q, qv, qa, ctrl, orc = states.quadruped_states(ref64.fields_from(sim.field), B, seed=22)
    sim.set_state(q, qv, qa)
    # nsub = 0: the derived quantities of the current state, dumped before the
    # constraint rows would take their space
    sim.step(ctrl, nsub=0, dump=True)
    nv = sim.nv
    for e in range(B):
        states.oracle_step(orc, q[e], qv[e], qa[e], ctrl[e])
        for name, tol in [("xpos", 1e-6), ("xquat", 1e-6), ("subtree_com", 1e-6), ("cinert", 1e-5),
                          ("cdof", 1e-6), ("cvel", 1e-4), ("qfrc_smooth", 1e-4), ("qacc_smooth", 1e-4)]:
            a = sim.dump(1, e, name)
            b = orc.get(name)
            assert np.max(np.abs(a - b)) / max(1.0, np.max(np.abs(b))) < tol, name
        M = sim.dump(1, e, "qM")[: nv * nv]
        assert np.max(np.abs(M - orc.get("qM"))) < 1e-5


def test_rollout_divergence_reported():
    """100 control steps from the same start: trajectory divergence between
    the FP32 kernel and the FP64 oracle (contact dynamics are chaotic, so this
    is reported with a loose bound, as the reference reports its own
    1-control-step joint-angle error, reference CONFORMANCE.md:74)."""
    xml = scenes.quadruped_xml()
    B = 8
    sim = emu.EmuSim(xml, B, maxcon=32, maxefc=96)
    orc = ref64.Oracle(ref64.fields_from(sim.field))
    rng = np.random.default_rng(9)
    q = scenes.quadruped_initial_qpos(B, 9).astype(np.float32)
    qv = np.zeros((B, sim.nv), np.float32)
    qa = np.zeros((B, sim.nv), np.float32)
    sim.set_state(q, qv, qa)
    q64, qv64, qa64 = q.astype(np.float64), qv.astype(np.float64), qa.astype(np.float64)
    first = None
    for k in range(100):
        c = scenes.quadruped_actions(B, rng).astype(np.float32)
        sim.step(c, nsub=4)
        orc.batch_step(q64, qv64, qa64, np.ascontiguousarray(c, dtype=np.float64), 4)
        if k == 0:
            first = np.max(np.abs(sim.get_state()[0][:, 7:] - q64[:, 7:]))
    qe = sim.get_state()[0]
    last = np.max(np.abs(qe[:, 7:] - q64[:, 7:]))
    print(f"joint-angle divergence after 1 control step {first:.2e} rad, after 100 {last:.2e} rad")
    assert first < 1e-4
    assert np.isfinite(qe).all()


@pytest.mark.parametrize("integrator", ["implicitfast", "Euler"])
def test_quadruped_step_integrators(integrator):
    """One full physics step (solve and integration) of the quadruped with the
    servos' velocity derivative folded into the integrator (implicitfast) and
    without it (Euler): qvel within 1e-4 relative, qpos within 1e-5 of the
    oracle's step from the same FP32 state."""
    xml = scenes.quadruped_xml(self_collision=True, integrator=integrator)
    B = 32
    sim = emu.EmuSim(xml, B, self_collision=1, maxcon=32, maxefc=96)
    q, qv, qa, ctrl, orc = states.quadruped_states(ref64.fields_from(sim.field), B, seed=5)
    sim.set_state(q, qv, qa)
    sim.step(ctrl, nsub=1)
    assert (sim.stats()["status"] & ~3 == 0).all()
    qk, vk, _ = sim.get_state()
    worst_v = worst_q = 0.0
    for e in range(B):
        orc.set_state(q[e].astype(np.float64), qv[e].astype(np.float64), qa[e].astype(np.float64),
                      ctrl[e].astype(np.float64))
        orc.step()
        qo, vo, _ = orc.get_state()
        worst_v = max(worst_v, np.max(np.abs(vk[e] - vo)) / max(1.0, np.max(np.abs(vo))))
        worst_q = max(worst_q, np.max(np.abs(qk[e] - qo)))
    assert worst_v < 1e-4 and worst_q < 1e-5, (worst_v, worst_q)


@pytest.mark.parametrize("scene", ["quadruped", "humanoid", "boxes6", "boxes14_cg", "convex"])
def test_results_independent_of_workspace_contents(scene):
    """No workspace word is read before the step writes it: the kernel source
    gives bit-identical states, statistics and sensors whether each env's
    workspace starts zeroed, filled with NaN bit patterns (positive and
    negative), with 1.0f, or left as the previous env wrote it (one workspace
    serving successive envs, as a persistent launch would)."""
    from test_convex_collision import DROP
    xml, nu = {"quadruped": (scenes.quadruped_xml(sensors=True), 12), "humanoid": (scenes.humanoid_xml(), 21),
               "boxes6": (scenes.boxes_xml(6, 3, solver="Newton"), 0), "boxes14_cg": (scenes.boxes_xml(14, 3), 0),
               "convex": (DROP, 0)}[scene]
    res = []
    for mode, pat in [(0, 0), (1, 0x7FC00000), (1, 0xFFC00000), (1, 0x3F800000), (2, 0x7FC00000)]:
        B = 6
        sim = emu.EmuSim(xml, B)
        rng = np.random.default_rng(0)
        if scene == "quadruped":
            q = scenes.quadruped_initial_qpos(B, 0).astype(np.float32)
            q[:, 2] -= 0.03
        else:
            q = np.tile(sim.field("qpos0"), (B, 1)).astype(np.float32)
        sim.set_state(q, np.zeros((B, sim.nv), np.float32), np.zeros((B, sim.nv), np.float32))
        sim.set_poison(mode, pat)
        outs = []
        for k in range(25):
            c = rng.uniform(-0.4, 0.4, (B, nu)).astype(np.float32) if nu else None
            sim.step(c, 2)
            outs.append((*sim.get_state(), sim.sensordata()))
        res.append((outs, sim.stats()))
    for outs, st in res[1:]:
        for a, b in zip(res[0][0], outs):
            for x, y in zip(a, b):
                assert np.array_equal(x, y)
        for k in st:
            assert np.array_equal(st[k], res[0][1][k])
    assert not res[0][1]["status"].any()


@pytest.mark.parametrize("scene", ["boxes14_newton", "boxes32_cg", "multi_geom_scene", "contact_stress"])
def test_compressed_rows_match_dense_rows(scene, monkeypatch):
    """Beyond 64 dofs the constraint rows are stored compressed (each row's
    chain dofs only) and J^T f is summed from per-body contact wrenches; the
    same step with dense rows (MJMLX_HIP_DENSE_ROWS=1) gives the same qacc and
    qfrc_constraint up to FP32 summation order, and both match the oracle."""
    from test_reference_models import model_xml
    xml = {"boxes14_newton": scenes.boxes_xml(14, 3, solver="Newton", iterations=50),
           "boxes32_cg": scenes.boxes_xml(32, 132), "multi_geom_scene": model_xml("multi_geom_scene"),
           "contact_stress": model_xml("contact_stress")}[scene]
    cm = emu.CompiledModel(xml)
    nfree = int(np.sum(cm.field("jnt_type") == 0))
    orc = ref64.Oracle(ref64.fields_from(cm.field))
    B = 3
    rng = np.random.default_rng(4)
    q = np.tile(cm.field("qpos0"), (B, 1))
    qv = rng.uniform(-0.1, 0.1, (B, orc.nv))
    qa = np.zeros((B, orc.nv))
    orc.batch_step(q, qv, qa, rng.uniform(-0.5, 0.5, (B, orc.nu)) if orc.nu else None,
                   {"boxes14_newton": 160, "boxes32_cg": 160, "multi_geom_scene": 160, "contact_stress": 360}[scene], 3)
    f32 = lambda a: np.ascontiguousarray(a, dtype=np.float32)
    q, qv, qa = f32(q), f32(qv), f32(qa)
    ctrl = f32(rng.uniform(-0.5, 0.5, (B, orc.nu))) if orc.nu else None
    out = []
    for dense in ("1", "0"):
        monkeypatch.setenv("MJMLX_HIP_DENSE_ROWS", dense)
        sim = emu.EmuSim(xml, B)
        sim.set_state(q, qv, qa)
        sim.step(ctrl, 1, dump=True)
        st = sim.stats()
        assert not (st["status"] & ~3).any()
        out.append((sim, st))
    (sd, std), (ss, sts) = out
    assert ss.total < sd.total  # compressed rows take less workspace
    maxcon = ss.layout("contact")[1] // 16
    if cm.field("nv")[0] > 32 or nfree > 1:
        orc.set_caps(con_cap=maxcon, efc_cap=0)
    ncon = 0
    for e in range(B):
        assert std["nefc"][e] == sts["nefc"][e] and std["ncon"][e] == sts["ncon"][e]
        ncon += sts["ncon"][e]
        states.oracle_step(orc, q[e], qv[e], qa[e], ctrl[e] if ctrl is not None else np.zeros(0, np.float32))
        sc = states.accel_scale(orc)
        a_d, a_s = sd.dump(2, e, "qacc"), ss.dump(2, e, "qacc")
        f_d, f_s = sd.dump(2, e, "qfrc_constraint"), ss.dump(2, e, "qfrc_constraint")
        tol = 1e-4 if scene != "boxes32_cg" else 2e-3  # CG at its 16-iteration budget amplifies rounding
        assert np.max(np.abs(a_d - a_s)) / sc < tol
        assert np.max(np.abs(f_d - f_s)) / max(1.0, np.max(np.abs(f_d))) < tol
        if scene in ("boxes14_newton", "contact_stress"):  # (flat GJK patches: see test_reference_models)
            eq, ef = states.qacc_force_errors(orc, a_s, ss.dump(2, e, "efc_force"), sts["nefc"][e])
            assert eq < 1e-4 and ef < 1e-4
    assert ncon > 0
