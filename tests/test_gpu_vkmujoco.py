"""mjmlx_hip.vkmujoco (the reference's RL binding interface, reference
python/rl_bindings.cpp:49-186) stepping on the GPU: the same results as the
C ABI called directly, the packed readback, and the reference's argument
checks."""
from __future__ import annotations

import numpy as np
import pytest

from mjmlx_hip import capi, scenes
from mjmlx_hip import vkmujoco

pytestmark = pytest.mark.gpu


def test_rl_loop_matches_c_abi(tmp_path):
    path = tmp_path / "quadruped.xml"
    path.write_text(scenes.quadruped_xml())
    B = 64
    model = vkmujoco.load_model(str(path))
    sim = vkmujoco.create_batched(model, B)
    ref = capi.BatchedSim(capi.Model(scenes.quadruped_xml()), B, integrator=0)
    q0 = scenes.quadruped_initial_qpos(B, 2).astype(np.float32)
    v0 = np.zeros((B, model.nv), np.float32)
    sim.set_state(q0, v0)
    ref.set_state(q0, v0)
    rng = np.random.default_rng(2)
    torso = model.name2id(1, "torso")
    stride = sim.set_pack_spec(np.array([torso], np.int32), 6, 12)
    assert stride == model.nq + model.nv + 1 + 2
    buf = np.zeros((B, stride), np.float32)
    for k in range(10):
        a = scenes.quadruped_actions(B, rng).astype(np.float32)
        sim.step(a, 4)
        ref.step(a, 4)
        sim.read_pack_into(buf)
    sim.sync()
    qp, qv = sim.state()
    rq, rv = ref.state()
    assert np.array_equal(qp, rq) and np.array_equal(qv, rv)
    assert np.array_equal(sim.qpos(), qp) and np.array_equal(sim.qvel(), qv)
    assert np.array_equal(buf[:, :model.nq], qp) and np.array_equal(buf, sim.read_pack())
    assert sim.xpos().shape == (B, 3 * model.nbody)
    assert sim.cfrc_ext().shape == (B, 6 * model.nbody)
    assert sim.qfrc_actuator().shape == (B, model.nv)
    assert sim.sensordata().shape == (B, 0)
    # reset puts masked envs back at qpos0 with zero velocity
    mask = np.zeros(B, np.int32)
    mask[::2] = 1
    sim.reset(mask)
    qp2, qv2 = sim.state()
    assert np.allclose(qp2[::2], model_qpos0(path)) and not qv2[::2].any()
    assert np.array_equal(qp2[1::2], qp[1::2])


def model_qpos0(path):
    return capi.Model(open(path).read()).field("qpos0").astype(np.float32)


def test_argument_checks_and_domain_rand(tmp_path):
    path = tmp_path / "quadruped.xml"
    path.write_text(scenes.quadruped_xml())
    model = vkmujoco.load_model(str(path))
    B = 16
    sim = vkmujoco.create_batched(model, B, domain_rand=True)
    with pytest.raises(RuntimeError, match="ctrl must have shape"):
        sim.step(np.zeros((B, model.nu + 1), np.float32))
    with pytest.raises(TypeError):
        sim.step(np.zeros((B, model.nu)))  # float64
    with pytest.raises(RuntimeError, match="qpos must have shape"):
        sim.set_state(np.zeros((B, 3), np.float32), np.zeros((B, model.nv), np.float32))
    with pytest.raises(RuntimeError, match="bad shape for mass_scale"):
        sim.set_env_params(mass_scale=np.ones((B, model.nbody + 1), np.float32))
    with pytest.raises(RuntimeError, match="call set_pack_spec first"):
        sim.read_pack()
    # shared rays (nray, 3) are the same rays for every env
    pnt = np.array([[0, 0, 2.0], [0.5, 0, 2.0]], np.float32)
    vec = np.array([[0, 0, -1.0], [0, 0, -1.0]], np.float32)
    d, g = sim.raycast(pnt, vec)
    d2, g2 = sim.raycast(np.ascontiguousarray(np.broadcast_to(pnt, (B, 2, 3))),
                         np.ascontiguousarray(np.broadcast_to(vec, (B, 2, 3))))
    assert d.shape == g.shape == (B, 2) and np.array_equal(d, d2) and np.array_equal(g, g2)
    assert (g >= 0).all()
    with pytest.raises(RuntimeError, match="pnt must have shape"):
        sim.raycast(np.zeros((2, 2), np.float32), np.ones((2, 3), np.float32))
    with pytest.raises(RuntimeError):
        vkmujoco.create_camera(sim, 64, 64)
    assert sim.num_free_joints == 1
    sim.set_env_params(np.arange(4, dtype=np.int32), friction_scale=np.full(4, 0.5, np.float32))
    sim.set_free_drag(np.full((B, 1), 0.1, np.float32))
    with pytest.raises(RuntimeError, match="set_env_params failed"):
        sim.set_env_params(np.array([B], np.int32), friction_scale=np.ones(1, np.float32))
    sim.step(np.zeros((B, model.nu), np.float32), 4)
    assert np.isfinite(sim.state()[0]).all()
    plain = vkmujoco.create_batched(model, B)
    with pytest.raises(RuntimeError, match="domain_rand=True"):
        plain.set_env_params(friction_scale=np.ones(B, np.float32))


def test_pack_columns_match_oracle(tmp_path):
    """Every column of the packed readback after a one-substep step against the
    FP64 oracle at the same state: qpos / qvel (the step's result), |contact
    force| of the chosen bodies (the oracle's contacts and row forces, in the
    cfrc_ext convention) and the sum of squares / of absolute values of
    qfrc_actuator over the chosen dofs (the oracle's actuator forces)."""
    import ref64
    import states
    path = tmp_path / "quadruped.xml"
    path.write_text(scenes.quadruped_xml())
    B = 64
    model = vkmujoco.load_model(str(path))
    sim = vkmujoco.create_batched(model, B)
    cm = capi.Model(scenes.quadruped_xml())
    q, qv, qa, ctrl, orc = states.quadruped_states(ref64.fields_from(cm.field), B, seed=29)
    sim.set_state(q, qv)
    assert capi.lib().mjmlx_batched_set_qacc_warmstart(sim.handle, capi._fp(qa)) == 0
    feet = [model.name2id(1, f"{leg}_calf") for leg in ("FL", "FR", "RL", "RR")]
    torso = model.name2id(1, "torso")
    bodies = np.array([torso, *feet], np.int32)
    lo, nd = 6, 12
    stride = sim.set_pack_spec(bodies, lo, nd)
    assert stride == model.nq + model.nv + len(bodies) + 2
    sim.step(ctrl, 1)
    buf = sim.read_pack()
    qp, qvv = sim.state()
    assert np.array_equal(buf[:, :model.nq], qp) and np.array_equal(buf[:, model.nq:model.nq + model.nv], qvv)
    b1, b2 = cm.field("cpair_b1").astype(int), cm.field("cpair_b2").astype(int)
    worst_f = worst_a = 0.0
    touching = 0
    for e in range(B):
        states.oracle_step(orc, q[e], qv[e], qa[e], ctrl[e])
        cf = states.oracle_cfrc_ext(orc, b1, b2, model.nbody)
        fn = np.linalg.norm(cf[bodies, 3:], axis=1)
        touching += int((fn > 0).any())
        got = buf[e, model.nq + model.nv: model.nq + model.nv + len(bodies)]
        worst_f = max(worst_f, np.max(np.abs(got - fn)) / max(1.0, np.abs(cf).max()))
        qfa = orc.get("qfrc_actuator")[lo:lo + nd]
        sq, ab = float(qfa @ qfa), float(np.abs(qfa).sum())
        g_sq, g_ab = buf[e, -2], buf[e, -1]
        worst_a = max(worst_a, abs(g_sq - sq) / max(1.0, sq), abs(g_ab - ab) / max(1.0, ab))
    print(f"pack: |contact force| worst {worst_f:.2e}, actuator sums worst {worst_a:.2e}, envs touching {touching}")
    assert touching > B // 2
    assert worst_f < 1e-4 and worst_a < 1e-4
