"""Domain randomisation: per-env mass scale, COM shift, friction scale,
actuator force scale and free-joint drag (reference src/batched.cpp:1348-1361,
1508-1545; src/compat/mkx_kernels/forward.hpp:110-115,147-173,427,455-465;
solver.hpp:146,435-436,530).

The oracle's randomised parameters are pinned by equivalence: an env with
parameters p behaves like the plain model whose fields carry p (masses and
inertias scaled, inertial frames shifted, geom friction and actuator gains
scaled), and the drag force has a closed form. The kernel source (serial
interpreter) is then checked against the oracle on randomised envs, to the
bars of the other parity tests.
"""
from __future__ import annotations

import copy

import numpy as np

import emu
import ref64
import states
from mjmlx_hip import scenes


def _fields(xml, **kw):
    return ref64.fields_from(emu.CompiledModel(xml, **kw).field)


def _rel(a, b):
    return float(np.max(np.abs(np.asarray(a) - np.asarray(b))) / max(1.0, float(np.max(np.abs(b)))))


def _random_params(rng, nbody, nu, nfree, B=None):
    shape = () if B is None else (B,)
    return dict(mass_scale=rng.uniform(0.7, 1.3, shape + (nbody,)),
                com_delta=rng.uniform(-0.02, 0.02, shape + (nbody, 3)),
                friction_scale=rng.uniform(0.5, 1.5, shape + (1,) if B is None else shape),
                gain_scale=rng.uniform(0.7, 1.3, shape + (nu,)),
                drag=rng.uniform(0.0, 0.5, shape + (nfree,)))


def test_identity_parameters_change_nothing():
    """Unit scales, zero shifts and zero drag reproduce the plain oracle step
    bit for bit."""
    f = _fields(scenes.quadruped_xml())
    q, qv, qa, ctrl, _ = states.quadruped_states(f, 4, seed=31)
    a = ref64.Oracle(f)
    b = ref64.Oracle(f)
    b.set_env_params(np.ones(b.nbody), np.zeros((b.nbody, 3)), [1.0], np.ones(b.nu), np.zeros(b.nfree()))
    for e in range(4):
        for o in (a, b):
            o.set_state(q[e].astype(np.float64), qv[e].astype(np.float64), qa[e].astype(np.float64),
                        ctrl[e].astype(np.float64))
            o.step()
        for x, y in zip(a.get_state(), b.get_state()):
            assert np.array_equal(x, y)


def test_mass_and_com_equal_a_model_with_those_inertials():
    """Mass scale s and COM shift d of each body give the smooth dynamics of
    the model with body_mass * s, body_inertia * s and body_ipos + d."""
    f = _fields(scenes.quadruped_xml())
    rng = np.random.default_rng(5)
    q, qv, qa, ctrl, _ = states.quadruped_states(f, 6, seed=32)
    for e in range(6):
        p = _random_params(rng, int(f["nbody"]), int(f["nu"]), 1)
        g = copy.deepcopy(f)
        g["body_mass"] = np.asarray(f["body_mass"]) * p["mass_scale"]
        g["body_inertia"] = (np.asarray(f["body_inertia"]).reshape(-1, 3) * p["mass_scale"][:, None]).ravel()
        g["body_ipos"] = (np.asarray(f["body_ipos"]).reshape(-1, 3) + p["com_delta"]).ravel()
        a = ref64.Oracle(f).set_env_params(mass_scale=p["mass_scale"], com_delta=p["com_delta"])
        b = ref64.Oracle(g)
        for o in (a, b):
            states.oracle_step(o, q[e], qv[e], qa[e], ctrl[e])
        for name in ("xipos", "subtree_com", "cinert", "qM", "qfrc_bias", "qfrc_smooth", "qacc_smooth"):
            assert _rel(a.get(name), b.get(name)) < 1e-12, name
        # the parameters move the smooth dynamics for real
        plain = states.oracle_step(ref64.Oracle(f), q[e], qv[e], qa[e], ctrl[e])
        assert _rel(a.get("qacc_smooth"), plain.get("qacc_smooth")) > 1e-3


def test_friction_scale_equals_a_model_with_scaled_friction():
    """Friction scale s on an env steps exactly like the model whose geom
    friction coefficients are all multiplied by s (pair mixing takes their
    maximum, which commutes with a positive scale)."""
    f = _fields(scenes.quadruped_xml())
    q, qv, qa, ctrl, _ = states.quadruped_states(f, 8, seed=33)
    moved = 0
    for e, s in enumerate(np.linspace(0.4, 1.6, 8)):
        g = copy.deepcopy(f)
        g["geom_friction"] = np.asarray(f["geom_friction"]) * s
        g["pair_friction"] = np.asarray(f["pair_friction"]) * s
        a = ref64.Oracle(f).set_env_params(friction_scale=[s])
        b = ref64.Oracle(g)
        for o in (a, b):
            states.oracle_step(o, q[e], qv[e], qa[e], ctrl[e])
        assert a.nefc() == b.nefc()
        assert _rel(a.get("qacc"), b.get("qacc")) < 1e-12
        c = states.oracle_step(ref64.Oracle(f), q[e], qv[e], qa[e], ctrl[e])
        moved += int(a.nefc() > 0 and abs(s - 1) > 0.1 and _rel(a.get("qacc"), c.get("qacc")) > 1e-6)
    assert moved >= 2  # states with contacts where the scale changes the solution


def test_gain_scale_equals_a_model_with_scaled_actuators():
    """Force scale s of an actuator acts like its gain and bias parameters
    multiplied by s."""
    f = _fields(scenes.quadruped_xml())
    rng = np.random.default_rng(6)
    q, qv, qa, ctrl, _ = states.quadruped_states(f, 6, seed=34)
    nu = int(f["nu"])
    for e in range(6):
        s = rng.uniform(0.5, 1.5, nu)
        g = copy.deepcopy(f)
        g["actuator_gainprm"] = (np.asarray(f["actuator_gainprm"]).reshape(nu, -1) * s[:, None]).ravel()
        g["actuator_biasprm"] = (np.asarray(f["actuator_biasprm"]).reshape(nu, -1) * s[:, None]).ravel()
        a = ref64.Oracle(f).set_env_params(gain_scale=s)
        b = ref64.Oracle(g)
        for o in (a, b):
            states.oracle_step(o, q[e], qv[e], qa[e], ctrl[e])
        for name in ("actuator_force", "qfrc_actuator", "qacc"):
            assert _rel(a.get(name), b.get(name)) < 1e-10, name


def test_free_joint_drag_closed_form():
    """Drag k on a free joint adds F = -k |v| v to its three translational dofs
    and nothing elsewhere."""
    f = _fields(scenes.boxes_xml(3, 2, solver="Newton"))
    nv = int(f["nv"])
    rng = np.random.default_rng(7)
    q = np.asarray(f["qpos0"], np.float64).copy()
    q[2::7] += 2.0  # boxes lifted clear of the floor and of each other
    q[0::7] += np.arange(3) * 1.0
    qv = rng.normal(size=nv)
    a = ref64.Oracle(f)
    a.set_state(q, qv, np.zeros(nv), np.zeros(0))
    a.forward()
    k = rng.uniform(0.1, 2.0, 3)
    b = ref64.Oracle(f).set_env_params(drag=k)
    assert b.nfree() == 3
    b.set_state(q, qv, np.zeros(nv), np.zeros(0))
    b.forward()
    want = np.zeros(nv)
    for j in range(3):
        v = qv[6 * j:6 * j + 3]
        want[6 * j:6 * j + 3] = -k[j] * np.linalg.norm(v) * v
    assert np.allclose(b.get("qfrc_smooth") - a.get("qfrc_smooth"), want, rtol=0, atol=1e-13)


def test_interpreter_matches_oracle_with_randomised_envs():
    """The kernel source with a parameter block per env against the oracle
    with the same parameters: identical contact pair lists, qacc and
    efc_force within 1e-4 relative."""
    xml = scenes.quadruped_xml()
    B = 24
    sim = emu.EmuSim(xml, B, maxcon=32, maxefc=96)
    f = ref64.fields_from(sim.field)
    q, qv, qa, ctrl, orc = states.quadruped_states(f, B, seed=35)
    rng = np.random.default_rng(8)
    p = _random_params(rng, orc.nbody, orc.nu, orc.nfree(), B)
    sim.set_rand(p["mass_scale"], p["com_delta"], p["friction_scale"], p["gain_scale"], p["drag"])
    sim.set_state(q, qv, qa)
    sim.step(ctrl, nsub=1, dump=True)
    st = sim.stats()
    assert (st["status"] & ~3 == 0).all()
    worst_q = worst_f = 0.0
    for e in range(B):
        orc.set_env_params(p["mass_scale"][e], p["com_delta"][e], [p["friction_scale"][e]], p["gain_scale"][e],
                           p["drag"][e])
        states.oracle_step(orc, q[e], qv[e], qa[e], ctrl[e])
        con = sim.dump(1, e, "contact").reshape(-1, 16)[: st["ncon"][e]]
        ok, msg = states.compare_contacts(con[:, 13].view(np.int32), con[:, 12], orc.contacts(), ncon=st["ncon"][e])
        assert ok, f"env {e}: {msg}"
        worst_q = max(worst_q, _rel(sim.dump(2, e, "qacc"), orc.get("qacc")))
        n = orc.nefc()
        assert st["nefc"][e] == n
        if n:
            worst_f = max(worst_f, _rel(sim.dump(2, e, "efc_force")[:n], orc.get("efc_force")))
    assert worst_q < 1e-4 and worst_f < 1e-4, (worst_q, worst_f)
