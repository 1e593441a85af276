"""Pins of the FP64 oracle (CPU).

Each check is independent of the oracle's own formula: the reference's known
answers (tests/golden/reference_kats.json), closed-form mechanics (free fall,
Euler's rigid-body equation, pendulum), brute force (M = sum J^T M_b J from
finite-difference Jacobians, segment distances by dense sampling, a
Lagrangian finite-difference check of RNE), invariants (L^T D L = M, KKT
optimality of the solver, Newton == CG, warm-start independence), exact
rotation for torque-free spin.
"""
from __future__ import annotations

import json
import math
import os

import numpy as np
import pytest

import ref64
from mjmlx_hip import scenes

import emu

HERE = os.path.dirname(os.path.abspath(__file__))


def oracle_for(xml: str):
    e = emu.CompiledModel(xml)
    return ref64.Oracle(ref64.fields_from(e.field)), e


@pytest.mark.parametrize("case", json.load(open(os.path.join(HERE, "golden", "reference_kats.json")))["cases"],
                         ids=lambda c: c["name"])
def test_reference_kats(case):
    out = ref64.math(case["op"], case["in"])
    assert np.allclose(out, case["out"], atol=case["tol"]), (out, case["out"])


def test_quat_integrate_exact_rotation():
    # q(t) = exp(t*omega/2) for constant body-frame omega: one step of size h
    w = np.array([0.3, -0.2, 0.9])
    h = 0.01
    q = ref64.math("quat_integrate", [1, 0, 0, 0, *w, h])
    ang = np.linalg.norm(w) * h
    ax = w / np.linalg.norm(w)
    assert np.allclose(q, [math.cos(ang / 2), *(ax * math.sin(ang / 2))], atol=1e-14)
    assert abs(np.linalg.norm(q) - 1) < 1e-14  # reference tests/test_math_full.cpp:146-155


def test_impedance_sigmoid():
    si = [0.9, 0.95, 0.001, 0.5, 2.0]
    assert ref64.math("impedance", si + [0.0])[0] == pytest.approx(0.9)
    assert ref64.math("impedance", si + [0.002])[0] == pytest.approx(0.95)
    # midpoint of the sigmoid: y = 0.5 -> dmin + 0.5 (dmax - dmin)
    assert ref64.math("impedance", si + [0.0005])[0] == pytest.approx(0.925)


def test_make_frame_orthonormal():
    rng = np.random.default_rng(0)
    for _ in range(50):
        n = rng.normal(size=3)
        f = ref64.math("make_frame", n).reshape(3, 3)
        assert np.allclose(f @ f.T, np.eye(3), atol=1e-12)
        assert np.allclose(f[0], n / np.linalg.norm(n), atol=1e-12)
        assert np.linalg.det(f) == pytest.approx(1.0)


FREE_BOX = """<mujoco><compiler angle="radian"/><option gravity="0 0 -9.81"/>
<worldbody><body pos="0 0 1"><freejoint/><geom type="box" size="0.1 0.2 0.3" density="700"/></body></worldbody></mujoco>"""


def test_free_fall_and_euler_equation():
    orc, e = oracle_for(FREE_BOX)
    q0 = e.field("qpos0")
    orc.set_state(q0, np.zeros(6), np.zeros(6), np.zeros(0))
    orc.forward()
    assert np.allclose(orc.get("qacc"), [0, 0, -9.81, 0, 0, 0], atol=1e-12)
    # torque-free spin without gravity: bias = [0; w x I w] in the body frame
    orc2, e2 = oracle_for(FREE_BOX.replace('gravity="0 0 -9.81"', 'gravity="0 0 0"'))
    w = np.array([0.7, -1.3, 2.1])
    orc2.set_state(q0, np.r_[0, 0, 0, w], np.zeros(6), np.zeros(0))
    orc2.forward()
    m = 700 * 8 * 0.1 * 0.2 * 0.3
    I = m / 3 * np.array([0.2 ** 2 + 0.3 ** 2, 0.1 ** 2 + 0.3 ** 2, 0.1 ** 2 + 0.2 ** 2])
    assert np.allclose(orc2.get("qfrc_bias"), np.r_[0, 0, 0, np.cross(w, I * w)], atol=1e-10)


PENDULUM = """<mujoco><compiler angle="radian"/><option gravity="0 0 -9.81"/>
<worldbody><body><joint type="hinge" axis="0 1 0"/>
<geom type="sphere" size="0.05" pos="0 0 -0.7" mass="2"/></body></worldbody></mujoco>"""


def test_pendulum_closed_form():
    orc, _ = oracle_for(PENDULUM)
    m, l, r = 2.0, 0.7, 0.05
    Izz = 0.4 * m * r * r + m * l * l
    for th, thd in [(0.3, 0.0), (-1.1, 2.0), (2.5, -0.7)]:
        orc.set_state([th], [thd], [0.0], np.zeros(0))
        orc.forward()
        assert orc.get("qM")[0] == pytest.approx(Izz, rel=1e-12)
        # rotation about +y by th moves the bob at (0,0,-l) to x = -l sin th, z = -l cos th
        # gravity torque about +y = r x F = (z*Fx - x*Fz) = -(-l sin th)(-m g) -> -m g l sin th
        assert orc.get("qfrc_bias")[0] == pytest.approx(m * 9.81 * l * math.sin(th), rel=1e-10)


CHAIN = """<mujoco><compiler angle="radian"/><option gravity="0 0 -9.81"/>
<worldbody><body pos="0 0 0"><joint type="hinge" axis="0 1 0"/><geom type="capsule" fromto="0 0 0 0.3 0 0" size="0.04"/>
<body pos="0.3 0 0"><joint type="hinge" axis="0 0 1"/><geom type="box" pos="0.15 0 0" size="0.15 0.03 0.02"/>
<body pos="0.3 0 0"><joint type="hinge" axis="1 0 0" pos="0.01 0.02 0"/><geom type="capsule" fromto="0 0 0 0.2 0.1 -0.1" size="0.03"/>
</body></body></body></worldbody></mujoco>"""


def _fk(orc, q):
    orc.set_state(q, np.zeros(orc.nv), np.zeros(orc.nv), np.zeros(0))
    orc.kinematics()
    return orc.get("xipos").reshape(-1, 3), orc.get("ximat").reshape(-1, 3, 3)


def test_mass_matrix_brute_force_and_ldl():
    orc, e = oracle_for(CHAIN)
    mass, inert = e.field("body_mass"), e.field("body_inertia").reshape(-1, 3)
    rng = np.random.default_rng(1)
    for _ in range(3):
        q = rng.uniform(-1.5, 1.5, size=3)
        x0, R0 = _fk(orc, q)
        orc.set_state(q, np.zeros(3), np.zeros(3), np.zeros(0))
        orc.forward()
        M = orc.get("qM").reshape(3, 3)
        LD = orc.get("qLD").reshape(3, 3)
        eps = 1e-6
        Jv, Jw = np.zeros((len(mass), 3, 3)), np.zeros((len(mass), 3, 3))
        for k in range(3):
            dq = np.zeros(3)
            dq[k] = eps
            xp, Rp = _fk(orc, q + dq)
            xm, Rm = _fk(orc, q - dq)
            Jv[:, :, k] = (xp - xm) / (2 * eps)
            for b in range(len(mass)):
                W = (Rp[b] - Rm[b]) / (2 * eps) @ R0[b].T  # skew(omega)
                Jw[b, :, k] = [W[2, 1], W[0, 2], W[1, 0]]
        Mb = np.zeros((3, 3))
        for b in range(1, len(mass)):
            Iw = R0[b] @ np.diag(inert[b]) @ R0[b].T
            Mb += mass[b] * Jv[b].T @ Jv[b] + Jw[b].T @ Iw @ Jw[b]
        assert np.allclose(M, Mb, rtol=1e-6, atol=1e-9)
        # L^T D L with L unit lower (chain: full pattern)
        L = np.tril(LD, -1) + np.eye(3)
        D = np.diag(np.diag(LD))
        assert np.allclose(L.T @ D @ L, M, atol=1e-12)


def test_rne_matches_lagrangian():
    """qfrc_bias = dM/dt qdot - dT/dq + dV/dq (finite differences of the
    oracle's own M(q) and of the potential from body COM heights)."""
    orc, e = oracle_for(CHAIN)
    mass = e.field("body_mass")
    rng = np.random.default_rng(2)
    q, qd = rng.uniform(-1, 1, 3), rng.uniform(-2, 2, 3)

    def M_at(qq):
        orc.set_state(qq, np.zeros(3), np.zeros(3), np.zeros(0))
        orc.forward()
        return orc.get("qM").reshape(3, 3)

    def V_at(qq):
        x, _ = _fk(orc, qq)
        return float(np.sum(mass * 9.81 * x[:, 2]))

    eps = 1e-6
    dM = []
    dV = np.zeros(3)
    for k in range(3):
        dq = np.zeros(3)
        dq[k] = eps
        dM.append((M_at(q + dq) - M_at(q - dq)) / (2 * eps))
        dV[k] = (V_at(q + dq) - V_at(q - dq)) / (2 * eps)
    Mdot = sum(dM[k] * qd[k] for k in range(3))
    dT = np.array([0.5 * qd @ dM[k] @ qd for k in range(3)])
    expect = Mdot @ qd - dT + dV
    orc.set_state(q, qd, np.zeros(3), np.zeros(0))
    orc.forward()
    assert np.allclose(orc.get("qfrc_bias"), expect, rtol=1e-6, atol=1e-7)


SPHERE_PLANE = """<mujoco><compiler angle="radian"/><worldbody><geom type="plane" size="0 0 1"/>
<body pos="0 0 0.1"><freejoint/><geom type="sphere" size="0.1"/></body>
<body pos="1 0 0.1"><freejoint/><geom type="box" size="0.1 0.2 0.05"/></body>
<body pos="-1 0 0.5"><freejoint/><geom type="capsule" size="0.05 0.2" quat="0.7071068 0.7071068 0 0"/></body>
</worldbody></mujoco>"""


def test_primitive_contacts_closed_form():
    orc, e = oracle_for(SPHERE_PLANE)
    q = e.field("qpos0").copy()
    q[2] = 0.095                       # sphere 5 mm into the floor
    q[9] = 0.04                        # box bottom 1 cm below the floor
    q[16] = 0.045                      # horizontal capsule 5 mm into the floor
    orc.set_state(q, np.zeros(18), np.zeros(18), np.zeros(0))
    orc.forward()
    cons = orc.contacts()
    by_geom = {}
    for c in cons:
        by_geom.setdefault(c["geom2"], []).append(c)
    sph, box, cap = by_geom[1], by_geom[2], by_geom[3]
    assert len(sph) == 1 and sph[0]["dist"] == pytest.approx(-0.005)
    assert np.allclose(sph[0]["frame"][:3], [0, 0, 1])
    assert sph[0]["pos"][2] == pytest.approx(-0.0025)
    assert len(box) == 4 and all(c["dist"] == pytest.approx(-0.01) for c in box)
    corners = sorted((round(c["pos"][0], 6), round(c["pos"][1], 6)) for c in box)
    assert corners == [(0.9, -0.2), (0.9, 0.2), (1.1, -0.2), (1.1, 0.2)]
    assert len(cap) == 2 and all(c["dist"] == pytest.approx(-0.005) for c in cap)


TWO_CAPS = """<mujoco><compiler angle="radian"/><option gravity="0 0 0"/><worldbody>
<body><freejoint/><geom type="capsule" size="0.05 0.3"/></body>
<body><freejoint/><geom type="capsule" size="0.07 0.2"/></body>
<body><freejoint/><geom type="sphere" size="0.06"/></body>
<body><freejoint/><geom type="box" size="0.1 0.15 0.2"/></body>
</worldbody></mujoco>"""


def _rand_quat(rng):
    q = rng.normal(size=4)
    return q / np.linalg.norm(q)


def _quat2mat(q):
    w, x, y, z = q
    return np.array([[w * w + x * x - y * y - z * z, 2 * (x * y - w * z), 2 * (x * z + w * y)],
                     [2 * (x * y + w * z), w * w - x * x + y * y - z * z, 2 * (y * z - w * x)],
                     [2 * (x * z - w * y), 2 * (y * z + w * x), w * w - x * x - y * y + z * z]])


def test_narrowphase_distances_brute_force():
    orc, e = oracle_for(TWO_CAPS)
    pairs = orc.pairs()
    idx = {(p[0], p[1]): i for i, p in enumerate(pairs)}
    rng = np.random.default_rng(3)
    s = np.linspace(-1, 1, 2001)
    for _ in range(40):
        # capsule-capsule: segment distance by dense sampling
        x1, x2 = rng.uniform(-0.2, 0.2, 3), rng.uniform(-0.2, 0.2, 3)
        R1, R2 = _quat2mat(_rand_quat(rng)), _quat2mat(_rand_quat(rng))
        a = x1[None] + s[:, None] * 0.3 * R1[:, 2][None]
        b = x2[None] + s[:, None] * 0.2 * R2[:, 2][None]
        dmin = np.min(np.linalg.norm(a[:, None, :] - b[None, :, :], axis=2)) - 0.05 - 0.07
        out = (ref64.OrcContact * 4)()
        n = ref64.lib().orc_collide_pair(orc.p, idx[(0, 1)], ref64._dp(np.ascontiguousarray(x1)),
                                         ref64._dp(np.ascontiguousarray(R1.ravel())),
                                         ref64._dp(np.ascontiguousarray(x2)),
                                         ref64._dp(np.ascontiguousarray(R2.ravel())), out, 4)
        if dmin < -1e-3:
            assert n == 1 and out[0].dist == pytest.approx(dmin, abs=2e-3)
        # sphere-box: distance to the box by dense surface sampling (outside) / depth (inside)
        c = rng.uniform(-0.3, 0.3, 3)
        Rb = _quat2mat(_rand_quat(rng))
        hs = np.array([0.1, 0.15, 0.2])
        pl = Rb.T @ c
        if np.all(np.abs(pl) <= hs):
            d_expect = -np.min(hs - np.abs(pl)) - 0.06
        else:
            g = np.stack(np.meshgrid(*[np.linspace(-h, h, 61) for h in hs], indexing="ij"), -1).reshape(-1, 3)
            d_expect = np.min(np.linalg.norm(g - pl[None], axis=1)) - 0.06
        n = ref64.lib().orc_collide_pair(orc.p, idx[(2, 3)], ref64._dp(np.ascontiguousarray(c)),
                                         ref64._dp(np.eye(3).ravel().copy()),
                                         ref64._dp(np.zeros(3)), ref64._dp(np.ascontiguousarray(Rb.ravel())), out, 4)
        if d_expect < -1e-3:
            assert n == 1 and out[0].dist == pytest.approx(d_expect, abs=6e-3)
        elif d_expect > 1e-3:
            assert n == 0


def _quad_oracle_state(seed, nenv=6):
    import states
    e = emu.EmuSim(scenes.quadruped_xml(), 1)
    return states.quadruped_states(ref64.fields_from(e.field), nenv, seed=seed), e


def test_solver_kkt_and_newton_equals_cg():
    (q, qv, qa, ctrl, orc), e = _quad_oracle_state(4)
    f = ref64.fields_from(e.field)
    f_cg = dict(f, solver=1, iterations=400)
    orc_cg = ref64.Oracle(f_cg)
    nv = orc.nv
    checked = 0
    for k in range(len(q)):
        for o in (orc, orc_cg):
            o.set_state(q[k].astype(float), qv[k].astype(float), qa[k].astype(float), ctrl[k].astype(float))
            o.forward()
        n = orc.nefc()
        if n == 0:
            continue
        checked += 1
        J = orc.get("efc_J").reshape(n, nv)
        F = orc.get("efc_force")
        M = orc.get("qM").reshape(nv, nv)
        a = orc.get("qacc")
        grad = M @ a - orc.get("qfrc_smooth") - J.T @ F
        assert np.linalg.norm(grad) < 1e-6 * max(1, np.linalg.norm(orc.get("qfrc_smooth")))
        jar = J @ a - orc.get("efc_aref")
        D = orc.get("efc_D")
        assert np.allclose(F, D * np.maximum(-jar, 0), atol=1e-9)  # one-sided rows
        assert np.all(F >= 0)
        a_cg = orc_cg.get("qacc")
        # CG stops on the scaled-improvement test (tolerance 1e-8), short of Newton's exact minimiser
        assert np.max(np.abs(a_cg - a)) / max(1, np.max(np.abs(a))) < 1e-5
    assert checked >= 3


def test_solution_independent_of_warm_start():
    (q, qv, qa, ctrl, orc), _ = _quad_oracle_state(5)
    for k in range(len(q)):
        res = []
        for ws in (np.zeros(orc.nv), qa[k].astype(float), np.full(orc.nv, 50.0)):
            orc.set_state(q[k].astype(float), qv[k].astype(float), ws, ctrl[k].astype(float))
            orc.forward()
            res.append(orc.get("qacc"))
        for r in res[1:]:
            assert np.max(np.abs(r - res[0])) / max(1, np.max(np.abs(res[0]))) < 1e-7


def test_joint_limit_rows():
    orc, e = oracle_for(scenes.quadruped_xml())
    q = e.field("qpos0").copy()
    q[7:] = scenes.QUAD_NOMINAL  # inside every range
    q[7 + 2] = -2.8     # FL calf below its lower limit -2.7
    q[7 + 4] = 3.5      # FR thigh above its upper limit 3.4
    orc.set_state(q, np.zeros(18), np.zeros(18), np.zeros(12))
    orc.forward()
    n = orc.nefc()
    J = orc.get("efc_J").reshape(n, 18)
    types = orc.get("efc_type")
    lim = J[types == 3]
    assert lim.shape[0] == 2
    assert lim[0, 6 + 2] == 1.0 and lim[1, 6 + 4] == -1.0
    assert np.allclose(orc.get("efc_pos")[types == 3], [-0.1, -0.1])


def test_torque_free_spin_integrates_exactly():
    orc, e = oracle_for(FREE_BOX.replace('gravity="0 0 -9.81"', 'gravity="0 0 0"'))
    w = np.array([0.0, 0.0, 3.0])  # principal axis: steady rotation
    q = e.field("qpos0").copy()
    qv, qa = np.r_[0, 0, 0, w], np.zeros(6)
    orc.set_state(q, qv, qa, np.zeros(0))
    for _ in range(100):
        orc.step()
    qp, qvel, _ = orc.get_state()
    ang = 3.0 * 100 * 0.002
    assert np.allclose(qp[3:], [math.cos(ang / 2), 0, 0, math.sin(ang / 2)], atol=1e-12)
    assert np.allclose(qvel, qv, atol=1e-12)


SLIDER_SERVO = """<mujoco model="slider_servo">
  <option timestep="0.01" gravity="0 0 0" integrator="{integrator}"/>
  <worldbody>
    <body name="b">
      <joint name="s" type="slide" axis="1 0 0" damping="{damping}"/>
      <geom type="sphere" size="0.1" mass="2.0" contype="0" conaffinity="0"/>
    </body>
  </worldbody>
  <actuator>
    <position joint="s" kp="20" kv="{kv}"/>
  </actuator>
</mujoco>"""


@pytest.mark.parametrize("integrator", ["Euler", "implicitfast", "implicit"])
@pytest.mark.parametrize("damping,kv", [(0.7, 0.5), (0.0, 3.0), (2.0, 0.0)])
def test_damped_servo_step_closed_form(integrator, damping, kv):
    """A 2 kg slider driven by a position servo (kp 20, kv) with joint damping
    d, no gravity or contacts. One step from (q, v) with ctrl c has the
    closed form a = (kp (c - q) - (d + kv) v) / (m + h d') with d' = d for
    Euler (implicit joint damping) and d' = d + kv for implicit/implicitfast
    (the servo's velocity derivative folded in as well); then v' = v + h a,
    q' = q + h v'. Dropping either fold, or its sign, fails the check."""
    orc, _ = oracle_for(SLIDER_SERVO.format(integrator=integrator, damping=damping, kv=kv))
    m, h, kp = 2.0, 0.01, 20.0
    q, v, c = 0.3, -1.2, 0.9
    orc.set_state(np.array([q]), np.array([v]), np.zeros(1), np.array([c]))
    orc.step()
    qp, qv, _ = orc.get_state()
    dd = damping + (kv if integrator != "Euler" else 0.0)
    a = (kp * (c - q) - (damping + kv) * v) / (m + h * dd)
    assert qv[0] == pytest.approx(v + h * a, rel=1e-12, abs=1e-14)
    assert qp[0] == pytest.approx(q + h * (v + h * a), rel=1e-12, abs=1e-14)
