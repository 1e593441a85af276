"""Elliptic friction cones in the FP64 oracle.

Pins: the cost's gradient and Hessian against central finite differences of
the cost (the primal cone cost is C1, its Hessian piecewise smooth), KKT
optimality of the converged forces (zero gradient, forces inside the
elliptic cone, on its surface when sliding), Newton == CG, and the
inclined-plane closed forms: a sliding block's friction is mu times its
normal force and a block below the friction angle sticks.

MuJoCo's exact elliptic details (its regularisation of mu with impratio and
the friction rows' R) are unpinned: MuJoCo is not available here.
"""
from __future__ import annotations

import math

import numpy as np
import pytest

import emu
import ref64

MIXED = """<mujoco><option cone="elliptic" solver="Newton" iterations="100" tolerance="1e-12" impratio="1"/>
<worldbody><geom type="plane" size="0 0 1" friction="0.8 0.02 0.002"/>
<body pos="0 0 0.2"><freejoint/><geom type="box" size="0.1 0.08 0.05" condim="3" friction="0.7 0.02 0.002"/></body>
<body pos="0.5 0 0.2"><freejoint/><geom type="capsule" size="0.04 0.15" condim="6" friction="0.9 0.05 0.01"/></body>
<body pos="-0.5 0 0.2"><freejoint/><geom type="sphere" size="0.06" condim="4" friction="0.6 0.03 0.002"/></body>
</worldbody></mujoco>"""


def oracle_for(xml, **over):
    e = emu.CompiledModel(xml)
    f = ref64.fields_from(e.field)
    f.update(over)
    return ref64.Oracle(f), e


def _mixed_state(rng, orc, e):
    q = e.field("qpos0").copy()
    # box resting slightly tilted, capsule lying, sphere touching; random velocities
    for b, (z, quat) in enumerate([(0.048, [1, 0.05, 0.02, 0]), (0.038, [0.7071, 0, 0.7071, 0]), (0.059, [1, 0, 0, 0])]):
        qa = 7 * b
        q[qa + 2] = z + rng.uniform(-0.003, 0.001)
        qq = np.array(quat, float) + rng.normal(0, 0.02, 4)
        q[qa + 3:qa + 7] = qq / np.linalg.norm(qq)
    qv = rng.normal(0, 0.5, orc.nv)
    return q, qv


def _forward(orc, q, qv):
    orc.set_state(q, qv, np.zeros(orc.nv), np.zeros(0))
    orc.forward()


def test_cost_derivatives_by_finite_differences():
    orc, e = oracle_for(MIXED)
    rng = np.random.default_rng(0)
    nv = orc.nv
    zones_seen = set()
    hess_ok = hess_n = 0
    for trial in range(12):
        q, qv = _mixed_state(rng, orc, e)
        _forward(orc, q, qv)
        n = orc.nefc()
        assert n > 0 and (orc.get("efc_type") == 7).all()
        a0 = orc.get("qacc")
        for _ in range(4):
            a = a0 + rng.normal(0, 2.0, nv)
            c, g, H = orc.cost(a)
            zones_seen.update(int(z) for z in orc.get("efc_active")[:n])
            h = 1e-6
            fd = np.array([(orc.cost(a + h * np.eye(nv)[i])[0] - orc.cost(a - h * np.eye(nv)[i])[0]) / (2 * h)
                           for i in range(nv)])
            assert np.max(np.abs(fd - g)) < 1e-5 * max(1.0, np.max(np.abs(g))), (fd - g)
            hh = 1e-7
            Hfd = np.array([(orc.cost(a + hh * np.eye(nv)[i])[1] - orc.cost(a - hh * np.eye(nv)[i])[1]) / (2 * hh)
                            for i in range(nv)])
            hess_n += 1
            hess_ok += np.max(np.abs(Hfd - H)) < 1e-4 * max(1.0, np.max(np.abs(H)))
            assert np.allclose(H, H.T, atol=1e-9 * max(1.0, np.abs(H).max()))
            assert np.linalg.eigvalsh(H).min() > 0
    assert zones_seen >= {0, 1, 2}, zones_seen  # top, bottom and middle zones all exercised
    assert hess_ok >= 0.9 * hess_n  # FD across a zone boundary is the only miss


def _cone_check(orc, e):
    """forces of each elliptic contact: f_n >= 0 and |f_t / friction| <= f_n"""
    n = orc.nefc()
    F = orc.get("efc_force")
    ids = orc.get("efc_id").astype(int)
    cons = orc.contacts()
    r = 0
    sliding = 0
    while r < n:
        c = cons[ids[r]]
        dim = c["dim"]
        fn = F[r]
        ft = F[r + 1:r + dim] / np.asarray(c["friction"][:dim - 1])
        assert fn >= -1e-12
        assert np.linalg.norm(ft) <= fn * (1 + 1e-9) + 1e-12
        sliding += abs(np.linalg.norm(ft) - fn) <= 1e-9 * max(1.0, fn) and fn > 0
        r += dim
    return sliding


def test_kkt_and_newton_equals_cg():
    orc, e = oracle_for(MIXED)
    orc_cg, _ = oracle_for(MIXED, solver=1, iterations=2000)
    rng = np.random.default_rng(1)
    nv = orc.nv
    sliding = 0
    for trial in range(10):
        q, qv = _mixed_state(rng, orc, e)
        _forward(orc, q, qv)
        _forward(orc_cg, q, qv)
        n = orc.nefc()
        a = orc.get("qacc")
        c, g, _ = orc.cost(a)
        assert np.linalg.norm(g) < 1e-6 * max(1.0, np.linalg.norm(orc.get("qfrc_smooth")))
        J = orc.get("efc_J").reshape(n, nv)
        M = orc.get("qM").reshape(nv, nv)
        assert np.allclose(M @ a - orc.get("qfrc_smooth") - J.T @ orc.get("efc_force"), g, atol=1e-9)
        sliding += _cone_check(orc, e)
        assert np.max(np.abs(orc_cg.get("qacc") - a)) / max(1.0, np.max(np.abs(a))) < 1e-4
    assert sliding > 0  # some contacts sit on the cone surface


SLOPE = """<mujoco><compiler angle="radian"/><option cone="elliptic" solver="Newton" iterations="100" tolerance="1e-12" impratio="1"/>
<worldbody><geom type="plane" size="0 0 1" euler="0 {th} 0" friction="{mu} 0.005 0.0001"/>
<body pos="{px} 0 {pz}" euler="0 {th} 0"><freejoint/>
<geom type="box" size="0.1 0.1 0.05" condim="3" friction="{mu} 0.005 0.0001"/></body>
</worldbody></mujoco>"""


@pytest.mark.parametrize("deg,mu", [(30.0, 0.3), (10.0, 0.6)])
def test_inclined_plane(deg, mu):
    th = math.radians(deg)
    n = np.array([math.sin(th), 0, math.cos(th)])
    c = n * (0.05 - 0.0005)
    orc, e = oracle_for(SLOPE.format(th=th, mu=mu, px=c[0], pz=c[2]))
    _forward(orc, e.field("qpos0"), np.zeros(6))
    a = orc.get("qacc")[:3]
    down = np.array([math.cos(th), 0, -math.sin(th)])
    F = orc.get("efc_force")
    ncon = len(F) // 3
    fn = F[0::3].sum()
    ft = np.hypot(F[1::3], F[2::3])
    m = float(e.field("body_mass")[1])
    g = 9.81
    if math.tan(th) > mu:
        # sliding: every contact on the cone surface, so the friction is mu times the normal force
        assert np.allclose(ft, mu * F[0::3], rtol=1e-9)
        assert a @ down == pytest.approx(g * math.sin(th) - mu * fn / m, rel=1e-6)
        assert a @ down > 0.5 * g * (math.sin(th) - mu * math.cos(th))
    else:
        # sticking: strictly inside the cone; the regularised (soft) friction rows
        # let the block creep by about m g sin(th) / (m + sum D) along the slope
        assert ncon == 4 and np.all(ft < mu * F[0::3])
        assert abs(a @ down) < 0.05 * g * math.sin(th)


def _settle(orc, q, qv, qa, steps, nsub=1):
    for _ in range(steps):
        qs, vs, as_ = q[None].copy(), qv[None].copy(), qa[None].copy()
        orc.batch_step(qs, vs, as_, np.zeros((1, orc.nu)), nsub)
        q, qv, qa = qs[0], vs[0], as_[0]
    return q, qv, qa


@pytest.mark.parametrize("scene", ["mixed", "humanoid"])
def test_interpreter_matches_oracle(scene):
    """The kernel source (serial interpreter, FP32) against the FP64 oracle on
    elliptic contacts (condim 3, 4 and 6)."""
    import states
    from mjmlx_hip import scenes
    xml = MIXED if scene == "mixed" else scenes.humanoid_xml()
    B = 12
    sim = emu.EmuSim(xml, B, maxcon=32, maxefc=128)
    orc = ref64.Oracle(ref64.fields_from(sim.field))
    rng = np.random.default_rng(5)
    q0 = sim.field("qpos0")
    qs, vs, as_ = [], [], []
    for e in range(B):
        q = q0.copy()
        qv = rng.normal(0, 0.3, sim.nv)
        if scene == "mixed":
            q, qv = _mixed_state(rng, orc, sim)
        q, qv, qa = _settle(orc, q, qv, np.zeros(sim.nv), int(rng.integers(0, 60)), 4)
        qs.append(q); vs.append(qv); as_.append(qa)
    f32 = lambda a: np.ascontiguousarray(np.array(a), dtype=np.float32)
    q, qv, qa = f32(qs), f32(vs), f32(as_)
    sim.set_state(q, qv, qa)
    sim.step(np.zeros((B, sim.nu), np.float32) if sim.nu else None, nsub=1, dump=True)
    st = sim.stats()
    assert not (st["status"] & ~3).any(), st["status"]
    worst = worstf = worstc = 0.0
    ncon = 0
    for e in range(B):
        states.oracle_step(orc, q[e], qv[e], qa[e], np.zeros(sim.nu, np.float32))
        con = sim.dump(1, e, "contact").reshape(-1, 16)[: st["ncon"][e]]
        ok, msg = states.compare_contacts(con[:, 13].view(np.int32), con[:, 12], orc.contacts(), margin_tol=1e-4, ncon=st["ncon"][e])
        assert ok, f"env {e}: {msg}"
        ncon += len(con)
        b = orc.get("qacc")
        n = orc.nefc()
        fb = orc.get("efc_force").copy() if n else None
        qk = sim.dump(2, e, "qacc")
        worst = max(worst, np.max(np.abs(qk - b)) / max(1.0, np.max(np.abs(b))))
        if n and st["nefc"][e] == n:
            worstf = max(worstf, np.max(np.abs(sim.dump(2, e, "efc_force")[:n] - fb)) / max(1.0, np.max(np.abs(fb))))
        # optimality: the kernel's qacc reaches the FP64 optimum's cost
        c_opt = orc.cost(b)[0]
        worstc = max(worstc, (orc.cost(qk.astype(np.float64))[0] - c_opt) / states.cost_scale(orc, c_opt))
    print(f"{scene}: contacts {ncon}, worst qacc rel err {worst:.2e}, efc_force {worstf:.2e}, cost gap {worstc:.2e}")
    assert ncon >= B
    # resting boxes make the optimum flat: FP32 rounding moves qacc along it by
    # up to a few 1e-4 while the cost stays optimal to 1e-9
    assert worst < 3e-4 and worstf < 3e-4
    assert worstc < 1e-8
