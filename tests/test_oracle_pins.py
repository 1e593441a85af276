"""Closed-form pins of the FP64 oracle's constraint model, setConst quantities
and solver tolerance, and of the compiled scene fields against the workload
definitions (CPU).

Every expected value here comes from rigid-body mechanics written out in the
test (hand-built mass matrices and contact Jacobians, the stationarity of the
quadratic constraint cost) and MuJoCo's published parameter formulas
(solref -> k, b; solimp -> impedance; R from invweight0 and the impedance;
the pyramidal and elliptic friction regularisers), never from the oracle's
own intermediate arrays. A dropped term, a wrong power or a missing
impratio in the oracle moves qacc away from the closed form.
"""
from __future__ import annotations

import re

import numpy as np
import pytest

import emu
import ref64
from mjmlx_hip import scenes


def oracle_for(xml: str):
    e = emu.CompiledModel(xml)
    return ref64.Oracle(ref64.fields_from(e.field)), e


# ---------------------------------------------------------------- invweight0

def _skew(c):
    return np.array([[0, -c[2], c[1]], [c[2], 0, -c[0]], [-c[1], c[0], 0]])


def test_free_body_invweight0_and_meaninertia_closed_form():
    """A free box whose COM sits off the body origin. Free-joint dofs are the
    world velocity of the body origin and the body-frame angular velocity, so
    with v_com = v - [c]x w the mass matrix is m A^T A + diag(0, I_c),
    A = [I, -[c]x]. dof_invweight0 is the mean of M^-1's translational and
    rotational diagonal, body_invweight0 that of J M^-1 J^T at the COM
    (= 1/m for the translation), meaninertia = trace(M)/nv."""
    hs = np.array([0.1, 0.05, 0.02])
    c = np.array([0.05, -0.02, 0.03])
    rho = 900.0
    xml = f"""<mujoco><compiler angle="radian"/><worldbody><body pos="0 0 1"><freejoint/>
      <geom type="box" size="{hs[0]} {hs[1]} {hs[2]}" pos="{c[0]} {c[1]} {c[2]}" density="{rho}"/></body></worldbody></mujoco>"""
    orc, _ = oracle_for(xml)
    m = rho * 8 * np.prod(hs)
    Ic = np.diag(m / 3 * np.array([hs[1] ** 2 + hs[2] ** 2, hs[0] ** 2 + hs[2] ** 2, hs[0] ** 2 + hs[1] ** 2]))
    A = np.hstack([np.eye(3), -_skew(c)])
    M = m * A.T @ A
    M[3:, 3:] += Ic
    Minv = np.linalg.inv(M)
    d = np.diag(Minv)
    assert np.allclose(orc.dof_invweight0(), [d[:3].mean()] * 3 + [d[3:].mean()] * 3, rtol=1e-10)
    assert orc.meaninertia() == pytest.approx(np.trace(M) / 6, rel=1e-12)
    Jr = np.hstack([np.zeros((3, 3)), np.eye(3)])
    tran = np.diag(A @ Minv @ A.T).mean()
    rot = np.diag(Jr @ Minv @ Jr.T).mean()
    assert tran == pytest.approx(1 / m, rel=1e-12)  # the COM decouples
    biw = orc.body_invweight0().reshape(-1, 2)
    assert np.allclose(biw[1], [tran, rot], rtol=1e-10)
    assert np.allclose(biw[0], 0)


def test_pendulum_invweight0_closed_form():
    """A point-like bob (sphere r, mass m) on a hinge at distance l: M = I_zz =
    2/5 m r^2 + m l^2; the COM's translational Jacobian is axis x (x_com -
    anchor) (one nonzero of size l), its rotational one the axis."""
    m, l, r = 2.0, 0.7, 0.05
    xml = f"""<mujoco><compiler angle="radian"/><worldbody><body><joint type="hinge" axis="0 1 0"/>
      <geom type="sphere" size="{r}" pos="0 0 {-l}" mass="{m}"/></body></worldbody></mujoco>"""
    orc, _ = oracle_for(xml)
    Izz = 0.4 * m * r * r + m * l * l
    assert orc.dof_invweight0()[0] == pytest.approx(1 / Izz, rel=1e-12)
    assert orc.meaninertia() == pytest.approx(Izz, rel=1e-12)
    assert np.allclose(orc.body_invweight0().reshape(-1, 2)[1], [l * l / (3 * Izz), 1 / (3 * Izz)], rtol=1e-10)


def test_two_link_invweight0_closed_form():
    """Two hinges about parallel axes (double pendulum of two point-like bobs):
    M from the textbook double-pendulum energy at the straight pose."""
    m1, m2, l1, l2, r = 1.5, 0.8, 0.4, 0.3, 0.01
    xml = f"""<mujoco><compiler angle="radian"/><worldbody>
      <body><joint type="hinge" axis="0 1 0"/><geom type="sphere" size="{r}" pos="0 0 {-l1}" mass="{m1}"/>
        <body pos="0 0 {-l1}"><joint type="hinge" axis="0 1 0"/><geom type="sphere" size="{r}" pos="0 0 {-l2}" mass="{m2}"/>
      </body></body></worldbody></mujoco>"""
    orc, _ = oracle_for(xml)
    i1, i2 = 0.4 * m1 * r * r, 0.4 * m2 * r * r
    # straight pose: both bobs on the same vertical line below the first hinge
    M = np.array([[i1 + i2 + m1 * l1 ** 2 + m2 * (l1 + l2) ** 2, i2 + m2 * l2 * (l1 + l2)],
                  [i2 + m2 * l2 * (l1 + l2), i2 + m2 * l2 ** 2]])
    Minv = np.linalg.inv(M)
    assert np.allclose(orc.dof_invweight0(), np.diag(Minv), rtol=1e-10)
    assert orc.meaninertia() == pytest.approx(np.trace(M) / 2, rel=1e-12)
    # body 2's COM: translational Jacobian columns (along x) l1 + l2 and l2
    jp = np.array([l1 + l2, l2])
    jr = np.array([1.0, 1.0])
    biw = orc.body_invweight0().reshape(-1, 2)
    assert biw[2][0] == pytest.approx(jp @ Minv @ jp / 3, rel=1e-10)
    assert biw[2][1] == pytest.approx(jr @ Minv @ jr / 3, rel=1e-10)
    assert biw[1][0] == pytest.approx(l1 * l1 * Minv[0, 0] / 3, rel=1e-10)


# ---------------------------------------------------- contact constraint model

SPHERE_ON_PLANE = """<mujoco><compiler angle="radian"/>
<option timestep="{dt}" gravity="{gx} 0 {gz}" cone="{cone}" impratio="{impratio}" iterations="100" tolerance="1e-14"/>
<default><geom friction="{mu} 0.005 0.0001" condim="{condim}" solref="{tc} {dr}" solimp="{si}"/></default>
<worldbody><geom type="plane" size="0 0 1"/>
<body pos="0 0 1"><freejoint/><geom type="sphere" size="{r}" mass="{m}"/></body></worldbody></mujoco>"""


def _sphere_case(dt=0.002, gx=0.0, gz=-9.81, cone="pyramidal", impratio=1.0, mu=0.8, condim=1,
                 tc=0.02, dr=1.0, si=(0.9, 0.9, 0.001, 0.5, 2.0), r=0.1, m=1.3):
    xml = SPHERE_ON_PLANE.format(dt=dt, gx=gx, gz=gz, cone=cone, impratio=impratio, mu=mu, condim=condim,
                                 tc=tc, dr=dr, si=" ".join(map(str, si)), r=r, m=m)
    orc, _ = oracle_for(xml)
    return orc, dict(dt=dt, g=np.array([gx, 0.0, gz]), cone=cone, impratio=impratio, mu=mu, condim=condim,
                     tc=tc, dr=dr, si=list(si), r=r, m=m)


def _k_b(p, imp_dmax):
    """MuJoCo's solref -> (stiffness, damping): time constant and damping
    ratio (time constant at least 2 dt), or direct -stiffness / -damping."""
    dmax = min(max(imp_dmax, 1e-4), 0.9999)
    tc, dr = p["tc"], p["dr"]
    if tc > 0:
        tc = max(tc, 2 * p["dt"])
        return 1 / (dmax * dmax * tc * tc * dr * dr), 2 / (dmax * tc)
    return -tc / (dmax * dmax), -dr / dmax


def _expected_sphere(p, z, v, w, tangents):
    """qacc and row forces of a free sphere (identity orientation, COM at the
    body origin) touching the z = 0 plane, every constraint row active; the
    friction rows along the given unit tangents (any orthonormal pair with the
    normal: the closed-form qacc does not depend on it, the row forces do)."""
    m, r = p["m"], p["r"]
    I = 0.4 * m * r * r
    M = np.diag([m, m, m, I, I, I])
    a_s = np.r_[p["g"], 0, 0, 0]              # no bias force for a sphere
    dist = z - r
    rho = r + dist / 2                         # contact point midway between the surfaces
    dvec = np.array([0, 0, -rho])
    n, (t1, t2) = np.eye(3)[2], tangents
    Jn = np.r_[n, np.cross(dvec, n)]
    Jt = [np.r_[t, np.cross(dvec, t)] for t in (t1, t2)]
    qv = np.r_[v, w]
    imp = ref64.math("impedance", p["si"] + [dist])[0]
    k, b = _k_b(p, p["si"][1])
    tran = 1 / m                               # sphere's invweight0 + the world's (0)
    mu = p["mu"]
    rows, D, aref = [], [], []
    if p["condim"] == 1:
        rows = [Jn]
        D = [1 / ((1 - imp) / imp * tran)]
        aref = [-b * Jn @ qv - k * imp * dist]
    elif p["cone"] == "elliptic":
        R0 = (1 - imp) / imp * tran
        R1 = R0 / p["impratio"]
        rows = [Jn, *Jt]
        D = [1 / R0, 1 / R1, 1 / R1]
        aref = [-b * Jn @ qv - k * imp * dist] + [-b * J @ qv for J in Jt]
    else:
        Rpy = 2 * mu * mu / p["impratio"] * (1 - imp) / imp * tran * (1 + mu * mu)
        for J in Jt:
            for s in (1, -1):
                Jr = Jn + s * mu * J
                rows.append(Jr)
                D.append(1 / Rpy)
                aref.append(-b * Jr @ qv - k * imp * dist)
    J = np.array(rows)
    D, aref = np.array(D), np.array(aref)
    H = M + J.T @ (D[:, None] * J)
    a = np.linalg.solve(H, M @ a_s + J.T @ (D * aref))
    x = J @ a - aref
    if p["condim"] == 1 or p["cone"] == "pyramidal":
        assert np.all(x < 0), "closed form assumes every row active"
    else:
        # elliptic bottom zone (every row quadratic): mu N + T <= 0
        U = np.r_[mu * x[0], mu * x[1:]]
        assert mu * U[0] + np.linalg.norm(U[1:]) <= 0, "closed form assumes the cone's bottom zone"
    return a, -D * x


def _oracle_sphere(orc, z, v, w):
    """qacc, efc_force and the contact frame's tangents (checked orthonormal,
    first row the floor normal)."""
    orc.set_state([0, 0, z, 1, 0, 0, 0], np.r_[v, w], np.zeros(6), np.zeros(0))
    orc.forward()
    (c,) = orc.contacts()
    fr = c["frame"].reshape(3, 3)
    assert np.allclose(fr @ fr.T, np.eye(3), atol=1e-12) and np.allclose(fr[0], [0, 0, 1], atol=1e-12)
    return orc.get("qacc"), orc.get("efc_force"), (fr[1], fr[2])


CASES = {
    "frictionless": dict(condim=1),
    "frictionless_direct_solref": dict(condim=1, tc=-2500.0, dr=-80.0),
    "frictionless_refsafe": dict(condim=1, tc=0.001, dt=0.004),
    "frictionless_sigmoid_imp": dict(condim=1, si=(0.8, 0.95, 0.01, 0.5, 2.0)),
    "elliptic_impratio1": dict(condim=3, cone="elliptic", gx=2.0),
    "elliptic_impratio3": dict(condim=3, cone="elliptic", gx=-1.5, impratio=3.0, mu=1.1),
    "pyramidal_mu08": dict(condim=3, cone="pyramidal", gx=1.0),
    "pyramidal_mu05_impratio4": dict(condim=3, cone="pyramidal", gx=-0.7, mu=0.5, impratio=4.0),
    "pyramidal_sigmoid": dict(condim=3, cone="pyramidal", gx=0.4, mu=1.2, si=(0.85, 0.97, 0.004, 0.3, 3.0)),
}


@pytest.mark.parametrize("name", list(CASES))
def test_sphere_contact_rows_closed_form(name):
    """qacc and efc_force of a sphere pressed into the floor, at rest and
    moving (pins k, b, the impedance's use, R = (1-imp)/imp * invweight, the
    pyramidal R = 2 mu^2 (1 + mu^2) R_n / impratio and the elliptic
    R_friction = R_n / impratio through the solver's optimum)."""
    orc, p = _sphere_case(**CASES[name])
    for z, v, w in [(p["r"] - 0.002, [0, 0, 0], [0, 0, 0]),
                    (p["r"] - 0.004, [0.02, -0.01, -0.05], [0.1, 0.2, -0.3]),
                    (p["r"] - 0.0015, [-0.01, 0.015, -0.2], [0.0, -0.1, 0.05])]:
        ga, gf, tang = _oracle_sphere(orc, z, v, w)
        a, f = _expected_sphere(p, z, np.array(v), np.array(w), tang)
        assert np.allclose(ga, a, rtol=1e-9, atol=1e-9 * 9.81), (name, z, ga, a)
        assert np.allclose(gf, f, rtol=1e-9, atol=1e-9 * p["m"] * 9.81), (name, z, gf, f)


@pytest.mark.parametrize("cone,condim", [("pyramidal", 1), ("pyramidal", 3), ("elliptic", 3)])
def test_sphere_rest_depth_closed_form(cone, condim):
    """At rest the contact carries the weight: with every row active and
    v = 0, qacc = 0 exactly at the depth where the rows' summed force equals
    m g. Frictionless / elliptic: m g = k imp |d| / R_n with R_n = (1-imp)/imp/m
    -> d* = -g (1-imp) / (k imp^2). Pyramidal: four rows of R_py share it ->
    d* = -m g R_py / (4 k imp). Stepping from 1 cm above settles there."""
    orc, p = _sphere_case(cone=cone, condim=condim, si=(0.9, 0.9, 0.001, 0.5, 2.0), tc=0.02, dr=1.0, mu=0.7,
                          impratio=2.0)
    g = 9.81
    imp = 0.9
    k, _ = _k_b(p, imp)
    if condim == 3 and cone == "pyramidal":
        mu = p["mu"]
        Rpy = 2 * mu * mu / p["impratio"] * (1 - imp) / imp / p["m"] * (1 + mu * mu)
        dstar = -p["m"] * g * Rpy / (4 * k * imp)
    else:
        dstar = -g * (1 - imp) / (k * imp * imp)
    a, _, _ = _oracle_sphere(orc, p["r"] + dstar, [0, 0, 0], [0, 0, 0])
    assert np.max(np.abs(a)) < 1e-9, a
    a2, _, _ = _oracle_sphere(orc, p["r"] + 1.2 * dstar, [0, 0, 0], [0, 0, 0])
    assert a2[2] > 1e-3  # deeper: pushed back up
    orc.set_state([0, 0, p["r"] + 0.01, 1, 0, 0, 0], np.zeros(6), np.zeros(6), np.zeros(0))
    for _ in range(2000):
        orc.step()
    qp, qv, _ = orc.get_state()
    assert qp[2] - p["r"] == pytest.approx(dstar, rel=1e-6)
    assert np.max(np.abs(qv)) < 1e-8


# ----------------------------------------------------------- solver tolerance

def _scale_density(xml: str, s: float) -> str:
    return re.sub(r'density="([0-9.eE+-]+)"', lambda mt: f'density="{float(mt.group(1)) * s!r}"', xml)


@pytest.mark.parametrize("solver", ["Newton", "CG"])
def test_solver_exit_invariant_to_mass_scale(solver):
    """Scaling every mass and inertia by s = 2^k scales M, the bias forces,
    the constraint regularisers' inverse (through invweight0) and the cost by
    s exactly, and leaves aref and the optimum unchanged. The solver's exit
    tests (improvement and gradient) are scaled by 1/meaninertia, so the
    iteration sequence must be the same: identical iteration counts and qacc."""
    base = scenes.boxes_xml(5, 3, solver=solver, iterations=40, tolerance=1e-8)
    orc0, _ = oracle_for(base)
    q = np.array(orc0._keep["qpos0"])
    # boxes resting 4 mm into the floor, tumbling
    rng = np.random.default_rng(5)
    qs = []
    for _ in range(4):
        qq = q.copy().reshape(-1, 7)
        qq[:, 2] = 0.03
        qq[:, 3:] = rng.normal(size=(len(qq), 4))
        qq[:, 3:] /= np.linalg.norm(qq[:, 3:], axis=1, keepdims=True)
        qs.append((qq.ravel(), rng.normal(scale=0.3, size=orc0.nv)))
    res = {}
    for s in (2.0 ** -10, 1.0, 2.0 ** 10):
        orc, _ = oracle_for(_scale_density(base, s))
        assert orc.meaninertia() == pytest.approx(s * orc0.meaninertia(), rel=1e-14)
        out = []
        for qq, vv in qs:
            orc.set_state(qq, vv, np.zeros(orc.nv), np.zeros(0))
            orc.forward()
            out.append((orc.niter(), orc.get("qacc"), orc.nefc()))
        res[s] = out
    for s in res:
        for (n0, a0, e0), (n1, a1, e1) in zip(res[1.0], res[s]):
            assert e0 == e1 and e0 > 0
            assert n0 == n1, (s, n0, n1)
            assert np.allclose(a1, a0, rtol=1e-12, atol=1e-12 * np.max(np.abs(a0)))


# ------------------------------------------- compiled scenes vs the workloads

def test_quadruped_fields_match_workload_definition():
    """The compiled quadruped against the workload's stated numbers (torso box
    0.19 x 0.045 x 0.03 m, 6.9 kg; hip ab/ad +-0.8, thigh -1.5..3.4, calf
    -2.7..-0.9; capsule links r 0.02, length 0.21; sphere feet r 0.022;
    condim 3, friction 0.6; position actuators kp 20 / kd 0.5), independent of
    the scene writer's constants."""
    cm = emu.CompiledModel(scenes.quadruped_xml(self_collision=False))
    f = cm.field
    gt = f("geom_type").astype(int)
    size = f("geom_size").reshape(-1, 3)
    PLANE, SPHERE, CAPSULE, BOX = 0, 2, 3, 6
    assert list(gt) == [PLANE, BOX] + [CAPSULE, CAPSULE, SPHERE] * 4
    assert np.allclose(size[1], [0.19, 0.045, 0.03])
    assert f("body_mass")[1] == pytest.approx(6.9)
    caps = size[gt == CAPSULE]
    assert np.allclose(caps[:, 0], 0.02) and np.allclose(caps[:, 1], 0.105)
    assert np.allclose(size[gt == SPHERE][:, 0], 0.022)
    gpos = f("geom_pos").reshape(-1, 3)
    assert np.allclose(gpos[gt == CAPSULE], [0, 0, -0.105])
    assert np.allclose(gpos[gt == SPHERE], [0, 0, -0.21])
    rng = f("jnt_range").reshape(-1, 2)[1:]
    assert np.allclose(rng, [[-0.8, 0.8], [-1.5, 3.4], [-2.7, -0.9]] * 4)
    assert list(f("jnt_limited").astype(int)) == [0] + [1] * 12
    ax = f("jnt_axis").reshape(-1, 3)[1:]
    assert np.allclose(ax, [[1, 0, 0], [0, 1, 0], [0, 1, 0]] * 4)
    bpos = f("body_pos").reshape(-1, 3)
    hips = bpos[2::3][:4]
    assert np.allclose(np.abs(hips), [[0.19, 0.045, 0]] * 4)
    assert np.allclose(bpos[4::3][:4], [[0, 0, -0.21]] * 4)   # calves one link below the thighs
    assert np.allclose(f("geom_friction").reshape(-1, 3)[:, 0], 0.6)
    assert np.all(f("geom_condim").astype(int) == 3)
    assert np.allclose(f("actuator_gainprm").reshape(12, -1)[:, 0], 20)
    assert np.allclose(f("actuator_biasprm").reshape(12, -1)[:, :3], [0, -20, -0.5])
    # no self-collision: the floor pairs only (torso + 12 leg geoms)
    assert len(f("cpair_g1")) == 13 and set(f("cpair_g1").astype(int)) == {0}


def test_boxes_fields_match_workload_definition():
    """Boxes: half-sizes in U(0.02, 0.1), densities in U(500, 2000) (mass =
    density * volume), friction in U(0.3, 1.0), unit orientations, drop
    heights U(0.1, 1.0) above the bounding radius; the draws of the scene's
    seeded generator reproduced here."""
    n, seed = 9, 4
    cm = emu.CompiledModel(scenes.boxes_xml(n, seed))
    f = cm.field
    rng = np.random.default_rng(seed)
    size = f("geom_size").reshape(-1, 3)[1:]
    mass = f("body_mass")[1:]
    fr = f("geom_friction").reshape(-1, 3)[1:, 0]
    bq = f("body_quat").reshape(-1, 4)[1:]
    bz = f("body_pos").reshape(-1, 3)[1:, 2]
    for i in range(n):
        hs = rng.uniform(0.02, 0.1, size=3)
        dens = rng.uniform(500.0, 2000.0)
        h = rng.uniform(0.1, 1.0) + np.linalg.norm(hs)
        quat = rng.normal(size=4)
        quat /= np.linalg.norm(quat)
        mu = rng.uniform(0.3, 1.0)
        rng.uniform(-0.05, 0.05), rng.uniform(-0.05, 0.05)
        assert np.allclose(size[i], hs, rtol=1e-12)
        assert mass[i] == pytest.approx(round(dens, 3) * 8 * np.prod(hs), rel=1e-9)
        assert fr[i] == pytest.approx(round(mu, 4))
        assert np.allclose(np.abs(bq[i] @ quat), 1, atol=1e-12)
        assert bz[i] == pytest.approx(h, abs=1e-6)
    assert np.all((size >= 0.02) & (size <= 0.1))


def test_humanoid_fields_match_workload_definition():
    """Humanoid: free joint + 21 hinges (nv 27), capsule limbs r 0.03-0.06,
    condim 6 elliptic cones, motor actuators with gear 100."""
    cm = emu.CompiledModel(scenes.humanoid_xml())
    f = cm.field
    assert [int(f(k)[0]) for k in ("nq", "nv", "nu")] == [28, 27, 21]
    jt = f("jnt_type").astype(int)
    assert jt[0] == 0 and np.all(jt[1:] == 3) and len(jt) == 22   # free, then hinges
    gt = f("geom_type").astype(int)
    caps = f("geom_size").reshape(-1, 3)[gt == 3][:, 0]
    assert caps.min() == pytest.approx(0.03) and caps.max() == pytest.approx(0.06)
    assert np.all(f("geom_condim").astype(int) == 6)
    assert int(f("cone")[0]) == 1
    assert np.allclose(f("actuator_gear").reshape(21, -1)[:, 0], 100)
    assert np.allclose(f("actuator_gainprm").reshape(21, -1)[:, 0], 1)
    assert np.allclose(f("actuator_biasprm").reshape(21, -1), 0)


def test_injected_contacts_reproduce_the_narrow_phase():
    """A forward on an injected contact list equal to the narrow phase's own
    gives bit-identical results (the hook used to solve another
    implementation's contact set in FP64); restoring the narrow phase too."""
    orc, _ = oracle_for(scenes.boxes_xml(5, 3, solver="Newton", iterations=40))
    q = np.array(orc._keep["qpos0"]).reshape(-1, 7)
    q[:, 2] = 0.03
    rng = np.random.default_rng(1)
    q[:, 3:] = rng.normal(size=(len(q), 4))
    q[:, 3:] /= np.linalg.norm(q[:, 3:], axis=1, keepdims=True)
    v = rng.normal(scale=0.2, size=orc.nv)
    orc.set_state(q.ravel(), v, np.zeros(orc.nv), np.zeros(0))
    orc.forward()
    cons = orc.contacts()
    assert len(cons) > 4
    a0, f0 = orc.get("qacc").copy(), orc.get("efc_force").copy()
    orc.inject_contacts([c["pair"] for c in cons], [c["dist"] for c in cons], [c["pos"] for c in cons],
                        [c["frame"][:3] for c in cons])
    orc.forward()
    assert np.array_equal(orc.get("qacc"), a0) and np.array_equal(orc.get("efc_force"), f0)
    orc.inject_contacts([], [], np.zeros((0, 3)), np.zeros((0, 3)))
    orc.forward()
    assert orc.nefc() == 0 or len(orc.contacts()) == 0
    orc.inject_contacts()
    orc.forward()
    assert np.array_equal(orc.get("qacc"), a0)
