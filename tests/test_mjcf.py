"""MJCF compiler (csrc/mjcf.cpp) checks.

Mass properties are pinned by numerical volume integration (not by the
closed forms the compiler uses); the pair table and setConst quantities by
the oracle's independent FP64 implementation; structure by the scene
definitions.
"""
from __future__ import annotations

import numpy as np
import pytest

import emu
import ref64
from mjmlx_hip import scenes


def compile_(xml, **kw):
    return emu.CompiledModel(xml, **kw)


def _grid_mass_props(inside, lo, hi, n=140):
    ax = [np.linspace(lo[i], hi[i], n, endpoint=False) + (hi[i] - lo[i]) / (2 * n) for i in range(3)]
    X, Y, Z = np.meshgrid(*ax, indexing="ij")
    P = np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1)
    m = inside(P)
    dv = np.prod([(hi[i] - lo[i]) / n for i in range(3)])
    P = P[m]
    vol = len(P) * dv
    c = P.mean(0)
    d = P - c
    I = np.array([[np.sum(d[:, 1] ** 2 + d[:, 2] ** 2), -np.sum(d[:, 0] * d[:, 1]), -np.sum(d[:, 0] * d[:, 2])],
                  [-np.sum(d[:, 0] * d[:, 1]), np.sum(d[:, 0] ** 2 + d[:, 2] ** 2), -np.sum(d[:, 1] * d[:, 2])],
                  [-np.sum(d[:, 0] * d[:, 2]), -np.sum(d[:, 1] * d[:, 2]), np.sum(d[:, 0] ** 2 + d[:, 1] ** 2)]]) * dv
    return vol, c, I


def _body_props(cm, b):
    mass = cm.field("body_mass")[b]
    ipos = cm.field("body_ipos").reshape(-1, 3)[b]
    q = cm.field("body_iquat").reshape(-1, 4)[b]
    w, x, y, z = q
    R = np.array([[w * w + x * x - y * y - z * z, 2 * (x * y - w * z), 2 * (x * z + w * y)],
                  [2 * (x * y + w * z), w * w - x * x + y * y - z * z, 2 * (y * z - w * x)],
                  [2 * (x * z - w * y), 2 * (y * z + w * x), w * w - x * x - y * y + z * z]])
    I = R @ np.diag(cm.field("body_inertia").reshape(-1, 3)[b]) @ R.T
    return mass, ipos, I


SHAPES = {
    "box": ('type="box" size="0.1 0.2 0.3"', lambda P: np.all(np.abs(P) <= [0.1, 0.2, 0.3], 1), 0.35),  # faces on cell edges
    "sphere": ('type="sphere" size="0.15"', lambda P: np.linalg.norm(P, axis=1) <= 0.15, 0.16),
    "capsule": ('type="capsule" size="0.1 0.2"',
                lambda P: np.linalg.norm(P - np.c_[0 * P[:, 0], 0 * P[:, 1], np.clip(P[:, 2], -0.2, 0.2)], axis=1) <= 0.1,
                0.31),
    "cylinder": ('type="cylinder" size="0.1 0.2"',
                 lambda P: (np.linalg.norm(P[:, :2], axis=1) <= 0.1) & (np.abs(P[:, 2]) <= 0.2), 0.21),
    "ellipsoid": ('type="ellipsoid" size="0.1 0.15 0.2"',
                  lambda P: np.sum((P / [0.1, 0.15, 0.2]) ** 2, 1) <= 1, 0.21),
}


@pytest.mark.parametrize("shape", list(SHAPES))
def test_geom_mass_properties_by_integration(shape):
    attr, inside, ext = SHAPES[shape]
    cm = compile_(f'<mujoco><worldbody><body><freejoint/><geom {attr} density="800"/></body></worldbody></mujoco>')
    vol, c, I = _grid_mass_props(inside, [-ext] * 3, [ext] * 3)
    mass, ipos, Ib = _body_props(cm, 1)
    assert mass == pytest.approx(800 * vol, rel=1e-2)
    assert np.allclose(ipos, c, atol=2e-3)
    assert np.allclose(Ib, 800 * I, rtol=3e-2, atol=3e-2 * np.max(np.abs(800 * I)))


def test_multi_geom_body_by_integration():
    xml = """<mujoco><compiler angle="radian"/><worldbody><body><freejoint/>
      <geom type="box" size="0.1 0.05 0.05" pos="0.1 0 0" euler="0 0 0.4"/>
      <geom type="sphere" size="0.08" pos="-0.1 0.05 0.02"/>
      <geom type="capsule" fromto="0 0 0 0 0.1 0.2" size="0.03"/></body></worldbody></mujoco>"""
    cm = compile_(xml)
    c4, s4 = np.cos(0.2), np.sin(0.2)
    Rz = np.array([[np.cos(0.4), -np.sin(0.4), 0], [np.sin(0.4), np.cos(0.4), 0], [0, 0, 1]])
    a = np.array([0, 0.1, 0.2])

    def inside(P):
        pb = (P - [0.1, 0, 0]) @ Rz
        box = np.all(np.abs(pb) <= [0.1, 0.05, 0.05], 1)
        sph = np.linalg.norm(P - [-0.1, 0.05, 0.02], axis=1) <= 0.08
        t = np.clip(P @ a / (a @ a), 0, 1)
        cap = np.linalg.norm(P - t[:, None] * a[None], axis=1) <= 0.03
        return box | sph | cap

    # overlapping geoms are summed by the compiler (as in MuJoCo), so integrate each separately
    vols = []
    for f in (lambda P: np.all(np.abs((P - [0.1, 0, 0]) @ Rz) <= [0.1, 0.05, 0.05], 1),
              lambda P: np.linalg.norm(P - [-0.1, 0.05, 0.02], axis=1) <= 0.08,
              lambda P: np.linalg.norm(P - np.clip(P @ a / (a @ a), 0, 1)[:, None] * a[None], axis=1) <= 0.03):
        vols.append(_grid_mass_props(f, [-0.25, -0.2, -0.15], [0.25, 0.2, 0.3], n=160))
    M = sum(1000 * v for v, _, _ in vols)
    C = sum(1000 * v * c for v, c, _ in vols) / M
    I = np.zeros((3, 3))
    for v, c, Ii in vols:
        d = c - C
        I += 1000 * Ii + 1000 * v * (d @ d * np.eye(3) - np.outer(d, d))
    mass, ipos, Ib = _body_props(cm, 1)
    assert mass == pytest.approx(M, rel=1.5e-2)
    assert np.allclose(ipos, C, atol=3e-3)
    assert np.allclose(Ib, I, rtol=5e-2, atol=5e-2 * np.max(np.abs(I)))


def test_quadruped_structure_and_pairs():
    for sc, npair in ((True, 67), (False, 13)):
        cm = compile_(scenes.quadruped_xml(self_collision=sc))
        assert [int(cm.field(k)[0]) for k in ("nq", "nv", "nu", "nbody", "ngeom")] == [19, 18, 12, 14, 14]
        g1, g2 = cm.field("cpair_g1").astype(int), cm.field("cpair_g2").astype(int)
        assert len(g1) == npair
        body = cm.field("geom_bodyid").astype(int)
        par = cm.field("body_parentid").astype(int)
        for a, b in zip(g1, g2):
            ba, bb = body[a], body[b]
            assert ba != bb
            if ba != 0 and bb != 0:  # the parent filter never applies to the world body
                assert par[ba] != bb and par[bb] != ba
        # the engine's table equals the oracle's independent restatement
        orc = ref64.Oracle(ref64.fields_from(cm.field))
        op = orc.pairs()
        assert [(p[0], p[1]) for p in op] == list(zip(g1, g2))
        fr = cm.field("cpair_friction").reshape(-1, 5)
        assert np.allclose(fr, [p[5] for p in op])
        assert np.allclose(fr[:, 0], 0.6)


def test_setconst_matches_oracle():
    for xml in (scenes.quadruped_xml(), scenes.humanoid_xml(), scenes.boxes_xml(6, 0)):
        cm = compile_(xml)
        orc = ref64.Oracle(ref64.fields_from(cm.field))
        assert cm.field("meaninertia")[0] == pytest.approx(orc.meaninertia(), rel=1e-10)
        assert np.allclose(cm.field("body_invweight0"), orc.body_invweight0(), rtol=1e-9, atol=1e-12)
        assert np.allclose(cm.field("dof_invweight0"), orc.dof_invweight0(), rtol=1e-9, atol=1e-12)


def test_scene_sizes():
    h = compile_(scenes.humanoid_xml())
    assert [int(h.field(k)[0]) for k in ("nq", "nv", "nu")] == [28, 27, 21]
    for n in (4, 17, 32):
        b = compile_(scenes.boxes_xml(n, n))
        assert [int(b.field(k)[0]) for k in ("nq", "nv", "nbody", "ngeom")] == [7 * n, 6 * n, n + 1, n + 1]
        assert len(b.field("cpair_g1")) == n + n * (n - 1) // 2


def test_defaults_fromto_and_options():
    xml = """<mujoco><compiler angle="degree"/>
      <default><joint damping="2"/><geom friction="0.7 0.01 0.001"/>
        <default class="leg"><joint damping="5" range="-30 30"/><geom condim="6"/></default></default>
      <worldbody><geom type="plane" size="0 0 1"/>
        <body childclass="leg" pos="0 0 1"><joint type="hinge" axis="0 1 0"/>
          <geom type="capsule" fromto="0 0 0 0 0 -0.4" size="0.05"/>
          <body pos="0 0 -0.4"><joint class="main" type="hinge"/><geom type="sphere" size="0.05"/></body>
        </body></worldbody></mujoco>"""
    cm = compile_(xml)
    assert list(cm.field("dof_damping")) == [5.0, 2.0]
    assert np.allclose(cm.field("jnt_range")[:2], np.radians([-30, 30]))
    assert list(cm.field("jnt_limited")) == [1, 0]
    assert list(cm.field("geom_condim")) == [3, 6, 6]  # childclass reaches nested bodies
    assert np.allclose(cm.field("geom_friction").reshape(-1, 3)[1], [0.7, 0.01, 0.001])
    size = cm.field("geom_size").reshape(-1, 3)[1]
    assert size[0] == pytest.approx(0.05) and size[1] == pytest.approx(0.2)
    assert np.allclose(cm.field("geom_pos").reshape(-1, 3)[1], [0, 0, -0.2])
    # load-time options (reference src/io.cpp:567-594): robot geoms only
    cm2 = compile_(xml, self_collision=0, max_condim=3)
    assert list(cm2.field("geom_conaffinity")) == [1, 0, 0]
    assert list(cm2.field("geom_condim")) == [3, 3, 3]


@pytest.mark.parametrize("bad", [
    "<mujoco><worldbody><body><geom type='sphere' size='0.1'/></worldbody></mujoco>",
    "<mujoco><worldbody><body><joint type='screw'/><geom type='sphere' size='0.1'/></body></worldbody></mujoco>",
    "<mujoco><worldbody><geom type='mesh' mesh='m'/></worldbody></mujoco>",
    "<notmujoco/>",
    "<mujoco><worldbody><body><freejoint/></body></worldbody></mujoco>",
])
def test_invalid_models_are_rejected(bad):
    with pytest.raises(RuntimeError):
        compile_(bad)
