"""mj_ray (reference src/batched.cpp:1739-1893): the oracle's ray-primitive
intersection pinned by closed forms and by brute force (the first sign change
of each primitive's signed distance along the ray, refined by bisection),
then the kernel source (serial interpreter, FP32) against the oracle, per
primitive and per env on a scene at rolled-out states."""
from __future__ import annotations

import numpy as np
import pytest

import emu
import ref64
import states
from mjmlx_hip import scenes

PLANE, SPHERE, CAPSULE, ELLIPSOID, CYLINDER, BOX = 0, 2, 3, 4, 5, 6


def sdf(t, s, q):
    """signed distance (or a function with the same sign) of points q [..., 3]
    to the primitive"""
    q = np.asarray(q, np.float64)
    x, y, z = q[..., 0], q[..., 1], q[..., 2]
    if t == SPHERE:
        return np.linalg.norm(q, axis=-1) - s[0]
    if t == ELLIPSOID:
        return (x / s[0]) ** 2 + (y / s[1]) ** 2 + (z / s[2]) ** 2 - 1
    if t == CAPSULE:
        return np.sqrt(x * x + y * y + (z - np.clip(z, -s[1], s[1])) ** 2) - s[0]
    if t == CYLINDER:
        return np.maximum(np.hypot(x, y) - s[0], np.abs(z) - s[1])
    if t == BOX:
        return np.maximum(np.maximum(np.abs(x) - s[0], np.abs(y) - s[1]), np.abs(z) - s[2])
    raise ValueError(t)


def brute(t, s, p, v, xmax=20.0, n=40000):
    """first crossing of the surface along p + x v, or -1; None when the ray
    comes within a hair of the surface without crossing (grazing)"""
    xs = np.linspace(0, xmax, n)
    f = sdf(t, s, p[None, :] + xs[:, None] * v[None, :])
    inside = f[0] < 0
    idx = np.nonzero((f > 0) if inside else (f < 0))[0]
    if len(idx) == 0:
        return None if np.min(np.abs(f)) < 1e-3 else -1.0
    lo, hi = xs[idx[0] - 1], xs[idx[0]]
    for _ in range(80):
        mid = 0.5 * (lo + hi)
        if (sdf(t, s, p + mid * v) < 0) == inside:
            lo = mid
        else:
            hi = mid
    # reject crossings right at a tangency, where sampling decides
    if min(abs(sdf(t, s, p + (lo - 1e-4) * v)), abs(sdf(t, s, p + (hi + 1e-4) * v))) < 1e-7:
        return None
    return 0.5 * (lo + hi)


def test_closed_forms():
    r = ref64.ray_geom
    assert r(SPHERE, [0.5, 0, 0], [0, 0, -5], [0, 0, 2]) == pytest.approx(2.25)  # (5 - r) / |v|
    assert r(SPHERE, [0.5, 0, 0], [0, 0, 0], [1, 0, 0]) == pytest.approx(0.5)    # from inside: exit
    assert r(SPHERE, [0.5, 0, 0], [0, 0, -5], [0, 0, -1]) == -1                  # pointing away
    assert r(PLANE, [0, 0, 0], [0.3, 0.2, 2.0], [0, 0, -1]) == pytest.approx(2.0)
    assert r(PLANE, [0, 0, 0], [0.3, 0.2, -2.0], [0, 0, 1]) == -1                # back side
    assert r(PLANE, [1, 1, 0], [3.0, 0.0, 2.0], [0, 0, -1]) == -1                # outside the extent
    assert r(PLANE, [1, 1, 0], [0.5, 0.0, 2.0], [0.2, 0, -1]) == pytest.approx(2.0)  # lands at x = 0.9
    assert r(BOX, [1, 2, 3], [-5, 0, 0], [1, 0, 0]) == pytest.approx(4.0)
    assert r(BOX, [1, 2, 3], [0, 0, 0], [0, 1, 0]) == pytest.approx(2.0)
    assert r(BOX, [1, 2, 3], [0, 5, 0], [1, 0, 0]) == -1
    assert r(CAPSULE, [0.2, 1.0, 0], [0, 0, -5], [0, 0, 1]) == pytest.approx(5 - 1.0 - 0.2)
    assert r(CAPSULE, [0.2, 1.0, 0], [-5, 0, 0.5], [1, 0, 0]) == pytest.approx(5 - 0.2)
    assert r(CYLINDER, [0.2, 1.0, 0], [0, 0, -5], [0, 0, 1]) == pytest.approx(4.0)
    assert r(CYLINDER, [0.2, 1.0, 0], [-5, 0, 0.5], [1, 0, 0]) == pytest.approx(4.8)
    assert r(CYLINDER, [0.2, 1.0, 0], [0.1, 0, -5], [0, 0, 1]) == pytest.approx(4.0)
    for k, a in enumerate([0.3, 0.6, 0.9]):
        p = np.zeros(3); p[k] = -5
        v = np.zeros(3); v[k] = 1
        assert r(ELLIPSOID, [0.3, 0.6, 0.9], p, v) == pytest.approx(5 - a)


@pytest.mark.parametrize("gtype", [SPHERE, CAPSULE, ELLIPSOID, CYLINDER, BOX])
def test_oracle_against_brute_force(gtype):
    rng = np.random.default_rng(gtype)
    checked = hits = 0
    for _ in range(80):
        s = rng.uniform(0.2, 1.0, 3)
        p = rng.uniform(-2.5, 2.5, 3)
        if rng.uniform() < 0.7:  # aim near the primitive
            v = rng.uniform(-0.6, 0.6, 3) - p
        else:
            v = rng.normal(size=3)
        v *= rng.uniform(0.5, 2.0) / np.linalg.norm(v)
        want = brute(gtype, s, p, v)
        if want is None:
            continue
        got = ref64.ray_geom(gtype, s, p, v)
        checked += 1
        if want < 0:
            assert got == -1, (s, p, v, got)
        else:
            hits += 1
            assert got == pytest.approx(want, abs=1e-7), (s, p, v, got, want)
    assert checked >= 65 and hits >= 25, (checked, hits)


@pytest.mark.parametrize("gtype", [PLANE, SPHERE, CAPSULE, ELLIPSOID, CYLINDER, BOX])
def test_kernel_ray_geom_matches_oracle(gtype):
    rng = np.random.default_rng(10 + gtype)
    hits = 0
    for _ in range(400):
        s = rng.uniform(0.2, 1.0, 3)
        if gtype == PLANE and rng.uniform() < 0.5:
            s[:2] = 0  # infinite plane
        p = rng.uniform(-2.5, 2.5, 3)
        v = (rng.uniform(-0.6, 0.6, 3) - p) if rng.uniform() < 0.6 else rng.normal(size=3)
        s32, p32, v32 = (np.asarray(a, np.float32) for a in (s, p, v))
        want = ref64.ray_geom(gtype, s32, p32, v32)
        got = emu.ray_geom(gtype, s32, p32, v32)
        if want < 0:
            if got >= 0:  # only a grazing ray may flip between FP32 and FP64
                assert ref64.ray_geom(gtype, s32 * 1.0001, p32, v32) >= 0 or \
                    ref64.ray_geom(gtype, s32 * 0.9999, p32, v32) >= 0
            continue
        hits += 1
        assert got >= 0 and abs(got - want) <= 1e-5 * max(1.0, want), (s, p, v, got, want)
    assert hits >= 100


def test_kernel_rays_match_oracle_on_scene():
    """Per-env rays through the interpreter's kinematics and geom loop against
    the oracle's kinematics and mj_ray, with a geom filter."""
    xml = scenes.quadruped_xml()
    B, nray = 6, 48
    sim = emu.EmuSim(xml, B, maxcon=32, maxefc=96)
    q, qv, qa, ctrl, orc = states.quadruped_states(ref64.fields_from(sim.field), B, seed=51)
    sim.set_state(q, qv, qa)
    ng = int(sim.field("ngeom")[0])
    rng = np.random.default_rng(51)
    rays = np.zeros((B, nray, 6), np.float32)
    for e in range(B):
        c = q[e, :3]
        rays[e, :, :3] = c + rng.uniform(-0.4, 0.4, (nray, 3)) + [0, 0, 0.5]
        rays[e, :, 3:] = rng.normal(size=(nray, 3)) * 0.3 + [0, 0, -1]
    for gmask in (np.ones(ng, np.int32), (np.arange(ng) != 1).astype(np.int32)):
        dist, gid = sim.rays(rays, gmask)
        hits = 0
        for e in range(B):
            orc.set_state(q[e].astype(np.float64))
            orc.kinematics()
            for r in range(nray):
                x, g = orc.ray(rays[e, r, :3], rays[e, r, 3:], gmask)
                if g != gid[e, r]:  # two surfaces within FP32 rounding of each other
                    assert x >= 0 and dist[e, r] >= 0 and abs(x - dist[e, r]) < 1e-4, (e, r, g, gid[e, r])
                    continue
                assert abs(x - dist[e, r]) <= 1e-5 * max(1.0, abs(x)), (e, r, x, dist[e, r])
                hits += g >= 0
        assert hits > B * nray // 2


def test_rays_onto_rotated_box_faces():
    """World-frame closed form for the rotated torso box: a ray started 0.3
    along a face's outward normal from a point of that face, pointing back,
    travels 0.3 (oracle and kernel; geom filter keeps the box alone)."""
    xml = scenes.quadruped_xml()
    B = 4
    sim = emu.EmuSim(xml, B, maxcon=32, maxefc=96)
    q, qv, qa, ctrl, orc = states.quadruped_states(ref64.fields_from(sim.field), B, seed=52, steps=(20, 40))
    sim.set_state(q, qv, qa)
    ng = int(sim.field("ngeom")[0])
    g = int(np.nonzero(sim.field("geom_type") == BOX)[0][0])
    size = sim.field("geom_size").reshape(-1, 3)[g]
    mask = (np.arange(ng) == g).astype(np.int32)
    rng = np.random.default_rng(52)
    rays = np.zeros((B, 6, 6), np.float32)
    for e in range(B):
        orc.set_state(q[e].astype(np.float64))
        orc.kinematics()
        x = orc.get("geom_xpos").reshape(-1, 3)[g]
        R = orc.get("geom_xmat").reshape(-1, 3, 3)[g]
        assert not np.allclose(R, np.eye(3), atol=1e-3)  # a rotated box
        for f in range(6):
            k, sgn = f // 2, 1 - 2 * (f % 2)
            loc = rng.uniform(-0.8, 0.8, 3) * size
            loc[k] = sgn * size[k]
            n = R @ (sgn * np.eye(3)[k])
            start = x + R @ loc + 0.3 * n
            xo, go = orc.ray(start, -n, mask)
            assert go == g and xo == pytest.approx(0.3, abs=1e-9)
            rays[e, f, :3], rays[e, f, 3:] = start, -n
    dist, gid = sim.rays(rays, mask)
    assert (gid == g).all() and np.allclose(dist, 0.3, atol=2e-5)
