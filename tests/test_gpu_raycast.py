"""mjmlx_batched_raycast on the GPU (reference src/batched.cpp:1808-1893)
against the oracle's mj_ray at each env's current state, and its geom
filters (bodyexclude, geomgroup, flg_static)."""
from __future__ import annotations

import numpy as np
import pytest

import ref64
import states
from mjmlx_hip import capi, scenes

pytestmark = pytest.mark.gpu


def _rays(q, nray, rng):
    B = q.shape[0]
    pnt = (q[:, None, :3] + rng.uniform(-0.4, 0.4, (B, nray, 3)) + [0, 0, 0.5]).astype(np.float32)
    vec = (rng.normal(size=(B, nray, 3)) * 0.3 + [0, 0, -1]).astype(np.float32)
    return pnt, vec


def test_raycast_matches_oracle_after_steps():
    model = capi.Model(scenes.quadruped_xml())
    B, nray = 64, 40
    q, qv, qa, ctrl, orc = states.quadruped_states(ref64.fields_from(model.field), B, seed=61)
    sim = capi.BatchedSim(model, B)
    sim.set_state(q, qv)
    sim.set_qacc_warmstart(qa)
    sim.step(ctrl, 4)  # rays see the state after the queued step
    qs, _ = sim.state()
    rng = np.random.default_rng(61)
    pnt, vec = _rays(qs, nray, rng)
    dist, gid = sim.raycast(pnt, vec)
    ng = int(model.field("ngeom")[0])
    ok = np.ones(ng, np.int32)
    hits = ties = 0
    for e in range(B):
        orc.set_state(qs[e].astype(np.float64))
        orc.kinematics()
        for r in range(nray):
            x, g = orc.ray(pnt[e, r], vec[e, r], ok)
            if g != gid[e, r]:
                ties += 1
                assert x >= 0 and dist[e, r] >= 0 and abs(x - dist[e, r]) < 1e-4, (e, r, x, g, dist[e, r], gid[e, r])
                continue
            assert abs(x - dist[e, r]) <= 2e-5 * max(1.0, abs(x)), (e, r, x, dist[e, r])
            hits += g >= 0
    print(f"raycast: {hits} hits of {B * nray} rays, {ties} near-ties")
    assert hits > B * nray // 2 and ties <= 2


def test_raycast_filters():
    model = capi.Model(scenes.quadruped_xml())
    B, nray = 16, 32
    sim = capi.BatchedSim(model, B)
    q = scenes.quadruped_initial_qpos(B, 62).astype(np.float32)
    sim.set_state(q, np.zeros((B, model.nv), np.float32))
    gbody = model.field("geom_bodyid").astype(int)
    weld = model.field("body_weldid").astype(int)
    # straight down from above the torso: the torso first
    pnt = np.repeat((q[:, None, :3] + [0, 0, 0.5]).astype(np.float32), nray, axis=1)
    pnt[:, :, :2] += np.random.default_rng(62).uniform(-0.02, 0.02, (B, nray, 2)).astype(np.float32)
    vec = np.zeros((B, nray, 3), np.float32)
    vec[:, :, 2] = -1
    d0, g0 = sim.raycast(pnt, vec)
    torso = 1
    assert (gbody[g0] == torso).all()
    d1, g1 = sim.raycast(pnt, vec, bodyexclude=torso)
    assert (g1 >= 0).all() and not (gbody[g1] == torso).any() and (d1 > d0).all()
    d2, g2 = sim.raycast(pnt, vec, geomgroup=np.zeros(6, np.uint8))
    assert (g2 == -1).all() and (d2 == -1).all()
    d3, g3 = sim.raycast(pnt, vec, bodyexclude=torso, flg_static=False)
    assert not ((g3 >= 0) & (weld[gbody[np.maximum(g3, 0)]] == 0)).any()  # nothing welded to the world
    # rays pointing up from below the floor: planes are one-sided
    d4, g4 = sim.raycast(pnt * [1, 1, 0] - [0, 0, 1], -vec)
    assert (g4 != 0).all() or (gbody[g4[g4 >= 0]] != 0).all()


def test_raycast_device_memory_workspace_boxes(monkeypatch):
    """Rays through the device-memory workspace variant on a scene of boxes:
    straight down onto each box's top face, against the oracle."""
    monkeypatch.setenv("MJMLX_HIP_GLOBAL_WORKSPACE", "1")
    model = capi.Model(scenes.boxes_xml(6, 3, solver="CG", iterations=16))
    B, nbox = 32, 6
    sim = capi.BatchedSim(model, B)
    assert sim.capacity()[2] > 0
    q = np.tile(model.field("qpos0"), (B, 1)).astype(np.float32)
    sim.set_state(q, np.zeros((B, model.nv), np.float32))
    for _ in range(10):
        sim.step(None, 4)
    qs, _ = sim.state()
    pnt = np.zeros((B, nbox, 3), np.float32)
    for k in range(nbox):
        pnt[:, k, :] = qs[:, 7 * k:7 * k + 3] + [0.003, -0.002, 2.0]
    vec = np.zeros((B, nbox, 3), np.float32)
    vec[:, :, 2] = -1
    dist, gid = sim.raycast(pnt, vec)
    orc = ref64.Oracle(ref64.fields_from(model.field))
    ok = np.ones(int(model.field("ngeom")[0]), np.int32)
    for e in range(B):
        orc.set_state(qs[e].astype(np.float64))
        orc.kinematics()
        for r in range(nbox):
            x, g = orc.ray(pnt[e, r], vec[e, r], ok)
            assert g == gid[e, r] and abs(x - dist[e, r]) <= 2e-5 * max(1.0, x), (e, r, x, g, dist[e, r], gid[e, r])
    assert (gid > 0).mean() > 0.9  # mostly boxes, not the floor
