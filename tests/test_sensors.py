"""Sensors of the batched path (reference src/sensor.cpp:112-359): the FP64
restatement (oracle/sensors64.py) pinned by finite differences, an
independent subtree COM and closed forms, then the kernel source (serial
interpreter) against it on rolled-out states."""
from __future__ import annotations

import numpy as np
import pytest

import emu
import ref64
import sensors64
import states
from mjmlx_hip import scenes


def _setup(B=6, seed=71):
    xml = scenes.quadruped_xml(sensors=True)
    sim = emu.EmuSim(xml, B, maxcon=32, maxefc=96)
    f = ref64.fields_from(sim.field)
    q, qv, qa, ctrl, orc = states.quadruped_states(f, B, seed=seed)
    return sim, f, q, qv, qa, ctrl, orc, sensors64.SensorModel(sim.field)


def _advance(q, qv, nq, h, free_adr=0):
    """positions after time h at constant qvel: hinges linearly, the free joint's
    position in world coordinates and its quaternion by its body-frame rate"""
    out = q.astype(np.float64).copy()
    out[free_adr:free_adr + 3] += h * qv[0:3]
    w = qv[3:6].astype(np.float64)
    th = np.linalg.norm(w) * h
    dq = np.r_[np.cos(th / 2), np.sin(th / 2) * w / max(np.linalg.norm(w), 1e-300)]
    out[3:7] = sensors64.quatmul(out[3:7], dq)
    out[7:] += h * qv[6:]
    return out


def test_frame_velocities_are_derivatives_of_frame_poses():
    sim, f, q, qv, qa, ctrl, orc, sm = _setup()
    h = 1e-7
    for e in range(6):
        states.oracle_step(orc, q[e], qv[e], qa[e], ctrl[e])
        now = [sm.frame(orc, i) for i in range(sm.nsensor) if sm.type[i] in (26, 27, 31, 32, 2, 3, 1)]
        orc.set_state(_advance(q[e], qv[e], orc.nq, h), qv[e].astype(np.float64))
        orc.forward()
        idx = [i for i in range(sm.nsensor) if sm.type[i] in (26, 27, 31, 32, 2, 3, 1)]
        for (p0, R0, q0, w0, v0), i in zip(now, idx):
            p1, R1, q1, _, _ = sm.frame(orc, i)
            assert np.allclose((p1 - p0) / h, v0, atol=2e-5 * max(1, np.abs(v0).max())), (e, i)
            dq = sensors64.quatmul(q1, q0 * [1, -1, -1, -1])
            assert np.allclose(2 * dq[1:] / h, w0, atol=2e-5 * max(1, np.abs(w0).max())), (e, i)


def test_closed_forms_and_subtree_com():
    sim, f, q, qv, qa, ctrl, orc, sm = _setup()
    mass = np.asarray(f["body_mass"])
    parent = np.asarray(f["body_parentid"]).astype(int)
    for e in range(6):
        states.oracle_step(orc, q[e], qv[e], qa[e], ctrl[e])
        y, _ = sm.sensordata(orc)
        qp, qvv = orc.get("qpos"), orc.get("qvel")
        assert y[0] == qp[7] and y[1] == qvv[6 + 5]  # FL hip angle, FR calf rate
        assert y[2] == qp[7 + 7] and y[3] == qvv[6 + 11]  # RL thigh, RR calf (gear 1)
        assert y[4] == orc.get("qfrc_actuator")[7]
        xipos = orc.get("xipos").reshape(-1, 3)
        for adr, body in ((42, 1), (45, 6)):  # torso, FR_thigh
            sub = [b for b in range(len(mass)) if b == body or body in _ancestors(b, parent)]
            com = (mass[sub, None] * xipos[sub]).sum(0) / mass[sub].sum()
            assert np.allclose(y[adr:adr + 3], com, atol=1e-12)
        R = sensors64.quat2mat(y[8:12])
        assert np.allclose(R[:, 0], y[12:15], atol=1e-12)  # framexaxis of the same site
        assert abs(np.linalg.norm(y[8:12]) - 1) < 1e-12 and y[8] >= 0


def _ancestors(b, parent):
    out = []
    while b > 0:
        b = parent[b]
        out.append(b)
    return out


def test_accelerometer_free_fall_and_rest():
    """Free fall (no contact, no velocity): the accelerometer reads zero after
    one substep; a body at rest reads -gravity in its frame."""
    sim, f, q, qv, qa, ctrl, orc, sm = _setup(B=1)
    acc = int(np.nonzero(sm.type == 1)[0][0])
    a = sm.adr[acc]
    q0 = q[0].astype(np.float64)
    q0[2] += 5.0  # far above the floor, joints inside their ranges, servos holding them
    hold = q0[7:].copy()
    orc.set_state(q0, np.zeros(orc.nv), np.zeros(orc.nv), hold)
    orc.forward()
    y0, lin0 = sm.sensordata(orc)
    R = sm.frame(orc, acc)[1]
    assert np.allclose(y0[a:a + 3], R.T @ -sm.gravity, atol=1e-12)  # no previous substep: rest reading
    orc.step()
    qs, vs, _ = orc.get_state()
    orc.set_state(qs, vs, np.zeros(orc.nv), hold)
    orc.forward()
    y1, _ = sm.sensordata(orc, lin0)
    assert np.abs(y1[a:a + 3]).max() < 1e-9  # free fall: no internal force, the body falls rigidly


def test_interpreter_sensors_match_oracle():
    """Two substeps in one call: the sensors of the second substep's start
    state, the accelerometer differencing the two."""
    sim, f, q, qv, qa, ctrl, orc, sm = _setup(B=12, seed=72)
    sim.set_state(q, qv, qa)
    sim.step(ctrl, nsub=2)
    got = sim.sensordata()
    for e in range(12):
        states.oracle_step(orc, q[e], qv[e], qa[e], ctrl[e])
        _, lin0 = sm.sensordata(orc)
        orc.step()
        qs, vs, ws = orc.get_state()
        orc.set_state(qs, vs, ws, ctrl[e].astype(np.float64))
        orc.forward()
        want, _ = sm.sensordata(orc, lin0)
        acc = [sm.adr[i] + k for i in range(sm.nsensor) if sm.type[i] == 1 for k in range(3)]
        rest = np.setdiff1d(np.arange(sm.nsensordata), acc)
        assert np.allclose(got[e, rest], want[rest], atol=2e-4, rtol=1e-3), (e, np.abs(got[e, rest] - want[rest]).max())
        # the accelerometer divides a velocity change by one substep
        assert np.allclose(got[e, acc], want[acc], atol=0.2, rtol=2e-2), (e, got[e, acc], want[acc])
