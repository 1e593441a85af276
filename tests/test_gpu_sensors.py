"""Sensors on the GPU through the C ABI (mjmlx_batched_get_sensordata,
reference src/batched.cpp:1635-1651, src/sensor.cpp) against the FP64
restatement, the model's sensor tables, and the accelerometer after a
reset."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

import ref64
import sensors64
import states
from mjmlx_hip import capi, scenes

pytestmark = pytest.mark.gpu


def test_sensors_match_oracle():
    model = capi.Model(scenes.quadruped_xml(sensors=True))
    B = 32
    q, qv, qa, ctrl, orc = states.quadruped_states(ref64.fields_from(model.field), B, seed=81)
    sm = sensors64.SensorModel(model.field)
    L = capi.lib()
    assert L.mjmlx_model_nsensordata(model.h) == sm.nsensordata == 48
    sim = capi.BatchedSim(model, B)
    sim.set_state(q, qv)
    sim.set_qacc_warmstart(qa)
    sim.step(ctrl, 2)
    got = sim.sensordata()
    acc = [sm.adr[i] + k for i in range(sm.nsensor) if sm.type[i] == 1 for k in range(3)]
    rest = np.setdiff1d(np.arange(sm.nsensordata), acc)
    worst = 0.0
    for e in range(B):
        states.oracle_step(orc, q[e], qv[e], qa[e], ctrl[e])
        _, lin0 = sm.sensordata(orc)
        orc.step()
        qs, vs, ws = orc.get_state()
        orc.set_state(qs, vs, ws, ctrl[e].astype(np.float64))
        orc.forward()
        want, _ = sm.sensordata(orc, lin0)
        worst = max(worst, float(np.max(np.abs(got[e, rest] - want[rest]))))
        assert np.allclose(got[e, rest], want[rest], atol=2e-4, rtol=1e-3), e
        assert np.allclose(got[e, acc], want[acc], atol=0.2, rtol=2e-2), (e, got[e, acc], want[acc])
    print(f"sensors: worst abs error {worst:.2e} (accelerometer excluded)")


def test_accelerometer_after_reset():
    """A reset env has no previous velocity: its next single-substep step
    reads -gravity in the site frame; the other envs keep differencing."""
    model = capi.Model(scenes.quadruped_xml(sensors=True))
    B = 8
    sim = capi.BatchedSim(model, B)
    sm = sensors64.SensorModel(model.field)
    i = int(np.nonzero(sm.type == 1)[0][0])
    a = sm.adr[i]
    q = scenes.quadruped_initial_qpos(B, 82).astype(np.float32)
    sim.set_state(q, np.zeros((B, model.nv), np.float32))
    for _ in range(3):
        sim.step(None, 1)
    mask = np.zeros(B, np.int32)
    mask[0] = 1
    sim.reset(mask)
    sim.step(None, 1)
    y = sim.sensordata()
    orc = ref64.Oracle(ref64.fields_from(model.field))
    orc.set_state(model.field("qpos0"), np.zeros(model.nv), np.zeros(model.nv), np.zeros(model.nu))
    orc.forward()
    want, _ = sm.sensordata(orc)
    assert np.allclose(y[0, a:a + 3], want[a:a + 3], atol=1e-4)
    assert not np.allclose(y[1:, a:a + 3], want[a:a + 3], atol=1e-2)  # falling envs difference their velocity


def test_autotune_keeps_sensor_history():
    """The autotuner restores what its timed steps wrote, the accelerometers'
    previous velocities and the aux outputs included: a tuned sim's sensors
    and cfrc_ext after further steps are bitwise those of an untuned one."""
    model = capi.Model(scenes.quadruped_xml(sensors=True))
    B = 64
    q = scenes.quadruped_initial_qpos(B, 83).astype(np.float32)
    ctrl = scenes.quadruped_actions(B, np.random.default_rng(83)).astype(np.float32)
    out = []
    for tune in (False, True):
        sim = capi.BatchedSim(model, B)
        sim.set_state(q, np.zeros((B, model.nv), np.float32))
        sim.cvel()  # aux outputs on
        for _ in range(3):
            sim.step(ctrl, 4)
        if tune:
            assert sim.autotune(2) in (0, 1, 2, 3, 4)
        before = (sim.sensordata().copy(), sim.cfrc_ext())
        sim.step(ctrl, 1)
        out.append((before, sim.sensordata().copy(), sim.state()))
        del sim
    (b0, s0, st0), (b1, s1, st1) = out
    assert np.array_equal(b0[0], b1[0]) and np.array_equal(b0[1], b1[1])
    assert np.array_equal(s0, s1)
    for a, b in zip(st0, st1):
        assert np.array_equal(a, b)
