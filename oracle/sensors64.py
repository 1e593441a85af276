"""FP64 restatement of the batched path's sensors (reference
src/sensor.cpp:112-359).

TEST INFRASTRUCTURE ONLY: the checker for the HIP engine's sensors; only
tests/ import it. It takes the FP64 oracle's quantities of one env's state
(body poses, cvel, subtree COM, actuator forces) and the compiled model's
sensor and site tables, and computes sensordata the way the reference does:
frame sensors from the body pose and cvel (point velocity about the root's
subtree COM), joint and actuator sensors from qpos / qvel / forces, and the
accelerometer as the one-substep difference of its frame's world velocity
minus gravity (zero difference without a previous substep), in the site
frame.

Pinned by tests/test_sensors.py: finite differences of frame positions and
orientations (linear and angular velocity), the subtree COM recomputed from
body masses, closed forms (joint and actuator values, free fall and rest of
the accelerometer). Sensor type ids are MuJoCo's mjtSensor values, which the
compiled model carries (parity unpinned: MuJoCo C is absent).
"""
from __future__ import annotations

import numpy as np

ACCELEROMETER, VELOCIMETER, GYRO, MAGNETOMETER = 1, 2, 3, 6
JOINTPOS, JOINTVEL, ACTUATORPOS, ACTUATORVEL, JOINTACTFRC, BALLQUAT, BALLANGVEL = 9, 10, 13, 14, 16, 18, 19
FRAMEPOS, FRAMEQUAT, FRAMEXAXIS, FRAMEYAXIS, FRAMEZAXIS, FRAMELINVEL, FRAMEANGVEL, SUBTREECOM = (
    26, 27, 28, 29, 30, 31, 32, 35)


def quat2mat(q):
    w, x, y, z = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                     [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                     [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])


def quatmul(a, b):
    return np.array([a[0] * b[0] - a[1] * b[1] - a[2] * b[2] - a[3] * b[3],
                     a[0] * b[1] + a[1] * b[0] + a[2] * b[3] - a[3] * b[2],
                     a[0] * b[2] - a[1] * b[3] + a[2] * b[0] + a[3] * b[1],
                     a[0] * b[3] + a[1] * b[2] - a[2] * b[1] + a[3] * b[0]])


class SensorModel:
    """Sensor, site and frame tables of a compiled model, read through its
    field getter (``getter(name) -> np.ndarray``)."""

    def __init__(self, getter):
        g = lambda n: np.asarray(getter(n))
        self.nsensor = int(g("nsensor")[0])
        self.nsensordata = int(g("nsensordata")[0])
        self.type = g("sensor_type").astype(int)
        self.objtype = g("sensor_objtype").astype(int)
        self.objid = g("sensor_objid").astype(int)
        self.adr = g("sensor_adr").astype(int)
        self.dim = g("sensor_dim").astype(int)
        self.site_bodyid = g("site_bodyid").astype(int)
        self.site_pos = g("site_pos").reshape(-1, 3)
        self.site_quat = g("site_quat").reshape(-1, 4)
        self.geom_bodyid = g("geom_bodyid").astype(int)
        self.geom_pos = g("geom_pos").reshape(-1, 3)
        self.geom_quat = g("geom_quat").reshape(-1, 4)
        self.body_ipos = g("body_ipos").reshape(-1, 3)
        self.body_iquat = g("body_iquat").reshape(-1, 4)
        self.body_rootid = g("body_rootid").astype(int)
        self.jnt_qposadr = g("jnt_qposadr").astype(int)
        self.jnt_dofadr = g("jnt_dofadr").astype(int)
        self.actuator_trnid = g("actuator_trnid").astype(int)
        self.actuator_gear = g("actuator_gear")
        self.magnetic = g("magnetic")
        self.gravity = g("gravity")
        self.timestep = float(g("timestep")[0])

    def frame_def(self, i):
        """(body, local position, local quaternion) of sensor i's frame
        (reference src/sensor.cpp:112-124)"""
        ot, oid = self.objtype[i], self.objid[i]
        if ot == 6:
            return self.site_bodyid[oid], self.site_pos[oid], self.site_quat[oid]
        if ot == 5:
            return self.geom_bodyid[oid], self.geom_pos[oid], self.geom_quat[oid]
        if ot == 1:
            return oid, self.body_ipos[oid], self.body_iquat[oid]
        return oid, np.zeros(3), np.array([1.0, 0, 0, 0])

    def frame(self, orc, i):
        """world (pos, R, quat, angular velocity, linear velocity) of sensor i's frame"""
        b, lp, lq = self.frame_def(i)
        xpos = orc.get("xpos").reshape(-1, 3)[b]
        xq = orc.get("xquat").reshape(-1, 4)[b]
        Rb = quat2mat(xq)
        pos = xpos + Rb @ lp
        R = Rb @ quat2mat(lq)
        q = quatmul(xq, lq)
        cv = orc.get("cvel").reshape(-1, 6)[b]
        c = orc.get("subtree_com").reshape(-1, 3)[self.body_rootid[b]]
        ang = cv[:3].copy()
        lin = cv[3:] + np.cross(ang, pos - c)
        return pos, R, q, ang, lin

    def sensordata(self, orc, prev_lin=None):
        """sensordata of the oracle's current state (after forward()); prev_lin:
        {sensor index: world velocity of the accelerometer's frame one
        substep earlier} or None. Returns (sensordata, {index: velocity now})."""
        out = np.zeros(self.nsensordata)
        lin_now = {}
        qpos, qvel = orc.get("qpos"), orc.get("qvel")
        qfa = orc.get("qfrc_actuator")
        for i in range(self.nsensor):
            t, oid, a = self.type[i], self.objid[i], self.adr[i]
            if t in (JOINTPOS, JOINTVEL, JOINTACTFRC, BALLQUAT, BALLANGVEL):
                qa, da = self.jnt_qposadr[oid], self.jnt_dofadr[oid]
            if t in (ACTUATORPOS, ACTUATORVEL):
                j = self.actuator_trnid[oid]
                qa, da, gear = self.jnt_qposadr[j], self.jnt_dofadr[j], self.actuator_gear[oid]
            if t == JOINTPOS:
                y = [qpos[qa]]
            elif t == JOINTVEL:
                y = [qvel[da]]
            elif t == JOINTACTFRC:
                y = [qfa[da]]
            elif t == ACTUATORPOS:
                y = [gear * qpos[qa]]
            elif t == ACTUATORVEL:
                y = [gear * qvel[da]]
            elif t == BALLQUAT:
                y = qpos[qa:qa + 4] / np.linalg.norm(qpos[qa:qa + 4])
            elif t == BALLANGVEL:
                y = qvel[da:da + 3]
            elif t == SUBTREECOM:
                y = orc.get("subtree_com").reshape(-1, 3)[oid]
            else:
                pos, R, q, ang, lin = self.frame(orc, i)
                if t == FRAMEPOS:
                    y = pos
                elif t == FRAMEQUAT:
                    y = q if q[0] >= 0 else -q
                elif t in (FRAMEXAXIS, FRAMEYAXIS, FRAMEZAXIS):
                    y = R[:, t - FRAMEXAXIS]
                elif t == FRAMELINVEL:
                    y = lin
                elif t == FRAMEANGVEL:
                    y = ang
                elif t == GYRO:
                    y = R.T @ ang
                elif t == VELOCIMETER:
                    y = R.T @ lin
                elif t == MAGNETOMETER:
                    y = R.T @ self.magnetic
                elif t == ACCELEROMETER:
                    acc = np.zeros(3) if prev_lin is None else (lin - prev_lin[i]) / self.timestep
                    y = R.T @ (acc - self.gravity)
                    lin_now[i] = lin
                else:
                    raise ValueError(f"sensor type {t}")
            out[a:a + self.dim[i]] = y
        return out, lin_now
