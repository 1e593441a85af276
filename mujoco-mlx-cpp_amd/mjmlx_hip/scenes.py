"""Procedural MJCF writers for the benchmark scenes.

Three scenes are written here, each as an MJCF string that the engine's own
MJCF compiler (csrc/mjcf.cpp) loads through ``mjmlx_load_model_from_string``:

* ``quadruped_xml``  - free-joint box torso, 4 legs x 3 hinges with capsule
  links and sphere feet, plane floor, position actuators (kp 20, kv 0.5).
* ``boxes_xml``      - a plane plus N free boxes with random sizes, densities,
  drop heights and orientations.
* ``humanoid_xml``   - free joint + 21 hinges, capsule limbs, condim 6 with
  elliptic cones, torque (motor) actuators with gear 100.

The reference has no generator of its own for these scenes (its benchmarks use
the network-fetched menagerie Go2, reference scripts/bench_physics_vs_others.py:41-65);
the sizes follow the workload definitions of the benchmark configurations.
"""
from __future__ import annotations

import math

import numpy as np

# Contact bit layout shared by the scenes:
#   floor:   contype 1, conaffinity 3  (collides with torso bit 1 and leg bit 2)
#   torso:   contype 1, conaffinity 0  (floor only)
#   legs:    contype 2, conaffinity 2 with self-collision, 0 without
QUAD_NOMINAL = np.array([0.0, 0.8, -1.5] * 4)
QUAD_HIP_RANGE = (-0.8, 0.8)
QUAD_THIGH_RANGE = (-1.5, 3.4)
QUAD_CALF_RANGE = (-2.7, -0.9)
QUAD_LINK = 0.21
QUAD_LINK_R = 0.02
QUAD_FOOT_R = 0.022
QUAD_TORSO = (0.19, 0.045, 0.03)


def _f(x) -> str:
    return " ".join(repr(float(v)) for v in np.atleast_1d(x))


def quadruped_xml(self_collision: bool = True, friction: float = 0.6,
                  solver: str = "Newton", iterations: int = 16,
                  tolerance: float = 1e-8, timestep: float = 0.005,
                  cone: str = "pyramidal", integrator: str = "Euler", sensors: bool = False) -> str:
    """Quadruped of the benchmark configurations: nq 19, nv 18, nu 12, 14 bodies, 14 geoms.
    sensors: an IMU site on the torso and a sensor of each kind the batched
    path computes (joint, actuator, frame, IMU, subtree)."""
    leg_aff = 2 if self_collision else 0
    hx, hy, hz = QUAD_TORSO
    legs = [("FL", +1, +1), ("FR", +1, -1), ("RL", -1, +1), ("RR", -1, -1)]
    nominal_h = QUAD_LINK * (math.cos(0.8) + math.cos(0.7)) + QUAD_FOOT_R + 0.01
    out = []
    out.append('<mujoco model="quadruped">')
    out.append('  <compiler angle="radian" autolimits="true"/>')
    out.append(f'  <option timestep="{timestep}" iterations="{iterations}" tolerance="{tolerance}" '
               f'solver="{solver}" cone="{cone}" integrator="{integrator}"/>')
    out.append('  <default>')
    out.append(f'    <geom friction="{friction} 0.005 0.0001" condim="3"/>')
    out.append('  </default>')
    out.append('  <worldbody>')
    out.append('    <geom name="floor" type="plane" size="0 0 0.05" contype="1" conaffinity="3"/>')
    out.append(f'    <body name="torso" pos="0 0 {nominal_h:.6f}">')
    out.append('      <freejoint name="root"/>')
    out.append(f'      <geom name="torso" type="box" size="{hx} {hy} {hz}" mass="6.9" contype="1" conaffinity="0"/>')
    if sensors:
        out.append('      <site name="imu" pos="0.05 0.01 0.02" euler="0.1 -0.2 0.3"/>')
    for name, sx, sy in legs:
        out.append(f'      <body name="{name}_hip" pos="{sx * hx} {sy * hy} 0">')
        out.append('        <inertial pos="0 0 0" mass="0.5" diaginertia="0.0004 0.0004 0.0004"/>')
        out.append(f'        <joint name="{name}_hip_joint" type="hinge" axis="1 0 0" range="{QUAD_HIP_RANGE[0]} {QUAD_HIP_RANGE[1]}"/>')
        out.append(f'        <body name="{name}_thigh" pos="0 {sy * 0.08} 0">')
        out.append(f'          <joint name="{name}_thigh_joint" type="hinge" axis="0 1 0" range="{QUAD_THIGH_RANGE[0]} {QUAD_THIGH_RANGE[1]}"/>')
        out.append(f'          <geom name="{name}_thigh" type="capsule" fromto="0 0 0 0 0 {-QUAD_LINK}" size="{QUAD_LINK_R}" '
                   f'contype="2" conaffinity="{leg_aff}"/>')
        out.append(f'          <body name="{name}_calf" pos="0 0 {-QUAD_LINK}">')
        out.append(f'            <joint name="{name}_calf_joint" type="hinge" axis="0 1 0" range="{QUAD_CALF_RANGE[0]} {QUAD_CALF_RANGE[1]}"/>')
        out.append(f'            <geom name="{name}_calf" type="capsule" fromto="0 0 0 0 0 {-QUAD_LINK}" size="{QUAD_LINK_R}" '
                   f'contype="2" conaffinity="{leg_aff}"/>')
        out.append(f'            <geom name="{name}_foot" type="sphere" pos="0 0 {-QUAD_LINK}" size="{QUAD_FOOT_R}" '
                   f'contype="2" conaffinity="{leg_aff}"/>')
        out.append('          </body>')
        out.append('        </body>')
        out.append('      </body>')
    out.append('    </body>')
    out.append('  </worldbody>')
    out.append('  <actuator>')
    ranges = [QUAD_HIP_RANGE, QUAD_THIGH_RANGE, QUAD_CALF_RANGE]
    for name, _, _ in legs:
        for part, rng in zip(("hip", "thigh", "calf"), ranges):
            out.append(f'    <position name="{name}_{part}" joint="{name}_{part}_joint" kp="20" kv="0.5" '
                       f'ctrlrange="{rng[0]} {rng[1]}"/>')
    out.append('  </actuator>')
    if sensors:
        out += ['  <sensor>',
                '    <jointpos name="FL_hip_pos" joint="FL_hip_joint"/>',
                '    <jointvel joint="FR_calf_joint"/>',
                '    <actuatorpos actuator="RL_thigh"/>',
                '    <actuatorvel actuator="RR_calf"/>',
                '    <jointactuatorfrc joint="FL_thigh_joint"/>',
                '    <framepos objtype="site" objname="imu"/>',
                '    <framequat objtype="site" objname="imu"/>',
                '    <framexaxis objtype="site" objname="imu"/>',
                '    <framezaxis objtype="body" objname="torso"/>',
                '    <framepos objtype="geom" objname="FL_foot"/>',
                '    <framelinvel objtype="geom" objname="RR_foot"/>',
                '    <frameangvel objtype="xbody" objname="FL_calf"/>',
                '    <frameyaxis objtype="xbody" objname="RL_thigh"/>',
                '    <gyro name="gyro" site="imu"/>',
                '    <velocimeter site="imu"/>',
                '    <accelerometer name="acc" site="imu"/>',
                '    <magnetometer site="imu"/>',
                '    <subtreecom body="torso"/>',
                '    <subtreecom body="FR_thigh"/>',
                '  </sensor>']
    out.append('</mujoco>')
    return "\n".join(out) + "\n"


def quadruped_initial_qpos(nenv: int, seed: int, noise: float = 0.05) -> np.ndarray:
    """Nominal standing pose plus U(+-noise) on every hinge (free joint at its qpos0)."""
    rng = np.random.default_rng(seed)
    h = QUAD_LINK * (math.cos(0.8) + math.cos(0.7)) + QUAD_FOOT_R + 0.01
    base = np.concatenate([[0.0, 0.0, h, 1.0, 0.0, 0.0, 0.0], QUAD_NOMINAL])
    q = np.tile(base, (nenv, 1))
    q[:, 7:] += rng.uniform(-noise, noise, size=(nenv, 12))
    return q


def quadruped_actions(nenv: int, rng: np.random.Generator, amp: float = 0.5) -> np.ndarray:
    """Nominal pose plus U(+-amp) position targets, one row per env."""
    return QUAD_NOMINAL[None, :] + rng.uniform(-amp, amp, size=(nenv, 12))


def _rand_quat(rng: np.random.Generator) -> np.ndarray:
    q = rng.normal(size=4)
    return q / np.linalg.norm(q)


def boxes_xml(nbox: int, seed: int, cone: str = "pyramidal", solver: str = "CG",
              iterations: int = 16, tolerance: float = 1e-8, timestep: float = 0.005,
              spacing: float = 0.3) -> str:
    """Plane plus ``nbox`` free boxes: half-size U(0.02,0.1), density U(500,2000),
    drop height U(0.1,1.0) (above the box's bounding radius), random orientation,
    friction U(0.3,1.0)."""
    rng = np.random.default_rng(seed)
    side = int(math.ceil(math.sqrt(nbox)))
    out = ['<mujoco model="boxes">', '  <compiler angle="radian"/>',
           f'  <option timestep="{timestep}" iterations="{iterations}" tolerance="{tolerance}" '
           f'solver="{solver}" cone="{cone}"/>',
           '  <worldbody>',
           '    <geom name="floor" type="plane" size="0 0 0.05" friction="0.3 0.005 0.0001"/>']
    for i in range(nbox):
        hs = rng.uniform(0.02, 0.1, size=3)
        dens = rng.uniform(500.0, 2000.0)
        h = rng.uniform(0.1, 1.0) + float(np.linalg.norm(hs))
        quat = _rand_quat(rng)
        mu = rng.uniform(0.3, 1.0)
        x = (i % side - (side - 1) / 2) * spacing + rng.uniform(-0.05, 0.05)
        y = (i // side - (side - 1) / 2) * spacing + rng.uniform(-0.05, 0.05)
        out.append(f'    <body name="box{i}" pos="{x:.6f} {y:.6f} {h:.6f}" quat="{_f(quat)}">')
        out.append('      <freejoint/>')
        out.append(f'      <geom type="box" size="{_f(hs)}" density="{dens:.3f}" friction="{mu:.4f} 0.005 0.0001"/>')
        out.append('    </body>')
    out += ['  </worldbody>', '</mujoco>']
    return "\n".join(out) + "\n"


def humanoid_xml(self_collision: bool = True, timestep: float = 0.005,
                 iterations: int = 16, tolerance: float = 1e-8) -> str:
    """Free joint + 21 hinges (nv 27), capsule limbs r 0.03-0.06, condim 6 elliptic cones,
    motor actuators gear 100."""
    aff = 1 if self_collision else 0
    L = []
    L.append('<mujoco model="humanoid">')
    L.append('  <compiler angle="radian" autolimits="true"/>')
    L.append(f'  <option timestep="{timestep}" iterations="{iterations}" tolerance="{tolerance}" '
             'solver="Newton" cone="elliptic" impratio="1"/>')
    L.append('  <default>')
    L.append('    <joint damping="0.5" armature="0.02"/>')
    L.append(f'    <geom condim="6" friction="0.9 0.02 0.002" contype="1" conaffinity="{aff}"/>')
    L.append('    <motor gear="100" ctrlrange="-1 1"/>')
    L.append('  </default>')
    L.append('  <worldbody>')
    L.append('    <geom name="floor" type="plane" size="0 0 0.05" contype="1" conaffinity="1"/>')
    L.append('    <body name="torso" pos="0 0 1.3">')
    L.append('      <freejoint name="root"/>')
    L.append('      <geom name="chest" type="capsule" fromto="0 -0.07 0 0 0.07 0" size="0.06"/>')
    L.append('      <geom name="head" type="sphere" pos="0 0 0.19" size="0.09"/>')
    L.append('      <body name="abdomen" pos="0 0 -0.16">')
    L.append('        <joint name="abdomen_z" axis="0 0 1" range="-0.8 0.8"/>')
    L.append('        <joint name="abdomen_y" axis="0 1 0" range="-1.3 0.5"/>')
    L.append('        <joint name="abdomen_x" axis="1 0 0" range="-0.6 0.6"/>')
    L.append('        <geom name="waist" type="capsule" fromto="-0.01 -0.06 0 -0.01 0.06 0" size="0.05"/>')
    L.append('        <body name="pelvis" pos="0 0 -0.16">')
    L.append('          <geom name="butt" type="capsule" fromto="-0.02 -0.07 0 -0.02 0.07 0" size="0.06"/>')
    for side, s in (("r", -1), ("l", 1)):
        L.append(f'          <body name="{side}_thigh" pos="0 {s * 0.1} -0.04">')
        L.append(f'            <joint name="{side}_hip_x" axis="1 0 0" range="-0.4 0.1"/>')
        L.append(f'            <joint name="{side}_hip_z" axis="0 0 1" range="-1.0 0.6"/>')
        L.append(f'            <joint name="{side}_hip_y" axis="0 1 0" range="-1.9 0.3"/>')
        L.append(f'            <geom name="{side}_thigh" type="capsule" fromto="0 0 0 0 {s * 0.01} -0.34" size="0.06"/>')
        L.append(f'            <body name="{side}_shin" pos="0 {s * 0.01} -0.403">')
        L.append(f'              <joint name="{side}_knee" axis="0 -1 0" range="-2.6 -0.02"/>')
        L.append(f'              <geom name="{side}_shin" type="capsule" fromto="0 0 0 0 0 -0.3" size="0.049"/>')
        L.append(f'              <body name="{side}_foot" pos="0 0 -0.39">')
        L.append(f'                <joint name="{side}_ankle" axis="0 1 0" range="-0.8 0.8"/>')
        L.append(f'                <joint name="{side}_ankle_x" axis="1 0 0" range="-0.5 0.5"/>')
        L.append(f'                <geom name="{side}_foot" type="capsule" fromto="-0.05 0 0 0.12 0 0" size="0.03"/>')
        L.append('              </body>')
        L.append('            </body>')
        L.append('          </body>')
    L.append('        </body>')
    L.append('      </body>')
    for side, s in (("r", -1), ("l", 1)):
        L.append(f'      <body name="{side}_upper_arm" pos="0 {s * 0.17} 0.06">')
        L.append(f'        <joint name="{side}_shoulder1" axis="2 {-s} 1" range="-1.5 1.0"/>')
        L.append(f'        <joint name="{side}_shoulder2" axis="0 {s * 1} 1" range="-1.5 1.0"/>')
        L.append(f'        <geom name="{side}_uarm" type="capsule" fromto="0 0 0 0.16 {s * 0.16} -0.16" size="0.04"/>')
        L.append(f'        <body name="{side}_lower_arm" pos="0.18 {s * 0.18} -0.18">')
        L.append(f'          <joint name="{side}_elbow" axis="0 {-s} 1" range="-1.5 0.8"/>')
        L.append(f'          <geom name="{side}_larm" type="capsule" fromto="0.01 {s * 0.01} 0.01 0.17 {s * 0.17} 0.17" size="0.031"/>')
        L.append('        </body>')
        L.append('      </body>')
    L.append('    </body>')
    L.append('  </worldbody>')
    L.append('  <contact>')
    L.append('    <exclude body1="torso" body2="r_thigh"/>')
    L.append('    <exclude body1="torso" body2="l_thigh"/>')
    L.append('  </contact>')
    L.append('  <actuator>')
    for j in ["abdomen_z", "abdomen_y", "abdomen_x",
              "r_hip_x", "r_hip_z", "r_hip_y", "r_knee", "r_ankle", "r_ankle_x",
              "l_hip_x", "l_hip_z", "l_hip_y", "l_knee", "l_ankle", "l_ankle_x",
              "r_shoulder1", "r_shoulder2", "r_elbow",
              "l_shoulder1", "l_shoulder2", "l_elbow"]:
        L.append(f'    <motor name="{j}" joint="{j}"/>')
    L.append('  </actuator>')
    L.append('</mujoco>')
    return "\n".join(L) + "\n"
