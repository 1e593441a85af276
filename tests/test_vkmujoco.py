"""The reference's RL binding interface (module vkmujoco, reference
python/rl_bindings.cpp:223-332) served by mjmlx_hip.vkmujoco over the C ABI:
module surface and model queries on the CPU; stepping is in
test_gpu_vkmujoco.py."""
from __future__ import annotations

import numpy as np
import pytest

from mjmlx_hip import capi, scenes
from mjmlx_hip import vkmujoco


def test_module_surface():
    for name in ("load_model", "create_batched", "create_camera", "version", "Model", "BatchedSim"):
        assert hasattr(vkmujoco, name), name
    for name in ("step", "reset", "set_env_qpos", "set_env_qvel", "set_state", "set_env_params", "set_free_drag",
                 "num_free_joints", "sync", "state", "qpos", "qvel", "xpos", "cfrc_ext", "sensordata",
                 "qfrc_actuator", "raycast", "set_pack_spec", "read_pack", "read_pack_into"):
        assert hasattr(vkmujoco.BatchedSim, name), name
    assert vkmujoco.__version__ == vkmujoco.version() != ""


def test_model_queries(tmp_path):
    path = tmp_path / "quadruped.xml"
    path.write_text(scenes.quadruped_xml())
    m = vkmujoco.load_model(str(path), self_collision=False)
    ref = capi.Model(scenes.quadruped_xml(), self_collision=False)
    assert (m.nq, m.nv, m.nu) == (ref.nq, ref.nv, ref.nu) == (19, 18, 12)
    assert m.nbody == int(ref.field("nbody")[0])
    tid = m.name2id(1, "torso")
    assert tid == 1 and m.name2id(1, "no such body") == -1
    assert m.body_mass(tid) == pytest.approx(float(ref.field("body_mass")[tid]), rel=1e-6)
    assert m.nsensordata == 0 and m.sensors() == []
    with pytest.raises(RuntimeError, match="failed to load model"):
        vkmujoco.load_model(str(tmp_path / "missing.xml"))


def test_no_cpu_fallback(tmp_path):
    """use_gpu=False asks for MuJoCo C, which the engine does not carry: the
    call fails instead of running something else."""
    path = tmp_path / "q.xml"
    path.write_text(scenes.quadruped_xml())
    m = vkmujoco.load_model(str(path))
    with pytest.raises(RuntimeError, match="failed to create batched sim"):
        vkmujoco.create_batched(m, 4, use_gpu=False)
    if capi.lib().mjmlx_hip_device_count() == 0:
        with pytest.raises(RuntimeError, match="failed to create batched sim"):
            vkmujoco.create_batched(m, 4)


def test_sensor_tables(tmp_path):
    """Model.sensors() lists (name, mjtSensor type, adr, dim) as the
    reference's binding does (python/rl_bindings.cpp:231-238)."""
    path = tmp_path / "q.xml"
    path.write_text(scenes.quadruped_xml(sensors=True))
    m = vkmujoco.load_model(str(path))
    s = m.sensors()
    assert m.nsensordata == 48 and len(s) == 19
    assert s[0] == ("FL_hip_pos", 9, 0, 1)        # jointpos
    assert s[6] == ("", 27, 8, 4)                 # framequat
    assert s[13] == ("gyro", 3, 30, 3)
    assert s[15] == ("acc", 1, 36, 3)             # accelerometer
    assert m.name2id(20, "gyro") == 13 and m.name2id(6, "imu") == 0
