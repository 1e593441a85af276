"""The C-ABI libraries load on a machine without a GPU and export every entry
point declared in include/mjmlx/*.h; host-only calls (model compile, field
access, struct layouts) work without touching the GPU."""
from __future__ import annotations

import ctypes as C
import os
import re

import pytest

from mjmlx_hip import capi, scenes

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INC = os.path.join(ROOT, "include", "mjmlx")
LIB = os.path.join(ROOT, "mujoco-mlx-cpp_amd", "lib")


def declared(header: str, macro: str):
    text = open(os.path.join(INC, header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(macro + r"\s+[\w\s\*]+?\b(\w+)\s*\(", text)))


@pytest.mark.parametrize("header,macro,lib", [
    ("mjmlx.h", "MJMLX_API", "libmjmlx.so"),
    ("mjmlx_hip.h", "MJMLX_API", "libmjmlx.so"),
    ("mjb.h", "MJB_API", "libmjb.so"),
])
def test_every_declared_symbol_is_exported(header, macro, lib):
    names = declared(header, macro)
    assert len(names) > 5
    L = C.CDLL(os.path.join(LIB, lib))
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_struct_layouts_match_reference_abi():
    # caller-allocated structs keep the reference's sizes (reference mjmlx_types.h:70-89, mjb_types.h:46-65)
    assert C.sizeof(capi.ModelInfo) == 9 * 4
    assert C.sizeof(capi.BatchedConfig) == 5 * 4
    assert C.sizeof(capi.ModelOptions) == 2 * 4


def test_host_only_calls_without_gpu():
    L = capi.lib()
    assert L.mjmlx_version().startswith(b"mjmlx-hip")
    o = L.mjmlx_default_model_options()
    assert (o.self_collision, o.max_condim) == (1, 0)
    m = capi.Model(scenes.quadruped_xml())
    assert (m.info.nq, m.info.nv, m.info.nu, m.info.nbody, m.info.ngeom) == (19, 18, 12, 14, 14)
    assert L.mjmlx_model_body_mass(m.h, 1) == pytest.approx(6.9)
    assert L.mjmlx_name2id(m.h, 1, b"torso") == 1
    assert L.mjmlx_name2id(m.h, 3, b"FL_calf_joint") == 3  # mjOBJ_JOINT
    assert L.mjmlx_name2id(m.h, 2, b"torso") == 1  # mjOBJ_XBODY
    assert L.mjmlx_name2id(m.h, 5, b"nope") == -1
    # id2name inverts name2id for every named body, joint and geom (mjb_id2name's source)
    for t, n in ((1, m.info.nbody), (3, m.info.njnt), (5, m.info.ngeom)):
        for i in range(n):
            nm = L.mjmlx_id2name(m.h, t, i)
            if nm is not None:
                assert L.mjmlx_name2id(m.h, t, nm) == i, (t, i, nm)
    assert L.mjmlx_id2name(m.h, 3, 3) == b"FL_calf_joint"
    assert L.mjmlx_id2name(m.h, 1, m.info.nbody) is None and L.mjmlx_id2name(m.h, 7, 0) is None
    assert L.mjmlx_model_opt_timestep(m.h) == pytest.approx(0.005)
    assert L.mjmlx_load_model_from_string(b"<mujoco><worldbody><geom type='mesh'/></worldbody></mujoco>") is None
    assert not L.mjmlx_get_mj_model(m.h)


def test_mjb_cpu_backend_reports_unavailable():
    L = C.CDLL(os.path.join(LIB, "libmjb.so"))
    L.mjb_create_backend.restype = C.c_void_p
    L.mjb_load_model_from_string.restype = C.c_void_p
    L.mjb_load_model_from_string.argtypes = [C.c_void_p, C.c_char_p]
    L.mjb_free_backend.argtypes = [C.c_void_p]
    b = L.mjb_create_backend(0)
    assert b
    assert not L.mjb_load_model_from_string(b, scenes.quadruped_xml().encode())
    L.mjb_free_backend(b)
    b = L.mjb_create_backend(1)
    L.mjb_free_model.argtypes = [C.c_void_p]
    m = L.mjb_load_model_from_string(b, scenes.quadruped_xml().encode())
    assert m
    L.mjb_id2name.restype = C.c_char_p
    L.mjb_id2name.argtypes = [C.c_void_p, C.c_int, C.c_int]
    assert L.mjb_id2name(m, 1, 1) == b"torso" and L.mjb_id2name(m, 3, 3) == b"FL_calf_joint"
    L.mjb_free_model(m)
    L.mjb_free_backend(b)
