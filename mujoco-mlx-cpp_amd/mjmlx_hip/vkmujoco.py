"""numpy-facing module with the interface of the reference's RL bindings
(module ``vkmujoco``, reference python/rl_bindings.cpp:223-332), served by the
HIP engine's C ABI (libmjmlx.so) through ctypes.

Same functions, classes, argument names, array shapes and error messages as
the reference, so ``import mjmlx_hip.vkmujoco as vkmujoco`` replaces
``import vkmujoco`` in callers such as the reference's Go2 example
(scripts/examples/go2/mjmlx_go2_env.py). Arrays are float32 / int32,
C-contiguous, shaped (num_envs, dim). Failures raise RuntimeError, as the
reference's do; there is no CPU fallback (``use_gpu=False`` fails, as
MuJoCo C is not part of this engine). Rendering is outside the engine:
``create_camera`` raises.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import capi

_F = C.POINTER(C.c_float)
_I = C.POINTER(C.c_int)
_U8 = C.POINTER(C.c_ubyte)

_extra = {
    "mjmlx_load_model_ex": (C.c_void_p, [C.c_char_p, C.POINTER(capi.ModelOptions)]),
    "mjmlx_model_nsensordata": (C.c_int, [C.c_void_p]),
    "mjmlx_model_nsensor": (C.c_int, [C.c_void_p]),
    "mjmlx_model_sensor_name": (C.c_char_p, [C.c_void_p, C.c_int]),
    "mjmlx_model_sensor_type": (C.c_int, [C.c_void_p, C.c_int]),
    "mjmlx_model_sensor_adr": (C.c_int, [C.c_void_p, C.c_int]),
    "mjmlx_model_sensor_dim": (C.c_int, [C.c_void_p, C.c_int]),
    "mjmlx_batched_nsensordata": (C.c_int, [C.c_void_p]),
    "mjmlx_batched_get_sensordata": (_F, [C.c_void_p, _I]),
    "mjmlx_batched_raycast": (C.c_int, [C.c_void_p, _F, _F, C.c_int, C.c_int, _U8, C.c_int, _F, _I]),
    "mjmlx_camera_create": (C.c_void_p, [C.c_void_p, C.c_int, C.c_int]),
}
_ready = False


def _lib():
    global _ready
    L = capi.lib()
    if not _ready:
        for name, (res, args) in _extra.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _ready = True
    return L


def _f32(a, name):
    a = np.asarray(a)
    if a.dtype != np.float32 or not a.flags.c_contiguous:
        raise TypeError(f"mjmlx_rl: {name} must be a C-contiguous float32 array")
    return a


def _i32(a, name):
    a = np.asarray(a)
    if a.dtype != np.int32 or not a.flags.c_contiguous:
        raise TypeError(f"mjmlx_rl: {name} must be a C-contiguous int32 array")
    return a


def _fp(a):
    return None if a is None else a.ctypes.data_as(_F)


def _ip(a):
    return None if a is None else a.ctypes.data_as(_I)


def _copy_out(p, n, num_envs):
    if num_envs <= 0 or n <= 0:
        raise RuntimeError("mjmlx_rl: empty state array")
    return np.ctypeslib.as_array(p, shape=(n,)).copy().reshape(num_envs, n // num_envs)


class Model:
    """Compiled model (reference RlModel, python/rl_bindings.cpp:44-47,226-239)."""

    def __init__(self, handle):
        self.handle = handle
        self._info = _lib().mjmlx_model_info(handle)

    def __del__(self):
        if getattr(self, "handle", None):
            _lib().mjmlx_free_model(self.handle)
            self.handle = None

    nq = property(lambda self: self._info.nq)
    nv = property(lambda self: self._info.nv)
    nu = property(lambda self: self._info.nu)
    nbody = property(lambda self: self._info.nbody)
    nsensordata = property(lambda self: _lib().mjmlx_model_nsensordata(self.handle))

    def sensors(self):
        """[(name, mjtSensor type, adr, dim), ...]"""
        L = _lib()
        out = []
        for i in range(L.mjmlx_model_nsensor(self.handle)):
            nm = L.mjmlx_model_sensor_name(self.handle, i)
            out.append(((nm or b"").decode(), L.mjmlx_model_sensor_type(self.handle, i),
                        L.mjmlx_model_sensor_adr(self.handle, i), L.mjmlx_model_sensor_dim(self.handle, i)))
        return out

    def name2id(self, obj_type: int, name: str) -> int:
        return _lib().mjmlx_name2id(self.handle, int(obj_type), name.encode())

    def body_mass(self, body_id: int) -> float:
        return _lib().mjmlx_model_body_mass(self.handle, int(body_id))


class BatchedSim:
    """num_envs copies of a model (reference RlBatchedSim,
    python/rl_bindings.cpp:49-186,241-267)."""

    def __init__(self, model: Model, handle, num_envs: int):
        self.model_ref = model  # the sim refers to the model: keep it alive
        self.handle = handle
        self.num_envs = num_envs
        self.nq, self.nv, self.nu, self.nbody = model.nq, model.nv, model.nu, model.nbody
        self.pack_stride = 0

    def __del__(self):
        if getattr(self, "handle", None):
            _lib().mjmlx_batched_free(self.handle)
            self.handle = None

    def step(self, ctrl, n: int = 1):
        if self.nu > 0:
            c = _f32(ctrl, "ctrl")
            if c.ndim != 2 or c.shape[0] != self.num_envs or c.shape[1] != self.nu:
                raise RuntimeError("mjmlx_rl: ctrl must have shape (num_envs, nu)")
            _lib().mjmlx_batched_step_n(self.handle, _fp(c), int(n))
        else:
            _lib().mjmlx_batched_step_n(self.handle, None, int(n))

    def reset(self, mask):
        m = _i32(mask, "mask")
        if m.ndim != 1 or m.shape[0] != self.num_envs:
            raise RuntimeError("mjmlx_rl: mask must have shape (num_envs,)")
        _lib().mjmlx_batched_reset(self.handle, _ip(m))

    def set_env_qpos(self, env_idx: int, qpos):
        q = _f32(qpos, "qpos")
        _lib().mjmlx_batched_set_env_qpos(self.handle, int(env_idx), _fp(q), int(q.shape[0]))

    def set_env_qvel(self, env_idx: int, qvel):
        v = _f32(qvel, "qvel")
        _lib().mjmlx_batched_set_env_qvel(self.handle, int(env_idx), _fp(v), int(v.shape[0]))

    def set_state(self, qpos, qvel):
        q, v = _f32(qpos, "qpos"), _f32(qvel, "qvel")
        if q.ndim != 2 or q.shape[0] != self.num_envs or q.shape[1] != self.nq:
            raise RuntimeError("mjmlx_rl: qpos must have shape (num_envs, nq)")
        if v.ndim != 2 or v.shape[0] != self.num_envs or v.shape[1] != self.nv:
            raise RuntimeError("mjmlx_rl: qvel must have shape (num_envs, nv)")
        _lib().mjmlx_batched_set_state(self.handle, _fp(q), _fp(v))

    def set_env_params(self, env_idx=None, mass_scale=None, com_delta=None, friction_scale=None, gain_scale=None):
        """Per-env physics randomisation (create_batched(domain_rand=True));
        env_idx None updates every env; None arguments are left unchanged."""
        idx = None if env_idx is None else _i32(env_idx, "env_idx")
        n = self.num_envs if idx is None else int(idx.shape[0])

        def check(a, shape, name):
            if a is None:
                return None
            a = _f32(a, name)
            if a.ndim != len(shape):
                raise RuntimeError(f"mjmlx_rl: bad ndim for {name}")
            if tuple(a.shape) != tuple(shape):
                raise RuntimeError(f"mjmlx_rl: bad shape for {name}")
            return a
        ms = check(mass_scale, (n, self.nbody), "mass_scale (n, nbody)")
        cd = check(com_delta, (n, self.nbody, 3), "com_delta (n, nbody, 3)")
        fs = check(friction_scale, (n,), "friction_scale (n,)")
        gs = check(gain_scale, (n, self.nu), "gain_scale (n, nu)")
        rc = _lib().mjmlx_batched_set_env_params(self.handle, _ip(idx), n, _fp(ms), _fp(cd), _fp(fs), _fp(gs))
        if rc != 0:
            raise RuntimeError("mjmlx_rl: set_env_params failed (create_batched(domain_rand=True) is required; "
                               "env_idx must be in range)")

    def set_free_drag(self, drag_k, env_idx=None):
        """Quadratic air drag k [N s^2/m^2] (F = -k|v|v) per free joint, drag_k (n, num_free_joints)."""
        nf = _lib().mjmlx_batched_num_free_joints(self.handle)
        idx = None if env_idx is None else _i32(env_idx, "env_idx")
        n = self.num_envs if idx is None else int(idx.shape[0])
        d = _f32(drag_k, "drag_k")
        if d.ndim != 2 or d.shape[0] != n or d.shape[1] != nf:
            raise RuntimeError("mjmlx_rl: drag_k must have shape (n, num_free_joints)")
        rc = _lib().mjmlx_batched_set_free_drag(self.handle, _ip(idx), n, _fp(d))
        if rc != 0:
            raise RuntimeError("mjmlx_rl: set_free_drag failed (create_batched(domain_rand=True) is required; "
                               "the model needs a free joint)")

    @property
    def num_free_joints(self) -> int:
        return _lib().mjmlx_batched_num_free_joints(self.handle)

    def sync(self):
        """Wait until the queued steps have finished on the GPU (no host copy)."""
        _lib().mjmlx_batched_eval_state(self.handle)

    def state(self):
        qp = np.zeros((self.num_envs, self.nq), np.float32)
        qv = np.zeros((self.num_envs, self.nv), np.float32)
        _lib().mjmlx_batched_get_state(self.handle, _fp(qp), _fp(qv), None, None)
        return qp, qv

    def _get(self, fn):
        n = C.c_int()
        p = getattr(_lib(), fn)(self.handle, C.byref(n))
        if not p:
            raise RuntimeError(f"mjmlx_rl: {fn} failed")
        return _copy_out(p, n.value, self.num_envs)

    def qpos(self):
        return self._get("mjmlx_batched_get_qpos")

    def qvel(self):
        return self._get("mjmlx_batched_get_qvel")

    def xpos(self):
        return self._get("mjmlx_batched_get_xpos")

    def cfrc_ext(self):
        return self._get("mjmlx_batched_get_cfrc_ext")

    def qfrc_actuator(self):
        return self._get("mjmlx_batched_get_qfrc_actuator")

    def sensordata(self):
        """(num_envs, nsensordata); (num_envs, 0) without sensors."""
        n = C.c_int()
        p = _lib().mjmlx_batched_get_sensordata(self.handle, C.byref(n))
        nsd = _lib().mjmlx_batched_nsensordata(self.handle)
        if not p or nsd == 0:
            return np.zeros((self.num_envs, 0), np.float32)
        return _copy_out(p, n.value, self.num_envs)

    def raycast(self, pnt, vec, bodyexclude: int = -1, geomgroup=None, flg_static: bool = True):
        """mj_ray per env and ray at the current state: pnt / vec (num_envs, nray, 3)
        or shared (nray, 3); returns (dist, geom id), each (num_envs, nray), -1 for no hit."""
        p, v = _f32(pnt, "pnt"), _f32(vec, "vec")

        def per_env(a, name):
            if a.ndim == 3 and a.shape[0] == self.num_envs and a.shape[2] == 3:
                return True
            if a.ndim == 2 and a.shape[1] == 3:
                return False
            raise RuntimeError(f"mjmlx_rl: {name} must have shape (num_envs, nray, 3) or (nray, 3)")
        nray = p.shape[1] if per_env(p, "pnt") else p.shape[0]
        if (v.shape[1] if per_env(v, "vec") else v.shape[0]) != nray:
            raise RuntimeError("mjmlx_rl: pnt and vec must have the same number of rays")
        pe = np.ascontiguousarray(np.broadcast_to(p, (self.num_envs, nray, 3)), np.float32)
        ve = np.ascontiguousarray(np.broadcast_to(v, (self.num_envs, nray, 3)), np.float32)
        dist = np.zeros((self.num_envs, nray), np.float32)
        gid = np.zeros((self.num_envs, nray), np.int32)
        gg = None if geomgroup is None else np.ascontiguousarray(geomgroup, np.uint8)
        rc = _lib().mjmlx_batched_raycast(self.handle, _fp(pe), _fp(ve), int(nray), int(bodyexclude),
                                          None if gg is None else gg.ctypes.data_as(_U8), int(bool(flg_static)),
                                          _fp(dist), _ip(gid))
        if rc != 0:
            raise RuntimeError("mjmlx_rl: raycast failed (GPU sim required)")
        return dist, gid

    def set_pack_spec(self, body_ids, dof_lo: int, ndof: int) -> int:
        b = _i32(body_ids, "body_ids")
        self.pack_stride = _lib().mjmlx_batched_set_pack_spec(self.handle, _ip(b), int(b.size), int(dof_lo), int(ndof))
        if self.pack_stride < 0:
            raise RuntimeError("mjmlx_rl: set_pack_spec failed (GPU sim and valid body/dof ids required)")
        return self.pack_stride

    def read_pack_into(self, out):
        """Reads straight into a caller-owned (e.g. pinned) buffer (num_envs, stride)."""
        if self.pack_stride <= 0:
            raise RuntimeError("mjmlx_rl: call set_pack_spec first")
        o = _f32(out, "out")
        if o.ndim != 2 or o.shape[0] != self.num_envs or o.shape[1] != self.pack_stride:
            raise RuntimeError("mjmlx_rl: out must have shape (num_envs, stride)")
        if _lib().mjmlx_batched_get_pack(self.handle, _fp(o)) != 0:
            raise RuntimeError("mjmlx_rl: read_pack failed")

    def read_pack(self):
        if self.pack_stride <= 0:
            raise RuntimeError("mjmlx_rl: call set_pack_spec first")
        out = np.zeros((self.num_envs, self.pack_stride), np.float32)
        self.read_pack_into(out)
        return out


def load_model(xml_path: str, self_collision: bool = True, max_condim: int = 0) -> Model:
    o = capi.ModelOptions(1 if self_collision else 0, int(max_condim))
    h = _lib().mjmlx_load_model_ex(xml_path.encode(), C.byref(o))
    if not h:
        raise RuntimeError("mjmlx_rl: failed to load model: " + xml_path)
    return Model(h)


def create_batched(model: Model, num_envs: int, frame_skip: int = 1, use_gpu: bool = True,
                   solver_iterations: int = 0, domain_rand: bool = False) -> BatchedSim:
    """Euler integrator, as the reference's binding sets it; frame skipping is
    done by calling step() repeatedly (frame_skip is accepted and ignored)."""
    cfg = capi.BatchedConfig(int(num_envs), 0, 1 if use_gpu else 0, int(solver_iterations), 1 if domain_rand else 0)
    h = _lib().mjmlx_batched_create(model.handle, C.byref(cfg))
    if not h:
        raise RuntimeError("mjmlx_rl: failed to create batched sim")
    return BatchedSim(model, h, int(num_envs))


def create_camera(sim: BatchedSim, width: int, height: int):
    h = _lib().mjmlx_camera_create(sim.handle, int(width), int(height))
    if not h:
        raise RuntimeError("mjmlx_rl: failed to create camera renderer (need GPU sim and <camera> in the model)")
    raise RuntimeError("mjmlx_rl: rendering is not part of the HIP engine")


def version() -> str:
    return _lib().mjmlx_version().decode()


def __getattr__(name):
    if name == "__version__":
        return version()
    raise AttributeError(name)
