"""ctypes binding of the engine's C ABI (lib/libmjmlx.so).

Mirrors the reference's Python caller surface (reference python/rl_bindings.cpp
NB_MODULE at :223-332: load_model, create_batched, step(ctrl, n), state,
set_state, sync) on top of the same C entry points a C caller uses
(include/mjmlx/mjmlx.h). Loading fails loudly when the library is missing:
there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

PKG = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIBDIR = os.path.join(PKG, "lib")
LIBPATH = os.path.join(LIBDIR, "libmjmlx.so")

_F = C.POINTER(C.c_float)
_I = C.POINTER(C.c_int)
_D = C.POINTER(C.c_double)


class ModelInfo(C.Structure):
    _fields_ = [(n, C.c_int) for n in ("nq", "nv", "nu", "na", "nbody", "njnt", "ngeom", "nsite", "ncon")]


class ModelOptions(C.Structure):
    _fields_ = [("self_collision", C.c_int), ("max_condim", C.c_int)]


class BatchedConfig(C.Structure):
    _fields_ = [("num_envs", C.c_int), ("integrator", C.c_int), ("use_gpu", C.c_int),
                ("solver_iterations", C.c_int), ("domain_rand", C.c_int)]


_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIBPATH):
            raise RuntimeError(f"HIP engine library not built: {LIBPATH} (run __graft_entry__.build())")
        L = C.CDLL(LIBPATH)
        sig = {
            "mjmlx_version": (C.c_char_p, []),
            "mjmlx_hip_device_count": (C.c_int, []),
            "mjmlx_load_model": (C.c_void_p, [C.c_char_p]),
            "mjmlx_load_model_from_string": (C.c_void_p, [C.c_char_p]),
            "mjmlx_load_model_from_string_ex": (C.c_void_p, [C.c_char_p, C.POINTER(ModelOptions)]),
            "mjmlx_default_model_options": (ModelOptions, []),
            "mjmlx_free_model": (None, [C.c_void_p]),
            "mjmlx_model_info": (ModelInfo, [C.c_void_p]),
            "mjmlx_model_opt_timestep": (C.c_float, [C.c_void_p]),
            "mjmlx_model_body_mass": (C.c_float, [C.c_void_p, C.c_int]),
            "mjmlx_name2id": (C.c_int, [C.c_void_p, C.c_int, C.c_char_p]),
            "mjmlx_id2name": (C.c_char_p, [C.c_void_p, C.c_int, C.c_int]),
            "mjmlx_model_field": (C.c_int, [C.c_void_p, C.c_char_p, _D, C.c_int]),
            "mjmlx_batched_create": (C.c_void_p, [C.c_void_p, C.POINTER(BatchedConfig)]),
            "mjmlx_batched_free": (None, [C.c_void_p]),
            "mjmlx_batched_step": (None, [C.c_void_p, _F]),
            "mjmlx_batched_step_n": (None, [C.c_void_p, _F, C.c_int]),
            "mjmlx_batched_reset": (None, [C.c_void_p, _I]),
            "mjmlx_batched_get_state": (None, [C.c_void_p, _F, _F, _I, _I]),
            "mjmlx_batched_set_state": (None, [C.c_void_p, _F, _F]),
            "mjmlx_batched_get_qpos": (_F, [C.c_void_p, _I]),
            "mjmlx_batched_get_qvel": (_F, [C.c_void_p, _I]),
            "mjmlx_batched_get_xpos": (_F, [C.c_void_p, _I]),
            "mjmlx_batched_get_xquat": (_F, [C.c_void_p, _I]),
            "mjmlx_batched_get_subtree_com": (_F, [C.c_void_p, _I]),
            "mjmlx_batched_get_cinert": (_F, [C.c_void_p, _I]),
            "mjmlx_batched_get_cvel": (_F, [C.c_void_p, _I]),
            "mjmlx_batched_get_qfrc_actuator": (_F, [C.c_void_p, _I]),
            "mjmlx_batched_get_cfrc_ext": (_F, [C.c_void_p, _I]),
            "mjmlx_batched_set_env_qpos": (None, [C.c_void_p, C.c_int, _F, C.c_int]),
            "mjmlx_batched_set_env_qvel": (None, [C.c_void_p, C.c_int, _F, C.c_int]),
            "mjmlx_batched_set_env_params": (C.c_int, [C.c_void_p, _I, C.c_int, _F, _F, _F, _F]),
            "mjmlx_batched_set_free_drag": (C.c_int, [C.c_void_p, _I, C.c_int, _F]),
            "mjmlx_batched_num_free_joints": (C.c_int, [C.c_void_p]),
            "mjmlx_batched_raycast": (C.c_int, [C.c_void_p, _F, _F, C.c_int, C.c_int, C.POINTER(C.c_ubyte), C.c_int,
                                                 _F, _I]),
            "mjmlx_get_mj_model": (C.c_void_p, [C.c_void_p]),
            "mjmlx_model_nsensor": (C.c_int, [C.c_void_p]),
            "mjmlx_model_nsensordata": (C.c_int, [C.c_void_p]),
            "mjmlx_batched_nsensordata": (C.c_int, [C.c_void_p]),
            "mjmlx_batched_get_sensordata": (_F, [C.c_void_p, _I]),
            "mjmlx_batched_set_pack_spec": (C.c_int, [C.c_void_p, _I, C.c_int, C.c_int, C.c_int]),
            "mjmlx_batched_get_pack": (C.c_int, [C.c_void_p, _F]),
            "mjmlx_batched_eval_state": (None, [C.c_void_p]),
            "mjmlx_batched_set_warmstart_mode": (C.c_int, [C.c_void_p, C.c_int]),
            "mjmlx_batched_get_qacc_warmstart": (C.c_int, [C.c_void_p, _F]),
            "mjmlx_batched_set_qacc_warmstart": (C.c_int, [C.c_void_p, _F]),
            "mjmlx_batched_solver_stats": (C.c_int, [C.c_void_p, _I, _I, _I, _I]),
            "mjmlx_batched_set_launch": (C.c_int, [C.c_void_p, C.c_int, C.c_int]),
            "mjmlx_batched_last_step_ms": (C.c_float, [C.c_void_p]),
            "mjmlx_batched_capacity": (C.c_int, [C.c_void_p, _I, _I, _I]),
            "mjmlx_batched_autotune": (C.c_int, [C.c_void_p, C.c_int, _I]),
            "mjmlx_batched_set_variant": (C.c_int, [C.c_void_p, C.c_int]),
            "mjmlx_batched_set_graph": (C.c_int, [C.c_void_p, C.c_int]),
            "mjmlx_batched_debug_step": (C.c_int, [C.c_void_p, _F, _F, _F]),
            "mjmlx_batched_layout": (C.c_int, [C.c_void_p, C.c_char_p, _I, _I, _I]),
            "mjmlx_mlp_create": (C.c_void_p, [C.c_int, C.c_int, C.c_int, C.c_int]),
            "mjmlx_mlp_free": (None, [C.c_void_p]),
            "mjmlx_mlp_dims": (C.c_int, [C.c_void_p, _I, _I, _I, _I]),
            "mjmlx_mlp_set_normalization": (C.c_int, [C.c_void_p, _F, _F, _F, _F]),
            "mjmlx_mlp_get_normalization": (C.c_int, [C.c_void_p, _F, _F, _F, _F]),
            "mjmlx_mlp_get_params": (C.c_int, [C.c_void_p, _F]),
            "mjmlx_mlp_set_params": (C.c_int, [C.c_void_p, _F]),
            "mjmlx_mlp_forward": (C.c_int, [C.c_void_p, _F, C.c_int, _F]),
            "mjmlx_mlp_grad": (C.c_float, [C.c_void_p, _F, _F, C.c_int, _F]),
            "mjmlx_mlp_train_step": (C.c_float, [C.c_void_p, _F, _F, C.c_int, C.c_float, C.c_float]),
            "mjmlx_mlp_dataset_reserve": (C.c_int, [C.c_void_p, C.c_long]),
            "mjmlx_mlp_dataset_size": (C.c_long, [C.c_void_p]),
            "mjmlx_mlp_dataset_append": (C.c_long, [C.c_void_p, _F, _F, C.c_long]),
            "mjmlx_mlp_dataset_get": (C.c_int, [C.c_void_p, C.c_long, C.c_long, _F, _F]),
            "mjmlx_mlp_fit_normalization": (C.c_int, [C.c_void_p]),
            "mjmlx_mlp_fit": (C.c_float, [C.c_void_p, C.c_int, C.c_int, C.c_float, C.c_float, C.c_int]),
            "mjmlx_batched_feature_dim": (C.c_int, [C.c_void_p]),
            "mjmlx_batched_set_feature_mode": (C.c_int, [C.c_void_p, C.c_int]),
            "mjmlx_batched_set_warmstart_mlp": (C.c_int, [C.c_void_p, C.c_void_p]),
            "mjmlx_batched_record": (C.c_long, [C.c_void_p, C.c_void_p, _F, C.c_int]),
            "mjmlx_batched_get_features": (C.c_int, [C.c_void_p, _F]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _fp(a):
    return None if a is None else a.ctypes.data_as(_F)


def _ip(a):
    return None if a is None else a.ctypes.data_as(_I)


class Model:
    """Compiled MJCF model (mjmlx_load_model_from_string_ex)."""

    def __init__(self, xml: str, self_collision: bool = True, max_condim: int = 0):
        o = ModelOptions(int(self_collision), int(max_condim))
        self.h = lib().mjmlx_load_model_from_string_ex(xml.encode(), C.byref(o))
        if not self.h:
            raise RuntimeError("mjmlx_load_model_from_string_ex failed (see stderr)")
        self.info = lib().mjmlx_model_info(self.h)

    def __del__(self):
        if getattr(self, "h", None):
            lib().mjmlx_free_model(self.h)
            self.h = None

    @property
    def nq(self):
        return self.info.nq

    @property
    def nv(self):
        return self.info.nv

    @property
    def nu(self):
        return self.info.nu

    def field(self, name: str) -> np.ndarray:
        n = lib().mjmlx_model_field(self.h, name.encode(), None, 0)
        if n < 0:
            raise KeyError(name)
        out = np.zeros(max(n, 1))
        lib().mjmlx_model_field(self.h, name.encode(), out.ctypes.data_as(_D), n)
        return out[:n]


INTEGRATORS = {"Euler": 0, "RK4": 1, "implicit": 2, "implicitfast": 3}


class BatchedSim:
    """num_envs copies of a model stepped on the GPU (mjmlx_batched_*)."""

    def __init__(self, model: Model, num_envs: int, integrator: int | None = None, solver_iterations: int = 0,
                 domain_rand: bool = False):
        if integrator is None:
            integrator = int(model.field("integrator")[0])
        cfg = BatchedConfig(num_envs, integrator, 1, solver_iterations, 1 if domain_rand else 0)
        self.model = model
        self.h = lib().mjmlx_batched_create(model.h, C.byref(cfg))
        if not self.h:
            raise RuntimeError("mjmlx_batched_create failed (see stderr)")
        self.B = num_envs
        self.nq, self.nv, self.nu = model.nq, model.nv, model.nu

    def __del__(self):
        if getattr(self, "h", None):
            lib().mjmlx_batched_free(self.h)
            self.h = None

    def step(self, ctrl=None, n: int = 1):
        c = None if ctrl is None else np.ascontiguousarray(ctrl, dtype=np.float32)
        if c is not None:
            assert c.size == self.B * self.nu, (c.shape, self.B, self.nu)
        lib().mjmlx_batched_step_n(self.h, _fp(c), n)

    def sync(self):
        lib().mjmlx_batched_eval_state(self.h)

    def set_env_params(self, env_idx=None, mass_scale=None, com_delta=None, friction_scale=None, gain_scale=None):
        """Per-env physical parameters (sim created with domain_rand); returns the C status."""
        idx = None if env_idx is None else np.ascontiguousarray(env_idx, dtype=np.int32)
        n = self.B if idx is None else len(idx)
        f = lambda a: None if a is None else np.ascontiguousarray(a, dtype=np.float32)
        ms, cd, fs, gs = f(mass_scale), f(com_delta), f(friction_scale), f(gain_scale)
        return lib().mjmlx_batched_set_env_params(self.h, _ip(idx), n, _fp(ms), _fp(cd), _fp(fs), _fp(gs))

    def set_free_drag(self, drag_k, env_idx=None):
        idx = None if env_idx is None else np.ascontiguousarray(env_idx, dtype=np.int32)
        n = self.B if idx is None else len(idx)
        d = np.ascontiguousarray(drag_k, dtype=np.float32)
        return lib().mjmlx_batched_set_free_drag(self.h, _ip(idx), n, _fp(d))

    def raycast(self, pnt, vec, bodyexclude: int = -1, geomgroup=None, flg_static: bool = True):
        """mj_ray per env and ray: pnt, vec [B, nray, 3]; returns (dist, geom id) [B, nray]."""
        p = np.ascontiguousarray(pnt, dtype=np.float32)
        v = np.ascontiguousarray(vec, dtype=np.float32)
        nray = p.shape[1]
        dist = np.zeros((self.B, nray), np.float32)
        gid = np.zeros((self.B, nray), np.int32)
        gg = None if geomgroup is None else np.ascontiguousarray(geomgroup, dtype=np.uint8)
        rc = lib().mjmlx_batched_raycast(self.h, _fp(p), _fp(v), nray, int(bodyexclude),
                                         None if gg is None else gg.ctypes.data_as(C.POINTER(C.c_ubyte)),
                                         int(bool(flg_static)), _fp(dist), _ip(gid))
        assert rc == 0, rc
        return dist, gid

    def sensordata(self) -> np.ndarray:
        """Sensor values [B, nsensordata] at the start of the last step's last substep."""
        nsd = lib().mjmlx_batched_nsensordata(self.h)
        if nsd == 0:
            return np.zeros((self.B, 0), np.float32)
        n = C.c_int()
        p = lib().mjmlx_batched_get_sensordata(self.h, C.byref(n))
        assert p and n.value == self.B * nsd
        return np.ctypeslib.as_array(p, shape=(n.value,)).copy().reshape(self.B, nsd)

    def num_free_joints(self) -> int:
        return lib().mjmlx_batched_num_free_joints(self.h)

    def state(self):
        qp = np.zeros((self.B, self.nq), np.float32)
        qv = np.zeros((self.B, self.nv), np.float32)
        lib().mjmlx_batched_get_state(self.h, _fp(qp), _fp(qv), None, None)
        return qp, qv

    def set_state(self, qpos=None, qvel=None):
        qp = None if qpos is None else np.ascontiguousarray(qpos, dtype=np.float32)
        qv = None if qvel is None else np.ascontiguousarray(qvel, dtype=np.float32)
        lib().mjmlx_batched_set_state(self.h, _fp(qp), _fp(qv))

    def qacc_warmstart(self):
        out = np.zeros((self.B, self.nv), np.float32)
        assert lib().mjmlx_batched_get_qacc_warmstart(self.h, _fp(out)) == 0
        return out

    def set_qacc_warmstart(self, a):
        a = np.ascontiguousarray(a, dtype=np.float32)
        assert lib().mjmlx_batched_set_qacc_warmstart(self.h, _fp(a)) == 0

    def reset(self, mask):
        m = np.ascontiguousarray(mask, dtype=np.int32)
        lib().mjmlx_batched_reset(self.h, m.ctypes.data_as(_I))

    def stats(self):
        out = [np.zeros(self.B, np.int32) for _ in range(4)]
        assert lib().mjmlx_batched_solver_stats(self.h, *(o.ctypes.data_as(_I) for o in out)) == 0
        return dict(niter=out[0], status=out[1], ncon=out[2], nefc=out[3])

    def set_launch(self, persistent: bool, grid: int = 0):
        assert lib().mjmlx_batched_set_launch(self.h, int(persistent), grid) == 0

    def set_warmstart_mode(self, mode: int):
        return lib().mjmlx_batched_set_warmstart_mode(self.h, mode)

    def last_step_ms(self) -> float:
        return lib().mjmlx_batched_last_step_ms(self.h)

    def set_graph(self, on: bool = True):
        """Replay the learned control step as a captured HIP graph."""
        assert lib().mjmlx_batched_set_graph(self.h, int(on)) == 0

    def autotune(self, steps: int = 3) -> int:
        """Time the step's kernel variants from the current state and keep the
        fastest; returns 0 (LDS workspace, 256 registers), 1 (device memory),
        2 (LDS workspace, 168 VGPRs), 3 (LDS workspace, 512 registers),
        4 (device memory, 168 VGPRs), 5 (LDS workspace, 32 lanes per env) or
        6 (the same at 168 VGPRs). The state is restored."""
        c = C.c_int()
        assert lib().mjmlx_batched_autotune(self.h, steps, C.byref(c)) == 0
        return c.value

    def set_variant(self, choice: int) -> bool:
        """Select the kernel variant by the autotuner's numbering; False when
        the sim cannot run it."""
        return lib().mjmlx_batched_set_variant(self.h, int(choice)) == 0

    def capacity(self):
        a, b, c = C.c_int(), C.c_int(), C.c_int()
        lib().mjmlx_batched_capacity(self.h, C.byref(a), C.byref(b), C.byref(c))
        return a.value, b.value, c.value

    def layout(self, name: str):
        off, ln, kind = C.c_int(), C.c_int(), C.c_int()
        if lib().mjmlx_batched_layout(self.h, name.encode(), C.byref(off), C.byref(ln), C.byref(kind)) != 0:
            raise KeyError(name)
        return off.value, ln.value, kind.value

    def debug_step(self, ctrl=None):
        total = self.layout("total")[1]
        d1 = np.zeros(self.B * total, np.float32)
        d2 = np.zeros(self.B * total, np.float32)
        c = None if ctrl is None else np.ascontiguousarray(ctrl, dtype=np.float32)
        assert lib().mjmlx_batched_debug_step(self.h, _fp(c), _fp(d1), _fp(d2)) == 0
        return DebugDump(self, d1, d2, total)

    def feature_dim(self) -> int:
        return lib().mjmlx_batched_feature_dim(self.h)

    def set_feature_mode(self, mode: int):
        assert lib().mjmlx_batched_set_feature_mode(self.h, mode) == 0

    def set_warmstart_mlp(self, mlp):
        self._mlp = mlp  # the sim borrows the network: keep it alive
        assert lib().mjmlx_batched_set_warmstart_mlp(self.h, None if mlp is None else mlp.h) == 0

    def record(self, mlp, ctrl=None, n: int = 1) -> int:
        c = None if ctrl is None else np.ascontiguousarray(ctrl, dtype=np.float32)
        r = lib().mjmlx_batched_record(self.h, mlp.h, _fp(c), n)
        assert r >= 0
        return r

    def features(self):
        out = np.zeros((self.B, self.feature_dim()), np.float32)
        assert lib().mjmlx_batched_get_features(self.h, _fp(out)) > 0
        return out

    def _getter(self, fn: str, shape):
        n = C.c_int()
        p = getattr(lib(), fn)(self.h, C.byref(n))
        if not p:
            raise RuntimeError(f"{fn} failed")
        return np.ctypeslib.as_array(p, shape=(n.value,)).copy().reshape(shape)

    def xpos(self):
        return self._getter("mjmlx_batched_get_xpos", (self.B, -1, 3))

    def xquat(self):
        return self._getter("mjmlx_batched_get_xquat", (self.B, -1, 4))

    def subtree_com(self):
        return self._getter("mjmlx_batched_get_subtree_com", (self.B, -1, 3))

    def cinert(self):
        return self._getter("mjmlx_batched_get_cinert", (self.B, -1, 10))

    def cvel(self):
        return self._getter("mjmlx_batched_get_cvel", (self.B, -1, 6))

    def qfrc_actuator(self):
        return self._getter("mjmlx_batched_get_qfrc_actuator", (self.B, -1))

    def cfrc_ext(self):
        return self._getter("mjmlx_batched_get_cfrc_ext", (self.B, -1, 6))


class DebugDump:
    def __init__(self, sim: BatchedSim, d1, d2, total):
        self.sim, self.d = sim, {1: d1, 2: d2}
        self.total = total

    def get(self, which: int, env: int, name: str) -> np.ndarray:
        off, ln, kind = self.sim.layout(name)
        seg = self.d[which][env * self.total + off: env * self.total + off + ln].copy()
        return seg.view(np.int32) if kind == 1 else seg


class MLP:
    """The learned warm-start network (mjmlx_mlp_*), FP32 master weights on the GPU."""

    def __init__(self, nin: int, nh: int, nout: int, seed: int = 0):
        self.h = lib().mjmlx_mlp_create(nin, nh, nout, seed)
        if not self.h:
            raise RuntimeError("mjmlx_mlp_create failed (see stderr)")
        self.nin, self.nh, self.nout = nin, nh, nout
        n = C.c_int()
        lib().mjmlx_mlp_dims(self.h, None, None, None, C.byref(n))
        self.nparam = n.value

    def __del__(self):
        if getattr(self, "h", None):
            lib().mjmlx_mlp_free(self.h)
            self.h = None

    def params(self):
        out = np.zeros(self.nparam, np.float32)
        assert lib().mjmlx_mlp_get_params(self.h, _fp(out)) == self.nparam
        return out

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.size == self.nparam
        assert lib().mjmlx_mlp_set_params(self.h, _fp(p)) == self.nparam

    def set_normalization(self, in_mean=None, in_std=None, out_mean=None, out_std=None):
        a = [None if v is None else np.ascontiguousarray(v, dtype=np.float32) for v in (in_mean, in_std, out_mean, out_std)]
        assert lib().mjmlx_mlp_set_normalization(self.h, *(_fp(v) for v in a)) == 0

    def normalization(self):
        a = [np.zeros(self.nin, np.float32), np.zeros(self.nin, np.float32),
             np.zeros(self.nout, np.float32), np.zeros(self.nout, np.float32)]
        assert lib().mjmlx_mlp_get_normalization(self.h, *(_fp(v) for v in a)) == 0
        return a

    def forward(self, x):
        x = np.ascontiguousarray(x, dtype=np.float32)
        y = np.zeros((x.shape[0], self.nout), np.float32)
        assert lib().mjmlx_mlp_forward(self.h, _fp(x), x.shape[0], _fp(y)) == 0
        return y

    def grad(self, x, y):
        x = np.ascontiguousarray(x, dtype=np.float32)
        y = np.ascontiguousarray(y, dtype=np.float32)
        g = np.zeros(self.nparam, np.float32)
        loss = lib().mjmlx_mlp_grad(self.h, _fp(x), _fp(y), x.shape[0], _fp(g))
        assert loss >= 0
        return loss, g

    def train_step(self, x, y, lr: float, wd: float = 0.0) -> float:
        x = np.ascontiguousarray(x, dtype=np.float32)
        y = np.ascontiguousarray(y, dtype=np.float32)
        loss = lib().mjmlx_mlp_train_step(self.h, _fp(x), _fp(y), x.shape[0], lr, wd)
        assert loss >= 0
        return loss

    def reserve(self, rows: int):
        assert lib().mjmlx_mlp_dataset_reserve(self.h, rows) == 0

    def dataset_size(self) -> int:
        return lib().mjmlx_mlp_dataset_size(self.h)

    def append(self, x, y) -> int:
        x = np.ascontiguousarray(x, dtype=np.float32)
        y = np.ascontiguousarray(y, dtype=np.float32)
        return lib().mjmlx_mlp_dataset_append(self.h, _fp(x), _fp(y), x.shape[0])

    def dataset(self, first: int, rows: int):
        x = np.zeros((rows, self.nin), np.float32)
        y = np.zeros((rows, self.nout), np.float32)
        assert lib().mjmlx_mlp_dataset_get(self.h, first, rows, _fp(x), _fp(y)) == 0
        return x, y

    def fit_normalization(self):
        assert lib().mjmlx_mlp_fit_normalization(self.h) == 0

    def fit(self, epochs: int, batch: int, lr: float = 1e-3, wd: float = 0.0, seed: int = 0) -> float:
        loss = lib().mjmlx_mlp_fit(self.h, epochs, batch, lr, wd, seed)
        assert loss >= 0, "mjmlx_mlp_fit failed (see stderr)"
        return loss
