"""ctypes binding of the test-only serial interpreter of the kernel source
(libmjmlx_emu.so, built from mujoco-mlx-cpp_amd/csrc/emu.cpp).

Used by the CPU test suite to check the kernel algorithm against the FP64
oracle without a GPU. Not part of the engine.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIBPATH = os.path.join(ROOT, "mujoco-mlx-cpp_amd", "lib", "libmjmlx_emu.so")

_F = C.POINTER(C.c_float)
_I = C.POINTER(C.c_int)
_D = C.POINTER(C.c_double)
_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIBPATH):
            raise RuntimeError(f"{LIBPATH} missing: run __graft_entry__.build() first")
        L = C.CDLL(LIBPATH)
        L.mjemu_create.restype = C.c_void_p
        L.mjemu_create.argtypes = [C.c_char_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]
        L.mjemu_free.argtypes = [C.c_void_p]
        L.mjemu_set_state.argtypes = [C.c_void_p, _F, _F, _F]
        L.mjemu_get_state.argtypes = [C.c_void_p, _F, _F, _F]
        L.mjemu_step.argtypes = [C.c_void_p, _F, C.c_int, C.c_int]
        L.mjemu_layout.argtypes = [C.c_void_p, C.c_char_p, _I, _I, _I]
        L.mjemu_dump.restype = _F
        L.mjemu_dump.argtypes = [C.c_void_p, C.c_int]
        L.mjemu_stats.argtypes = [C.c_void_p, _I, _I, _I, _I]
        L.mjemu_model_field.argtypes = [C.c_void_p, C.c_char_p, _D, C.c_int]
        L.mjemu_features.argtypes = [C.c_void_p, _F]
        L.mjemu_set_pred.argtypes = [C.c_void_p, _F]
        L.mjemu_set_rand.argtypes = [C.c_void_p, _F]
        L.mjemu_ray_geom.restype = C.c_float
        L.mjemu_ray_geom.argtypes = [C.c_int, _F, _F, _F]
        L.mjemu_rays.argtypes = [C.c_void_p, _F, C.c_int, _I, _F, _I]
        L.mjemu_sensordata.argtypes = [C.c_void_p, _F]
        L.mjemu_narrow.argtypes = [C.c_int, C.c_int, _F, _F, _F, _F, _F, _F, C.c_float, _F]
        L.mjemu_set_poison.argtypes = [C.c_void_p, C.c_int, C.c_uint]
        L.mjemu_compile.restype = C.c_void_p
        L.mjemu_compile.argtypes = [C.c_char_p, C.c_int, C.c_int]
        _lib = L
    return _lib


def _fp(a):
    return None if a is None else a.ctypes.data_as(_F)


class EmuSim:
    def __init__(self, xml: str, nenv: int, self_collision: int = 1, max_condim: int = 0, maxcon: int = 0,
                 maxefc: int = 0):
        self.h = lib().mjemu_create(xml.encode(), self_collision, max_condim, nenv, maxcon, maxefc)
        if not self.h:
            raise RuntimeError("mjemu_create failed")
        self.B = nenv
        self.nq = int(self.field("nq")[0])
        self.nv = int(self.field("nv")[0])
        self.nu = int(self.field("nu")[0])
        self.total = self.layout("total")[1]

    def __del__(self):
        if getattr(self, "h", None):
            lib().mjemu_free(self.h)

    def field(self, name: str) -> np.ndarray:
        n = lib().mjemu_model_field(self.h, name.encode(), None, 0)
        if n < 0:
            raise KeyError(name)
        out = np.zeros(max(n, 1))
        lib().mjemu_model_field(self.h, name.encode(), out.ctypes.data_as(_D), n)
        return out[:n]

    def layout(self, name: str):
        off, ln, kind = C.c_int(), C.c_int(), C.c_int()
        if lib().mjemu_layout(self.h, name.encode(), C.byref(off), C.byref(ln), C.byref(kind)) != 0:
            raise KeyError(name)
        return off.value, ln.value, kind.value

    def set_state(self, qpos=None, qvel=None, qacc_ws=None):
        a = [None if x is None else np.ascontiguousarray(x, dtype=np.float32) for x in (qpos, qvel, qacc_ws)]
        lib().mjemu_set_state(self.h, *(_fp(x) for x in a))

    def get_state(self):
        qp = np.zeros((self.B, self.nq), np.float32)
        qv = np.zeros((self.B, self.nv), np.float32)
        qa = np.zeros((self.B, self.nv), np.float32)
        lib().mjemu_get_state(self.h, _fp(qp), _fp(qv), _fp(qa))
        return qp, qv, qa

    def step(self, ctrl=None, nsub: int = 1, dump: bool = False) -> int:
        c = None if ctrl is None else np.ascontiguousarray(ctrl, dtype=np.float32)
        return lib().mjemu_step(self.h, _fp(c), nsub, int(dump))

    def set_poison(self, mode: int, pattern: int = 0):
        """Workspace before each env: 0 zeroed, 1 filled with the 32-bit
        pattern, 2 the pattern once and then each env reuses the previous
        env's workspace (as one persistent workgroup serving successive envs)."""
        lib().mjemu_set_poison(self.h, int(mode), int(pattern) & 0xFFFFFFFF)

    def stats(self):
        out = [np.zeros(self.B, np.int32) for _ in range(4)]
        lib().mjemu_stats(self.h, *(o.ctypes.data_as(_I) for o in out))
        return dict(niter=out[0], status=out[1], ncon=out[2], nefc=out[3])

    def features(self) -> np.ndarray:
        """Learned warm-start features written by the last step, [B, nfeat]."""
        nf = lib().mjemu_features(self.h, None)
        out = np.zeros((self.B, nf), np.float32)
        lib().mjemu_features(self.h, _fp(out))
        return out

    def set_pred(self, pred=None):
        """Predicted qacc [B, nv] offered to the solver as a warm-start candidate (None: none)."""
        self._pred = None if pred is None else np.ascontiguousarray(pred, dtype=np.float32)
        lib().mjemu_set_pred(self.h, _fp(self._pred))

    def set_rand(self, mass_scale, com_delta, friction_scale, gain_scale, drag):
        """Domain randomisation of every env: mass_scale [B, nbody], com_delta
        [B, nbody, 3], friction_scale [B], gain_scale [B, nu], drag [B, nfree]
        (packed into the engine's per-env parameter blocks)."""
        cols = [np.asarray(a, np.float32).reshape(self.B, -1)
                for a in (mass_scale, com_delta, friction_scale, gain_scale, drag)]
        self._rand = np.ascontiguousarray(np.concatenate(cols, axis=1))
        n = lib().mjemu_set_rand(self.h, _fp(self._rand))
        assert n == self._rand.shape[1], (n, self._rand.shape)
        self.total = self.layout("total")[1]

    def sensordata(self) -> np.ndarray:
        """Sensor values written by the last step, [B, nsensordata]."""
        nsd = lib().mjemu_sensordata(self.h, None)
        out = np.zeros((self.B, nsd), np.float32)
        if nsd:
            lib().mjemu_sensordata(self.h, _fp(out))
        return out

    def rays(self, rays, gmask):
        """The kernel's per-env rays at the current qpos: rays [B, nray, 6]
        (pnt, vec), gmask [ngeom]; returns dist, gid [B, nray]."""
        r = np.ascontiguousarray(rays, dtype=np.float32)
        nray = r.shape[1]
        g = np.ascontiguousarray(gmask, dtype=np.int32)
        dist = np.zeros((self.B, nray), np.float32)
        gid = np.zeros((self.B, nray), np.int32)
        lib().mjemu_rays(self.h, _fp(r), nray, g.ctypes.data_as(_I), _fp(dist), gid.ctypes.data_as(_I))
        return dist, gid

    def dump(self, which: int, env: int, name: str) -> np.ndarray:
        off, ln, kind = self.layout(name)
        p = lib().mjemu_dump(self.h, which)
        base = np.ctypeslib.as_array(p, shape=(self.B * self.total,))
        seg = base[env * self.total + off: env * self.total + off + ln].copy()
        if kind == 1:
            return seg.view(np.int32)
        return seg


def narrow(t1, t2, s1, s2, x1, R1, x2, R2, margin):
    """The kernel's narrow phase for one pair: list of (dist, pos, normal)."""
    f = lambda a: np.ascontiguousarray(np.asarray(a, np.float32).ravel())
    args = [f(a) for a in (s1, s2, x1, R1, x2, R2)]
    out = np.zeros(56, np.float32)
    n = lib().mjemu_narrow(int(t1), int(t2), *(_fp(a) for a in args), float(margin), _fp(out))
    if n < 0:
        raise ValueError("unsupported pair")
    return [(float(out[7 * k]), out[7 * k + 1:7 * k + 4].copy(), out[7 * k + 4:7 * k + 7].copy()) for k in range(n)]


def ray_geom(gtype, size, p, v) -> float:
    """The kernel's ray-primitive intersection (FP32) in the geom frame."""
    f = lambda a: np.ascontiguousarray(np.asarray(a, np.float32).ravel())
    a = [f(x) for x in (size, p, v)]
    return float(lib().mjemu_ray_geom(int(gtype), *(_fp(x) for x in a)))


class CompiledModel:
    """Model fields of an MJCF compiled by the engine's compiler, without
    building kernel tables (accepts pair types the kernels do not support)."""

    def __init__(self, xml: str, self_collision: int = 1, max_condim: int = 0):
        self.h = lib().mjemu_compile(xml.encode(), self_collision, max_condim)
        if not self.h:
            raise RuntimeError("mjemu_compile failed")

    def __del__(self):
        if getattr(self, "h", None):
            lib().mjemu_free(self.h)

    field = EmuSim.field
