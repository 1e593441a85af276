"""ctypes binding of the FP64 oracle (oracle/ref64.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker. The engine never imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "lib", "libref64.so")

_I = C.POINTER(C.c_int)
_D = C.POINTER(C.c_double)

INT_FIELDS = [
    "body_parentid", "body_rootid", "body_weldid", "body_jntnum", "body_jntadr", "body_dofnum", "body_dofadr",
]
MODEL_LAYOUT = [
    ("nq", C.c_int), ("nv", C.c_int), ("nu", C.c_int), ("nbody", C.c_int), ("njnt", C.c_int), ("ngeom", C.c_int),
    ("timestep", C.c_double), ("gravity", C.c_double * 3), ("tolerance", C.c_double), ("ls_tolerance", C.c_double),
    ("impratio", C.c_double),
    ("iterations", C.c_int), ("ls_iterations", C.c_int), ("solver", C.c_int), ("cone", C.c_int),
    ("integrator", C.c_int), ("disableflags", C.c_int),
    ("body_parentid", _I), ("body_rootid", _I), ("body_weldid", _I), ("body_jntnum", _I), ("body_jntadr", _I),
    ("body_dofnum", _I), ("body_dofadr", _I),
    ("body_pos", _D), ("body_quat", _D), ("body_ipos", _D), ("body_iquat", _D), ("body_mass", _D), ("body_inertia", _D),
    ("jnt_type", _I), ("jnt_qposadr", _I), ("jnt_dofadr", _I), ("jnt_bodyid", _I), ("jnt_limited", _I),
    ("jnt_pos", _D), ("jnt_axis", _D), ("jnt_range", _D), ("jnt_margin", _D), ("jnt_solref", _D), ("jnt_solimp", _D),
    ("jnt_stiffness", _D),
    ("dof_bodyid", _I), ("dof_jntid", _I), ("dof_parentid", _I),
    ("dof_armature", _D), ("dof_damping", _D), ("dof_frictionloss", _D), ("dof_solref", _D), ("dof_solimp", _D),
    ("geom_type", _I), ("geom_contype", _I), ("geom_conaffinity", _I), ("geom_condim", _I), ("geom_bodyid", _I),
    ("geom_priority", _I),
    ("geom_size", _D), ("geom_pos", _D), ("geom_quat", _D), ("geom_friction", _D), ("geom_solref", _D),
    ("geom_solimp", _D), ("geom_margin", _D), ("geom_gap", _D), ("geom_solmix", _D), ("geom_rbound", _D),
    ("actuator_trnid", _I), ("actuator_ctrllimited", _I), ("actuator_forcelimited", _I),
    ("actuator_gear", _D), ("actuator_gainprm", _D), ("actuator_biasprm", _D), ("actuator_ctrlrange", _D),
    ("actuator_forcerange", _D),
    ("nexclude", C.c_int), ("exclude_signature", _I),
    ("npair", C.c_int), ("pair_geom1", _I), ("pair_geom2", _I), ("pair_dim", _I),
    ("pair_friction", _D), ("pair_solref", _D), ("pair_solimp", _D), ("pair_margin", _D), ("pair_gap", _D),
    ("qpos0", _D), ("qpos_spring", _D),
]


class OrcModel(C.Structure):
    _fields_ = MODEL_LAYOUT


class OrcPair(C.Structure):
    _fields_ = [("g1", C.c_int), ("g2", C.c_int), ("b1", C.c_int), ("b2", C.c_int), ("condim", C.c_int),
                ("friction", C.c_double * 5), ("solref", C.c_double * 2), ("solimp", C.c_double * 5),
                ("margin", C.c_double), ("gap", C.c_double)]


class OrcContact(C.Structure):
    _fields_ = [("dist", C.c_double), ("pos", C.c_double * 3), ("frame", C.c_double * 9),
                ("geom1", C.c_int), ("geom2", C.c_int), ("pair", C.c_int), ("dim", C.c_int),
                ("friction", C.c_double * 5), ("solref", C.c_double * 2), ("solimp", C.c_double * 5),
                ("includemargin", C.c_double)]


SCALARS = ["nq", "nv", "nu", "nbody", "njnt", "ngeom", "timestep", "tolerance", "ls_tolerance", "impratio",
           "iterations", "ls_iterations", "solver", "cone", "integrator", "disableflags", "npair", "nexclude"]


def build_lib() -> str:
    """Compile the oracle with its Makefile (gcc)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIBPATH


_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIBPATH):
            build_lib()
        L = C.CDLL(LIBPATH)
        L.orc_prepare.restype = C.c_void_p
        L.orc_prepare.argtypes = [C.POINTER(OrcModel)]
        L.orc_free.argtypes = [C.c_void_p]
        L.orc_set_caps.argtypes = [C.c_void_p, C.c_int, C.c_int]
        L.orc_npair.argtypes = [C.c_void_p]
        L.orc_pairs.restype = C.POINTER(OrcPair)
        L.orc_pairs.argtypes = [C.c_void_p]
        L.orc_body_invweight0.restype = _D
        L.orc_body_invweight0.argtypes = [C.c_void_p]
        L.orc_dof_invweight0.restype = _D
        L.orc_dof_invweight0.argtypes = [C.c_void_p]
        L.orc_meaninertia.restype = C.c_double
        L.orc_meaninertia.argtypes = [C.c_void_p]
        L.orc_data_new.restype = C.c_void_p
        L.orc_data_new.argtypes = [C.c_void_p]
        L.orc_data_free.argtypes = [C.c_void_p]
        L.orc_set_state.argtypes = [C.c_void_p, _D, _D, _D, _D]
        L.orc_get_state.argtypes = [C.c_void_p, _D, _D, _D]
        L.orc_set_env_params.argtypes = [C.c_void_p, _D, _D, _D, _D, _D]
        L.orc_nfree.argtypes = [C.c_void_p]
        L.orc_inject_contacts.argtypes = [C.c_void_p, C.c_void_p, C.c_int, _I, _D, _D, _D]
        L.orc_inject_contacts.restype = None
        L.orc_ray_geom.restype = C.c_double
        L.orc_ray_geom.argtypes = [C.c_int, _D, _D, _D]
        L.orc_ray.restype = C.c_double
        L.orc_ray.argtypes = [C.c_void_p, C.c_void_p, _D, _D, _I, C.POINTER(C.c_int)]
        L.orc_step.argtypes = [C.c_void_p, C.c_void_p]
        L.orc_forward.argtypes = [C.c_void_p, C.c_void_p]
        L.orc_kinematics.argtypes = [C.c_void_p, C.c_void_p]
        L.orc_get.argtypes = [C.c_void_p, C.c_char_p, _D, C.c_int]
        L.orc_ncon.argtypes = [C.c_void_p]
        L.orc_contacts.restype = C.POINTER(OrcContact)
        L.orc_contacts.argtypes = [C.c_void_p]
        L.orc_nefc.argtypes = [C.c_void_p]
        L.orc_solver_niter.argtypes = [C.c_void_p]
        L.orc_cost.argtypes = [C.c_void_p, C.c_void_p, _D, _D, _D, _D]
        L.orc_collide_pair.argtypes = [C.c_void_p, C.c_int, _D, _D, _D, _D, C.POINTER(OrcContact), C.c_int]
        L.orc_batch_step.argtypes = [C.c_void_p, C.c_int, _D, _D, _D, _D, C.c_int, C.c_int, _I]
        _lib = L
    return _lib


def _dp(a: np.ndarray):
    return a.ctypes.data_as(_D)


def _ip(a: np.ndarray):
    return a.ctypes.data_as(_I)


class Oracle:
    """FP64 reference step for one compiled model (raw fields from the engine's
    MJCF compiler; pair table, invweight0 and meaninertia recomputed here)."""

    def __init__(self, fields: dict):
        self._keep = {}
        m = OrcModel()
        for name, ctype in MODEL_LAYOUT:
            if name == "gravity":
                g = np.asarray(fields["gravity"], dtype=np.float64)
                m.gravity = (C.c_double * 3)(*g)
            elif ctype is _I:
                a = np.ascontiguousarray(np.asarray(fields.get(name, []), dtype=np.int32).ravel())
                if a.size == 0:
                    a = np.zeros(1, dtype=np.int32)
                self._keep[name] = a
                setattr(m, name, _ip(a))
            elif ctype is _D:
                a = np.ascontiguousarray(np.asarray(fields.get(name, []), dtype=np.float64).ravel())
                if a.size == 0:
                    a = np.zeros(1, dtype=np.float64)
                self._keep[name] = a
                setattr(m, name, _dp(a))
            else:
                v = fields[name]
                setattr(m, name, ctype(int(v) if ctype is C.c_int else float(v)))
        self._m = m
        self.nq, self.nv, self.nu = int(fields["nq"]), int(fields["nv"]), int(fields["nu"])
        self.nbody, self.ngeom = int(fields["nbody"]), int(fields["ngeom"])
        self.p = lib().orc_prepare(C.byref(self._m))
        if not self.p:
            raise RuntimeError("oracle: model preparation failed (mass matrix not positive definite)")
        self.d = lib().orc_data_new(self.p)

    def __del__(self):
        try:
            L = lib()
            if getattr(self, "d", None):
                L.orc_data_free(self.d)
            if getattr(self, "p", None):
                L.orc_free(self.p)
        except Exception:
            pass

    def set_caps(self, con_cap: int = 128, efc_cap: int | None = None):
        """Per-env contact cap and row budget (0: unbounded); the defaults are
        the reference's: 128 contacts, 72 rows for nv <= 32, else 256."""
        if efc_cap is None:
            efc_cap = 72 if self.nv <= 32 else 256
        lib().orc_set_caps(self.p, int(con_cap), int(efc_cap))
        return self

    # ---- derived tables
    def pairs(self):
        n = lib().orc_npair(self.p)
        ps = lib().orc_pairs(self.p)
        return [(ps[i].g1, ps[i].g2, ps[i].b1, ps[i].b2, ps[i].condim, list(ps[i].friction), list(ps[i].solref),
                 list(ps[i].solimp), ps[i].margin, ps[i].gap) for i in range(n)]

    def body_invweight0(self):
        return np.ctypeslib.as_array(lib().orc_body_invweight0(self.p), shape=(2 * self.nbody,)).copy()

    def dof_invweight0(self):
        return np.ctypeslib.as_array(lib().orc_dof_invweight0(self.p), shape=(max(self.nv, 1),))[: self.nv].copy()

    def meaninertia(self):
        return lib().orc_meaninertia(self.p)

    # ---- stepping
    def set_state(self, qpos, qvel=None, qacc_ws=None, ctrl=None):
        def arr(x, n):
            if x is None:
                return None
            a = np.ascontiguousarray(np.asarray(x, dtype=np.float64).ravel())
            assert a.size == n, (a.size, n)
            return a
        qp, qv, qa, ct = arr(qpos, self.nq), arr(qvel, self.nv), arr(qacc_ws, self.nv), arr(ctrl, self.nu)
        self._st = (qp, qv, qa, ct)
        lib().orc_set_state(self.d, *(None if a is None else _dp(a) for a in (qp, qv, qa, ct)))

    def set_env_params(self, mass_scale=None, com_delta=None, friction_scale=None, gain_scale=None, drag=None):
        """Domain randomisation of the oracle's env (reference
        src/batched.cpp:1508-1538): mass_scale [nbody], com_delta [nbody, 3],
        friction_scale (scalar), gain_scale [nu], drag [nfree]; None keeps a
        parameter (initially the identity)."""
        nfree = lib().orc_nfree(self.d)
        def arr(x, n):
            if x is None:
                return None
            a = np.ascontiguousarray(np.asarray(x, dtype=np.float64).ravel())
            assert a.size == n, (a.size, n)
            return a
        args = (arr(mass_scale, self.nbody), arr(com_delta, 3 * self.nbody), arr(friction_scale, 1),
                arr(gain_scale, self.nu), arr(drag, nfree))
        self._rand = args
        lib().orc_set_env_params(self.d, *(None if a is None else _dp(a) for a in args))
        return self

    def ray(self, pnt, vec, geom_ok):
        """mj_ray at the current geom poses (call kinematics() first): (x, geom id)
        of the nearest allowed geom along pnt + x vec, or (-1, -1)."""
        p = np.ascontiguousarray(np.asarray(pnt, np.float64).ravel())
        v = np.ascontiguousarray(np.asarray(vec, np.float64).ravel())
        ok = np.ascontiguousarray(np.asarray(geom_ok, np.int32).ravel())
        assert ok.size == self.ngeom
        g = C.c_int()
        x = lib().orc_ray(self.p, self.d, _dp(p), _dp(v), _ip(ok), C.byref(g))
        return x, g.value

    def inject_contacts(self, pairs=None, dist=None, pos=None, normal=None):
        """Following forwards use this contact list (pair index, distance,
        position [n,3], normal [n,3] from geom1 to geom2) instead of the
        narrow phase; no arguments restores the narrow phase."""
        L = lib()
        if pairs is None:
            L.orc_inject_contacts(self.p, self.d, -1, None, None, None, None)
            return
        pr = np.ascontiguousarray(np.asarray(pairs, np.int32).ravel())
        f = lambda a: np.ascontiguousarray(np.asarray(a, np.float64).ravel())
        dd, pp, nn = f(dist), f(pos), f(normal)
        assert dd.size == pr.size and pp.size == 3 * pr.size and nn.size == 3 * pr.size
        self._inj = (pr, dd, pp, nn)
        L.orc_inject_contacts(self.p, self.d, pr.size, _ip(pr), _dp(dd), _dp(pp), _dp(nn))

    def nfree(self) -> int:
        return lib().orc_nfree(self.d)

    def get_state(self):
        qp, qv, qa = np.zeros(self.nq), np.zeros(max(self.nv, 1)), np.zeros(max(self.nv, 1))
        lib().orc_get_state(self.d, _dp(qp), _dp(qv), _dp(qa))
        return qp, qv[: self.nv], qa[: self.nv]

    def step(self) -> int:
        return lib().orc_step(self.p, self.d)

    def forward(self) -> int:
        return lib().orc_forward(self.p, self.d)

    def kinematics(self):
        lib().orc_kinematics(self.p, self.d)

    def get(self, name: str) -> np.ndarray:
        n = lib().orc_get(self.d, name.encode(), None, 0)
        if n < 0:
            raise KeyError(name)
        out = np.zeros(max(n, 1))
        lib().orc_get(self.d, name.encode(), _dp(out), n)
        return out[:n]

    def contacts(self):
        n = lib().orc_ncon(self.d)
        cs = lib().orc_contacts(self.d)
        return [dict(dist=cs[i].dist, pos=np.array(cs[i].pos), frame=np.array(cs[i].frame), geom1=cs[i].geom1,
                     geom2=cs[i].geom2, pair=cs[i].pair, dim=cs[i].dim, friction=np.array(cs[i].friction))
                for i in range(n)]

    def nefc(self) -> int:
        return lib().orc_nefc(self.d)

    def niter(self) -> int:
        return lib().orc_solver_niter(self.d)

    def cost(self, qacc):
        """(cost, gradient, Hessian) of the last forward's constraint problem at qacc."""
        q = np.ascontiguousarray(qacc, dtype=np.float64)
        c = C.c_double()
        g = np.zeros(self.nv)
        H = np.zeros((self.nv, self.nv))
        lib().orc_cost(self.p, self.d, _dp(q), C.byref(c), _dp(g), _dp(H))
        return c.value, g, H

    def batch_step(self, qpos, qvel, qacc_ws, ctrl, nsub: int, nthreads: int = 1):
        """Step env-major arrays in place; returns per-env solver iterations."""
        B = qpos.shape[0]
        for a in (qpos, qvel, qacc_ws):
            assert a.dtype == np.float64 and a.flags.c_contiguous
        ct = None if ctrl is None else np.ascontiguousarray(ctrl, dtype=np.float64)
        nit = np.zeros(B, dtype=np.int32)
        bad = lib().orc_batch_step(self.p, B, _dp(qpos), _dp(qvel), _dp(qacc_ws),
                                   None if ct is None else _dp(ct), nsub, nthreads, _ip(nit))
        return nit, bad


RAW_FIELDS = [name for name, _ in MODEL_LAYOUT if name not in ("gravity",)] + ["gravity"]


def fields_from(getter) -> dict:
    """Collect the raw model fields through ``getter(name) -> np.ndarray``."""
    out = {}
    for name in RAW_FIELDS:
        v = getter(name)
        if name in SCALARS:
            out[name] = v[0]
        else:
            out[name] = v
    return out


def ray_geom(gtype: int, size, p, v) -> float:
    """mj_ray against one primitive in its own frame: x along p + x v, or -1."""
    f = lambda a: np.ascontiguousarray(np.asarray(a, np.float64).ravel())
    s, pp, vv = f(size), f(p), f(v)
    return lib().orc_ray_geom(int(gtype), _dp(s), _dp(pp), _dp(vv))


def math(op: str, inp, nout: int = 64) -> np.ndarray:
    """Call one of the oracle's math routines (orc_math) on a flat input."""
    L = lib()
    if not hasattr(L, "_math_set"):
        L.orc_math.argtypes = [C.c_char_p, _D, _D]
        L.orc_math.restype = C.c_int
        L._math_set = True
    a = np.ascontiguousarray(np.asarray(inp, dtype=np.float64).ravel())
    out = np.zeros(max(nout, 64) + a.size)
    n = L.orc_math(op.encode(), _dp(a), _dp(out))
    if n < 0:
        raise ValueError(f"orc_math({op}) failed")
    return out[:n].copy()
