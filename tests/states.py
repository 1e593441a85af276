"""Seeded quadruped / boxes states for parity tests.

States are produced by the FP64 oracle rolling the scene forward from its
initial pose under random actions, so they cover feet in contact, fallen
bodies, active joint limits and self-contacts. They are rounded to FP32 (the
engine's state type) and the oracle is restarted from the rounded state, so
both sides see bit-identical inputs.
"""
from __future__ import annotations

import numpy as np

import ref64
from mjmlx_hip import scenes


def quadruped_states(fields: dict, nenv: int, seed: int, steps=(0, 60), nsub: int = 4):
    """Per-env rollout lengths drawn from U{steps}; returns float32 qpos, qvel,
    qacc_warmstart and the next control, plus the oracle used."""
    rng = np.random.default_rng(seed)
    orc = ref64.Oracle(fields)
    q = scenes.quadruped_initial_qpos(nenv, seed)
    qv = np.zeros((nenv, orc.nv))
    qa = np.zeros((nenv, orc.nv))
    lens = rng.integers(steps[0], steps[1] + 1, size=nenv)
    for k in range(int(lens.max())):
        ctrl = scenes.quadruped_actions(nenv, rng)
        active = lens > k
        if not active.any():
            break
        idx = np.nonzero(active)[0]
        qs, vs, as_ = q[idx].copy(), qv[idx].copy(), qa[idx].copy()
        orc.batch_step(qs, vs, as_, np.ascontiguousarray(ctrl[idx]), nsub)
        q[idx], qv[idx], qa[idx] = qs, vs, as_
    ctrl = scenes.quadruped_actions(nenv, rng)
    f32 = lambda a: np.ascontiguousarray(a, dtype=np.float32)
    return f32(q), f32(qv), f32(qa), f32(ctrl), orc


def oracle_step(orc, q, qv, qa, ctrl):
    """One oracle physics step from FP32-rounded state (forward fields kept)."""
    orc.set_state(q.astype(np.float64), qv.astype(np.float64), qa.astype(np.float64), ctrl.astype(np.float64))
    orc.forward()
    return orc


def sorted_pairs(pairs, dists):
    return sorted(zip([int(p) for p in pairs], [float(d) for d in dists]))


def compare_contacts(gpu_pairs, gpu_dist, orc_contacts, margin_tol=1e-5, ncon=None):
    """Contact pair lists must agree exactly except for contacts whose distance
    sits within margin_tol of the (zero) margin, where FP32 and FP64 may round
    to different sides. The kernel stores only the contacts the row budget can
    admit: when ``ncon`` (its full count) exceeds the stored list, the count
    must match the oracle's and the stored list its prefix. Returns (ok, message)."""
    if ncon is not None and int(ncon) > len(gpu_pairs):
        n = len(gpu_pairs)
        pre = [c["pair"] for c in orc_contacts[:n]]
        ok = int(ncon) == len(orc_contacts) and [int(p) for p in gpu_pairs] == pre
        return ok, f"truncated: gpu ncon={int(ncon)} stored={list(map(int, gpu_pairs))} oracle n={len(orc_contacts)} prefix={pre}"
    a = sorted_pairs(gpu_pairs, gpu_dist)
    b = sorted_pairs([c["pair"] for c in orc_contacts], [c["dist"] for c in orc_contacts])
    if [p for p, _ in a] == [p for p, _ in b]:
        return True, ""
    ra = [x for x in a if abs(x[1]) > margin_tol]
    rb = [x for x in b if abs(x[1]) > margin_tol]
    ok = [p for p, _ in ra] == [p for p, _ in rb]
    return ok, f"gpu={a} oracle={b}"


def cost_scale(orc, c_opt: float) -> float:
    """Scale for relative gaps of the constraint cost after an oracle step:
    the larger of 1, |optimal cost| and the Gauss energy of the unconstrained
    acceleration 0.5 a^T M a (an optimum at a = qacc_smooth has cost ~0, so
    FP32 rounding of that acceleration is measured against its energy)."""
    a = orc.get("qacc_smooth")
    M = orc.get("qM").reshape(orc.nv, orc.nv)
    return max(1.0, abs(c_opt), 0.5 * float(a @ M @ a))


def cg_progress_ok(orc_cg16, orc_opt, q, qv, qa, ctrl, qk, slack: float = 1.25):
    """CG at a fixed iteration budget (unconverged): the kernel's remaining
    cost gap to the optimum is at most ``slack`` times the FP64 CG's at the
    same budget, plus 1e-6 of the cost scale (FP32 resolution of the cost).
    Returns (ok, kernel gap, FP64 CG gap, scale); orc_cg16 is left at its step."""
    oracle_step(orc_opt, q, qv, qa, ctrl)
    a_opt = orc_opt.get("qacc").copy()
    oracle_step(orc_cg16, q, qv, qa, ctrl)
    a16 = orc_cg16.get("qacc").copy()
    c = lambda a: orc_cg16.cost(np.asarray(a, np.float64))[0]
    c_opt = c(a_opt)
    sc = cost_scale(orc_cg16, c_opt)
    gk, g16 = c(qk) - c_opt, c(a16) - c_opt
    oracle_step(orc_cg16, q, qv, qa, ctrl)
    return gk <= slack * max(g16, 0.0) + 1e-6 * sc, gk, g16, sc


def accel_scale(orc) -> float:
    """Scale of the accelerations that balance in the oracle's last step: the
    larger of 1, |qacc| and |qacc_smooth| (max norms). A resting body has
    qacc ~ 0 while its contact forces cancel gravity's g, so qacc's error is
    measured against the accelerations being balanced."""
    return max(1.0, float(np.max(np.abs(orc.get("qacc")))), float(np.max(np.abs(orc.get("qacc_smooth")))))


def qacc_force_errors(orc, qacc_k, force_k=None, nefc_k=None):
    """(qacc error relative to accel_scale, efc_force error relative to the
    oracle's largest row force floored at 1, or None when the row counts
    differ or there are no rows) of the kernel's qacc / efc_force against the
    oracle's last step. Rows are in the same order on both sides (contacts in
    pair order, each contact's rows in the same order)."""
    b = orc.get("qacc")
    eq = float(np.max(np.abs(np.asarray(qacc_k, np.float64) - b))) / accel_scale(orc)
    ef = None
    n = orc.nefc()
    if force_k is not None and n and (nefc_k is None or int(nefc_k) == n):
        f = orc.get("efc_force")
        ef = float(np.max(np.abs(np.asarray(force_k[:n], np.float64) - f))) / max(1.0, float(np.max(np.abs(f))))
    return eq, ef


def input_sensitivity(orc, q, qv, qa, ctrl, seed: int = 0, rel: float = 6e-8) -> float:
    """The FP64 oracle's own qacc change, relative to accel_scale, when qpos is
    perturbed by FP32 rounding-level noise (rel): the error floor that FP32
    state arithmetic cannot go below. Leaves the oracle at the perturbed step."""
    oracle_step(orc, q, qv, qa, ctrl)
    b, sc = orc.get("qacc").copy(), accel_scale(orc)
    rng = np.random.default_rng(seed)
    qp = q.astype(np.float64) * (1 + rng.uniform(-rel, rel, q.size))
    orc.set_state(qp, qv.astype(np.float64), qa.astype(np.float64), ctrl.astype(np.float64))
    orc.forward()
    return float(np.max(np.abs(orc.get("qacc") - b))) / sc


def oracle_cfrc_ext(orc, pair_b1, pair_b2, nbody):
    """Per-body contact force of the oracle's last step in the engine's
    cfrc_ext convention (reference src/compat/mkx_kernels/solver.hpp:906-929):
    each contact's normal force along its normal, + on body2 and - on body1,
    zero torque. The normal force is the sum of a pyramidal contact's edge
    forces, the normal row's force otherwise."""
    out = np.zeros((nbody, 6))
    ty, ids, f = orc.get("efc_type").astype(int), orc.get("efc_id").astype(int), orc.get("efc_force")
    cons = orc.contacts()
    fn = np.zeros(len(cons))
    seen = set()
    for r in range(orc.nefc()):
        if ty[r] == 6:          # pyramidal edge
            fn[ids[r]] += f[r]
        elif ty[r] in (5, 7) and ids[r] not in seen:   # frictionless / elliptic normal row
            fn[ids[r]] = f[r]
            seen.add(ids[r])
    for i, c in enumerate(cons):
        n = c["frame"][:3]
        out[pair_b2[c["pair"]], 3:] += fn[i] * n
        out[pair_b1[c["pair"]], 3:] -= fn[i] * n
    return out
