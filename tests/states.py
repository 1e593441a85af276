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
