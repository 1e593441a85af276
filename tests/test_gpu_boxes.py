"""Box scenes on the GPU (the step kernel built with the capsule-box and
box-box narrow phases) against the FP64 oracle, through the C ABI."""
from __future__ import annotations

import numpy as np
import pytest

import ref64
import states
from mjmlx_hip import capi, scenes

pytestmark = pytest.mark.gpu


def _caps(model):
    """The engine's capacities for a scene of several free bodies: no row
    budget, min(512, the pairs' contact maxima) contacts."""
    sim = capi.BatchedSim(model, 1)
    maxcon = sim.capacity()[0]
    del sim
    return dict(con_cap=maxcon, efc_cap=0)


def _settled_states(model, B, seed, nbox):
    orc = ref64.Oracle(ref64.fields_from(model.field)).set_caps(**_caps(model))
    rng = np.random.default_rng(seed)
    q = np.tile(model.field("qpos0"), (B, 1))
    qv = np.zeros((B, model.nv))
    qa = np.zeros((B, model.nv))
    lens, nsub = (rng.integers(20, 120, size=B), 4) if nbox <= 14 else (rng.integers(30, 50, size=B), 4)
    for k in range(int(lens.max())):
        idx = np.nonzero(lens > k)[0]
        qs, vs, as_ = q[idx].copy(), qv[idx].copy(), qa[idx].copy()
        orc.batch_step(qs, vs, as_, np.zeros((len(idx), 0)), nsub, 8)
        q[idx], qv[idx], qa[idx] = qs, vs, as_
    assert np.isfinite(qv).all()
    f32 = lambda a: np.ascontiguousarray(a, dtype=np.float32)
    return f32(q), f32(qv), f32(qa), orc


def _newton_optimum(model):
    f = ref64.fields_from(model.field)
    f.update(solver=2, iterations=200, tolerance=1e-14)
    return ref64.Oracle(f).set_caps(**_caps(model))


@pytest.mark.parametrize("nbox,solver", [(6, "Newton"), (6, "CG"), (14, "Newton"), (32, "CG")])
def test_box_scene_single_step_parity(nbox, solver):
    """6 boxes run from LDS; 14 (nv 84) and 32 (nv 192, landed: ~470 rows,
    beyond the reference's 256-row budget) from the device-memory workspace.
    No row budget: every contact gets its rows, here and in the oracle."""
    model = capi.Model(scenes.boxes_xml(nbox, 3, solver=solver, iterations=50 if solver == "Newton" else 16))
    B = 32 if nbox <= 14 else 16
    q, qv, qa, orc = _settled_states(model, B, 4, nbox)
    orc_opt = _newton_optimum(model)
    worstp = 0.0
    sim = capi.BatchedSim(model, B)
    sim.set_state(q, qv)
    sim.set_qacc_warmstart(qa)
    dump = sim.debug_step(None)
    st = sim.stats()
    assert not st["status"].any(), st["status"]  # no overflow either
    ncon = 0
    worst = worstf = worsts = 0.0
    for e in range(B):
        states.oracle_step(orc, q[e], qv[e], qa[e], np.zeros(0, np.float32))
        con = dump.get(1, e, "contact").reshape(-1, 16)[: st["ncon"][e]]
        ok, msg = states.compare_contacts(con[:, 13].view(np.int32), con[:, 12], orc.contacts(), margin_tol=1e-4, ncon=st["ncon"][e])
        assert ok, f"env {e}: {msg}"
        ncon += len(con)
        assert st["nefc"][e] == orc.nefc(), (e, st["nefc"][e], orc.nefc())
        eq, ef = states.qacc_force_errors(orc, dump.get(2, e, "qacc"), dump.get(2, e, "efc_force"), st["nefc"][e])
        worst, worstf = max(worst, eq), max(worstf, ef or 0.0)
        if e % 8 == 0:
            worsts = max(worsts, states.input_sensitivity(orc, q[e], qv[e], qa[e], np.zeros(0, np.float32), seed=e))
        if solver == "CG":
            ok, gk, g16, sc = states.cg_progress_ok(orc, orc_opt, q[e], qv[e], qa[e], np.zeros(0, np.float32),
                                                    dump.get(2, e, "qacc").astype(np.float64))
            worstp = max(worstp, gk / max(g16, 1e-6 * sc))
            assert ok, (e, gk, g16, sc)
    print(f"nbox {nbox} {solver}: contacts {ncon}, worst qacc rel err {worst:.2e}, efc_force {worstf:.2e}, "
          f"oracle's own FP32-input sensitivity {worsts:.2e}, "
          f"mean iterations {st['niter'].mean():.2f}, lds {sim.capacity()}"
          + (f", worst remaining-gap ratio to FP64 CG {worstp:.2f}" if solver == "CG" else ""))
    assert ncon > B
    # Newton converges to the 1e-8 scaled residual: qacc within 1e-4 of the
    # accelerations that balance (resting boxes have qacc ~ 0 while contact
    # forces cancel g), row forces within 1e-4 of the largest. CG at its
    # 16-iteration budget is unconverged: FP32 and FP64 CG end at different
    # iterates of about equal cost, judged by the remaining-gap ratio above
    if solver == "Newton":
        assert worst < 1e-4 and worstf < 1e-4


def test_device_workspace_matches_lds(monkeypatch):
    """The same kernel with its per-env workspace in device memory instead of
    LDS, and in LDS at each register budget (the variants the autotuner
    picks from), runs the same arithmetic: 20 steps of the quadruped agree bit
    for bit."""
    model = capi.Model(scenes.quadruped_xml())
    B = 256
    q0 = scenes.quadruped_initial_qpos(B, 3)
    rng = np.random.default_rng(3)
    acts = [scenes.quadruped_actions(B, rng).astype(np.float32) for _ in range(20)]
    out = []
    for gws, occ, big in (("0", "1", "0"), ("1", "0", "0"), ("0", "0", "0"), ("0", "0", "1"), ("1", "1", "0")):
        monkeypatch.setenv("MJMLX_HIP_GLOBAL_WORKSPACE", gws)
        monkeypatch.setenv("MJMLX_HIP_HIGH_OCC", occ)
        monkeypatch.setenv("MJMLX_HIP_BIG_REGS", big)
        sim = capi.BatchedSim(model, B)
        sim.set_state(q0.astype(np.float32), np.zeros((B, model.nv), np.float32))
        for a in acts:
            sim.step(a)
        out.append(sim.state())
        st = sim.stats()
        assert not (st["status"] & ~3).any()
        del sim
    for o in out[1:]:
        for a, b in zip(out[0], o):
            assert np.array_equal(a, b)


def test_autotune_keeps_state():
    """The workspace autotuner times both placements from the current state and
    restores it: the tuned sim then steps bit-identically to an untuned one."""
    model = capi.Model(scenes.boxes_xml(6, 3, solver="CG", iterations=16))
    B = 128
    q = np.tile(model.field("qpos0"), (B, 1)).astype(np.float32)
    out = []
    for tune in (False, True):
        sim = capi.BatchedSim(model, B)
        sim.set_state(q, np.zeros((B, model.nv), np.float32))
        for _ in range(20):
            sim.step(None, 4)
        if tune:
            choice = sim.autotune(2)
            assert choice in (0, 1, 2, 3, 4, 5, 6)
        for _ in range(5):
            sim.step(None, 4)
        out.append(sim.state())
        del sim
    for a, b in zip(out[0], out[1]):
        assert np.array_equal(a, b)


def test_mixed_box_counts_batch():
    """The uneven-contact workload as the bench runs it, at a size the oracle
    covers: 96 envs with N in U{4..32} boxes, grouped one sim per N, queued
    together, 30 control steps of landing, then one substep of every group
    compared env by env with the FP64 oracle (same capacities: no row
    budget): identical contact pairs and row counts, no overflow, and CG's
    remaining cost gap to the optimum at most 1.25x the FP64 CG's at the same
    16-iteration budget."""
    rng = np.random.default_rng(9)
    ns = rng.integers(4, 33, size=96)
    groups = []
    for n in sorted(set(ns.tolist())):
        nb = int((ns == n).sum())
        model = capi.Model(scenes.boxes_xml(n, 100 + n, solver="CG", iterations=16))
        q = np.tile(model.field("qpos0"), (nb, 1))
        hs = model.field("geom_size").reshape(-1, 3)[1:]
        for b in range(n):
            a = 7 * b
            q[:, a + 2] = rng.uniform(0.1, 1.0, nb) + np.linalg.norm(hs[b])
            qq = rng.normal(size=(nb, 4))
            q[:, a + 3:a + 7] = qq / np.linalg.norm(qq, axis=1, keepdims=True)
        sim = capi.BatchedSim(model, nb)
        sim.set_state(q.astype(np.float32), np.zeros((nb, model.nv), np.float32))
        groups.append((n, nb, model, sim))
    for _ in range(30):
        for _, _, _, sim in groups:
            sim.step(None, 4)
    for _, _, _, sim in groups:
        sim.sync()
    ncon = nenv = 0
    worst_ratio = 0.0
    for n, nb, model, sim in groups:
        q, qv = sim.state()
        qa = sim.qacc_warmstart()
        caps = _caps(model)
        orc = ref64.Oracle(ref64.fields_from(model.field)).set_caps(**caps)
        opt = _newton_optimum(model)
        dump = sim.debug_step(None)
        st = sim.stats()
        assert not st["status"].any(), (n, st["status"])
        for e in range(nb):
            states.oracle_step(orc, q[e], qv[e], qa[e], np.zeros(0, np.float32))
            con = dump.get(1, e, "contact").reshape(-1, 16)[: st["ncon"][e]]
            ok, msg = states.compare_contacts(con[:, 13].view(np.int32), con[:, 12], orc.contacts(), margin_tol=1e-4,
                                              ncon=st["ncon"][e])
            assert ok, f"N {n} env {e}: {msg}"
            assert st["nefc"][e] == orc.nefc(), (n, e, st["nefc"][e], orc.nefc())
            ok, gk, g16, sc = states.cg_progress_ok(orc, opt, q[e], qv[e], qa[e], np.zeros(0, np.float32),
                                                    dump.get(2, e, "qacc").astype(np.float64))
            assert ok, (n, e, gk, g16, sc)
            worst_ratio = max(worst_ratio, gk / max(g16, 1e-6 * sc))
            ncon += len(con)
            nenv += 1
    print(f"mixed boxes: {len(groups)} groups, {nenv} envs, {ncon} contacts, worst remaining-gap ratio to FP64 CG "
          f"{worst_ratio:.2f}")
    assert nenv == 96 and ncon > 96


@pytest.mark.parametrize("nbox", [8, 24])
def test_boxes_500_step_rollout(nbox):
    """The boxes workload's rollout length (500 control steps, CG 16): 64 envs
    dropped from rest stay finite, never overflow a capacity, and settle on
    the floor (every box centre above the floor by at least its smallest
    half-size minus 1 cm, and the boxes' mean speed below 0.05 m/s at the end)."""
    model = capi.Model(scenes.boxes_xml(nbox, 7))
    B = 64
    sim = capi.BatchedSim(model, B)
    for _ in range(500):
        sim.step(None, 4)
    q, v = sim.state()
    st = sim.stats()
    assert np.isfinite(q).all() and np.isfinite(v).all()
    assert not (st["status"] & ~0).any(), st["status"]
    size = model.field("geom_size").reshape(-1, 3)[1:]
    z = q.reshape(B, nbox, 7)[:, :, 2]
    assert (z > size.min(axis=1)[None, :] - 0.01).all(), z.min()
    speed = np.linalg.norm(v.reshape(B, nbox, 6)[:, :, :3], axis=2)
    print(f"{nbox} boxes, 500 steps: lowest centre {z.min():.3f} m, mean speed {speed.mean():.4f} m/s, "
          f"mean contacts {st['ncon'].mean():.1f}, iterations {st['niter'].mean():.1f}")
    assert speed.mean() < 0.05
