"""The step with 32 lanes per env (two envs per 64-lane wave, variants 5 and 6
of the autotuner) against the FP64 oracle, through the C ABI.

Odd batch sizes leave the last wave's second half idle. Bars as the 64-lane
kernels': contact pair lists identical, qacc and efc_force within 1e-4 of
the oracle (qacc relative to the balanced accelerations, efc_force to the
largest row force); an env gives the same bits in either half of a wave, in
another wave, and stepped alone.
"""
from __future__ import annotations

import numpy as np
import pytest

import ref64
import states
from mjmlx_hip import capi, scenes

pytestmark = pytest.mark.gpu

VARIANTS = [5, 6]  # 256 registers per lane; 168 VGPRs


@pytest.fixture(scope="module")
def quadruped():
    if capi.lib().mjmlx_hip_device_count() < 1:
        pytest.fail("no HIP device visible to the engine")
    return capi.Model(scenes.quadruped_xml(self_collision=True))


def _humanoid_states(model, B, seed):
    orc = ref64.Oracle(ref64.fields_from(model.field))
    rng = np.random.default_rng(seed)
    q = np.tile(model.field("qpos0"), (B, 1))
    qv = rng.normal(0, 0.3, (B, model.nv))
    qa = np.zeros((B, model.nv))
    lens = rng.integers(5, 80, size=B)
    for k in range(int(lens.max())):
        idx = np.nonzero(lens > k)[0]
        ctrl = rng.uniform(-0.4, 0.4, (len(idx), model.nu))
        qs, vs, as_ = q[idx].copy(), qv[idx].copy(), qa[idx].copy()
        orc.batch_step(qs, vs, as_, ctrl, 4)
        q[idx], qv[idx], qa[idx] = qs, vs, as_
    f32 = lambda a: np.ascontiguousarray(a, dtype=np.float32)
    return f32(q), f32(qv), f32(qa), f32(rng.uniform(-0.4, 0.4, (B, model.nu))), orc


def _parity(model, q, qv, qa, ctrl, orc, variant):
    B = q.shape[0]
    sim = capi.BatchedSim(model, B)
    assert sim.set_variant(variant)
    sim.set_state(q, qv)
    sim.set_qacc_warmstart(qa)
    dump = sim.debug_step(ctrl)
    st = sim.stats()
    assert not (st["status"] & ~3).any(), st["status"]
    worst_q = worst_f = 0.0
    ncon = 0
    for e in range(B):
        states.oracle_step(orc, q[e], qv[e], qa[e], ctrl[e])
        con = dump.get(1, e, "contact").reshape(-1, 16)[: st["ncon"][e]]
        ok, msg = states.compare_contacts(con[:, 13].view(np.int32), con[:, 12], orc.contacts(), margin_tol=1e-4,
                                          ncon=st["ncon"][e])
        assert ok, f"env {e}: {msg}"
        ncon += len(con)
        assert st["nefc"][e] == orc.nefc(), (e, st["nefc"][e], orc.nefc())
        eq, ef = states.qacc_force_errors(orc, dump.get(2, e, "qacc"), dump.get(2, e, "efc_force"), st["nefc"][e])
        worst_q, worst_f = max(worst_q, eq), max(worst_f, ef or 0.0)
    return worst_q, worst_f, ncon


@pytest.mark.parametrize("variant", VARIANTS)
def test_quadruped_parity(quadruped, variant):
    B = 97
    q, qv, qa, ctrl, orc = states.quadruped_states(ref64.fields_from(quadruped.field), B, seed=3)
    wq, wf, ncon = _parity(quadruped, q, qv, qa, ctrl, orc, variant)
    print(f"quadruped, variant {variant}: contacts {ncon}, worst qacc rel err {wq:.2e}, efc_force {wf:.2e}")
    assert ncon > B
    assert wq < 1e-4 and wf < 1e-4


@pytest.mark.parametrize("variant", VARIANTS)
def test_humanoid_parity(variant):
    model = capi.Model(scenes.humanoid_xml())
    B = 33
    q, qv, qa, ctrl, orc = _humanoid_states(model, B, 2)
    wq, wf, ncon = _parity(model, q, qv, qa, ctrl, orc, variant)
    print(f"humanoid, variant {variant}: contacts {ncon}, worst qacc rel err {wq:.2e}, efc_force {wf:.2e}")
    assert ncon > B // 2
    assert wq < 1e-4 and wf < 1e-4


@pytest.mark.parametrize("variant", VARIANTS)
def test_same_bits_in_either_half(quadruped, variant):
    """Envs 0 and 1 share a wave (lanes 0-31 and 32-63); envs 2k and 2k+1
    hold the same state swapped, and the batch's last env (B odd) is alone in
    its wave. Ten control steps of 4 substeps with per-env controls."""
    B = 9
    q, qv, qa, ctrl, _ = states.quadruped_states(ref64.fields_from(quadruped.field), 3, seed=5)
    order = [0, 1, 1, 0, 2, 0, 0, 2, 1]  # source state of each env
    qb, vb = q[order], qv[order]
    rng = np.random.default_rng(7)
    cs = rng.uniform(-0.5, 0.5, (10, 3, quadruped.nu)).astype(np.float32)
    sim = capi.BatchedSim(quadruped, B)
    assert sim.set_variant(variant)
    sim.set_state(qb, vb)
    for k in range(10):
        sim.step(cs[k][order], 4)
    qk, vk = sim.state()
    assert not (sim.stats()["status"] & ~3).any()
    for e in range(B):
        f = order.index(order[e])
        assert np.array_equal(qk[e], qk[f]) and np.array_equal(vk[e], vk[f]), (e, f)
    # the same env stepped alone (B = 1: one half of one wave)
    one = capi.BatchedSim(quadruped, 1)
    assert one.set_variant(variant)
    one.set_state(qb[8:9], vb[8:9])
    for k in range(10):
        one.step(cs[k][order[8]:order[8] + 1], 4)
    q1, v1 = one.state()
    assert np.array_equal(q1[0], qk[8]) and np.array_equal(v1[0], vk[8])


def test_variant_selection_rules(quadruped):
    """32-lane variants only for models of at most 32 dofs."""
    sim = capi.BatchedSim(quadruped, 4)
    for v in range(7):
        assert sim.set_variant(v), v
    assert not sim.set_variant(7)
    big = capi.Model(scenes.boxes_xml(16, 3))
    sb = capi.BatchedSim(big, 2)
    assert not sb.set_variant(5) and not sb.set_variant(6)
    assert sb.set_variant(1)
