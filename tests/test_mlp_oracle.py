"""Learned warm start on the CPU: the FP64 network oracle is pinned by finite
differences and the Adam closed form; the kernel's feature writer and its
prediction candidate are checked through the serial interpreter.
"""
from __future__ import annotations

import numpy as np
import pytest

import emu
import mlp64
import ref64
import states
from mjmlx_hip import scenes


def _rand_net(rng, nin, nh, nout, scale=0.5):
    return np.concatenate([rng.normal(0, scale, int(np.prod(s))) for s in mlp64.shapes(nin, nh, nout)])


def test_oracle_backward_matches_central_differences():
    rng = np.random.default_rng(0)
    nin, nh, nout, n = 5, 6, 3, 7
    p = _rand_net(rng, nin, nh, nout)
    x = rng.normal(size=(n, nin))
    t = rng.normal(size=(n, nout))
    g = mlp64.grad(p, x, t, nin, nh, nout)
    h = 1e-6
    fd = np.zeros_like(p)
    for i in range(p.size):
        e = np.zeros_like(p)
        e[i] = h
        fd[i] = (mlp64.loss(p + e, x, t, nin, nh, nout) - mlp64.loss(p - e, x, t, nin, nh, nout)) / (2 * h)
    assert np.max(np.abs(g - fd)) < 1e-7 * max(1.0, np.max(np.abs(fd)))
    # a dropped SiLU derivative term or a transposed weight would fail this by orders of magnitude
    assert np.max(np.abs(fd)) > 1e-3


def test_oracle_adamw_closed_form():
    rng = np.random.default_rng(1)
    p0 = rng.normal(size=50)
    g1, g2 = rng.normal(size=50), rng.normal(size=50)
    lr, wd = 1e-3, 0.01
    m = v = np.zeros(50)
    p1, m, v = mlp64.adamw(p0, g1, m, v, 1, lr, wd)
    # first step: bias-corrected moments are g and g^2, so the step is lr * sign(g) (+ decay)
    assert np.allclose(p1, p0 - lr * (g1 / (np.abs(g1) + 1e-8) + wd * p0), atol=1e-15)
    p2, _, _ = mlp64.adamw(p1, g2, m, v, 2, lr, wd)
    mh = (0.9 * 0.1 * g1 + 0.1 * g2) / (1 - 0.81)
    vh = (0.999 * 0.001 * g1 ** 2 + 0.001 * g2 ** 2) / (1 - 0.999 ** 2)
    assert np.allclose(p2, p1 - lr * (mh / (np.sqrt(vh) + 1e-8) + wd * p1), atol=1e-15)


def test_bf16_rounding_is_nearest_even():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -2.5, 1.0 + 2 ** -9, 3.0e38])
    r = mlp64.bf16(x)
    assert list(r[:4]) == [1.0, 1.0, 1.0 + 2 ** -6, -2.5]  # ties go to the even mantissa
    assert r[4] == 1.0 and np.isfinite(r[5])


def _limb_class(field):
    """Independent restatement: the limb of a geom is the rank of the root
    body's child whose subtree holds it (mod 4), -1 on the world / root."""
    par = field("body_parentid").astype(int)
    gb = field("geom_bodyid").astype(int)
    limb = {}
    for b in range(1, len(par)):
        if par[b] == 0:
            limb[b] = -1
        elif par[par[b]] == 0:
            limb[b] = sum(1 for c in range(1, b) if par[c] == par[b]) % 4
        else:
            limb[b] = limb[par[b]]
    return [limb.get(b, -1) if b else -1 for b in gb]


def test_features_layout_matches_state_and_contacts():
    xml = scenes.quadruped_xml(self_collision=True)
    B = 24
    sim = emu.EmuSim(xml, B, maxcon=32, maxefc=96)
    q, qv, qa, ctrl, orc = states.quadruped_states(ref64.fields_from(sim.field), B, seed=5)
    sim.set_state(q, qv, qa)
    sim.step(ctrl, nsub=1, dump=True)
    f = sim.features()
    nq, nv, nu = sim.nq, sim.nv, sim.nu
    assert f.shape == (B, nq + nv + nu + 16 * 8)
    qp, qvel, _ = sim.get_state()
    assert np.array_equal(f[:, :nq], qp) and np.array_equal(f[:, nq:nq + nv], qvel)
    assert not f[:, nq + nv:nq + nv + nu].any()  # ctrl is inserted later, by the caller
    cls = _limb_class(sim.field)
    g1, g2 = sim.field("cpair_g1").astype(int), sim.field("cpair_g2").astype(int)
    st = sim.stats()
    saw_contact = 0
    for e in range(B):
        con = sim.dump(1, e, "contact").reshape(-1, 16)[: st["ncon"][e]]
        blocks = f[e, nq + nv + nu:].reshape(16, 8)
        k = min(len(con), 16)
        saw_contact += k > 0
        assert not blocks[k:].any()
        for c in range(k):
            n3, depth, hot = blocks[c, :3], blocks[c, 3], blocks[c, 4:]
            assert np.array_equal(n3, con[c, 3:6]) and depth == -con[c, 12]
            assert abs(np.linalg.norm(n3) - 1) < 1e-5
            p = int(con[c, 13:14].view(np.int32)[0])
            want = np.zeros(4)
            for g in (g1[p], g2[p]):
                if cls[g] >= 0:
                    want[cls[g]] = 1
            assert np.array_equal(hot, want), (e, c, hot, want)
    assert saw_contact >= B // 2
    # the four feet carry four different limb classes
    foot = [i for i in range(len(cls)) if cls[i] >= 0]
    assert sorted(set(cls[i] for i in foot)) == [0, 1, 2, 3]


def test_prediction_candidate_keeps_results():
    xml = scenes.quadruped_xml(self_collision=True)
    B = 32
    sim = emu.EmuSim(xml, B, maxcon=32, maxefc=96)
    q, qv, qa, ctrl, _ = states.quadruped_states(ref64.fields_from(sim.field), B, seed=9)
    # plain step
    sim.set_state(q, qv, qa)
    sim.step(ctrl, nsub=1)
    plain = sim.get_state()
    it_plain = sim.stats()["niter"].copy()
    solution = plain[2].copy()  # converged qacc of this step
    # a useless prediction loses the cost comparison: results are bitwise unchanged
    rng = np.random.default_rng(3)
    sim.set_pred(rng.normal(0, 1e3, size=(B, sim.nv)))
    sim.set_state(q, qv, qa)
    sim.step(ctrl, nsub=1)
    for a, b in zip(sim.get_state(), plain):
        assert np.array_equal(a, b)
    # the converged qacc as the prediction: same answer to solver tolerance, fewer iterations
    sim.set_pred(solution)
    sim.set_state(q, qv, qa)
    sim.step(ctrl, nsub=1)
    got = sim.get_state()
    it_pred = sim.stats()["niter"]
    scale = np.maximum(1.0, np.abs(plain[2]).max(1, keepdims=True))
    assert np.max(np.abs(got[2] - plain[2]) / scale) < 1e-4
    assert np.max(np.abs(got[1] - plain[1])) < 1e-4 * max(1.0, np.abs(plain[1]).max())
    assert (it_pred <= it_plain).all() and it_pred.mean() < it_plain.mean()
    assert (it_pred == 1).mean() > 0.8
