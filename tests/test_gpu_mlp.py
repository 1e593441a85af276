"""Learned warm start on the GPU, through the C ABI: the BF16 MFMA network
against the FP64 oracle (oracle/mlp64.py), the AdamW update against its closed
form, GPU training, the feature writer against the serial interpreter, and the
warm start in the solver (unchanged results, iteration counts reported).
"""
from __future__ import annotations

import numpy as np
import pytest

import emu
import mlp64
import ref64
import states
from mjmlx_hip import capi, scenes

pytestmark = pytest.mark.gpu

NIN, NH, NOUT = 177, 256, 18


def _net(seed=1):
    return capi.MLP(NIN, NH, NOUT, seed=seed)


def test_forward_matches_oracle():
    m = _net()
    rng = np.random.default_rng(0)
    x = rng.normal(size=(300, NIN)).astype(np.float32)
    y = m.forward(x)
    yref, _ = mlp64.forward(m.params().astype(np.float64), x, NIN, NH, NOUT, round_bf16=True)
    err = np.max(np.abs(y - yref))
    print(f"forward max abs err {err:.3e} (output scale {np.max(np.abs(yref)):.3e})")
    assert err < 2e-3 * (1 + np.max(np.abs(yref)))
    # the unrounded FP64 network is close too (BF16 storage error only)
    yfp, _ = mlp64.forward(m.params().astype(np.float64), x, NIN, NH, NOUT)
    assert np.max(np.abs(y - yfp)) < 3e-2 * (1 + np.max(np.abs(yfp)))


def test_gradient_matches_oracle():
    m = _net(2)
    rng = np.random.default_rng(1)
    rows = 100  # not a multiple of the 32-row GEMM step: exercises the masked padding rows
    x = rng.normal(size=(rows, NIN)).astype(np.float32)
    t = rng.normal(size=(rows, NOUT)).astype(np.float32)
    loss, g = m.grad(x, t)
    p = m.params().astype(np.float64)
    gref = mlp64.grad(p, x, t, NIN, NH, NOUT, round_bf16=True)
    assert loss == pytest.approx(mlp64.loss(p, x, t, NIN, NH, NOUT), rel=1e-2)
    o = 0
    for s in mlp64.shapes(NIN, NH, NOUT):
        k = int(np.prod(s))
        a, b = g[o:o + k].astype(np.float64), gref[o:o + k]
        cos = a @ b / (np.linalg.norm(a) * np.linalg.norm(b))
        rel = np.linalg.norm(a - b) / np.linalg.norm(b)
        print(f"block {s}: cos {cos:.6f} rel {rel:.2e}")
        assert cos > 0.999 and rel < 3e-2, (s, cos, rel)
        o += k


def test_adamw_step_closed_form():
    m = _net(3)
    rng = np.random.default_rng(2)
    x = rng.normal(size=(64, NIN)).astype(np.float32)
    t = rng.normal(size=(64, NOUT)).astype(np.float32)
    p0 = m.params().astype(np.float64)
    _, g = m.grad(x, t)
    lr, wd = 1e-3, 0.01
    m.train_step(x, t, lr, wd)
    p1 = m.params().astype(np.float64)
    want, _, _ = mlp64.adamw(p0, g.astype(np.float64), 0 * p0, 0 * p0, 1, lr, wd)
    assert np.max(np.abs(p1 - want)) < 1e-6


def test_fit_on_device_dataset():
    nin, nh, nout = 40, 64, 8
    rng = np.random.default_rng(4)
    teacher = np.concatenate([rng.normal(0, 0.4, int(np.prod(s))) for s in mlp64.shapes(nin, nh, nout)])
    x = rng.normal(size=(8192, nin))
    y, _ = mlp64.forward(teacher, x, nin, nh, nout)
    m = capi.MLP(nin, nh, nout, seed=5)
    m.reserve(8192)
    assert m.append(x * 3.0 + 1.0, y) == 8192  # raw inputs; the fitted normalisation undoes the affine map
    assert m.dataset_size() == 8192
    xs, ys = m.dataset(100, 4)
    assert np.allclose(xs, (x[100:104] * 3 + 1).astype(np.float32)) and np.allclose(ys, y[100:104].astype(np.float32))
    m.fit_normalization()
    mu, sd, omu, osd = m.normalization()
    assert np.allclose(mu, (x * 3 + 1).mean(0), atol=0.05) and np.allclose(sd, 3 * x.std(0), rtol=0.05)
    assert np.allclose(omu, y.mean(0), atol=1e-3) and np.allclose(osd, y.std(0), rtol=1e-2)
    l1 = m.fit(1, 256, lr=1e-3, seed=0)
    l2 = m.fit(30, 256, lr=1e-3, seed=1)
    print(f"fit loss after 1 epoch {l1:.4f}, after 31 {l2:.4f}")
    assert l2 < 0.3 * l1


def _quadruped_sim(B, seed):
    model = capi.Model(scenes.quadruped_xml(self_collision=True))
    sim = capi.BatchedSim(model, B)
    sim.set_state(scenes.quadruped_initial_qpos(B, seed).astype(np.float32), np.zeros((B, model.nv), np.float32))
    return model, sim


def test_features_match_interpreter():
    B = 64
    model, sim = _quadruped_sim(B, 0)
    emu_sim = emu.EmuSim(scenes.quadruped_xml(self_collision=True), B, maxcon=32, maxefc=96)
    q, qv, qa, ctrl, _ = states.quadruped_states(ref64.fields_from(emu_sim.field), B, seed=13)
    sim.set_state(q, qv)
    sim.set_qacc_warmstart(qa)
    net = capi.MLP(sim.feature_dim(), NH, model.nv)
    net.reserve(B)
    assert sim.record(net, ctrl, 1) == 0  # no features before the first substep after set_state
    fg = sim.features()
    emu_sim.set_state(q, qv, qa)
    emu_sim.step(ctrl, nsub=1)
    fe = emu_sim.features()
    assert fg.shape == fe.shape == (B, model.nq + model.nv + model.nu + 128)
    ncon_g = sim.stats()["ncon"]
    assert np.array_equal(ncon_g, emu_sim.stats()["ncon"])
    assert np.max(np.abs(fg - fe)) < 1e-4 * max(1.0, np.max(np.abs(fe)))
    assert sim.record(net, ctrl, 2) == B and net.dataset_size() == B  # capacity-limited append


def test_learned_warm_start_keeps_results_and_reports_iterations():
    B, nsteps = 1024, 16
    model, sim = _quadruped_sim(B, 1)
    rng = np.random.default_rng(7)
    net = capi.MLP(sim.feature_dim(), NH, model.nv, seed=0)
    net.reserve(B * 4 * nsteps)
    rows = 0
    for _ in range(nsteps):
        rows += sim.record(net, scenes.quadruped_actions(B, rng).astype(np.float32), 4)
    assert rows == B * (4 * nsteps - 1) == net.dataset_size()
    net.fit_normalization()
    loss = net.fit(6, 1024, lr=1e-3, seed=0)
    print(f"trained on {rows} pairs, loss {loss:.4f} (normalised units)")
    assert 0 <= loss < 1.0  # better than predicting the mean
    q, qv = sim.state()
    qa = sim.qacc_warmstart()
    plain = capi.BatchedSim(model, B)
    learned = capi.BatchedSim(model, B)
    learned.set_warmstart_mlp(net)
    assert learned.set_warmstart_mode(2) == 0
    assert plain.set_warmstart_mode(2) == -1  # no network installed
    it_p = it_l = 0
    for k, s in enumerate((plain, learned)):
        s.set_state(q, qv)
        s.set_qacc_warmstart(qa)
    for k in range(4):
        c = scenes.quadruped_actions(B, rng).astype(np.float32)
        plain.step(c, 4)
        learned.step(c, 4)
        sp, sl = plain.stats(), learned.stats()
        assert not (sp["status"] & ~3).any() and not (sl["status"] & ~3).any()
        if k == 0:
            (qp, vp), (ql, vl) = plain.state(), learned.state()
            dv = np.max(np.abs(vp - vl)) / max(1.0, np.max(np.abs(vp)))
            dq = np.max(np.abs(qp - ql))
            print(f"after one control step: max |dqvel| rel {dv:.2e}, max |dqpos| {dq:.2e}")
            # both runs stop at the same 1e-8 scaled residual
            assert dv < 1e-3 and dq < 1e-4
        if k >= 1:  # the first learned substep has no features yet
            it_p += sp["niter"].mean()
            it_l += sl["niter"].mean()
    print(f"mean solver iterations per env-step: plain {it_p / 3:.3f}, learned {it_l / 3:.3f}")
    # held-out states (later in the rollout, actions from another stream): the
    # trained network saves iterations
    assert it_l < it_p
    # mode 3: the prediction only at each call's first substep, one fused launch
    first = capi.BatchedSim(model, B)
    first.set_warmstart_mlp(net)
    assert first.set_warmstart_mode(3) == 0
    ref = capi.BatchedSim(model, B)
    for s in (first, ref):
        s.set_state(q, qv)
        s.set_qacc_warmstart(qa)
    rng = np.random.default_rng(8)
    it_r = it_f = 0
    for k in range(4):
        c = scenes.quadruped_actions(B, rng).astype(np.float32)
        first.step(c, 4)
        ref.step(c, 4)
        sf, sr = first.stats(), ref.stats()
        assert not (sf["status"] & ~3).any()
        if k == 0:  # no features yet: the first call is the plain step, bit for bit
            for a, b in zip(first.state(), ref.state()):
                assert np.array_equal(a, b)
        else:
            it_f += sf["niter"].mean()
            it_r += sr["niter"].mean()
        if k == 1:
            (qr, vr), (qf, vf) = ref.state(), first.state()
            assert np.max(np.abs(vr - vf)) / max(1.0, np.max(np.abs(vr))) < 1e-3 and np.max(np.abs(qr - qf)) < 1e-4
    print(f"first-substep mode: iterations per env-step plain {it_r / 3:.3f}, learned {it_f / 3:.3f}")
    assert it_f < it_r


def test_graph_replay_matches_eager():
    """The learned control step replayed as a captured HIP graph gives the
    eager launches' results bit for bit, on every replay, with eager work
    (state readback, statistics) and synchronisations between the replays,
    and after a change of feature-buffer parity forces a new capture."""
    B = 512
    model, src = _quadruped_sim(B, 3)
    net = capi.MLP(src.feature_dim(), NH, model.nv, seed=2)
    q, qv = src.state()
    qa = src.qacc_warmstart()
    sims = []
    for graph in (False, True):
        s = capi.BatchedSim(model, B)
        s.set_warmstart_mlp(net)
        assert s.set_warmstart_mode(2) == 0
        s.set_graph(graph)
        s.set_state(q, qv)
        s.set_qacc_warmstart(qa)
        sims.append(s)
    rng = np.random.default_rng(11)
    for k in range(8):
        c = scenes.quadruped_actions(B, rng).astype(np.float32)
        nsub = 3 if k == 5 else 4  # an odd count flips the feature parity
        for s in sims:
            s.step(c, nsub)
            s.sync()  # the two sims share the network's activation buffers
        (qe, ve), (qg, vg) = sims[0].state(), sims[1].state()
        assert np.array_equal(qe, qg) and np.array_equal(ve, vg), k
        se, sg = sims[0].stats(), sims[1].stats()
        assert np.array_equal(se["niter"], sg["niter"]), k


def test_humanoid_learned_warm_start():
    """The humanoid's leg of the learned warm start: a network trained on the
    humanoid's own solver states (features 204, output 27) keeps the results
    of the plain warm start and its iteration counts are reported."""
    B, nsteps = 1024, 12
    model = capi.Model(scenes.humanoid_xml())
    sim = capi.BatchedSim(model, B)
    rng = np.random.default_rng(3)
    q0 = np.tile(model.field("qpos0"), (B, 1)).astype(np.float32)
    sim.set_state(q0, rng.normal(0, 0.3, (B, model.nv)).astype(np.float32))
    nf = sim.feature_dim()
    assert nf == model.nq + model.nv + model.nu + 128
    net = capi.MLP(nf, NH, model.nv, seed=1)
    net.reserve(B * 4 * nsteps)
    rows = 0
    for _ in range(nsteps):
        rows += sim.record(net, rng.uniform(-0.4, 0.4, (B, model.nu)).astype(np.float32), 4)
    net.fit_normalization()
    loss = net.fit(6, 1024, lr=1e-3, seed=0)
    print(f"humanoid: trained on {rows} pairs, loss {loss:.4f} (normalised units)")
    assert 0 <= loss < 1.0
    q, qv = sim.state()
    qa = sim.qacc_warmstart()
    plain = capi.BatchedSim(model, B)
    learned = capi.BatchedSim(model, B)
    learned.set_warmstart_mlp(net)
    assert learned.set_warmstart_mode(2) == 0
    for s in (plain, learned):
        s.set_state(q, qv)
        s.set_qacc_warmstart(qa)
    it_p = it_l = 0.0
    for k in range(4):
        c = rng.uniform(-0.4, 0.4, (B, model.nu)).astype(np.float32)
        plain.step(c, 4)
        learned.step(c, 4)
        sp, sl = plain.stats(), learned.stats()
        assert not (sp["status"] & ~3).any() and not (sl["status"] & ~3).any()
        if k == 0:
            (qp, vp), (ql, vl) = plain.state(), learned.state()
            dv = np.max(np.abs(vp - vl)) / max(1.0, np.max(np.abs(vp)))
            dq = np.max(np.abs(qp - ql))
            print(f"humanoid after one control step: max |dqvel| rel {dv:.2e}, max |dqpos| {dq:.2e}")
            assert dv < 1e-3 and dq < 1e-4
        else:
            it_p += sp["niter"].mean()
            it_l += sl["niter"].mean()
    print(f"humanoid mean solver iterations per env-step: plain {it_p / 3:.3f}, learned {it_l / 3:.3f}")
