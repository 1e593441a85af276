"""The allocation-free step (mjmlx_batched_step_n_noalloc, workspaces in the
library's fixed device array) against mjmlx_batched_step_n with its
library-allocated device-memory workspaces, through the C ABI: bit-identical
states on two calls in a row, the second on a smaller sim; sizes beyond the
array and the learned modes rejected."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from mjmlx_hip import capi, scenes

pytestmark = pytest.mark.gpu


def _bind():
    L = capi.lib()
    L.mjmlx_batched_step_n_noalloc.restype = C.c_int
    L.mjmlx_batched_step_n_noalloc.argtypes = [C.c_void_p, C.POINTER(C.c_float), C.c_int]
    L.mjmlx_batched_noalloc_scratch_bytes.restype = C.c_size_t
    L.mjmlx_batched_noalloc_scratch_bytes.argtypes = []
    return L


def _sims(model, B, seed, monkeypatch):
    q = scenes.quadruped_initial_qpos(B, seed).astype(np.float32)
    q[:, 2] -= 0.02
    v = np.zeros((B, model.nv), np.float32)
    monkeypatch.setenv("MJMLX_HIP_GLOBAL_WORKSPACE", "1")
    ref = capi.BatchedSim(model, B)
    monkeypatch.delenv("MJMLX_HIP_GLOBAL_WORKSPACE")
    sim = capi.BatchedSim(model, B)
    for s in (ref, sim):
        s.set_state(q, v)
    return ref, sim


def test_noalloc_step_matches_step_n(monkeypatch):
    if capi.lib().mjmlx_hip_device_count() < 1:
        pytest.fail("no HIP device visible to the engine")
    L = _bind()
    cap = L.mjmlx_batched_noalloc_scratch_bytes()
    assert cap == 64 << 20
    model = capi.Model(scenes.quadruped_xml(self_collision=True))
    rng = np.random.default_rng(5)
    for B, nsub, steps in ((1024, 4, 3), (96, 2, 2)):  # the second call's sim is smaller
        ref, sim = _sims(model, B, B, monkeypatch)
        ws = sim.capacity()[2]
        assert B * ws <= cap
        for _ in range(steps):
            c = scenes.quadruped_actions(B, rng).astype(np.float32)
            ref.step(c, nsub)
            assert L.mjmlx_batched_step_n_noalloc(sim.h, capi._fp(c), nsub) == 0
            (q1, v1), (q2, v2) = ref.state(), sim.state()
            assert np.array_equal(q1, q2) and np.array_equal(v1, v2)
            s1, s2 = ref.stats(), sim.stats()
            assert np.array_equal(s1["niter"], s2["niter"]) and np.array_equal(s1["ncon"], s2["ncon"])
            assert not (s2["status"] & ~3).any()
        del ref, sim


def test_noalloc_rejects_oversize_and_learned_modes():
    L = _bind()
    cap = L.mjmlx_batched_noalloc_scratch_bytes()
    model = capi.Model(scenes.quadruped_xml(self_collision=True))
    small = capi.BatchedSim(model, 8)
    ws = small.capacity()[2]
    B = cap // ws + 1
    big = capi.BatchedSim(model, B)
    q0, v0 = big.state()
    assert L.mjmlx_batched_step_n_noalloc(big.h, None, 1) == -1
    q1, v1 = big.state()
    assert np.array_equal(q0, q1) and np.array_equal(v0, v1)  # nothing ran
    net = capi.MLP(small.feature_dim(), 32, model.nv)
    small.set_warmstart_mlp(net)
    assert small.set_warmstart_mode(2) == 0
    assert L.mjmlx_batched_step_n_noalloc(small.h, None, 1) == -1
    assert small.set_warmstart_mode(0) == 0
    assert L.mjmlx_batched_step_n_noalloc(small.h, None, 1) == 0
    assert L.mjmlx_batched_step_n_noalloc(None, None, 1) == -1
