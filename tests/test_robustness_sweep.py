"""Robustness sweep of the batched step over hinge chains (the reference's
tests/test_robustness.cpp:32-55,128-161, restated): chains of nv hinges
(nv from 1 to 33, including sizes that are not multiples of 8 or 32) under
four constraint configurations - free (no constraint), limit (joint limits,
started beyond them), contact (the chain's tip starts on the floor) and
floss (dof friction loss).

Per case: states finite and bounded; duplicated envs bit-identical; two runs
bit-identical; an env stepped alone (B = 1) identical to the same env in the
batch; qpos after the steps within the reference's 5e-3 of the FP64 oracle.
The interpreter of the kernel source runs the sweep on the CPU, the HIP
engine through the C ABI on the GPU. (The reference's fifth configuration,
a fixed tendon, is outside this engine: tendons are not built.)
"""
from __future__ import annotations

import math

import numpy as np
import pytest

import emu
import ref64

B = 8          # the second half duplicates the first
STEPS = 10
NVS = [1, 3, 4, 7, 8, 9, 16, 17, 31, 32, 33]
CFGS = ["free", "limit", "contact", "floss"]


def chain_xml(nv: int, cfg: str) -> str:
    """nv hinge links (capsules 0.1 m long, r 0.02, density 500, axes
    alternating y / x, damping 0.02) hanging from height nv * 0.1 + 0.05."""
    ln = 0.1
    top = nv * ln + 0.05
    out = ['<mujoco><compiler angle="radian"/>',
           '<option timestep="0.002" gravity="0 0 -9.81" solver="Newton" iterations="50"/><worldbody>']
    if cfg == "free":
        out.append('<geom type="plane" size="5 5 0.1" contype="0" conaffinity="0"/>')
    elif cfg == "contact":
        # the tip of the chain swung 1.2 rad about the first hinge touches the floor
        out.append(f'<geom type="plane" size="5 5 0.1" pos="0 0 {top - nv * ln * math.cos(1.2) - 0.004:.6f}"/>')
    else:
        out.append('<geom type="plane" size="5 5 0.1" pos="0 0 -50"/>')
    extra = {"limit": ' limited="true" range="-0.35 0.35"', "floss": ' frictionloss="0.05"'}.get(cfg, "")
    nocol = ' contype="0" conaffinity="0"' if cfg == "free" else ""
    for i in range(nv):
        axis = "1 0 0" if i % 2 else "0 1 0"
        out.append(f'<body pos="0 0 {top if i == 0 else -ln}"><joint name="j{i}" type="hinge" axis="{axis}" '
                   f'damping="0.02"{extra}/><geom type="capsule" fromto="0 0 0 0 0 {-ln}" size="0.02" '
                   f'density="500"{nocol}/>')
    out.append("</body>" * nv)
    out.append("</worldbody></mujoco>")
    return "".join(out)


def chain_states(nv: int, cfg: str, seed: int):
    """First half random (angles beyond the limits included), second half
    copies; contact: the first hinge at 1.2 rad, the rest nearly straight."""
    rng = np.random.default_rng(seed)
    q = np.zeros((B, nv), np.float32)
    v = np.zeros((B, nv), np.float32)
    h = B // 2
    q[:h] = 0.5 * rng.uniform(-1, 1, (h, nv)) * (0.01 if cfg == "contact" else 1.0)
    if cfg == "contact":
        q[:h, 0] = 1.2
    v[:h] = 0.5 * rng.uniform(-1, 1, (h, nv))
    q[h:], v[h:] = q[:h], v[:h]
    return q, v


def oracle_rollout(xml_field, q0, v0, steps):
    orc = ref64.Oracle(ref64.fields_from(xml_field))
    q, v = q0.astype(np.float64).copy(), v0.astype(np.float64).copy()
    a = np.zeros_like(v)
    orc.batch_step(q, v, a, None, steps)
    return q, v


def check_runs(runs, single, e_single, q_orc):
    """runs: two (q, v) batch results of the same inputs; single: (q, v) of env
    e_single stepped alone; q_orc: the oracle's qpos. Returns max |dq| vs the oracle."""
    (q, v), (q2, v2) = runs
    assert np.isfinite(q).all() and np.isfinite(v).all()
    assert np.abs(q).max() < 1e3 and np.abs(v).max() < 1e3
    h = B // 2
    assert np.array_equal(q[:h], q[h:]) and np.array_equal(v[:h], v[h:])      # duplicated envs
    assert np.array_equal(q, q2) and np.array_equal(v, v2)                      # run to run
    assert np.array_equal(single[0][0], q[e_single]) and np.array_equal(single[1][0], v[e_single])  # B = 1
    return float(np.abs(q - q_orc).max())


@pytest.mark.parametrize("cfg", CFGS)
@pytest.mark.parametrize("nv", NVS)
def test_chain_sweep_interpreter(nv, cfg):
    xml = chain_xml(nv, cfg)
    q0, v0 = chain_states(nv, cfg, 1234 + nv)

    def run(qq, vv):
        sim = emu.EmuSim(xml, qq.shape[0])
        sim.set_state(qq, vv, np.zeros_like(vv))
        for _ in range(STEPS):
            sim.step(None, 1)
        st = sim.stats()
        assert not (st["status"] & ~3).any(), st["status"]
        qq2, vv2, _ = sim.get_state()
        return qq2, vv2

    runs = [run(q0, v0), run(q0, v0)]
    e = B // 2 + 1
    single = run(q0[e:e + 1], v0[e:e + 1])
    cm = emu.CompiledModel(xml)
    q_orc, _ = oracle_rollout(cm.field, q0, v0, STEPS)
    dq = check_runs(runs, single, e, q_orc)
    assert dq < 5e-3, dq


@pytest.mark.gpu
def test_chain_sweep_gpu():
    """Every (nv, configuration) case on the HIP engine through the C ABI."""
    from mjmlx_hip import capi
    if capi.lib().mjmlx_hip_device_count() < 1:
        pytest.fail("no HIP device visible to the engine")
    worst = 0.0
    for cfg in CFGS:
        for nv in NVS:
            xml = chain_xml(nv, cfg)
            model = capi.Model(xml)
            q0, v0 = chain_states(nv, cfg, 1234 + nv)

            def run(qq, vv):
                sim = capi.BatchedSim(model, qq.shape[0])
                sim.set_state(qq, vv)
                for _ in range(STEPS):
                    sim.step(None, 1)
                st = sim.stats()
                assert not (st["status"] & ~3).any(), (cfg, nv, st["status"])
                return sim.state()

            runs = [run(q0, v0), run(q0, v0)]
            e = B // 2 + 1
            single = run(q0[e:e + 1], v0[e:e + 1])
            q_orc, _ = oracle_rollout(model.field, q0, v0, STEPS)
            dq = check_runs(runs, single, e, q_orc)
            assert dq < 5e-3, (cfg, nv, dq)
            worst = max(worst, dq)
    print(f"chain sweep: {len(CFGS) * len(NVS)} cases, worst |qpos - oracle| after {STEPS} steps {worst:.2e}")
