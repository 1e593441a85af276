"""The reference's own test models (tests/golden/models/, copied from the
reference's tests/models/ as fixtures) through the kernel source against the
FP64 oracle: contact_stress.xml (50 free spheres, nv 300), high_dof_tree.xml
(free joint + 48 limited hinges, nv 54, 48 motors), stiff_springs.xml (slide
joints with stiff springs under implicitfast) and multi_geom_scene.xml
(spheres, capsules, boxes, cylinders and ellipsoids, nv 66).

States come from oracle rollouts. One step from each state: contact pair
lists identical, qacc within 1e-4 of the balanced accelerations and
efc_force within 1e-4 of the largest row force; where the GJK/EPA pairs meet
on flat patches (multi_geom_scene) the constraint solve is compared on the
kernel's own contact set (depths and normals checked against the oracle's).
The CPU interpreter runs them here; the HIP engine on the GPU.
"""
from __future__ import annotations

import os

import numpy as np
import pytest

import emu
import ref64
import states

HERE = os.path.dirname(os.path.abspath(__file__))
MODELS = ["contact_stress", "high_dof_tree", "stiff_springs", "multi_geom_scene"]
# rollout before the compared step: (control steps, substeps each, envs)
ROLL = {"contact_stress": (90, 4, 4), "high_dof_tree": (40, 4, 6), "stiff_springs": (20, 4, 6),
        "multi_geom_scene": (40, 4, 6)}


def model_xml(name: str) -> str:
    with open(os.path.join(HERE, "golden", "models", name + ".xml")) as f:
        return f.read()


def model_states(field, name: str, seed: int):
    """Oracle rollout from qpos0 (velocities U(+-0.2), motors U(+-0.5) per
    step); returns float32 qpos, qvel, qacc_ws, ctrl of the compared step."""
    nsteps, nsub, B = ROLL[name]
    orc = ref64.Oracle(ref64.fields_from(field))
    rng = np.random.default_rng(seed)
    q = np.tile(field("qpos0"), (B, 1))
    qv = rng.uniform(-0.2, 0.2, (B, orc.nv))
    qa = np.zeros((B, orc.nv))
    lens = rng.integers(nsteps // 2, nsteps + 1, size=B)
    for k in range(int(lens.max())):
        idx = np.nonzero(lens > k)[0]
        ctrl = rng.uniform(-0.5, 0.5, (len(idx), orc.nu))
        qs, vs, as_ = q[idx].copy(), qv[idx].copy(), qa[idx].copy()
        orc.batch_step(qs, vs, as_, ctrl, nsub, 8)
        q[idx], qv[idx], qa[idx] = qs, vs, as_
    f32 = lambda a: np.ascontiguousarray(a, dtype=np.float32)
    return f32(q), f32(qv), f32(qa), f32(rng.uniform(-0.5, 0.5, (B, orc.nu))), orc


def compare_step(name, orc, q, qv, qa, ctrl, ncon_k, nefc_k, status_k, getter):
    """getter(which, env, field) reads the kernel's dump. Returns the worst
    (qacc, efc_force) errors and the contact count."""
    assert not (status_k & ~3).any(), status_k
    flat = name == "multi_geom_scene"
    wq = wf = 0.0
    ncon = 0
    for e in range(q.shape[0]):
        states.oracle_step(orc, q[e], qv[e], qa[e], ctrl[e])
        con = getter(1, e, "contact").reshape(-1, 16)[: ncon_k[e]]
        ok, msg = states.compare_contacts(con[:, 13].view(np.int32), con[:, 12], orc.contacts(), margin_tol=1e-4,
                                          ncon=ncon_k[e])
        assert ok, f"{name} env {e}: {msg}"
        ncon += len(con)
        if flat and len(con):
            oc = sorted((c["pair"], c["dist"]) for c in orc.contacts())
            kc = sorted((int(c[13:14].view(np.int32)[0]), float(c[12])) for c in con)
            assert all(p1 == p2 and abs(d1 - d2) < 1e-5 for (p1, d1), (p2, d2) in zip(oc, kc))
            c64 = con.astype(np.float64)
            orc.inject_contacts(con[:, 13].view(np.int32), c64[:, 12], c64[:, 0:3], c64[:, 3:6])
            states.oracle_step(orc, q[e], qv[e], qa[e], ctrl[e])
        assert orc.nefc() == nefc_k[e], (name, e, orc.nefc(), nefc_k[e])
        eq, ef = states.qacc_force_errors(orc, getter(2, e, "qacc"), getter(2, e, "efc_force"), nefc_k[e])
        wq, wf = max(wq, eq), max(wf, ef or 0.0)
        orc.inject_contacts()
    return wq, wf, ncon


def caps_for(field, maxcon):
    """The engine's capacities for the oracle: scenes of several free bodies
    (or more than 32 dofs) run without a row budget."""
    nfree = int(np.sum(field("jnt_type") == 0))
    if int(field("nv")[0]) > 32 or nfree > 1:
        return dict(con_cap=maxcon, efc_cap=0)
    return {}


@pytest.mark.parametrize("name", MODELS)
def test_reference_model_interpreter(name):
    xml = model_xml(name)
    B = ROLL[name][2]
    sim = emu.EmuSim(xml, B)
    q, qv, qa, ctrl, orc = model_states(sim.field, name, 11)
    orc.set_caps(**caps_for(sim.field, sim.layout("contact")[1] // 16))
    sim.set_state(q, qv, qa)
    sim.step(ctrl if sim.nu else None, 1, dump=True)
    st = sim.stats()
    wq, wf, ncon = compare_step(name, orc, q, qv, qa, ctrl, st["ncon"], st["nefc"], st["status"], sim.dump)
    print(f"{name}: contacts {ncon}, worst qacc rel err {wq:.2e}, efc_force {wf:.2e}")
    assert wq < 1e-4 and wf < 1e-4
    if name != "stiff_springs":
        assert ncon > 0


@pytest.mark.gpu
def test_reference_models_gpu():
    from mjmlx_hip import capi
    if capi.lib().mjmlx_hip_device_count() < 1:
        pytest.fail("no HIP device visible to the engine")
    for name in MODELS:
        xml = model_xml(name)
        model = capi.Model(xml)
        B = ROLL[name][2]
        q, qv, qa, ctrl, orc = model_states(model.field, name, 11)
        sim = capi.BatchedSim(model, B)
        orc.set_caps(**caps_for(model.field, sim.capacity()[0]))
        sim.set_state(q, qv)
        sim.set_qacc_warmstart(qa)
        dump = sim.debug_step(ctrl if model.nu else None)
        st = sim.stats()
        wq, wf, ncon = compare_step(name, orc, q, qv, qa, ctrl, st["ncon"], st["nefc"], st["status"], dump.get)
        print(f"{name}: contacts {ncon}, worst qacc rel err {wq:.2e}, efc_force {wf:.2e}")
        assert wq < 1e-4 and wf < 1e-4
        # a 250-step rollout (1000 substeps) stays finite
        roll = capi.BatchedSim(model, 64)
        rng = np.random.default_rng(1)
        for _ in range(250):
            roll.step(rng.uniform(-0.5, 0.5, (64, model.nu)).astype(np.float32) if model.nu else None, 4)
        qr, vr = roll.state()
        assert np.isfinite(qr).all() and np.isfinite(vr).all(), name
        assert not (roll.stats()["status"] & ~3).any(), name
