"""Benchmark: physics env-steps/s of the batched quadruped step on one MI355X.

Workload: procedurally written quadruped (free-joint torso + 12 hinges,
capsule legs, sphere feet, self-collision on, condim 3 pyramidal, position
actuators kp 20 / kv 0.5), Newton solver, 16 iterations, tolerance 1e-8,
dt 0.005. One env-step = one control step = 4 physics substeps with the same
ctrl (the reference's physics-only unit, reference README.md:50). Actions are
nominal + U(+-0.5 rad), synthetic, from a pre-generated pool.

Two warm starts are measured from the same state: the plain previous-step
warm start (one fused launch per control step) and the learned one (an MLP
trained here, on the GPU, on ~2M (features, converged qacc) pairs recorded
from the plain solver; one launch per substep plus the network). Both iterate
to the same 1e-8 scaled residual.

`value` is the faster of the two, from HIP-event time per control step with
the actions already in device memory; the wall-clock rate (which also pays
the per-step host-to-device copy of the actions) is reported beside it.

Prints one JSON line with both throughputs, the roofline of the step kernel
(HIP-event kernel time over the timed region) and an FP64 CPU baseline (the
project's oracle on the host cores, bounded sample).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "mujoco-mlx-cpp_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

HBM_PEAK_GBS = 8000.0


def bytes_per_env_step(nq: int, nv: int, nu: int) -> int:
    # read qpos, qvel, ctrl, qacc_warmstart; write qpos, qvel, qacc_warmstart (FP32)
    return 4 * (2 * nq + 4 * nv + nu)


def cpu_baseline_leg(model_nosc_fields, model_fields, check):
    """The CPU leg, where the FP64 oracle (oracle/ref64.c) runs, as baseline
    and as checker:

    * baseline: the CPU workload of the benchmark configurations as written
      (the quadruped without self-collision, 64 envs, seed 0: nominal pose +
      U(+-0.05), actions nominal + U(+-0.5); 1000 control steps of 4 substeps
      after 25 warm-up steps; min(threads, 64 envs) threads spawned per call,
      as reference src/mjb.cpp:826-838 does; median of per-step wall times, as
      reference scripts/bench_physics_vs_others.py:55-65). Threads are capped
      at 16, the host share of one GPU on the measurement machine.
    * check: the FP64 oracle's qacc / efc_force against the kernel's on envs
      sampled from the timed state (one substep from that state, outside the
      timed region); relative to the balanced accelerations max(1, |qacc|,
      |qacc_smooth|) and to the largest row force."""
    import ref64
    from mjmlx_hip import scenes

    nproc = os.cpu_count() or 1
    B = 64
    nthreads = max(1, min(16, nproc, B))
    orc = ref64.Oracle(model_nosc_fields)
    rng = np.random.default_rng(0)
    q = scenes.quadruped_initial_qpos(B, 0)
    qv = np.zeros((B, orc.nv))
    qa = np.zeros((B, orc.nv))
    for _ in range(25):
        orc.batch_step(q, qv, qa, scenes.quadruped_actions(B, rng), 4, nthreads)
    ctrls = [np.ascontiguousarray(scenes.quadruped_actions(B, rng)) for _ in range(1000)]
    per = np.zeros(1000)
    t_all = time.perf_counter()
    for k in range(1000):
        t0 = time.perf_counter()
        orc.batch_step(q, qv, qa, ctrls[k], 4, nthreads)
        per[k] = time.perf_counter() - t0
    t_all = time.perf_counter() - t_all
    med = float(np.median(per))
    base = {"value": B / med, "unit": "env-steps/s", "cores": nthreads, "nproc": nproc, "kind": "port",
            "mean_value": B * 1000 / t_all,
            "sample": f"own FP64 restatement (oracle/ref64.c; MuJoCo C is not available): quadruped without "
                      f"self-collision, {B} envs x 1000 control steps (4 substeps each, Newton 16 it, tol 1e-8) "
                      f"after 25 warm-up steps, {nthreads} threads per call, median step {med * 1e3:.3f} ms, "
                      f"{t_all:.2f} s in all"}
    chk = None
    if check is not None:
        sim, q32, v32, a32, c32 = check
        dump = sim.debug_step(c32)
        st = sim.stats()
        o = ref64.Oracle(model_fields)
        wq = wf = 0.0
        ncon = 0
        for e in range(q32.shape[0]):
            o.set_state(q32[e].astype(np.float64), v32[e].astype(np.float64), a32[e].astype(np.float64),
                        c32[e].astype(np.float64))
            o.forward()
            b = o.get("qacc")
            sc = max(1.0, float(np.max(np.abs(b))), float(np.max(np.abs(o.get("qacc_smooth")))))
            wq = max(wq, float(np.max(np.abs(dump.get(2, e, "qacc") - b))) / sc)
            n = o.nefc()
            if n and int(st["nefc"][e]) == n:
                f = o.get("efc_force")
                wf = max(wf, float(np.max(np.abs(dump.get(2, e, "efc_force")[:n] - f))) / max(1.0, float(np.max(np.abs(f)))))
            ncon += len(o.contacts())
        chk = {"envs": int(q32.shape[0]), "contacts": ncon, "max_qacc_rel_err": wq, "max_efc_force_rel_err": wf,
               "against": "FP64 oracle (oracle/ref64.c), one substep from the timed state"}
    return base, chk


def timed_run(sim, pool, warmup, steps, offset=0):
    """Warm-up then `steps` control steps; returns (event ms per step, wall s, mean iterations per env-step)."""
    for k in range(warmup):
        sim.step(pool[(offset + k) % len(pool)], 4)
    sim.sync()
    kernel_ms, iters = [], []
    t0 = time.perf_counter()
    for k in range(steps):
        sim.step(pool[(offset + warmup + k) % len(pool)], 4)
        kernel_ms.append(sim.last_step_ms())  # waits for the step
        if k % 10 == 0:
            iters.append(float(sim.stats()["niter"].mean()))
    sim.sync()
    dt = time.perf_counter() - t0
    return float(np.mean(kernel_ms)), dt, float(np.mean(iters))


def run_boxes(B, warmup, steps):
    """Uneven-contact batching: every env is a plane plus N in U{4..32} free
    boxes (half-sizes U(0.02, 0.1), densities U(500, 2000), friction
    U(0.3, 1.0), drop heights U(0.1, 1.0) above the box's bounding radius,
    random orientations), condim 3 pyramidal, CG 16 iterations. Envs with the
    same N share one model (one sim and stream per N); the groups' launches
    are queued together, so value = all envs x steps / wall time of the timed
    steps with everything resident."""
    from mjmlx_hip import capi, scenes

    if capi.lib().mjmlx_hip_device_count() < 1:
        raise SystemExit("bench: no HIP device visible")
    rng = np.random.default_rng(3)
    ns = rng.integers(4, 33, size=B)
    groups = []
    for n in range(4, 33):
        nb = int((ns == n).sum())
        if nb == 0:
            continue
        model = capi.Model(scenes.boxes_xml(n, 100 + n, solver="CG", iterations=16))
        q = np.tile(model.field("qpos0"), (nb, 1))
        hs = model.field("geom_size").reshape(-1, 3)[1:]
        for b in range(n):
            a = 7 * b
            q[:, a] += rng.uniform(-0.05, 0.05, nb)
            q[:, a + 1] += rng.uniform(-0.05, 0.05, nb)
            q[:, a + 2] = rng.uniform(0.1, 1.0, nb) + np.linalg.norm(hs[b])
            qq = rng.normal(size=(nb, 4))
            q[:, a + 3:a + 7] = qq / np.linalg.norm(qq, axis=1, keepdims=True)
        sim = capi.BatchedSim(model, nb)
        sim.set_state(q.astype(np.float32), np.zeros((nb, model.nv), np.float32))
        groups.append((n, nb, model, sim))
    # the boxes land within the warm-up; the autotuner then picks each group's
    # kernel variant (workspace in LDS or device memory, register budget)
    # from its landed state
    for k in range(warmup):
        for _, _, _, sim in groups:
            sim.step(None, 4)
    placement = {}
    for n, _, _, sim in groups:
        placement[n] = ("LDS", "device memory", "LDS, 168 VGPRs", "LDS, 512 registers",
                        "device memory, 168 VGPRs", "LDS, 32 lanes per env",
                        "LDS, 32 lanes per env, 168 VGPRs")[sim.autotune(2)]
    for k in range(2):
        for _, _, _, sim in groups:
            sim.step(None, 4)
    for _, _, _, sim in groups:
        sim.sync()
    t0 = time.perf_counter()
    for k in range(steps):
        for _, _, _, sim in groups:
            sim.step(None, 4)
    for _, _, _, sim in groups:
        sim.sync()
    wall = time.perf_counter() - t0
    # per-group kernel time and statistics, each group timed alone
    per = []
    for n, nb, model, sim in groups:
        ms = []
        for k in range(5):
            sim.step(None, 4)
            ms.append(sim.last_step_ms())
        st = sim.stats()
        per.append({"n_boxes": n, "envs": nb, "nv": model.nv, "ms_per_step_alone": float(np.mean(ms)),
                    "workspace": placement[n],
                    "lds_bytes_per_env": sim.capacity()[2], "mean_ncon": float(st["ncon"].mean()),
                    "mean_nefc": float(st["nefc"].mean()), "row_budget_overflow_envs": int(((st["status"] & 2) > 0).sum()),
                    "mean_iters": float(st["niter"].mean())})
    value = B * steps / wall
    bpe = float(np.mean([bytes_per_env_step(7 * n, 6 * n, 0) for n in ns]))
    sum_ms = float(sum(g["ms_per_step_alone"] for g in per))
    traffic = _traffic("boxes", B)
    out = {
        "metric": "physics env-steps/s @16384 envs, plane + U{4..32} boxes",
        "value": value, "unit": "env-steps/s", "n_gpus": 1, "steps": steps, "warmup": warmup,
        "ms_per_step": wall / steps * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": "plane + N in U{4..32} free boxes, condim 3 pyramidal, CG 16 it, tol 1e-8, "
                               "dt 0.005 x 4 substeps; no row budget (up to 512 contacts per env, all with rows)",
                   "num_envs": B, "substeps_per_env_step": 4, "groups": len(groups),
                   "parallelism": "single GPU, 1 wavefront per env, one sim (stream) per N",
                   "hw_queues": int(os.environ.get("GPU_MAX_HW_QUEUES", "4")),
                   "workspace": "LDS or device memory per group, picked by the autotuner (mjmlx_batched_autotune)"},
        # the groups' launches overlap on their streams: the control step's
        # time is the wall time per step of the queued groups
        "roofline": {"bound": "hbm", "achieved": B * bpe / (wall / steps) / 1e9, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": B * bpe / (wall / steps) / 1e9 / HBM_PEAK_GBS,
                     "traffic": traffic, "traffic_unit": "HBM bytes per control step, all groups",
                     "kernel": "k_step (one launch per group)", "kernel_ms_sum_alone": sum_ms,
                     "bytes_per_env_step": bpe},
        "groups": per,
        "cpu_baseline": None,
    }
    return out


STAGES = [("kin", "kinematics"), ("compos", "subtree com"), ("crb", "CRB"), ("factor", "LDL of qM"),
          ("comvel", "com velocities"), ("smooth", "smooth forces (RNE, passive, actuation)"),
          ("coll", "collision"), ("constr", "constraint rows"), ("solve_init", "solver set-up"),
          ("ls", "line search"), ("costgrad", "cost / gradient"), ("dir", "search direction (LDL^T of H; CG: M^-1 solve)"),
          ("solve_rest", "other solver work"), ("integrate", "integration"), ("io", "state in/out"),
          ("hbuild", "Newton Hessian build"), ("nsolve", "Newton direction substitution")]


def run_stages(args):
    """Per-stage cycle shares from the profiling build (libmjmlx_prof.so:
    s_memtime counters around each stage of the step kernel), one 4-substep
    control step of every env after the warm-up, mean per env."""
    import ctypes as C
    from mjmlx_hip import capi, scenes
    capi.LIBPATH = os.path.join(capi.LIBDIR, "libmjmlx_prof.so")
    L = capi.lib()
    L.mjmlx_batched_profile_step.argtypes = [C.c_void_p, C.POINTER(C.c_float), C.c_int, C.POINTER(C.c_ulonglong)]
    rng = np.random.default_rng(1)
    B = args.envs
    if args.scene == "boxes":
        model = capi.Model(scenes.boxes_xml(args.nbox, 100 + args.nbox, solver="CG", iterations=16))
        pool = [None]
        q0 = np.tile(model.field("qpos0"), (B, 1)).astype(np.float32)
    else:
        model, pool, q0 = scene_setup(args.scene, B, rng)
    sim = capi.BatchedSim(model, B)
    sim.set_state(q0, np.zeros((B, model.nv), np.float32))
    for k in range(args.warmup):
        sim.step(pool[k % len(pool)], 4)
    sim.sync()
    out = np.zeros((B, len(STAGES)), np.uint64)
    c = pool[0]
    cp = None if c is None else c.ctypes.data_as(C.POINTER(C.c_float))
    assert L.mjmlx_batched_profile_step(sim.h, cp, 4, out.ctypes.data_as(C.POINTER(C.c_ulonglong))) == 0
    st = sim.stats()
    tot = out.sum(1).astype(np.float64)
    name = args.scene if args.scene != "boxes" else f"boxes N={args.nbox}"
    print(f"{name}, {B} envs: mean solver iterations per control step {st['niter'].mean():.2f}, "
          f"mean contacts {st['ncon'].mean():.1f}, rows {st['nefc'].mean():.1f}, "
          f"{tot.mean():.0f} cycles per env-step")
    frac = out.sum(0) / max(1, out.sum())
    print("| stage | share | cycles per substep |\n|---|---|---|")
    for i in np.argsort(-frac):
        if frac[i] > 0.001:
            print(f"| {STAGES[i][1]} | {frac[i] * 100:.1f}% | {out[:, i].mean() / 4:.0f} |")


def copy_calibration():
    """A plain device copy of known size (3 x 256 MiB, fp32) whose FETCH_SIZE /
    WRITE_SIZE counters calibrate those of the step kernel."""
    import torch
    n = 256 * 1024 * 1024 // 4
    a = torch.ones(n, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print("copied 3 x 256 MiB", flush=True)


SCENES = {
    "quadruped": "quadruped nq19 nv18 nu12, self-collision, condim 3 pyramidal, Newton 16 it, tol 1e-8, "
                 "dt 0.005 x 4 substeps",
    "humanoid": "humanoid nq28 nv27 nu21 (free joint + 21 hinges), self-collision, condim 6 elliptic, Newton 16 it, "
                "tol 1e-8, dt 0.005 x 4 substeps, motor ctrl U(+-0.4)",
    "boxes": "plane + N in U{4..32} free boxes, CG 16 it",
}


def scene_setup(name, B, rng):
    """model, action pool and initial state of a bench scene"""
    from mjmlx_hip import capi, scenes
    if name == "quadruped":
        model = capi.Model(scenes.quadruped_xml(self_collision=True))
        pool = [scenes.quadruped_actions(B, rng).astype(np.float32) for _ in range(16)]
        q0 = scenes.quadruped_initial_qpos(B, 0).astype(np.float32)
    else:
        model = capi.Model(scenes.humanoid_xml())
        pool = [rng.uniform(-0.4, 0.4, (B, model.nu)).astype(np.float32) for _ in range(16)]
        q0 = np.tile(model.field("qpos0"), (B, 1)).astype(np.float32)
    return model, pool, q0


def _traffic(scene, B):
    """Calibrated HBM bytes per launch (per control step for the boxes'
    groups) from the committed PMC summary of that workload, or None."""
    tpath = os.path.join(ROOT, "profiles", f"traffic_{scene}{B}.txt")
    if os.path.exists(tpath):
        with open(tpath) as f:
            return json.load(f).get("bytes_per_launch")
    return None


def _roofline(B, bpe, ms, traffic, kernel):
    achieved = B * bpe / (ms / 1e3) / 1e9
    return {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "kernel": kernel, "kernel_ms_avg": ms,
            "bytes_per_env_step": bpe}


def run_plain_scene(scene, B, warmup, steps):
    """Plain warm start on one scene: HIP-event kernel time per control step
    after `warmup` steps, with its roofline (sub-result of the default line)."""
    from mjmlx_hip import capi
    rng = np.random.default_rng(1)
    model, pool, q0 = scene_setup(scene, B, rng)
    sim = capi.BatchedSim(model, B)
    sim.set_state(q0, np.zeros((B, model.nv), np.float32))
    ms, wall, it = timed_run(sim, pool, warmup, steps)
    st = sim.stats()
    bpe = bytes_per_env_step(model.nq, model.nv, model.nu)
    return {"workload": SCENES[scene], "num_envs": B, "warmup": warmup, "steps": steps,
            "env_steps_per_s": B / ms * 1e3, "ms_per_step": ms, "wall_env_steps_per_s": B * steps / wall,
            "mean_solver_iters_per_env_step": it, "overflow_envs": int(((st["status"] & 3) > 0).sum()),
            "roofline": _roofline(B, bpe, ms, _traffic(scene, B), "k_step (fused 4-substep step)")}


def run(args):
    from mjmlx_hip import capi

    if capi.lib().mjmlx_hip_device_count() < 1:
        raise SystemExit("bench: no HIP device visible")
    B = args.envs
    rng = np.random.default_rng(1)
    model, pool, q0 = scene_setup(args.scene, B, rng)
    src = capi.BatchedSim(model, B)
    src.set_state(q0, np.zeros((B, model.nv), np.float32))
    for k in range(args.warmup):
        src.step(pool[k % len(pool)], 4)
    learned = None
    if not args.no_learned:
        # training data: (features, converged qacc) of every env and substep of plain rollouts
        net = capi.MLP(src.feature_dim(), 256, model.nv, seed=0)
        nrec = max(1, -(-args.pairs // (4 * B)))
        net.reserve(4 * B * nrec)
        t0 = time.perf_counter()
        rows = 0
        for k in range(nrec):
            rows += src.record(net, pool[(args.warmup + k) % len(pool)], 4)
        src.sync()
        t_rec = time.perf_counter() - t0
        t0 = time.perf_counter()
        net.fit_normalization()
        loss = net.fit(args.epochs, 4096, lr=1e-3, wd=1e-4, seed=0)
        t_fit = time.perf_counter() - t0
        learned = {"pairs": rows, "record_s": t_rec, "train_s": t_fit, "epochs": args.epochs, "batch": 4096,
                   "optimizer": "AdamW lr 1e-3 wd 1e-4", "final_loss_normalised": loss,
                   "network": f"{src.feature_dim()}-256-256-{model.nv} SiLU, BF16 MFMA, FP32 accumulate"}
    q, qv = src.state()
    qa = src.qacc_warmstart()
    results = {}
    check = None
    preroll = None
    for mode in ("plain", "learned", "learned_first"):
        if mode != "plain" and learned is None:
            continue
        sim = capi.BatchedSim(model, B)
        if mode == "learned":
            sim.set_warmstart_mlp(net)
            assert sim.set_warmstart_mode(2) == 0
            sim.set_graph(True)  # the per-substep launches replayed as one captured graph
        elif mode == "learned_first":
            # one network pass per control step, one fused 4-substep launch
            sim.set_warmstart_mlp(net)
            assert sim.set_warmstart_mode(3) == 0
        sim.set_state(q, qv)
        sim.set_qacc_warmstart(qa)
        ms, wall, it = timed_run(sim, pool, 5, args.steps, offset=3)
        st = sim.stats()
        results[mode] = {"env_steps_per_s": B / ms * 1e3, "ms_per_step": ms,
                         "wall_env_steps_per_s": B * args.steps / wall,
                         "mean_solver_iters_per_env_step": it,
                         "overflow_envs": int(((st["status"] & 3) > 0).sum())}
        if mode == "plain":
            # envs sampled from the timed state for the oracle check (run in the CPU leg)
            idx = np.linspace(0, B - 1, 64).astype(int)
            qs, vs = sim.state()
            as_ = sim.qacc_warmstart()
            cs = pool[(3 + 5 + args.steps) % len(pool)]
            csim = capi.BatchedSim(model, len(idx))
            q32, v32, a32, c32 = (np.ascontiguousarray(x[idx]) for x in (qs, vs, as_, cs))
            csim.set_state(q32, v32)
            csim.set_qacc_warmstart(a32)
            check = (csim, q32, v32, a32, c32)
            # a second window after a 200-step untimed pre-roll (the robots fall and settle)
            nroll = 200
            for k in range(nroll):
                sim.step(pool[k % len(pool)], 4)
            ms2, wall2, it2 = timed_run(sim, pool, 0, args.steps, offset=7)
            st2 = sim.stats()
            preroll = {"preroll_steps": 5 + args.steps + nroll, "env_steps_per_s": B / ms2 * 1e3,
                       "ms_per_step": ms2, "wall_env_steps_per_s": B * args.steps / wall2,
                       "mean_solver_iters_per_env_step": it2,
                       "overflow_envs": int(((st2["status"] & 3) > 0).sum())}
        del sim
    plain = results["plain"]
    plain["after_preroll"] = preroll
    if learned is not None:
        learned.update(results["learned"])
        learned["speedup_vs_plain"] = learned["env_steps_per_s"] / plain["env_steps_per_s"]
        learned["iteration_ratio_plain_over_learned"] = (plain["mean_solver_iters_per_env_step"]
                                                         / learned["mean_solver_iters_per_env_step"])
        first = results["learned_first"]
        learned["first_substep_only"] = dict(first, speedup_vs_plain=first["env_steps_per_s"] / plain["env_steps_per_s"],
                                             iteration_ratio_plain_over_learned=(
                                                 plain["mean_solver_iters_per_env_step"]
                                                 / first["mean_solver_iters_per_env_step"]))
    best = max(results, key=lambda k: results[k]["env_steps_per_s"])
    head = results[best]
    bpe = bytes_per_env_step(model.nq, model.nv, model.nu)
    out = {
        "metric": "physics env-steps/s @16384 envs, plain vs learned warm-start at 1e-8 residual",
        "value": head["env_steps_per_s"],
        "unit": "env-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (FP64 Hessian Cholesky and free-joint quaternion integration)",
        "data": "synthetic",
        "config": {"workload": SCENES[args.scene],
                   "num_envs": B, "substeps_per_env_step": 4, "parallelism": "single GPU, 1 wavefront per env",
                   "warm_start": best, "launch": "one workgroup per env"},
        "plain": plain,
        "learned": learned,
        # roofline of the fused step kernel (plain path: one k_step launch per control step)
        "roofline": _roofline(B, bpe, plain["ms_per_step"], _traffic(args.scene, B),
                              "k_step (fused 4-substep step; k_step_occ, its 168-VGPR variant, for small workspaces)"),
        "cpu_baseline": None,
        "oracle_check": None,
    }
    if not args.no_cpu and args.scene == "quadruped":
        import ref64
        from mjmlx_hip import scenes
        nosc = capi.Model(scenes.quadruped_xml(self_collision=False))
        out["cpu_baseline"], out["oracle_check"] = cpu_baseline_leg(ref64.fields_from(nosc.field),
                                                                    ref64.fields_from(model.field), check)
    if args.scene == "quadruped" and not args.headline_only:
        out["humanoid"] = run_plain_scene("humanoid", B, 25, 20)
        out["boxes"] = run_boxes(B, 30, 10)
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=25)
    ap.add_argument("--envs", type=int, default=16384)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--scene", choices=sorted(SCENES), default="quadruped",
                    help="quadruped is the headline workload; humanoid is the high-DoF elliptic-cone one; "
                         "boxes the uneven-contact one")
    ap.add_argument("--no-learned", action="store_true")
    ap.add_argument("--headline-only", action="store_true",
                    help="skip the humanoid and boxes sub-results of the default line")
    ap.add_argument("--pairs", type=int, default=2_000_000, help="training pairs for the learned warm start")
    ap.add_argument("--epochs", type=int, default=8)
    ap.add_argument("--stages", action="store_true", help="per-stage cycle table from the profiling build")
    ap.add_argument("--nbox", type=int, default=16, help="boxes per env for --stages --scene boxes")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run only a 3 x 256 MiB device copy (HBM counter calibration)")
    ap.add_argument("--hw-queues", type=int, default=16,
                    help="GPU_MAX_HW_QUEUES for this process (HIP's default is 4): the boxes workload's 29 "
                         "sims step on 29 streams, which HIP maps onto this many hardware queues")
    args = ap.parse_args()
    # read by the HIP runtime when it starts, which is after this point (the
    # engine library starts it at its first call)
    os.environ["GPU_MAX_HW_QUEUES"] = str(args.hw_queues)
    if args.copy_calibration:
        copy_calibration()
        return
    if args.gpus != 1:
        raise SystemExit("bench.py measures one GPU only: the batched step has no multi-GPU path "
                         "(envs are independent, so each GPU would step its own share of them)")
    if args.stages:
        run_stages(args)
    elif args.scene == "boxes":
        print(json.dumps(run_boxes(args.envs, args.warmup, args.steps)), flush=True)
    else:
        run(args)


if __name__ == "__main__":
    main()
