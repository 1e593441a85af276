"""Benchmark: physics env-steps/s of the batched quadruped step on one MI355X.

Workload: procedurally written quadruped (free-joint torso + 12 hinges,
capsule legs, sphere feet, self-collision on, condim 3 pyramidal, position
actuators kp 20 / kv 0.5), Newton solver, 16 iterations, tolerance 1e-8,
dt 0.005. One env-step = one control step = 4 physics substeps with the same
ctrl (the reference's physics-only unit, reference README.md:50). Actions are
nominal + U(+-0.5 rad), synthetic, from a pre-generated pool.

Prints one JSON line with the throughput, the roofline of the step kernel
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


def cpu_baseline(xml_fields, nenv=1024, warm=5, steps=20):
    """FP64 oracle on the host cores: env-steps/s on a bounded sample."""
    import ref64
    from mjmlx_hip import scenes

    nthreads = max(1, min(16, os.cpu_count() or 1))
    orc = ref64.Oracle(xml_fields)
    rng = np.random.default_rng(0)
    q = scenes.quadruped_initial_qpos(nenv, 0)
    qv = np.zeros((nenv, orc.nv))
    qa = np.zeros((nenv, orc.nv))
    for _ in range(warm):
        orc.batch_step(q, qv, qa, scenes.quadruped_actions(nenv, rng), 4, nthreads)
    ctrls = [np.ascontiguousarray(scenes.quadruped_actions(nenv, rng)) for _ in range(steps)]
    t0 = time.perf_counter()
    for c in ctrls:
        orc.batch_step(q, qv, qa, c, 4, nthreads)
    dt = time.perf_counter() - t0
    return {"value": nenv * steps / dt, "unit": "env-steps/s", "cores": nthreads, "kind": "port",
            "sample": f"own FP64 restatement (oracle/ref64.c, not MuJoCo C): {nenv} envs x {steps} control steps "
                      f"(4 substeps each) after {warm} warm-up steps, {nthreads} threads, {dt:.2f} s"}


def run(args):
    from mjmlx_hip import capi, scenes

    if capi.lib().mjmlx_hip_device_count() < 1:
        raise SystemExit("bench: no HIP device visible")
    B = args.envs
    model = capi.Model(scenes.quadruped_xml(self_collision=True))
    sim = capi.BatchedSim(model, B)
    if args.persistent:
        sim.set_launch(True, args.grid)
    rng = np.random.default_rng(1)
    sim.set_state(scenes.quadruped_initial_qpos(B, 0).astype(np.float32), np.zeros((B, model.nv), np.float32))
    pool = [scenes.quadruped_actions(B, rng).astype(np.float32) for _ in range(16)]
    for k in range(args.warmup):
        sim.step(pool[k % len(pool)], 4)
    sim.sync()
    kernel_ms = []
    iters = []
    t0 = time.perf_counter()
    for k in range(args.steps):
        sim.step(pool[k % len(pool)], 4)
        kernel_ms.append(sim.last_step_ms())  # waits for the step's kernel
    sim.sync()
    dt = time.perf_counter() - t0
    st = sim.stats()
    iters.append(float(st["niter"].mean()))
    env_steps_s = B * args.steps / dt
    avg_kernel_s = float(np.mean(kernel_ms)) / 1e3
    bpe = bytes_per_env_step(model.nq, model.nv, model.nu)
    achieved = B * bpe / avg_kernel_s / 1e9
    traffic = None
    tpath = os.path.join(ROOT, "profiles", f"traffic_quadruped{B}.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            traffic = json.load(f).get("bytes_per_launch")
    out = {
        "metric": "physics env-steps/s @16384 envs, plain vs learned warm-start at 1e-8 residual",
        "value": env_steps_s,
        "unit": "env-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dt / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (FP64 Hessian Cholesky and free-joint quaternion integration)",
        "data": "synthetic",
        "config": {"workload": "quadruped nq19 nv18 nu12, self-collision, condim 3 pyramidal, Newton 16 it, "
                               "tol 1e-8, dt 0.005 x 4 substeps, plain (previous-step) warm start",
                   "num_envs": B, "substeps_per_env_step": 4, "parallelism": "single GPU, 1 wavefront per env",
                   "launch": "persistent work queue" if args.persistent else "one workgroup per env"},
        "plain": {"env_steps_per_s": env_steps_s, "mean_solver_iters_per_env_step": iters[-1]},
        "learned": None,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": "k_step", "kernel_ms_avg": avg_kernel_s * 1e3,
                     "bytes_per_env_step": bpe},
        "cpu_baseline": None,
    }
    if not args.no_cpu:
        import ref64
        out["cpu_baseline"] = cpu_baseline(ref64.fields_from(model.field))
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=25)
    ap.add_argument("--envs", type=int, default=16384)
    ap.add_argument("--persistent", action="store_true")
    ap.add_argument("--grid", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        raise SystemExit("bench.py measures one GPU only: the batched step has no multi-GPU path "
                         "(envs are independent; see DESIGN.md for how it would shard)")
    run(args)


if __name__ == "__main__":
    main()
