"""Throughput of the batched FP64 physics step on one MI355X.

Workload (the headline configuration): Ant (pippin_amd/assets/ant.xml: free torso +
8 hinge legs, nv = 14, capsules on a ground plane), 16384 envs, 4 substeps per
action with the control held across them, Newton solver, FP64. Initial state
qpos0 + N(0, 0.01) (quaternion renormalised), controls i.i.d. U[-1, 1] per env,
actuator and action, seed 0. One bench "step" is one action: a control update
followed by 4 physics steps of every env (Physics::set + Physics::step, as
AsyncEnv drives a backend). `value` counts physics steps: env-steps/s =
envs x substeps x actions / s. The controls for all actions are resident in HBM
before the timed region; each action copies its slice device-to-device.

The timed window starts at action 250 of the 1000-action run (physics step
1000), whatever --steps and --warmup are: the first 250 actions are an untimed
pre-roll (the launch-shape autotuner runs inside it), so the Ant has left its
start pose and walks with steady ground contact, then come --warmup untimed and
--steps timed actions. The line checks itself: after the timed actions the
state of all 16384 envs is compared with the oracle stepped over the same
actions from the same start (`max_abs_qpos_err`, `ncon_mismatches`), and the
run exits with status 1 when the error exceeds the metric's 1e-9.

Printed extras: `roofline` for the step kernel (counted FP64 operations per
env-step of the reference algorithm, measured by the oracle's op-counting build,
over the kernel time from hipEvents on the library's stream) and `cpu_baseline`
(the oracle, a C++ restatement of Pippin's CPU engine, stepping a sample of the
same workload on the host cores with Pippin's chunked thread pool).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "env-steps/s on 1 MI355X (Ant 16k envs, contact) at ≤1e-9 err; Newton iters/step"
FP64_PEAK_TFLOPS = 78.6
# reference FP64 ops per env-step on this workload, used only with --no-cpu
OPS_PER_ENV_STEP_DEFAULT = 25000.0
HBM_PEAK_GBS = 8000.0
# HBM bytes per k_step launch of the default workload (16384 envs x 4 substeps,
# the steady-contact window), from rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE
# scaled by a known-size copy (profiles/k_step_ws_ant16k.md); other workloads
# report null.
TRAFFIC_BYTES_PER_LAUNCH = 51.2e6
TRAFFIC_SOURCE = "profiles/k_step_ws_ant16k.md"
PREROLL_ACTIONS = 250  # untimed: the timed window starts at action 250 of 1000 (steady contact)
ERR_BOUND = 1e-9


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def initial_qpos(qpos0: np.ndarray, n: int, seed: int) -> np.ndarray:
    rng = np.random.default_rng(seed)
    q = np.tile(qpos0, (n, 1)) + rng.normal(0.0, 0.01, size=(n, qpos0.size))
    q[:, 3:7] /= np.linalg.norm(q[:, 3:7], axis=1, keepdims=True)
    return q


def controls(n: int, nu: int, nact: int, seed: int) -> np.ndarray:
    rng = np.random.default_rng(seed + 1)
    return rng.uniform(-1.0, 1.0, size=(nact, n, nu))


def cpu_baseline(arrays, q0, ctrl, substeps: int, t_begin: int, sample: int):
    """Time the oracle on a bounded sample of the same workload on the host cores,
    and return its end state for the self-check of the GPU's.

    The first `sample` envs (all 16384 by default) from the same start with the
    same controls: actions [0, t_begin) untimed, then the timed GPU window's
    actions, step-synchronous over the range with Pippin's chunked thread pool
    (`value`); again in Batch::rollout mode (`rollout_value`, which must end in
    the same state). Also counts the FP64 operations of the reference
    algorithm on the window (op-counting build, 256 envs), the roofline's
    numerator.
    """
    from oracle import pyoracle
    from oracle.pyoracle import OracleSim, ops_per_env_step

    tuned = pyoracle.use_host_tuned_build()
    threads = len(os.sched_getaffinity(0))
    nact, n, nu = ctrl.shape
    sample = min(n, sample)
    sim = OracleSim(arrays, sample)
    sim.set("qpos", 0, sample, q0[:sample])
    for a in range(t_begin):
        sim.set("ctrl", 0, sample, ctrl[a, :sample])
        sim.step(0, sample, substeps, threads=threads)
    t0 = time.perf_counter()
    for a in range(t_begin, nact):
        sim.set("ctrl", 0, sample, ctrl[a, :sample])
        sim.step(0, sample, substeps, threads=threads)
    dt = time.perf_counter() - t0
    end = {"qpos": sim.get("qpos"), "qvel": sim.get("qvel"), "ncon": sim.ncon()}
    env_steps = sample * substeps * (nact - t_begin)
    # Batch::rollout mode (batch.rs:46-59) over the same window: each thread
    # takes its chunk of envs through every step, controls per physics step
    ro = OracleSim(arrays, sample)
    ro.set("qpos", 0, sample, q0[:sample])
    per_step = np.repeat(np.transpose(ctrl[:, :sample], (1, 0, 2)), substeps, axis=1)  # envs x steps x nu
    ro.rollout(np.ascontiguousarray(per_step[:, :t_begin * substeps]), threads=threads, want_qpos=False)
    r0 = time.perf_counter()
    ro.rollout(np.ascontiguousarray(per_step[:, t_begin * substeps:]), threads=threads, want_qpos=False)
    rdt = time.perf_counter() - r0
    same = bool(np.array_equal(ro.get("qpos"), end["qpos"]))
    ops = ops_per_env_step(arrays, q0[:256], ctrl[:, :256], substeps, skip=t_begin)
    return {
        "value": env_steps / dt,
        "unit": "env-steps/s",
        "cores": threads,
        "rollout_value": env_steps / rdt,
        "rollout_equals_step_loop": same,
        "kind": "port",
        "sample": f"Ant, envs 0..{sample} of {n}, the {nact - t_begin} timed actions x {substeps} substeps "
                  f"after the same {t_begin} untimed actions ({env_steps} env-steps in {dt:.1f} s), "
                  f"Physics::step over the range, chunk = n/(4 x {threads} threads), "
                  f"oracle built {'-march=x86-64-v3' if tuned else 'for plain x86-64'} -O3 -ffp-contract=off",
    }, ops, end


def copy_calibration(gib: float) -> None:
    """Five device-to-device copies of `gib` GiB: the known-size copy that HBM
    byte counters (FETCH_SIZE, WRITE_SIZE) are scaled against."""
    import torch

    n = int(gib * (1 << 30)) // 8
    src = torch.arange(n, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    for _ in range(5):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes_each_way": n * 8, "copies": 5}), flush=True)


def warm_start_study(model: str, n: int, sub: int, record: int, train_iters: int, measure: int, seed: int,
                     arrays=None, prepare=None, sim_kw=None):
    """Mean Newton iterations per physics step with the reference warm start and
    with the learned candidate (PH_MLP_WARM), from the same states, and the step
    kernel time of each. The network is trained online on contacts recorded by
    the same sim (PH_MLP_LOG) before the measurement. `model` names an asset
    file, or labels `arrays` (a loaded model) whose sims `prepare(sim)` sets up."""
    import torch
    from pippin_amd.physics import HipPhysics, load_model

    if arrays is None:
        arrays = load_model(ROOT / "pippin_amd" / "assets" / model)
    ref, mlp = HipPhysics(arrays, n, **(sim_kw or {})), HipPhysics(arrays, n, **(sim_kw or {}))
    mlp.mlp_init(hidden=64, seed=seed, max_batch=8192)
    rng = np.random.default_rng(seed)
    if prepare is not None:
        for s in (ref, mlp):
            prepare(s)
    else:
        q = ref.get("qpos")
        nfree = [j for j in range(int(arrays["njnt"])) if int(np.asarray(arrays["jnt_type"])[j]) == 0]
        if nfree and model == "ant.xml":
            q = initial_qpos(np.asarray(arrays["qpos0"]), n, seed)
        else:  # move each free body in the plane only
            adr = np.asarray(arrays["jnt_qposadr"])
            for j in nfree:
                q[:, adr[j]:adr[j] + 2] += rng.normal(0.0, 0.01, size=(n, 2))
        for s in (ref, mlp):
            s.set("qpos", 0, n, q)
    ctrl = rng.uniform(-1.0, 1.0, size=(record + measure, n, max(ref.nu, 1)))
    mlp.mlp_mode(HipPhysics.MLP_LOG)
    for a in range(record):
        for s in (ref, mlp):
            if ref.nu:
                s.set("ctrl", 0, n, ctrl[a])
            s.step(0, n, sub)
        stored = mlp.mlp_collect()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    loss = mlp.mlp_train(train_iters, batch=8192, lr=2e-3, seed=seed)
    torch.cuda.synchronize()
    train_s = time.perf_counter() - t0
    it_ref, it_mlp = [], []
    mlp.mlp_mode(HipPhysics.MLP_WARM)
    ref.kernel_timing(reset=True)
    mlp.kernel_timing(reset=True)
    for a in range(record, record + measure):
        mlp.set("qpos", 0, n, ref.get("qpos"))
        mlp.set("qvel", 0, n, ref.get("qvel"))
        for s in (ref, mlp):
            if ref.nu:
                s.set("ctrl", 0, n, ctrl[a])
        for _ in range(sub):
            ref.step(0, n, 1)
            mlp.step(0, n, 1)
            it_ref.append(ref.solver_iterations())
            it_mlp.append(mlp.solver_iterations())
    kr, lr_ = ref.kernel_timing(reset=True)
    km, lm = mlp.kernel_timing(reset=True)
    a, b = np.concatenate(it_ref), np.concatenate(it_mlp)
    stats = lambda x: {"mean": float(x.mean()), "p50": float(np.percentile(x, 50)),
                       "p99": float(np.percentile(x, 99))}
    return {"model": model, "envs": n, "recorded_contacts": int(stored), "train_steps": train_iters,
            "train_s": train_s, "loss_first": float(loss[:10].mean()), "loss_last": float(loss[-10:].mean()),
            "iters_reference": stats(a), "iters_learned": stats(b),
            "iters_reduction": float((a.mean() - b.mean()) / a.mean()),
            "kernel_ms_per_step_reference": kr / lr_, "kernel_ms_per_step_learned": km / lm}


def warm_start_pgs_study(n: int, record: int, train_iters: int, measure: int, seed: int, pile_seed: int = 29):
    """The learned warm start for PGS on a six-body pile: the network, trained
    online on the solved forces of the sim's own contacts (PH_MLP_LOG), warm-
    starts the contacts that the previous-step matching leaves at zero
    (PH_MLP_WARM). Recorded over the first `record` steps after the drop, then
    measured over the next `measure` steps, while contacts still come and go:
    sweeps per physics step with the reference warm start and with the learned
    one, each step taken from the same saved state (positions, velocities and
    contact history), and the step kernel time of each."""
    import torch
    from pippin_amd.physics import HipPhysics, load_model
    from pippin_amd.workloads import random_pile_xml

    arrays = load_model(random_pile_xml(pile_seed), is_xml=True)
    arrays["solver_kind"] = 1
    ref, mlp = HipPhysics(arrays, n), HipPhysics(arrays, n)
    for s in (ref, mlp):
        s.mlp_init(hidden=64, seed=seed, max_batch=8192)
    rng = np.random.default_rng(pile_seed)
    q = ref.get("qpos")
    for b in range(int(arrays["nq"]) // 7):
        q[:, 7 * b:7 * b + 2] += rng.normal(size=(n, 2)) * 0.005
    for s in (ref, mlp):
        s.set("qpos", 0, n, q)
    ref.mlp_mode(HipPhysics.MLP_LOG)  # the reference warm start; sweep counts reported
    mlp.mlp_mode(HipPhysics.MLP_LOG)
    stored = 0
    for _ in range(record):
        for s in (ref, mlp):
            s.step(0, n, 1)
        stored = mlp.mlp_collect()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    loss = mlp.mlp_train(train_iters, batch=8192, lr=2e-3, seed=seed)
    torch.cuda.synchronize()
    train_s = time.perf_counter() - t0
    mlp.mlp_mode(HipPhysics.MLP_WARM)
    sw_ref, sw_mlp = [], []
    ref.kernel_timing(reset=True)
    mlp.kernel_timing(reset=True)
    for _ in range(measure):
        mlp.load_state(ref.save_state())
        ref.step(0, n, 1)
        mlp.step(0, n, 1)
        sw_ref.append(ref.solver_iterations())
        sw_mlp.append(mlp.solver_iterations())
    kr, lr_ = ref.kernel_timing(reset=True)
    km, lm = mlp.kernel_timing(reset=True)
    a, b = np.concatenate(sw_ref), np.concatenate(sw_mlp)
    stats = lambda x: {"mean": float(x.mean()), "p50": float(np.percentile(x, 50)),
                       "p99": float(np.percentile(x, 99))}
    return {"model": f"pile seed {pile_seed} (6 bodies, nv 36), PGS", "envs": n, "recorded_contacts": int(stored),
            "train_steps": train_iters, "train_s": train_s, "loss_first": float(loss[:10].mean()),
            "loss_last": float(loss[-10:].mean()), "sweeps_reference": stats(a), "sweeps_learned": stats(b),
            "sweeps_reduction": float((a.mean() - b.mean()) / a.mean()),
            "kernel_ms_per_step_reference": kr / lr_, "kernel_ms_per_step_learned": km / lm}


def pile_study(n: int, steps: int, seed: int, cpu_envs: int, general: bool = False, solvers=("newton", "pgs")):
    """Tabletop piles (pippin_amd/workloads.py random_pile_xml: 2 to 6 free
    primitives dropped onto a plane), after the reference's tumble scenes:
    env-steps/s on the GPU (4 physics steps per launch, bodies settling and
    resting) and of the oracle on the host cores over a sample of the same
    envs and steps, for both solvers. general=True lets cylinders meet every
    body (the general convex narrow phase)."""
    import torch
    from oracle.pyoracle import OracleSim
    from pippin_amd.physics import HipError, HipPhysics, load_model
    from pippin_amd.workloads import random_pile_xml

    threads = len(os.sched_getaffinity(0))
    failures = []

    def step(sim, nstep, where):
        # An env whose step fails (the reference panics: a mass matrix that is
        # not positive definite after a PGS blow-up, which the oracle reproduces
        # step for step) is frozen by the library and reported; the others go on.
        try:
            sim.step(0, n, nstep)
        except HipError as ex:
            failures.append(f"{where}: {ex}")
    arrays = load_model(random_pile_xml(seed, general), is_xml=True)
    out = []
    for solver in solvers:
        sim = HipPhysics(arrays, n, solver=solver)
        rng = np.random.default_rng(seed)
        q = sim.get("qpos")
        nb = int(arrays["nq"]) // 7
        for b in range(nb):
            q[:, 7 * b:7 * b + 2] += rng.normal(size=(n, 2)) * 0.005
        sim.set("qpos", 0, n, q)
        failures.clear()
        step(sim, 40, "drop")
        # autotune the launch shape on the settling pile (untimed; results do not depend on it);
        # three timed rounds, so that each shape's median spans the settling
        sim.autotune(rounds=3)
        tuned = 0
        while sim.autotune_pending() and tuned < 200:
            step(sim, 4, f"autotune launch {tuned}")
            tuned += 1
        torch.cuda.synchronize()
        sim.kernel_timing(reset=True)
        t0 = time.perf_counter()
        for li in range(steps // 4):
            step(sim, 4, f"timed launch {li}")
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        kms, launches = sim.kernel_timing(reset=True)
        m = min(cpu_envs, n)
        cdt, same = None, None
        if m > 0:
            # the host runs the same solver (the model file names Newton)
            ref = OracleSim({**arrays, "solver_kind": 0 if solver == "newton" else 1}, m)
            ref.set("qpos", 0, m, q[:m])
            ref.step(0, m, 40 + 4 * tuned, threads=threads)
            c0 = time.perf_counter()
            ref.step(0, m, steps, threads=threads)
            cdt = time.perf_counter() - c0
            same = {"envs": m, "qpos": bool(np.array_equal(ref.get("qpos"), sim.get("qpos", 0, m))),
                    "ncon": bool(np.array_equal(ref.ncon(), sim.ncon(0, m)))}
        kind = ", cylinders meet every body" if general else ""
        out.append({"workload": f"pile seed {seed} ({nb} bodies{kind}), {solver}", "envs": n, "steps": steps,
                    "gpu_env_steps_per_s": n * steps / dt, "kernel_ms_per_launch": kms / launches,
                    "cpu_env_steps_per_s": m * steps / cdt if cdt else None, "cpu_envs": m, "cpu_threads": threads,
                    "equal_to_oracle": same,
                    "newton_iters_mean": float(sim.solver_iterations().mean()),
                    "launch_shape": sim.launch_shape(),
                    "env_failures": {"first": failures[0] if failures else None,
                                     "launches_reporting_one": len(failures)},
                    "contacts_mean": float(sim.ncon().mean())})
        sim.close()
    return out


def mfma_study(n: int, steps: int, seed: int, cpu_envs: int):
    """Newton's dense Hessian J'DJ on the FP64 matrix cores (ph_options.hessian_mfma)
    against the reference's row-by-row sums, on a 6-body pile (nv 36; islands over
    16 dofs take the dense solver), one env per wavefront so that the matrix cores
    serve each env: step kernel time with and without, and how far the trajectory
    of the first `cpu_envs` envs moves from the oracle's (which the plain path
    matches bit for bit), every 20 steps."""
    import torch
    from oracle.pyoracle import OracleSim
    from pippin_amd.physics import HipPhysics, load_model
    from pippin_amd.workloads import random_pile_xml

    arrays = load_model(random_pile_xml(seed), is_xml=True)
    rng = np.random.default_rng(seed)
    probe = HipPhysics(arrays, 1)
    q1 = probe.get("qpos")[0]
    probe.close()
    q = np.tile(q1, (n, 1))
    nb = int(arrays["nq"]) // 7
    for b in range(nb):
        q[:, 7 * b:7 * b + 2] += rng.normal(size=(n, 2)) * 0.005
    out = {"workload": f"pile seed {seed} ({nb} bodies, nv {int(arrays['nv'])}), Newton, {n} envs, "
                       f"launch shape (1, generic, ws_bucket)", "steps": steps}
    for flag in (False, True):
        sim = HipPhysics(arrays, n, hessian_mfma=flag)
        sim.set_launch_shape(1, 0, "ws_bucket")
        sim.set("qpos", 0, n, q)
        sim.step(0, n, 40)
        torch.cuda.synchronize()
        sim.kernel_timing(reset=True)
        t0 = time.perf_counter()
        for _ in range(steps // 4):
            sim.step(0, n, 4)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        kms, launches = sim.kernel_timing(reset=True)
        key = "mfma" if flag else "plain"
        out[f"{key}_env_steps_per_s"] = n * steps / dt
        out[f"{key}_kernel_ms_per_launch"] = kms / launches
        out[f"{key}_newton_iters_mean"] = float(sim.solver_iterations().mean())
        sim.close()
    m = min(cpu_envs, n)
    ref = OracleSim(arrays, m)
    dev = {f: HipPhysics(arrays, m, hessian_mfma=f) for f in (False, True)}
    for f, d in dev.items():
        d.set_launch_shape(1, 0, "ws_bucket")
        d.set("qpos", 0, m, q[:m])
    ref.set("qpos", 0, m, q[:m])
    growth = {"plain": [], "mfma": []}
    for t in range(0, 40 + steps, 20):
        ref.step(0, m, 20)
        for f, d in dev.items():
            d.step(0, m, 20)
            growth["mfma" if f else "plain"].append(float(np.abs(d.get("qpos") - ref.get("qpos")).max()))
    out["max_abs_qpos_err_every_20_steps"] = growth
    out["err_bound"] = ERR_BOUND
    out["mfma_within_bound"] = bool(max(growth["mfma"]) <= ERR_BOUND)
    return out


def tumble_study(n: int, steps: int, seed: int, cpu_envs: int):
    """The tabletop tumble configuration (pippin_amd/workloads.py
    tabletop_tumble): per env 2 to 6 objects from boxes, spheres, cylinders,
    capsules and 12-vertex hull wedges, per-env sizes, densities U[500, 1500]
    and frictions U[0.5, 1.0] set through per-env parameters, dropped from
    non-overlapping poses 0.05 to 0.2 m above a plane; `steps` physics steps
    from the drop, both solvers. The launch shape is autotuned on a separate
    sim of the same envs first (the choice is cached and taken by the timed
    sim). The oracle steps the first `cpu_envs` envs over the same steps on the
    host cores; its end state must equal the GPU's bit for bit."""
    import tempfile

    import torch
    from oracle.pyoracle import OracleSim
    from pippin_amd.physics import HipError, HipPhysics, load_model
    from pippin_amd.workloads import tabletop_tumble

    threads = len(os.sched_getaffinity(0))
    t0 = time.perf_counter()
    plan = tabletop_tumble(n, seed)
    gen_s = time.perf_counter() - t0
    arrays = load_model(plan.xml, is_xml=True)
    cache = os.path.join(tempfile.mkdtemp(), "shapes.txt")
    out = []
    for solver in ("newton", "pgs"):
        failures = []

        def step(sim, nstep, where):
            try:
                sim.step(0, n, nstep)
            except HipError as ex:
                failures.append(f"{where}: {ex}")

        tuner = HipPhysics(arrays, n, solver=solver, max_contacts=96)
        plan.apply(tuner)
        tuner.autotune(rounds=3, cache_path=cache)
        tuned = 0
        while tuner.autotune_pending() and tuned < 200:
            step(tuner, 4, f"autotune launch {tuned}")
            tuned += 1
        tuner.close()
        sim = HipPhysics(arrays, n, solver=solver, max_contacts=96)
        a0 = time.perf_counter()
        plan.apply(sim)
        torch.cuda.synchronize()
        apply_s = time.perf_counter() - a0
        sim.autotune(rounds=1, cache_path=cache)  # takes the cached shape
        failures.clear()
        sim.kernel_timing(reset=True)
        t0 = time.perf_counter()
        for li in range(steps // 4):
            step(sim, 4, f"timed launch {li}")
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        kms, launches = sim.kernel_timing(reset=True)
        m = min(cpu_envs, n)
        # the host runs the same solver (the model names Newton)
        ref = OracleSim({**arrays, "solver_kind": 0 if solver == "newton" else 1}, m)
        plan.apply(ref, 0, m)
        c0 = time.perf_counter()
        ref.step(0, m, steps, threads=threads)
        cdt = time.perf_counter() - c0
        same_q = bool(np.array_equal(ref.get("qpos"), sim.get("qpos", 0, m)))
        same_n = bool(np.array_equal(ref.ncon(), sim.ncon(0, m)))
        out.append({"workload": f"tabletop tumble (seed {seed}), {solver}", "envs": n, "steps": steps,
                    "gpu_env_steps_per_s": n * steps / dt, "kernel_ms_per_launch": kms / launches,
                    "cpu_env_steps_per_s": m * steps / cdt if cdt else None, "cpu_envs": m, "cpu_threads": threads,
                    "equal_to_oracle": {"envs": m, "qpos": same_q, "ncon": same_n},
                    "objects_per_env_mean": float(plan.counts.mean()),
                    "wedges": int((plan.kinds == 4).sum()),
                    "contacts_mean": float(sim.ncon().mean()),
                    "newton_iters_mean": float(sim.solver_iterations().mean()),
                    "launch_shape": sim.launch_shape(), "param_apply_s": apply_s, "plan_s": gen_s,
                    "env_failures": {"first": failures[0] if failures else None,
                                     "launches_reporting_one": len(failures)}})
        sim.close()
    return out


def arm_study(n: int, steps: int, seed: int, cpu_envs: int):
    """The arm configuration (pippin_amd/workloads.py arm_gripper_xml: 7-DoF
    arm, tendon-coupled two-finger gripper, cube on a table, implicitfast),
    drawn per env through per-env parameters: link masses U[0.5, 3] kg (inertia
    scaled with the mass), cube half-extent U[0.02, 0.03] m resting on the table
    and cube friction U[0.6, 1.0]; joint targets on smooth random waypoints every
    step. Batch::rollout on the device (per-step
    controls resident in HBM, 50 steps per launch) against the oracle's
    rollout of a sample of the same envs on the host cores."""
    import torch
    from oracle.pyoracle import OracleSim
    from pippin_amd.physics import HipPhysics, load_model
    from pippin_amd.workloads import arm_env_params, arm_gripper_xml, arm_targets

    threads = len(os.sched_getaffinity(0))
    arrays = load_model(arm_gripper_xml(seed), is_xml=True)
    nu, cube = int(arrays["nu"]), int(arrays["ngeom"]) - 1
    warm, chunk, tune_steps = 50, 50, 400
    tg = arm_targets(seed, n, tune_steps + warm + steps, nu)  # (steps, envs, nu)
    sim = HipPhysics(arrays, n)
    ref = OracleSim(arrays, min(cpu_envs, n))
    params, qpos0 = arm_env_params(arrays, n, seed)
    for s in (sim, ref):
        for name, gid, values in params:
            s.set_param(name, gid, 0, s.n, np.ascontiguousarray(values[:s.n]))
        s.set_param("qpos0", 0, 0, s.n, np.ascontiguousarray(qpos0[:s.n]))
        s.reset(np.arange(s.n))
    ctrl = np.ascontiguousarray(np.transpose(tg, (1, 0, 2)))  # envs x steps x nu
    dctrl = torch.from_numpy(ctrl).to("cuda")
    # autotune the launch shape on 10-step rollouts (untimed; results do not depend on it)
    sim.autotune(rounds=1)
    c0 = 0
    while sim.autotune_pending() and c0 + 10 <= tune_steps:
        part = dctrl[:, c0:c0 + 10].contiguous()
        sim.rollout_device(part.data_ptr(), None, 10)
        c0 += 10
    tuned = c0
    for c0 in range(tuned, tuned + warm, chunk):  # warmup (untimed)
        part = dctrl[:, c0:c0 + chunk].contiguous()
        sim.rollout_device(part.data_ptr(), None, part.shape[1])
    t_begin = tuned + warm
    parts = [dctrl[:, c0:c0 + chunk].contiguous() for c0 in range(t_begin, t_begin + steps, chunk)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for part in parts:
        sim.rollout_device(part.data_ptr(), None, part.shape[1])
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    ncon = float(sim.ncon().mean())
    iters = float(sim.solver_iterations().mean())
    shape = sim.launch_shape()
    sim.close()
    m = ref.n
    ref.rollout(ctrl[:m, :t_begin], threads=threads, want_qpos=False)
    c0 = time.perf_counter()
    ref.rollout(ctrl[:m, t_begin:t_begin + steps], threads=threads, want_qpos=False)
    cdt = time.perf_counter() - c0
    return {"workload": f"7-DoF arm + tendon gripper + cube (seed {seed}), implicitfast, Newton, rollout with "
                        f"per-step targets", "envs": n, "steps": steps,
            "gpu_env_steps_per_s": n * steps / dt, "cpu_env_steps_per_s": m * steps / cdt, "cpu_envs": m,
            "cpu_threads": threads, "contacts_mean": ncon, "newton_iters_mean": iters, "launch_shape": shape,
            "autotune_steps": tuned}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=250, help="timed actions")
    ap.add_argument("--warmup", type=int, default=25, help="untimed actions")
    ap.add_argument("--envs", type=int, default=16384)
    ap.add_argument("--substeps", type=int, default=4)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--cpu-envs", type=int, default=16384, help="envs in the CPU-baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--iter-stats", type=int, default=20,
                    help="extra (untimed) actions stepped one physics step at a time for the Newton "
                         "iteration and contact statistics; 0 skips them (profiling runs)")
    ap.add_argument("--copy-calibration", type=float, default=0.0, metavar="GIB",
                    help="only run a plain device-to-device copy of this many GiB five times "
                         "(reference kernel for scaling HBM byte counters), then exit")
    ap.add_argument("--pile-gpu", default="",
                    help="SEED,SOLVER: one pile study (8192 envs, 200 steps) on the GPU only, for profiling")
    ap.add_argument("--warm-start-pgs", action="store_true",
                    help="the learned warm start for PGS on the six-body pile: sweeps per step with and without")
    ap.add_argument("--warm-start-wide", action="store_true",
                    help="the learned warm-start study on the six-body pile and the tabletop tumble (Newton)")
    ap.add_argument("--warm-start", action="store_true",
                    help="only run the learned warm-start study (Ant and box_drop: mean Newton iterations "
                         "with and without the candidate) and print one JSON line per model")
    ap.add_argument("--piles", action="store_true",
                    help="only run the tabletop pile study (8192 envs, both solvers, GPU and host oracle) "
                         "and print one JSON line per solver")
    ap.add_argument("--shape", default="",
                    help="launch shape 'K,specialised,schedule' (schedule: grid, ws, ws_bucket) instead of "
                         "autotuning it on the first actions")
    ap.add_argument("--tune-rounds", type=int, default=2, help="timed rounds over the launch shapes")
    ap.add_argument("--no-graphs", action="store_true", help="launch every step eagerly instead of graph replay")
    ap.add_argument("--arm", action="store_true",
                    help="only run the arm study (4096 envs, rollout with per-step targets, GPU and host oracle)")
    ap.add_argument("--tumble", action="store_true",
                    help="only run the tabletop tumble study (8192 envs, 500 steps, both solvers, per-env object "
                         "sets with hull wedges, GPU and host oracle) and print one JSON line per solver")
    ap.add_argument("--mfma-study", action="store_true",
                    help="only run the study of Newton's Hessian on the FP64 matrix cores (6-body pile, 8192 envs): "
                         "speed and distance from the oracle, one JSON line")
    ap.add_argument("--piles-general", action="store_true",
                    help="as --piles, on piles whose cylinders meet every body (general convex pairs)")
    args = ap.parse_args()
    if args.arm:
        print(json.dumps(arm_study(4096, 200, 0, 512)), flush=True)
        return
    if args.mfma_study:
        print(json.dumps(mfma_study(8192, 200, 29, 256)), flush=True)
        return
    if args.tumble:
        for line in tumble_study(args.envs if args.envs != 16384 else 8192, 500, args.seed, 1024):
            print(json.dumps(line), flush=True)
        return
    if args.pile_gpu:
        seed, solver = args.pile_gpu.split(",")
        for line in pile_study(8192, 200, int(seed), 0, solvers=(solver,)):
            print(json.dumps(line), flush=True)
        return
    if args.piles or args.piles_general:
        seeds = (13, 18) if args.piles_general else (11, 13, 29)
        for seed in seeds:
            for line in pile_study(8192, 200, seed, 1024, general=args.piles_general):
                print(json.dumps(line), flush=True)
        return
    if args.warm_start:
        for model, n, sub, rec, meas in (("ant.xml", args.envs, args.substeps, 100, 25),
                                          ("box_drop.xml", args.envs, 1, 400, 100)):
            log(f"warm-start study: {model}, {n} envs")
            print(json.dumps(warm_start_study(model, n, sub, rec, 1500, meas, args.seed)), flush=True)
        return
    if args.warm_start_pgs:
        log("warm-start study: PGS pile seed 29, 8192 envs")
        print(json.dumps(warm_start_pgs_study(8192, 100, 1500, 100, args.seed)), flush=True)
        return
    if args.warm_start_wide:
        # the wide Newton workloads: the six-body pile and the tabletop tumble, 8192 envs
        from pippin_amd.physics import load_model
        from pippin_amd.workloads import random_pile_xml, tabletop_tumble
        pile = load_model(random_pile_xml(29), is_xml=True)
        log("warm-start study: pile seed 29, 8192 envs")
        print(json.dumps(warm_start_study("pile seed 29 (6 bodies, nv 36)", 8192, 1, 200, 1500, 100, args.seed,
                                          arrays=pile)), flush=True)
        plan = tabletop_tumble(8192, args.seed)
        log("warm-start study: tabletop tumble, 8192 envs")
        print(json.dumps(warm_start_study("tabletop tumble (seed 0)", 8192, 1, 200, 1500, 100, args.seed,
                                          arrays=load_model(plan.xml, is_xml=True), prepare=plan.apply,
                                          sim_kw={"max_contacts": 96})), flush=True)
        return
    if args.copy_calibration > 0:
        copy_calibration(args.copy_calibration)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU: every box here has one MI355X and the environments are "
              f"independent (zero exchange while stepping), so --gpus {args.gpus} is not run.", file=sys.stderr)
        sys.exit(2)

    import torch
    from pippin_amd.physics import HipPhysics, load_model

    arrays = load_model(ROOT / "pippin_amd" / "assets" / "ant.xml")
    n, sub = args.envs, args.substeps
    pre = PREROLL_ACTIONS
    nact = pre + args.warmup + args.steps
    sim = HipPhysics(arrays, n)
    if args.shape:
        k, sp, sc = args.shape.split(",")
        sim.set_launch_shape(int(k), int(sp), sc)
    sim.use_graphs(not args.no_graphs)
    nu = sim.nu
    q0 = initial_qpos(np.asarray(arrays["qpos0"]), n, args.seed)
    ctrl = controls(n, nu, nact, args.seed)
    sim.set("qpos", 0, n, q0)
    dctrl = torch.from_numpy(ctrl).to("cuda")  # resident in HBM before timing
    base = dctrl.data_ptr()
    stride = n * nu * 8
    log(f"bench: {n} envs x {sub} substeps, {pre} pre-roll + {args.warmup} warmup + {args.steps} timed actions")

    # Untimed pre-roll to steady contact. The autotuner tries every launch shape
    # on its first live actions; every shape gives the same trajectories, so
    # this is pre-roll like any other.
    a0 = 0
    if not args.shape:
        sim.autotune(rounds=args.tune_rounds)
        while sim.autotune_pending() and a0 < pre:
            sim.set_device("ctrl", 0, n, base + a0 * stride)
            sim.step(0, n, sub)
            a0 += 1
        log(f"autotuned after {a0} actions: launch shape {sim.launch_shape()}")
    tuned_actions = a0
    for a in range(a0, pre + args.warmup):
        sim.set_device("ctrl", 0, n, base + a * stride)
        sim.step(0, n, sub)
    t_begin = pre + args.warmup
    torch.cuda.synchronize()
    sim.kernel_timing(reset=True)
    t0 = time.perf_counter()
    for a in range(t_begin, t_begin + args.steps):
        sim.set_device("ctrl", 0, n, base + a * stride)
        sim.step(0, n, sub)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    kms, launches = sim.kernel_timing(reset=True)
    env_steps = n * sub * args.steps
    value = env_steps / dt
    log(f"timed {args.steps} actions in {dt:.3f} s; kernel {kms / launches:.3f} ms/launch")
    t_end = t_begin + args.steps
    gpu_end = {"qpos": sim.get("qpos"), "qvel": sim.get("qvel"), "ncon": sim.ncon()}

    # Newton iterations per step and contacts, over a few more (untimed) actions
    iters, ncon = [], []
    for a in range(t_end - min(args.iter_stats, t_end), t_end):
        sim.set_device("ctrl", 0, n, base + a * stride)
        for _ in range(sub):
            sim.step(0, n, 1)
            iters.append(sim.solver_iterations())
            ncon.append(sim.ncon())
    iters = np.concatenate(iters) if iters else np.zeros(1)
    ncon = np.concatenate(ncon) if ncon else np.zeros(1, dtype=np.int32)

    check = None
    if args.no_cpu:
        cb, ops = None, OPS_PER_ENV_STEP_DEFAULT
    else:
        cb, ops, cpu_end = cpu_baseline(arrays, q0, ctrl[:t_end], sub, t_begin, args.cpu_envs)
        m = cpu_end["qpos"].shape[0]
        check = {"envs_compared": m, "bound": ERR_BOUND,
                 "max_abs_qpos_err": float(np.abs(gpu_end["qpos"][:m] - cpu_end["qpos"]).max()),
                 "max_abs_qvel_err": float(np.abs(gpu_end["qvel"][:m] - cpu_end["qvel"]).max()),
                 "ncon_mismatches": int((gpu_end["ncon"][:m] != cpu_end["ncon"]).sum()),
                 "against": "the oracle stepped over the same actions from the same start"}
    # roofline of the step kernel: FP64 operations of the reference algorithm per env-step
    # (op-counting oracle on the same window) over the kernel's own time
    kernel_s = kms / 1e3 / launches
    achieved = ops * n * sub / kernel_s / 1e12
    line = {
        "metric": METRIC,
        "value": value,
        "unit": "env-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dt / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"Ant (assets/ant.xml) 16384 envs x 4 substeps per action, Newton, FP64, "
                               f"ctrl U[-1,1] per action, qpos0 + N(0,0.01); timed window = actions "
                               f"{t_begin}..{t_end} of the 1000-action run, after an untimed {pre}-action "
                               f"pre-roll to steady contact and {args.warmup} warmup actions",
                   "envs": n, "substeps": sub, "actions": args.steps, "window_first_action": t_begin,
                   "launch_shape": dict(zip(("envs_per_wave", "specialised", "schedule"), sim.launch_shape())),
                   "autotune_actions": tuned_actions},
        "newton_iters_per_step": {"mean": float(iters.mean()), "p50": float(np.percentile(iters, 50)),
                                  "p99": float(np.percentile(iters, 99))},
        "contacts_per_env": {"mean": float(ncon.mean()), "max": int(ncon.max())},
        "roofline": {"bound": "fp64_valu", "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / FP64_PEAK_TFLOPS,
                     "traffic": TRAFFIC_BYTES_PER_LAUNCH if (n, sub) == (16384, 4) else None,
                     "traffic_source": f"{TRAFFIC_SOURCE} (rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE of this "
                                       f"command, copy-calibrated); not measured inside this run",
                     "note": f"the step kernel issues no MFMA: it is bound by FP64 VALU issue and the latency "
                             f"of dependent chains (MI355X FP64 peak, vector = matrix); {ops:.0f} reference "
                             f"FP64 ops per env-step; step kernel {kernel_s * 1e3:.3f} ms per launch "
                             f"({n} envs x {sub} steps)"},
    }
    if check is not None:
        line["max_abs_qpos_err"] = check["max_abs_qpos_err"]
        line["ncon_mismatches"] = check["ncon_mismatches"]
        line["self_check"] = check
    if cb is not None:
        line["cpu_baseline"] = cb
        line["gpu_over_cpu"] = value / cb["value"]
    print(json.dumps(line), flush=True)
    if check is not None and (check["max_abs_qpos_err"] > ERR_BOUND or check["ncon_mismatches"] > 0):
        log(f"self-check failed: qpos error {check['max_abs_qpos_err']:.3e} (bound {ERR_BOUND}), "
            f"{check['ncon_mismatches']} envs with other contact counts")
        sys.exit(1)


if __name__ == "__main__":
    main()
