"""Benchmark: full Barnes-Hut step throughput (body-steps/s) on one MI355X.

Workload (the reference's maximum scale, 800k bodies): 2D exponential disk,
radial scale length 0.2, circular velocities from the enclosed mass plus
N(0, 0.05 v_c) dispersion, lognormal(sigma=0.5) masses normalised to 1,
G=1, constant softening eps=1e-3, theta=0.7, dt=5e-5, 30-bit Morton keys
over the per-step bounding box, seed 1.  A step is the whole pipeline of
src/simulation.ts:393-421: bbox + Morton encode, radix sort, LBVH build,
mass/COM fill, tree walk with kick, drift.

Timed region: K steps replayed from a captured HIP graph with the bodies
already resident in HBM, bracketed by stream synchronisation.  Afterwards,
untimed: per-phase device times from HIP events on the engine's stream
(eager launches), walk counters for the flop model, and the CPU baseline
(the oracle's OpenMP restatement of the same step on a bounded sample).

Prints ONE JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for p in (ROOT, os.path.join(ROOT, "webgpu-fluid-sim_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)

METRIC = "body-steps/s, full BH step, N=800k θ=0.7; ensemble & adjoint vs plain"
N_BODIES = 800_000
THETA = 0.7
DT = 5e-5
EPS = 1e-3
SEED = 1
PEAK_FP32_TFLOPS = 157.3   # MI355X published FP32 vector peak
PMC_PROFILE = os.path.join(ROOT, "profiles", "step_800k_pmc.md")
PEAK_HBM_GBS = 8000.0
FLOPS_ACCEPT = 18          # per accepted interaction (constant softening, rsqrt = 1)
FLOPS_OPEN = 9             # per opened node
QUAD_EXTRA = 24            # quadrupole term per accepted interaction: inv^2 1, inv^5 2, Q r 6, r.Q r 3,
                           # k 6, g 3, k1 1, a -= g 2
DESCRIPTIONS = {"plain_800k": "plain BH step, 2D exponential disk N=800000, theta=0.7, "
                              "eps=1e-3, dt=5e-5, bbox30 keys, seed 1"}
# The walk kernel this workload dispatches: pair walk, plain (one simulation),
# no counters, reference softening off, own-leaf term exactly zero, 32-bit
# record offsets, monopole (the quadrupole expansion's has true last).
WALK_KERNEL = "walk_pair_kernel<false, false, false, false, true, false>"
WALK_KERNEL_QUAD = "walk_pair_kernel<false, false, false, false, true, true>"


def walk_traffic():
    """HBM bytes per walk launch at this workload, from the committed rocprofv3
    PMC summary (FETCH_SIZE and WRITE_SIZE, scaled by the copy-kernel
    calibration written there); (None, None) if the summary is absent."""
    try:
        with open(PMC_PROFILE) as f:
            for line in f:
                if line.startswith(WALK_KERNEL + " HBM bytes per launch:"):
                    return float(line.rsplit(":", 1)[1]), os.path.relpath(PMC_PROFILE, ROOT)
    except OSError:
        pass
    return None, None


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def workload(name="plain_800k", expansion="monopole", theta=None):
    """(workload, config, mass, pos, vel) of the timed leg."""
    from nbody_sim import workloads as W

    w = W.WORKLOADS[name]
    if name == "plain_800k":
        assert (w.n, w.theta, w.dt, w.eps, w.seed) == (N_BODIES, THETA, DT, EPS, SEED)
    mass, pos, vel = w.bodies()
    cfg = w.config("fast").with_(expansion=expansion, bhTheta=w.theta if theta is None else theta)
    return w, cfg, mass, pos, vel


def cpu_threads():
    """Host threads for the CPU baseline: the CPUs this process may run on,
    capped by OMP_NUM_THREADS when the environment sets it (the GPU box
    allots a share of its host cores per GPU and sets it to that share)."""
    avail = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        avail = min(avail, int(env))
    return max(1, avail)


def cpu_baseline(cfg, mass, pos, vel, budget_s=15.0):
    """The oracle's restatement of one full step, OpenMP over the host cores,
    repeated until ~budget_s of work (at least one step)."""
    from oracle import oracle as O

    threads = cpu_threads()
    P = O.Params(grav_constant=1.0, delta_time=cfg.deltaTime, bh_theta=cfg.bhTheta,
                 epsilon=cfg.epsilon, softening_mode=O.SOFTEN_CONSTANT,
                 morton_mode=O.MORTON_BBOX30)
    steps = 0
    t0 = time.perf_counter()
    p, v = pos, vel
    while True:
        p, v = O.step(mass, p, v, P, nsteps=1, threads=threads)
        steps += 1
        el = time.perf_counter() - t0
        log(f"cpu baseline step {steps}: {el:.1f}s")
        if el >= budget_s:
            break
    n = mass.shape[0]
    return {"value": n * steps / el, "unit": "body-steps/s", "cores": threads,
            "host_logical_cpus": os.cpu_count(),
            "omp_num_threads_env": os.environ.get("OMP_NUM_THREADS"),
            "kind": "port",
            "sample": f"{steps} full step(s) of the same {n}-body workload from its initial "
                      f"state, oracle C restatement, OpenMP {threads} threads, {el:.1f} s"}


# Steps captured per HIP graph in every timed leg (nbody_use_graph): one graph
# launch per 10 steps instead of per step (800k: 0.390 -> 0.385 ms per step,
# 16k: 121 -> 114 us); bit-identical to single-step graphs and to eager steps.
GRAPH_STEPS = 10


def timed_steps(sim, steps):
    sim.synchronize()
    t0 = time.perf_counter()
    sim.step(steps)
    sim.synchronize()
    return time.perf_counter() - t0


def plain_config(name, steps, math="fast", phases=False):
    """body-steps/s of a plain workload (graph replay, bodies resident);
    phases: also the per-phase device times of eager steps (HIP events)."""
    import nbody_sim
    from nbody_sim import workloads as W

    w = W.WORKLOADS[name]
    mass, pos, vel = w.bodies()
    sim = nbody_sim.Simulation(w.config(math))
    sim.setBodies(mass, pos, vel)
    sim.useGraph(True, GRAPH_STEPS)
    sim.step(GRAPH_STEPS)
    sim.prepareGraph()
    el = timed_steps(sim, steps)
    out = {"n": w.n, "steps": steps, "math_mode": math, "value": w.n * steps / el,
           "ms_per_step": 1e3 * el / steps}
    if phases:
        sim.useGraph(False)
        sim.setProfiling(True)
        sim.phaseTimes(reset=True)
        sim.step(20)
        sim.synchronize()
        ph, n = sim.phaseTimes(reset=True)
        out["phases_ms"] = {k: round(v / n, 5) for k, v in ph.items()}
    sim.close()
    return out


def ensemble_config(steps=200, tune=True):
    """256 independent sims (N log-uniform in [1024, 65536], uniform disk or
    Plummer) stepped as one ensemble, with the launch tunables picked by the
    autotuner on the ensemble itself, vs a loop of plain per-sim runs."""
    import nbody_sim
    from nbody_sim import autotune
    from nbody_sim import workloads as W

    sizes, _, parts = W.ensemble_members()
    cfg = W.ensemble_config("fast")
    total = int(sizes.sum())
    ens = nbody_sim.Ensemble(cfg, sizes)
    ens.setBodies(*W.concat(parts))
    tuning = None
    if tune:
        t0 = time.perf_counter()
        best, table = autotune.autotune(ens, steps=2, variants=(2, 4, 1), blocks=(128, 256, 512),
                                        stacks=(64, 128), xcd=(1, 0), lane_bodies=(1, 2),
                                        radix=(10, 8, 11))
        tuning = ens.getTuning()
        tuning["autotune_s"] = round(time.perf_counter() - t0, 2)
        tuning["candidates"] = len(table)
        log(f"ensemble autotune: {tuning}")
    ens.useGraph(True, GRAPH_STEPS)
    ens.step(GRAPH_STEPS)
    el_ens = timed_steps(ens, steps)
    ens.close()
    sims = []
    for p in parts:
        s = nbody_sim.Simulation(cfg)
        s.setBodies(*p)
        s.useGraph(True, GRAPH_STEPS)
        s.step(1)
        s.synchronize()
        sims.append(s)
    t0 = time.perf_counter()
    for s in sims:  # one plain run after the other
        s.step(steps)
        s.synchronize()
    el_loop = time.perf_counter() - t0
    for s in sims:
        s.close()
    return {"num_sims": len(sizes), "total_bodies": total, "key_bits": 38, "steps": steps,
            "value": total * steps / el_ens, "plain_loop_value": total * steps / el_loop,
            "speedup_vs_plain_loop": el_loop / el_ens, "tuning": tuning}


def adjoint_config(steps=200, every=20):
    """Adjoint (reverse sweep incl. segment recomputation) vs the forward run,
    both warmed up; the forward is timed eagerly (as the adjoint's own
    recomputation runs) and replayed from a HIP graph (the production step)."""
    import numpy as np

    import nbody_sim
    from nbody_sim import workloads as W

    w = W.WORKLOADS["adjoint_65k"]
    mass, pos, vel = w.bodies()
    sim = nbody_sim.Simulation(w.config("fast"))
    sim.setBodies(mass, pos, vel)
    sim.step(5)                                   # warm-up
    sim.setBodies(mass, pos, vel)
    el_fwd = timed_steps(sim, steps)              # eager forward
    sim.setBodies(mass, pos, vel)
    sim.useGraph(True, GRAPH_STEPS)
    sim.step(GRAPH_STEPS)
    sim.setBodies(mass, pos, vel)
    el_graph = timed_steps(sim, steps)            # graph-replayed forward
    sim.useGraph(False)
    sim.setBodies(mass, pos, vel)
    sim.adjointForward(10, every)                 # warm-up of the adjoint path
    sim.adjointBackward(np.ones((w.n, 2), np.float32), None)
    sim.setBodies(mass, pos, vel)
    sim.synchronize()
    t0 = time.perf_counter()
    sim.adjointForward(steps, every)
    sim.synchronize()
    el_cp = time.perf_counter() - t0
    _, xT, _ = sim.getBodies()
    g = np.random.default_rng(0).normal(size=xT.shape).astype(np.float32)
    t0 = time.perf_counter()
    gx, gv, gm = sim.adjointBackward(g, None)
    el_bwd = time.perf_counter() - t0
    sim.close()
    ok = bool(np.all(np.isfinite(gx)) and np.all(np.isfinite(gv)) and np.all(np.isfinite(gm)))
    return {"n": w.n, "steps": steps, "checkpoint_every": every,
            "forward_ms_per_step": 1e3 * el_fwd / steps,
            "forward_graph_ms_per_step": 1e3 * el_graph / steps,
            "checkpointed_forward_ms_per_step": 1e3 * el_cp / steps,
            "adjoint_ms_per_step": 1e3 * el_bwd / steps,
            "adjoint_over_forward": el_bwd / el_fwd,
            "adjoint_over_graph_forward": el_bwd / el_graph, "gradients_finite": ok}


def force_error(nt=4096):
    """Median, p99 and max relative force error of the benchmarked walk (fast
    mode, one step from each workload's initial state) against the library's
    binary64 direct sum, over `nt` seeded sample bodies per workload (the
    ensemble: every 16th of its 256 simulations)."""
    import numpy as np

    import nbody_sim
    from nbody_sim import workloads as W

    def errs(cfg, mass, pos, vel, seed):
        n = mass.shape[0]
        tg = np.sort(np.random.default_rng(seed).choice(n, min(n, nt), replace=False))
        sim = nbody_sim.Simulation(cfg)
        sim.setBodies(mass, pos, vel)
        ref = sim.directAccelF64(tg)
        sim.setDebug(accel=True)
        sim.step(1)
        acc = sim.debugBuffer("accel")[tg]
        sim.close()
        return np.linalg.norm(acc - ref, axis=1) / np.linalg.norm(ref, axis=1)

    def summary(e, **kw):
        return {"median": float(np.median(e)), "p99": float(np.percentile(e, 99)),
                "max": float(e.max()), "samples": int(e.size), **kw}

    out = {}
    for name in ("plain_16k", "plain_800k", "collision_4M", "adjoint_65k"):
        w = W.WORKLOADS[name]
        out[name] = summary(errs(w.config("fast"), *w.bodies(), 0), theta=w.theta, n=w.n)
    sizes, _, parts = W.ensemble_members()
    e = np.concatenate([errs(W.ensemble_config("fast"), *parts[s], s)
                        for s in range(0, len(sizes), 16)])
    out["ensemble_256"] = summary(e, theta=W.ENSEMBLE["theta"], members=list(range(0, 256, 16)))
    out["reference"] = "binary64 softened direct sum (nbody_direct_accel_f64)"
    return out


def accuracy_target(steps=20, nt=4096):
    """The 800k workload run where the force error meets p99 <= 1e-2 against
    the binary64 direct sum: the reference's monopole walk needs theta 0.18;
    the quadrupole expansion (opt-in) meets it at theta 0.4.  body-steps/s
    from graph replay (bodies resident), error over nt sampled bodies."""
    import numpy as np

    import nbody_sim
    from nbody_sim import workloads as W

    w = W.WORKLOADS["plain_800k"]
    mass, pos, vel = w.bodies()
    tg = np.sort(np.random.default_rng(0).choice(w.n, nt, replace=False))
    out = {"target_p99": 1e-2}
    ref = None
    for name, expn, theta in (("monopole", "monopole", 0.18), ("quadrupole", "quadrupole", 0.4)):
        sim = nbody_sim.Simulation(w.config("fast").with_(bhTheta=theta, expansion=expn))
        sim.setBodies(mass, pos, vel)
        if ref is None:
            ref = sim.directAccelF64(tg)
        sim.setDebug(accel=True)
        sim.step(1)
        acc = sim.debugBuffer("accel")[tg]
        sim.setDebug()
        e = np.linalg.norm(acc - ref, axis=1) / np.linalg.norm(ref, axis=1)
        sim.setBodies(mass, pos, vel)
        sim.useGraph(True, GRAPH_STEPS)
        sim.step(GRAPH_STEPS)
        el = timed_steps(sim, steps)
        sim.close()
        out[name] = {"theta": theta, "value": w.n * steps / el, "ms_per_step": 1e3 * el / steps,
                     "median": float(np.median(e)), "p99": float(np.percentile(e, 99)),
                     "max": float(e.max())}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--profile-steps", type=int, default=50)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cpu-budget", type=float, default=15.0)
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the other configurations, the ensemble, adjoint and error legs")
    ap.add_argument("--workload", default="plain_800k",
                    help="timed leg: plain_800k (default), plain_16k, collision_4M, adjoint_65k")
    ap.add_argument("--expansion", default="monopole", choices=("monopole", "quadrupole"))
    ap.add_argument("--theta", type=float, default=None, help="override the workload's theta")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              "(the path runs on a single MI355X).", file=sys.stderr)
        sys.exit(2)

    import nbody_sim

    w, cfg, mass, pos, vel = workload(args.workload, args.expansion, args.theta)
    default_leg = (args.workload, args.expansion, args.theta) == ("plain_800k", "monopole", None)
    log(f"workload generated: {w.name}, n={w.n}, {args.expansion}, theta={cfg.bhTheta}")
    sim = nbody_sim.Simulation(cfg, device=0)
    sim.setBodies(mass, pos, vel)
    sim.useGraph(True, GRAPH_STEPS)
    sim.step(args.warmup)
    sim.prepareGraph()  # capture (untimed) the graph of the warm step
    sim.synchronize()
    log(f"warmup {args.warmup} steps done")

    sim.synchronize()
    t0 = time.perf_counter()
    sim.step(args.steps)
    sim.synchronize()
    el = time.perf_counter() - t0
    ms_per_step = 1e3 * el / args.steps
    value = w.n * args.steps / el
    log(f"timed {args.steps} steps: {el:.3f}s, {ms_per_step:.4f} ms/step")

    # Untimed: per-phase event timing and walk counters over the same steps.
    _, p_state, v_state = sim.getBodies()
    sim.useGraph(False)
    sim.setProfiling(True)
    sim.phaseTimes(reset=True)
    sim.step(args.profile_steps)
    sim.synchronize()
    phases, psteps = sim.phaseTimes(reset=True)
    sim.setProfiling(False)
    sim.setBodies(mass, p_state, v_state)
    sim.setDebug(stats=True)
    sim.walkStats(reset=True)
    sim.step(args.profile_steps)
    sim.synchronize()
    st = sim.walkStats(reset=True)
    sim.setDebug()
    sim.close()
    phase_avg = {k: v / psteps for k, v in phases.items()}
    walk_ms = phase_avg["walk"]
    flops = (FLOPS_ACCEPT * st["accepted"] + FLOPS_OPEN * st["opened"]) / args.profile_steps
    achieved = flops / (walk_ms * 1e-3) / 1e12
    log(f"phases ms/step: {phase_avg}; walk flops/launch {flops:.3e}")

    kernel = WALK_KERNEL_QUAD if args.expansion == "quadrupole" else WALK_KERNEL
    # the committed PMC summary covers the default leg only
    traffic, traffic_src = walk_traffic() if default_leg else (None, None)
    flop_model = "18 per accepted interaction + 9 per opened node"
    if args.expansion == "quadrupole":
        flops += QUAD_EXTRA * st["accepted"] / args.profile_steps
        achieved = flops / (walk_ms * 1e-3) / 1e12
        flop_model += f" + {QUAD_EXTRA} per accepted interaction (quadrupole term)"
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "body-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": DESCRIPTIONS.get(w.name, w.name) if default_leg else
                   f"{w.name}: {w.scenario} N={w.n}, {args.expansion}, theta={cfg.bhTheta}",
                   "n_bodies": w.n, "theta": cfg.bhTheta, "dt": w.dt, "eps": w.eps,
                   "math_mode": "fast", "expansion": args.expansion, "graph": True,
                   "steps_per_graph": GRAPH_STEPS},
        "phases_ms": {k: round(v, 5) for k, v in phase_avg.items()},
        "roofline": {
            "kernel": kernel + " (tree walk + kick + drift)",
            "bound": "valu-fp32",
            "achieved": achieved,
            "peak": PEAK_FP32_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / PEAK_FP32_TFLOPS,
            "traffic": traffic,
            "traffic_source": traffic_src,
            "flops_per_launch": flops,
            "flop_model": flop_model,
            "walk_ms": walk_ms,
            "interactions_per_step": st["accepted"] / args.profile_steps,
            "opened_per_step": st["opened"] / args.profile_steps,
        },
    }
    if not args.no_extras and default_leg:
        log("exact-mode 800k")
        out["plain_800k_exact"] = plain_config("plain_800k", 20, math="exact", phases=True)
        log("small plain config")
        out["plain_small_16k"] = plain_config("plain_16k", 100, phases=True)
        log("collision config")
        out["plain_collision_4M"] = plain_config("collision_4M", 100)
        log("ensemble")
        out["ensemble_256"] = ensemble_config()
        log("adjoint")
        out["adjoint_65k"] = adjoint_config()
        log("force error")
        out["force_error"] = force_error()
        log("accuracy target")
        out["accuracy_target"] = accuracy_target()
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(cfg, mass, pos, vel, args.cpu_budget)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
