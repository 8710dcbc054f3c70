"""Shared helpers: run one step on the GPU through the C ABI and on the oracle."""
from __future__ import annotations

import numpy as np

import nbody_sim
from oracle import oracle as O

BUFS = ("keys", "body_indices", "children", "parent", "node_cm", "node_aabb", "accel", "bbox")


def params(theta=0.5, dt=1e-4, eps=1e-3, G=1.0, eps_mult=1.0, soft="constant", morton="bbox30"):
    """Matching (Config, oracle Params) pair."""
    cfg = nbody_sim.Config(gravConstant=G, bhTheta=theta, deltaTime=dt, epsilonMultiplier=eps_mult,
                           epsilon=eps, softeningMode=soft, mortonMode=morton)
    P = O.Params(grav_constant=G, delta_time=dt, epsilon_multiplier=eps_mult, bh_theta=theta,
                 epsilon=eps,
                 softening_mode=O.SOFTEN_CONSTANT if soft == "constant" else O.SOFTEN_REFERENCE,
                 morton_mode=O.MORTON_BBOX30 if morton == "bbox30" else O.MORTON_REFERENCE)
    return cfg, P


def gpu_step(cfg, mass, pos, vel, nsteps=1, math="exact", graph=False, stats=False,
             fill_climb=False):
    """Run nsteps on the GPU; return the debug buffers of the last step + state."""
    sim = nbody_sim.Simulation(cfg.with_(mathMode=math), device=0)
    try:
        sim.setBodies(mass, pos, vel)
        sim.setDebug(accel=True, stats=stats, fill_climb=fill_climb)
        if graph:
            sim.useGraph(True)
        sim.step(nsteps)
        sim.synchronize()
        out = {k: sim.debugBuffer(k) for k in BUFS}
        _, out["pos"], out["vel"] = sim.getBodies()
        if stats:
            out["stats"] = sim.walkStats()
        return out
    finally:
        sim.close()


def oracle_step(P, mass, pos, vel, nsteps=1):
    """Oracle state after nsteps-1 steps, then a debug step."""
    if nsteps > 1:
        pos, vel = O.step(mass, pos, vel, P, nsteps=nsteps - 1)
    return O.step_debug(mass, pos, vel, P)


def rel_err(a, b):
    """Per-body relative error |a-b| / |b| of (n,2) vectors."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    num = np.linalg.norm(a - b, axis=1)
    den = np.linalg.norm(b, axis=1)
    return num / np.maximum(den, 1e-300)
