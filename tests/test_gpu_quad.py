"""The quadrupole expansion (constant softening, fast arithmetic) against the
oracle: the nodes' second moments (bit for bit for nodes of at most 8 leaves,
summed directly on both sides; within 1e-5 of the trace above, where the GPU
differences binary64 prefix sums), the walk body by body (same opening
decisions and counters as the oracle except near-theta flips; accelerations
within 1e-5, or rounding of the terms), and the force error against the FP64
direct sum at the bench workload (p99 <= 1e-2 at theta 0.4)."""
import numpy as np
import pytest

import nbody_sim
from frozen_bh import node_ranges
from helpers import params, rel_err
from oracle import oracle as O

pytestmark = pytest.mark.gpu

ROUND_BOUND = 1e-6
FLIP_MARGIN = 2e-6
NUDGE = 1e-6


def quad_step(cfg, mass, pos, vel, stats=False):
    sim = nbody_sim.Simulation(cfg.with_(mathMode="fast", expansion="quadrupole"))
    sim.setBodies(mass, pos, vel)
    sim.setDebug(accel=True, stats=stats)
    sim.step(1)
    out = {k: sim.debugBuffer(k) for k in ("node_quad", "accel", "body_indices")}
    if stats:
        out["counts"] = sim.debugBuffer("walk_counts").astype(np.int64)
    sim.close()
    return out


@pytest.mark.parametrize("scenario,n,theta", [("exp-disk", 200000, 0.5),
                                              ("plummer", 20000, 0.4),
                                              ("uniform-disk", 5, 0.5),
                                              ("uniform-disk", 1, 0.5)])
def test_quad_walk_against_oracle(scenario, n, theta):
    mass, pos, vel = nbody_sim.generate(scenario, n, seed=17, dispersion=0.05, mass_sigma=0.5)
    cfg, P = params(theta=theta, dt=1e-4, eps=1e-3)
    g = quad_step(cfg, mass, pos, vel, stats=True)
    g_plain = quad_step(cfg, mass, pos, vel)
    assert np.array_equal(g["accel"], g_plain["accel"]), "stats-mode walk changed the sums"
    r = O.step_debug(mass, pos, vel, P)
    q = O.node_quad(r["idx"], mass, pos, r["left"], r["right"])
    # moments
    lo, hi = node_ranges(r["left"], r["right"], n)
    small = (hi - lo + 1) <= 8
    assert np.array_equal(g["node_quad"][small], q[small])
    tr = np.abs(q[:, 0]) + np.abs(q[:, 2])
    assert np.all(np.abs(g["node_quad"] - q) <= 1e-5 * tr[:, None] + 1e-30)
    if n == 1:
        assert np.all(g["accel"] == 0)
        return
    # walk, body by body
    tree = (r["idx"], r["node_cm"], r["left"], r["right"], mass, pos)
    acc, st, margin, tnorm = O.walk(*tree, P, margin=True, quad=q)
    flip = np.any(g["counts"] != st[:, :2], axis=1)
    e = rel_err(g["accel"], acc)
    dabs = np.linalg.norm(g["accel"].astype(np.float64) - acc, axis=1)
    print(f"quadrupole {scenario} n={n}: median rel {np.median(e):.2e}, {flip.sum()} flips, "
          f"max same-decision |da|/sum|term| {(dabs[~flip] / np.maximum(tnorm[~flip], 1e-30)).max():.2e}")
    assert np.median(e) < 1e-6
    assert flip.mean() < 1e-3
    assert np.all(dabs[~flip] <= np.maximum(1e-5 * np.linalg.norm(acc[~flip], axis=1),
                                            ROUND_BOUND * tnorm[~flip]))
    assert np.all(margin[flip] < FLIP_MARGIN)
    if flip.any():
        ok = np.zeros(flip.sum(), bool)
        for sign in (-1.0, 1.0):
            Pn = O.Params(**{**P.__dict__, "bh_theta": float(np.float32(theta) * (1 + sign * NUDGE))})
            acc_n, st_n = O.walk(*tree, Pn, quad=q)
            d = np.linalg.norm(g["accel"][flip].astype(np.float64) - acc_n[flip], axis=1)
            ok |= ((d <= np.maximum(1e-5 * np.linalg.norm(acc_n[flip], axis=1),
                                    ROUND_BOUND * tnorm[flip]))
                   & np.all(g["counts"][flip] == st_n[flip][:, :2], axis=1))
        assert np.all(ok)


def test_quad_meets_force_error_target_800k():
    """The 800k-body exponential disk: at theta 0.4 the quadrupole expansion's
    p99 force error against the FP64 direct sum is below 1e-2 (the monopole
    needs theta ~0.19 for that), and at theta 0.7 it is 4x below the
    monopole's."""
    from nbody_sim import workloads as W

    w = W.WORKLOADS["plain_800k"]
    mass, pos, vel = w.bodies()
    tg = np.sort(np.random.default_rng(0).choice(w.n, 4096, replace=False))
    P = O.Params(grav_constant=1.0, delta_time=w.dt, bh_theta=w.theta, epsilon=w.eps,
                 softening_mode=O.SOFTEN_CONSTANT)
    ref = O.direct_sum_f64(pos, mass, P, targets=tg)
    p99 = {}
    for expn, th in (("quadrupole", 0.4), ("quadrupole", 0.7), ("monopole", 0.7)):
        sim = nbody_sim.Simulation(w.config("fast").with_(bhTheta=th, expansion=expn))
        sim.setBodies(mass, pos, vel)
        sim.setDebug(accel=True)
        sim.step(1)
        acc = sim.debugBuffer("accel")[tg]
        sim.close()
        e = np.linalg.norm(acc - ref, axis=1) / np.linalg.norm(ref, axis=1)
        p99[(expn, th)] = np.percentile(e, 99)
        print(f"{expn} theta {th}: median {np.median(e):.2e} p99 {p99[(expn, th)]:.2e}")
    assert p99[("quadrupole", 0.4)] <= 1e-2
    assert p99[("quadrupole", 0.7)] < p99[("monopole", 0.7)] / 4


def test_quad_graph_replay_and_ensemble():
    """Graph replay gives the eager steps' bits; an ensemble member equals its
    plain run within 1e-5 (the moments of nodes above 8 leaves come from
    prefix sums over the whole batch)."""
    sizes = [3000, 700, 12000]
    parts = [nbody_sim.generate("plummer", n, seed=90 + i) for i, n in enumerate(sizes)]
    cfg, _ = params(theta=0.5)
    cfg = cfg.with_(mathMode="fast", expansion="quadrupole")
    out = []
    for graph in (0, 3):
        sim = nbody_sim.Simulation(cfg)
        sim.setBodies(*parts[2])
        sim.useGraph(graph > 0, steps=max(graph, 1))
        sim.step(6)
        out.append(sim.getBodies()[1])
        sim.close()
    assert np.array_equal(out[0], out[1])
    ens = nbody_sim.Ensemble(cfg, sizes)
    ens.setBodies(*[np.concatenate([p[i] for p in parts]) for i in range(3)])
    ens.setDebug(accel=True)
    ens.step(1)
    acc = ens.debugBuffer("accel")
    ens.close()
    off = np.concatenate([[0], np.cumsum(sizes)])
    for s in range(3):
        sim = nbody_sim.Simulation(cfg)
        sim.setBodies(*parts[s])
        sim.setDebug(accel=True)
        sim.step(1)
        a = sim.debugBuffer("accel")
        sim.close()
        e = rel_err(acc[off[s]:off[s + 1]], a)
        assert np.percentile(e, 99.9) < 1e-5, s


def test_quad_config_rules():
    """The quadrupole expansion needs constant softening and the fast mode;
    the adjoint and the allocation-free step keep the monopole."""
    with pytest.raises(nbody_sim.NbodyError):
        nbody_sim.Simulation(nbody_sim.Config(expansion="quadrupole"))  # reference softening
    with pytest.raises(nbody_sim.NbodyError):
        nbody_sim.Simulation(nbody_sim.Config(expansion="quadrupole", softeningMode="constant",
                                              mathMode="exact"))
    sim = nbody_sim.Simulation(nbody_sim.Config(expansion="quadrupole", softeningMode="constant"))
    sim.setBodies(*nbody_sim.generate("uniform-disk", 100, seed=1))
    with pytest.raises(nbody_sim.NbodyError):
        sim.adjointForward(2, 1)
    sim.close()
