"""Ensemble batching: many independent simulations in one sort, one tree
build and one walk per step.  Keys, permutation, topology (under the
sub-tree mapping), accelerations and trajectories must equal plain per-sim
runs bit for bit."""
import numpy as np
import pytest

import nbody_sim
from helpers import params

pytestmark = pytest.mark.gpu


def make_ensemble(sizes, seed):
    rng = np.random.default_rng(seed)
    parts = []
    for s, n in enumerate(sizes):
        scen = "plummer" if rng.random() < 0.5 else "uniform-disk"
        parts.append(nbody_sim.generate(scen, int(n), seed=1000 * seed + s))
    return parts


def plain_run(cfg, m, p, v, steps):
    sim = nbody_sim.Simulation(cfg)
    sim.setBodies(m, p, v)
    sim.setDebug(accel=True)
    sim.step(steps)
    sim.synchronize()
    out = {k: sim.debugBuffer(k) for k in ("keys", "body_indices", "children", "accel")}
    _, out["pos"], out["vel"] = sim.getBodies()
    sim.close()
    return out


@pytest.mark.parametrize("math", ["exact", "fast"])
def test_ensemble_equals_plain_runs(math):
    sizes = [1000, 3, 64, 2500, 700, 1, 4096, 130, 65, 2000]
    parts = make_ensemble(sizes, seed=3)
    cfg, _ = params(theta=0.6, dt=1e-4, eps=1e-3)
    cfg = cfg.with_(mathMode=math)
    steps = 3
    ens = nbody_sim.Ensemble(cfg, sizes)
    ens.setBodies(np.concatenate([p[0] for p in parts]), np.concatenate([p[1] for p in parts]),
                  np.concatenate([p[2] for p in parts]))
    ens.setDebug(accel=True)
    ens.step(steps)
    ens.synchronize()
    keys = ens.debugBuffer("keys")
    bidx = ens.debugBuffer("body_indices")
    roots = ens.debugBuffer("sim_roots")
    acc = ens.debugBuffer("accel")
    _, pos, vel = ens.getBodies()
    ens.close()
    off = np.concatenate([[0], np.cumsum(sizes)])
    for s, n in enumerate(sizes):
        r = plain_run(cfg, *parts[s], steps)
        lo, hi = off[s], off[s + 1]
        assert np.all((keys[lo:hi] >> 30) == s)
        assert np.array_equal((keys[lo:hi] & ((1 << 30) - 1)).astype(np.uint32), r["keys"]), s
        assert np.array_equal(bidx[lo:hi] - lo, r["body_indices"]), s
        assert np.array_equal(acc[lo:hi], r["accel"]), s
        assert np.array_equal(pos[lo:hi], r["pos"]) and np.array_equal(vel[lo:hi], r["vel"]), s
        if n == 1:
            assert roots[s] == off[-1] - 1 + lo
        else:
            assert roots[s] in (lo, hi - 1)


def test_ensemble_topology_maps_to_plain_trees():
    sizes = [300, 512, 77, 1200]
    parts = make_ensemble(sizes, seed=5)
    cfg, _ = params(theta=0.6)
    ens = nbody_sim.Ensemble(cfg.with_(mathMode="exact"), sizes)
    ens.setBodies(*[np.concatenate([p[i] for p in parts]) for i in range(3)])
    ens.step(1)
    ens.synchronize()
    ch = ens.debugBuffer("children")
    roots = ens.debugBuffer("sim_roots")
    ens.close()
    N = sum(sizes)
    off = np.concatenate([[0], np.cumsum(sizes)])
    for s, n in enumerate(sizes):
        r = plain_run(cfg.with_(mathMode="exact"), *parts[s], 0 + 1)
        lo = off[s]

        def to_ens(c):  # plain node id -> ensemble node id
            if c >= n - 1:
                return N - 1 + lo + (c - (n - 1))
            return roots[s] if c == 0 else lo + c

        for k in range(n - 1):
            e = to_ens(k)
            assert tuple(ch[e]) == (to_ens(r["children"][k, 0]), to_ens(r["children"][k, 1])), (s, k)
