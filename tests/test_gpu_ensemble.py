"""Ensemble batching: many independent simulations in one sort, one tree
build and one walk per step.  Keys, permutation, topology (under the
sub-tree mapping), accelerations and trajectories must equal plain per-sim
runs bit for bit."""
import numpy as np
import pytest

import nbody_sim
from helpers import params
from nbody_sim import workloads as W

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


def _c4_ensemble(math, radix_bits=None, steps=2):
    sizes, _, parts = W.ensemble_members()
    ens = nbody_sim.Ensemble(W.ensemble_config(math), sizes)
    ens.setBodies(*W.concat(parts))
    if radix_bits is not None:
        ens.setTuning(radix_bits=radix_bits)
    ens.setDebug(accel=True)
    ens.step(steps)
    ens.synchronize()
    out = {k: ens.debugBuffer(k) for k in ("keys", "body_indices", "accel", "sim_roots")}
    _, out["pos"], out["vel"] = ens.getBodies()
    ens.close()
    return sizes, parts, out


@pytest.mark.parametrize("math", ["fast", "exact"])
def test_c4_ensemble_256_sims_equals_plain_runs(math):
    """The benchmarked ensemble itself: 256 sims of 1,024-65,536 bodies
    (4.65 M in all), 38-bit keys (8 sim-id bits above the 30-bit Morton
    code; 4 passes of 10 bits), 2 steps: sorted keys, permutation,
    accelerations and state equal 256 plain runs bit for bit."""
    sizes, parts, e = _c4_ensemble(math)
    keys = e["keys"]
    assert keys.dtype == np.uint64 and int(keys.max() >> 30) == 255  # 38-bit keys
    assert np.all(np.diff(keys.astype(np.uint64)) >= 0)
    assert np.array_equal(np.sort(e["body_indices"]), np.arange(keys.shape[0], dtype=np.uint32))
    off = np.concatenate([[0], np.cumsum(sizes)])
    cfg = W.ensemble_config(math)
    for s in range(len(sizes)):
        r = plain_run(cfg, *parts[s], 2)
        lo, hi = off[s], off[s + 1]
        assert np.all((keys[lo:hi] >> 30) == s), s
        assert np.array_equal((keys[lo:hi] & ((1 << 30) - 1)).astype(np.uint32), r["keys"]), s
        assert np.array_equal(e["body_indices"][lo:hi] - lo, r["body_indices"]), s
        assert np.array_equal(e["accel"][lo:hi], r["accel"]), s
        assert np.array_equal(e["pos"][lo:hi], r["pos"]), s
        assert np.array_equal(e["vel"][lo:hi], r["vel"]), s
        assert e["sim_roots"][s] in (lo, hi - 1), s


def test_c4_ensemble_radix_widths_identical():
    """38-bit ensemble keys sorted with 8-, 10- and 11-bit digits (5, 4 and 4
    passes) give the same keys, permutation and state."""
    ref = None
    for rb in (10, 8, 11):
        _, _, e = _c4_ensemble("fast", radix_bits=rb, steps=1)
        got = [e["keys"], e["body_indices"], e["pos"], e["vel"]]
        if ref is None:
            ref = got
        else:
            for a, b in zip(ref, got):
                assert np.array_equal(a, b), rb
