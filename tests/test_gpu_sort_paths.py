"""The sort's paths give one permutation: a plain handle sorts up to 4096
bodies in one workgroup, larger ones between the previous step's splitters
(buckets; a bucket too large for workgroup memory goes through memory), and
the radix passes otherwise (the first step after setBodies, ensembles,
NBODY_DEBUG_SORT_RADIX).  Sorted keys, body order and every later state are
compared bit for bit with the radix passes forced; the radix passes are the
ones the oracle parity tests pin."""
from __future__ import annotations

import numpy as np
import pytest

import nbody_sim
from helpers import params

pytestmark = pytest.mark.gpu


def run(cfg, bodies, chunks, radix, graph=0, reset=None):
    """Snapshots (keys, body order, pos, vel) after each chunk of steps;
    reset: bodies set again before the last chunk."""
    sim = nbody_sim.Simulation(cfg, device=0)
    try:
        sim.setBodies(*bodies)
        sim.setDebug(sort_radix=radix)
        if graph:
            sim.useGraph(True, graph)
        out = []
        for c, k in enumerate(chunks):
            if reset is not None and c == len(chunks) - 1:
                sim.setBodies(*reset)
            sim.step(k)
            sim.synchronize()
            _, pos, vel = sim.getBodies()
            out.append((sim.debugBuffer("keys"), sim.debugBuffer("body_indices"), pos, vel))
        return out
    finally:
        sim.close()


def same(a, b):
    assert len(a) == len(b)
    for c, (x, y) in enumerate(zip(a, b)):
        for name, u, v in zip(("keys", "body_indices", "pos", "vel"), x, y):
            assert np.array_equal(u, v), (c, name)


# bucket spacings (992, 1984 or 3968 positions, the narrowest giving at most
# 256 buckets, then 3968 up to 1024 buckets) on either side of each change,
# with the encode's 1024-body tiles below 2^18 keys and 4096 from there
@pytest.mark.parametrize("n", [2, 3, 1000, 4096, 4097, 16384, 4096 * 5 + 1, 50000, 131079,
                               253952, 253953, 262143, 262144, 507904, 507905, 524288])
def test_sort_paths_identical(n):
    bodies = nbody_sim.generate("exp-disk", n, seed=7)
    cfg, _ = params(theta=0.7, dt=5e-5)
    cfg = cfg.with_(mathMode="fast")
    chunks = [1, 1, 3]
    same(run(cfg, bodies, chunks, radix=False), run(cfg, bodies, chunks, radix=True))


@pytest.mark.parametrize("morton", ["bbox30", "reference"])
def test_sort_paths_heavy_duplicates(morton):
    """25000 coincident bodies at rest stay coincident (equal pulls), so
    their equal keys fill one bucket far larger than workgroup memory every
    step: the in-memory bucket path."""
    mass, pos, vel = nbody_sim.generate("uniform-disk", 60000, seed=3)
    pos = pos.copy()
    vel = vel.copy()
    pos[:25000] = (0.1, 0.2)
    vel[:25000] = 0.0
    cfg, _ = params(theta=0.5, dt=1e-4, morton=morton)
    cfg = cfg.with_(mathMode="fast")
    chunks = [1, 1, 2]
    a = run(cfg, (mass, pos, vel), chunks, radix=False)
    same(a, run(cfg, (mass, pos, vel), chunks, radix=True))
    keys = a[-1][0]
    assert np.max(np.unique(keys, return_counts=True)[1]) >= 25000


def test_sort_paths_small_clusters():
    """Clusters of 100 coincident bodies: bins of more than 32 equal keys,
    which the in-LDS bucket sort ranks by its LSD passes, not by counting."""
    mass, pos, vel = nbody_sim.generate("uniform-disk", 60000, seed=5)
    pos = pos.copy()
    vel = vel.copy()
    for c in range(50):
        pos[c * 1000:c * 1000 + 100] = pos[c * 1000]
        vel[c * 1000:c * 1000 + 100] = 0.0
    cfg, _ = params(theta=0.5, dt=1e-4)
    cfg = cfg.with_(mathMode="fast")
    a = run(cfg, (mass, pos, vel), [1, 1, 2], radix=False)
    same(a, run(cfg, (mass, pos, vel), [1, 1, 2], radix=True))
    assert np.max(np.unique(a[-1][0], return_counts=True)[1]) >= 100


def test_sort_paths_all_bodies_coincide():
    mass = np.full(20000, 1e-4, np.float32)
    pos = np.zeros((20000, 2), np.float32)
    vel = np.zeros((20000, 2), np.float32)
    cfg, _ = params(theta=0.5, dt=1e-4)
    cfg = cfg.with_(mathMode="fast")
    same(run(cfg, (mass, pos, vel), [1, 2], radix=False),
         run(cfg, (mass, pos, vel), [1, 2], radix=True))


def test_sort_paths_graph_and_new_bodies():
    """Graphs (captured before any splitters exist, then again once they
    do) and bodies set anew between steps keep the radix permutation."""
    bodies = nbody_sim.generate("exp-disk", 50000, seed=11)
    other = nbody_sim.generate("uniform-disk", 50000, seed=12)
    cfg, _ = params(theta=0.7, dt=5e-5)
    cfg = cfg.with_(mathMode="fast")
    chunks = [4, 8, 4]
    ref = run(cfg, bodies, chunks, radix=True, reset=other)
    same(run(cfg, bodies, chunks, radix=False, graph=4, reset=other), ref)
    same(run(cfg, bodies, chunks, radix=False, reset=other), ref)
