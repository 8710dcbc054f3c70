"""Host-surface features around the step: radius multipliers (the reference's
radiusMultiplier buffer, simulation.ts:17, 87, 240-241, 321-339, 382-384),
state snapshots (save, load, resume), and multi-step graph replay."""
import json

import numpy as np
import pytest

import nbody_sim
from helpers import params

pytestmark = pytest.mark.gpu


def test_radius_multiplier_follows_set_scenario():
    sim = nbody_sim.Simulation(nbody_sim.getDefaultConfig())
    sim.setNumBodies(1000)
    sim.setScenario("planetary-disk")
    r = sim.getBuffers()["radiusMultiplier"]
    assert r[:4].tolist() == [12.0, 5.0, 5.0, 1.0] and np.all(r[3:] == 1.0)
    sim.setScenario("default")
    assert np.all(sim.getBuffers()["radiusMultiplier"] == 1.0)
    custom = np.linspace(0.5, 2.0, 1000).astype(np.float32)
    sim.setRadiusMultiplier(custom)
    sim.step(2)
    assert np.array_equal(sim.getRadiusMultiplier(), custom)  # the step leaves them alone
    sim.setBodies(*nbody_sim.generate("uniform-disk", 500, seed=1))
    assert np.all(sim.getRadiusMultiplier() == 1.0)  # new bodies: reset to 1
    with pytest.raises(ValueError):
        sim.setRadiusMultiplier(np.ones(3, np.float32))
    sim.close()


@pytest.mark.parametrize("math,graph", [("fast", 5), ("exact", 0)])
def test_snapshot_resume_equals_continuous_run(tmp_path, math, graph):
    """k steps, snapshot, load into a fresh handle, k more steps: bit-equal to
    2k continuous steps; the files are plain .npy arrays."""
    k = 6
    bodies = nbody_sim.generate("exp-disk", 30000, seed=8, dispersion=0.05, mass_sigma=0.5)
    cfg, _ = params(theta=0.7, dt=5e-5)
    cfg = cfg.with_(mathMode=math)
    a = nbody_sim.Simulation(cfg)
    a.setBodies(*bodies)
    a.setRadiusMultiplier(np.full(30000, 2.5, np.float32))
    a.useGraph(graph > 0, steps=max(graph, 1))
    a.step(k)
    a.synchronize()
    snap = tmp_path / "snap"
    a.saveSnapshot(snap)
    perm_a = a.debugBuffer("body_indices")
    a.step(k)
    _, pos_a, vel_a = a.getBodies()
    a.close()

    m = np.load(snap / "mass.npy")
    assert m.dtype == np.float32 and np.array_equal(m, bodies[0])
    assert np.load(snap / "pos.npy").shape == (30000, 2)
    assert np.array_equal(np.load(snap / "perm.npy"), perm_a)
    meta = json.loads((snap / "meta.json").read_text())
    assert meta["n"] == 30000 and meta["steps"] == k and meta["n_per_sim"] == [30000]

    b = nbody_sim.Simulation(nbody_sim.getDefaultConfig())  # config comes from the snapshot
    b.loadSnapshot(snap)
    assert (b.config.mathMode, b.config.softeningMode, b.config.mortonMode) == (math, "constant",
                                                                                 "bbox30")
    assert np.float32(b.config.bhTheta) == np.float32(0.7)
    assert np.float32(b.config.deltaTime) == np.float32(5e-5)
    assert b.getStepCount() == k
    assert np.array_equal(b.debugBuffer("body_indices"), perm_a)
    assert np.all(b.getRadiusMultiplier() == 2.5)
    b.step(k)
    _, pos_b, vel_b = b.getBodies()
    assert b.getStepCount() == 2 * k
    b.close()
    assert np.array_equal(pos_a, pos_b) and np.array_equal(vel_a, vel_b)


def test_ensemble_snapshot_layout_checked(tmp_path):
    sizes = [300, 1200, 64]
    parts = [nbody_sim.generate("plummer", n, seed=60 + i) for i, n in enumerate(sizes)]
    cfg, _ = params(theta=0.6)
    e = nbody_sim.Ensemble(cfg, sizes)
    e.setBodies(*[np.concatenate([p[i] for p in parts]) for i in range(3)])
    e.step(3)
    e.saveSnapshot(tmp_path / "ens")
    e.step(2)
    _, pos_a, _ = e.getBodies()
    e.close()
    f = nbody_sim.Ensemble(cfg, sizes)
    f.loadSnapshot(tmp_path / "ens")
    f.step(2)
    _, pos_b, _ = f.getBodies()
    f.close()
    assert np.array_equal(pos_a, pos_b)
    g = nbody_sim.Ensemble(cfg, [1200, 300, 64])  # same total, other layout
    with pytest.raises(nbody_sim.NbodyError):
        g.loadSnapshot(tmp_path / "ens")
    g.close()
    s = nbody_sim.Simulation(cfg)
    with pytest.raises(nbody_sim.NbodyError):
        s.loadSnapshot(tmp_path / "nowhere")
    s.close()


@pytest.mark.parametrize("per_graph,steps", [(3, 7), (4, 8), (1, 5), (100, 100)])
def test_multi_step_graphs_equal_eager(per_graph, steps):
    """Graphs of several steps (and the remainder launched one by one) give
    the eager steps' bits."""
    bodies = nbody_sim.generate("uniform-disk", 16384, seed=1)
    cfg, _ = params(theta=0.5)
    out = []
    for g in (0, per_graph):
        sim = nbody_sim.Simulation(cfg.with_(mathMode="fast"))
        sim.setBodies(*bodies)
        sim.useGraph(g > 0, steps=max(g, 1))
        sim.step(steps)
        sim.synchronize()
        out.append(sim.getBodies()[1:] + (sim.getStepCount(),))
        sim.close()
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])
    assert out[0][2] == out[1][2] == steps


@pytest.mark.parametrize("n", [16384, 200000])
def test_prepared_graph_equals_eager(n):
    """Warm-up steps launched one by one, the graph captured by prepareGraph
    (after the sort's splitters exist), then replays: the eager steps' bits."""
    bodies = nbody_sim.generate("exp-disk", n, seed=2)
    cfg, _ = params(theta=0.7)
    out = []
    for g in (0, 10):
        sim = nbody_sim.Simulation(cfg.with_(mathMode="fast"))
        sim.setBodies(*bodies)
        sim.useGraph(g > 0, steps=max(g, 1))
        sim.step(3)
        sim.prepareGraph()
        sim.prepareGraph()  # still valid: a no-op
        sim.step(20)
        sim.synchronize()
        out.append(sim.getBodies()[1:] + (sim.getStepCount(),))
        sim.close()
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])
    assert out[0][2] == out[1][2] == 23
