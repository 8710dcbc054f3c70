"""Launch tunables (walk variant, workgroup size, stack depth, XCD order,
bodies per lane) never change results, and the autotuner picks a
configuration that runs."""
import numpy as np
import pytest

import nbody_sim
from helpers import params
from nbody_sim.autotune import autotune

pytestmark = pytest.mark.gpu


def run(cfg, bodies, tuning, steps=3, ensemble_sizes=None):
    if ensemble_sizes is None:
        sim = nbody_sim.Simulation(cfg)
    else:
        sim = nbody_sim.Ensemble(cfg, ensemble_sizes)
    sim.setBodies(*bodies)
    sim.setTuning(*tuning)
    sim.setDebug(accel=True, stats=True)
    sim.step(steps)
    sim.synchronize()
    acc = sim.debugBuffer("accel")
    st = sim.walkStats()
    _, pos, vel = sim.getBodies()
    sim.close()
    return acc, pos, vel, st


@pytest.mark.parametrize("math", ["exact", "fast"])
def test_walk_variants_identical(math):
    bodies = nbody_sim.generate("exp-disk", 30000, seed=12, dispersion=0.05, mass_sigma=0.5)
    cfg, _ = params(theta=0.7, dt=5e-5)
    cfg = cfg.with_(mathMode=math)
    ref = run(cfg, bodies, (256, 128, 0))
    for tuning in [(64, 64, 0, 1), (1024, 256, 0, 0), (256, 128, 1, 0, 1), (64, 64, 1, 1, 1),
                   (512, 128, 1, 1, 2), (256, 128, 1, 0, 2), (128, 64, 1, 1, 4),
                   (1024, 128, 1, 0, 4)]:
        got = run(cfg, bodies, tuning)
        assert np.array_equal(got[0], ref[0]), tuning
        assert np.array_equal(got[1], ref[1]) and np.array_equal(got[2], ref[2]), tuning
        assert got[3]["accepted"] == ref[3]["accepted"] and got[3]["opened"] == ref[3]["opened"]


def test_register_stack_walk_ensemble_and_reference_modes():
    sizes = [700, 1, 2048, 333, 64]
    parts = [nbody_sim.generate("plummer", n, seed=i) for i, n in enumerate(sizes)]
    bodies = [np.concatenate([p[i] for p in parts]) for i in range(3)]
    cfg, _ = params(theta=0.6)
    a = run(cfg.with_(mathMode="exact"), bodies, (256, 128, 0), ensemble_sizes=sizes)
    for lb in (1, 2, 4):
        b = run(cfg.with_(mathMode="exact"), bodies, (128, 128, 1, 1, lb), ensemble_sizes=sizes)
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]), lb
    bodies = nbody_sim.generate("two-clusters", 5000, seed=3)
    cfg, _ = params(theta=0.6, dt=0.1 / 60, soft="reference", morton="reference")
    a = run(cfg, bodies, (256, 128, 0))
    for lb in (1, 2, 4):
        b = run(cfg, bodies, (256, 128, 1, 1, lb))
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]), lb


def test_autotune_picks_a_working_configuration():
    bodies = nbody_sim.generate("exp-disk", 50000, seed=13, dispersion=0.05, mass_sigma=0.5)
    cfg, _ = params(theta=0.7, dt=5e-5)
    sim = nbody_sim.Simulation(cfg)
    sim.setBodies(*bodies)
    best, table = autotune(sim, steps=2, blocks=(128, 256), stacks=(128,), xcd=(1,),
                           lane_bodies=(1, 2))
    assert best in table and all(v > 0 for v in table.values())
    _, p0, _ = sim.getBodies()
    assert np.array_equal(p0, bodies[1])  # state restored
    sim.step(2)
    sim.synchronize()
    sim.close()
