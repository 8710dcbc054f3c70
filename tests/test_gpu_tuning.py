"""Launch tunables (workgroup size, stack depth, XCD order, bodies per lane)
never change results; the walks that keep the reference's term order
(variants 0 and 1) agree bit for bit, the pair walk (variant 2, fast mode)
takes the same decisions and sums the same terms in its own fixed order; and
the autotuner picks a configuration that runs."""
import numpy as np
import pytest

import nbody_sim
from helpers import params, rel_err
from nbody_sim.autotune import autotune

pytestmark = pytest.mark.gpu


def run(cfg, bodies, tuning, steps=3, ensemble_sizes=None, stats=True):
    if ensemble_sizes is None:
        sim = nbody_sim.Simulation(cfg)
    else:
        sim = nbody_sim.Ensemble(cfg, ensemble_sizes)
    sim.setBodies(*bodies)
    sim.setTuning(*tuning)
    sim.setDebug(accel=True, stats=stats)
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


def test_pair_walk_same_decisions_own_order():
    """Fast mode: the pair walk (variant 2) takes every decision of the
    reference-order walk (identical counters) and its sums agree to
    rounding; its own order is fixed, so every launch setting and leaving
    the own leaf in as an exact zero (no counters) give identical bits."""
    bodies = nbody_sim.generate("exp-disk", 30000, seed=12, dispersion=0.05, mass_sigma=0.5)
    cfg, _ = params(theta=0.7, dt=5e-5)
    cfg = cfg.with_(mathMode="fast")
    ref = run(cfg, bodies, (256, 128, 1, 0, 1), steps=1)
    pair = run(cfg, bodies, (256, 128, 2, 0, 1), steps=1)
    assert pair[3]["accepted"] == ref[3]["accepted"] and pair[3]["opened"] == ref[3]["opened"]
    assert pair[3]["overflowed"] == 0
    e = rel_err(pair[0], ref[0])
    print(f"pair vs reference order: median rel {np.median(e):.2e}, max {e.max():.2e}")
    # the same terms in another order: rounding-level differences only
    assert np.median(e) < 1e-6 and np.mean(e <= 1e-5) >= 0.999
    for tuning in [(64, 64, 2, 1, 1), (1024, 128, 2, 0, 4), (512, 64, 2, 1, 2)]:
        got = run(cfg, bodies, tuning, steps=1)
        assert np.array_equal(got[0], pair[0]), tuning
        assert np.array_equal(got[1], pair[1]) and np.array_equal(got[2], pair[2]), tuning
    # without counters the own leaf is accepted as an exact zero term
    quiet = run(cfg, bodies, (256, 128, 2, 0, 1), steps=1, stats=False)
    assert np.array_equal(quiet[0], pair[0]) and np.array_equal(quiet[1], pair[1])
    # the exact mode keeps the reference order: variant 2 runs the register-stack walk
    ex1 = run(cfg.with_(mathMode="exact"), bodies, (256, 128, 1, 0, 1), steps=1)
    ex2 = run(cfg.with_(mathMode="exact"), bodies, (256, 128, 2, 0, 1), steps=1)
    assert np.array_equal(ex1[0], ex2[0]) and np.array_equal(ex1[1], ex2[1])


def test_queue_walk_equals_pair_walk():
    """Variant 4, the persistent work-queue walk, runs each 64-body chunk with
    the pair walk's own code, in queue order: the same bits as variant 2,
    plain and ensemble (largest simulations first), with and without
    counters, eager and from a graph."""
    bodies = nbody_sim.generate("exp-disk", 30000, seed=12, dispersion=0.05, mass_sigma=0.5)
    cfg, _ = params(theta=0.7, dt=5e-5)
    cfg = cfg.with_(mathMode="fast")
    for stats in (False, True):
        a = run(cfg, bodies, (256, 128, 2, 1, 1), stats=stats)
        b = run(cfg, bodies, (256, 128, 4, 1, 1), stats=stats)
        for x, y in zip(a[:3], b[:3]):
            assert np.array_equal(x, y), stats
        assert a[3] == b[3]
    sizes = [700, 1, 2048, 333, 64, 9000, 5]
    parts = [nbody_sim.generate("plummer", n, seed=60 + i) for i, n in enumerate(sizes)]
    ens = [np.concatenate([p[i] for p in parts]) for i in range(3)]
    cfg2, _ = params(theta=0.6)
    a = run(cfg2, ens, (256, 128, 2, 1, 1), ensemble_sizes=sizes, stats=False)
    b = run(cfg2, ens, (256, 128, 4, 1, 1), ensemble_sizes=sizes, stats=False)
    for x, y in zip(a[:3], b[:3]):
        assert np.array_equal(x, y)
    out = []
    for graph in (False, True):
        sim = nbody_sim.Simulation(cfg)
        sim.setBodies(*bodies)
        sim.setTuning(walk_variant=4)
        sim.useGraph(graph)
        sim.step(4)
        sim.synchronize()
        out.append(sim.getBodies()[1:])
        sim.close()
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])


def test_pair_walk_reference_softening_and_ensemble():
    """Variant 2 in the reference softening mode (own-leaf compare kept) and
    over an ensemble: the same decisions as the reference-order walk."""
    bodies = nbody_sim.generate("two-clusters", 5000, seed=3)
    cfg, _ = params(theta=0.6, dt=0.1 / 60, soft="reference", morton="reference")
    a = run(cfg, bodies, (256, 128, 1, 0, 1))
    b = run(cfg, bodies, (256, 128, 2, 0, 1))
    assert a[3]["accepted"] == b[3]["accepted"] and a[3]["opened"] == b[3]["opened"]
    assert np.median(rel_err(b[0], a[0])) < 1e-6
    sizes = [700, 1, 2048, 333, 64]
    parts = [nbody_sim.generate("plummer", n, seed=i) for i, n in enumerate(sizes)]
    bodies = [np.concatenate([p[i] for p in parts]) for i in range(3)]
    cfg, _ = params(theta=0.6)
    a = run(cfg, bodies, (256, 128, 1, 0, 1), ensemble_sizes=sizes)
    b = run(cfg, bodies, (128, 128, 2, 1, 1), ensemble_sizes=sizes)
    assert a[3]["accepted"] == b[3]["accepted"] and a[3]["opened"] == b[3]["opened"]
    assert np.median(rel_err(b[0], a[0])) < 1e-6


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
                           lane_bodies=(1, 2), radix=(10, 8))
    assert best[:6] in table and ("radix_bits", best[6]) in table
    assert all(v > 0 for v in table.values())
    t = sim.getTuning()
    assert (t["walk_variant"], t["walk_block"], t["wave_bodies"], t["radix_bits"]) == \
        (best[0], best[1], best[5], best[6])
    _, p0, _ = sim.getBodies()
    assert np.array_equal(p0, bodies[1])  # state restored
    sim.step(2)
    sim.synchronize()
    sim.close()


@pytest.mark.parametrize("morton", ["bbox30", "reference"])
def test_radix_widths_identical(morton):
    """The sort is stable for every digit width (8, 10, 11 bits: 4, 3, 3
    passes for 30-bit keys, 4, 4, 3 for 32-bit keys): the same keys,
    permutation and therefore the same steps."""
    bodies = nbody_sim.generate("exp-disk", 40000, seed=21, dispersion=0.05, mass_sigma=0.5)
    cfg, _ = params(theta=0.7, dt=5e-5, morton=morton)
    out = {}
    for rb in (8, 10, 11):
        sim = nbody_sim.Simulation(cfg)
        sim.setBodies(*bodies)
        sim.setTuning(radix_bits=rb)
        sim.setDebug(accel=True)
        sim.step(3)
        sim.synchronize()
        out[rb] = [sim.debugBuffer(k) for k in ("keys", "body_indices", "accel")]
        out[rb] += list(sim.getBodies()[1:])
        sim.close()
    for rb in (10, 11):
        for a, b in zip(out[8], out[rb]):
            assert np.array_equal(a, b), rb


def test_radix_widths_ensemble():
    """33-bit ensemble keys (5 sims: 3 sim-id bits above the 30-bit Morton code):
    5, 4 and 3 passes at 8, 10 and 11 bits (38-bit keys: test_gpu_ensemble.py)."""
    sizes = [900, 5000, 1, 3000, 64]
    parts = [nbody_sim.generate("plummer", n, seed=40 + i) for i, n in enumerate(sizes)]
    bodies = [np.concatenate([p[i] for p in parts]) for i in range(3)]
    cfg, _ = params(theta=0.6)
    ref = None
    for rb in (8, 10, 11):
        ens = nbody_sim.Ensemble(cfg, sizes)
        ens.setBodies(*bodies)
        ens.setTuning(radix_bits=rb)
        ens.step(2)
        ens.synchronize()
        got = [ens.debugBuffer("keys"), ens.debugBuffer("body_indices"), ens.getBodies()[1]]
        ens.close()
        if ref is None:
            ref = got
        else:
            for a, b in zip(ref, got):
                assert np.array_equal(a, b), rb


def test_bbox_carried_between_steps():
    """Plain bbox30 handles take each step's bounding box from the previous
    walk's accumulators: after eager steps, graph replays, new bodies and an
    adjoint sweep (which moves the positions outside a step) the box equals
    the reduction of the current positions."""
    from oracle import oracle as O

    bodies = nbody_sim.generate("exp-disk", 20000, seed=22, dispersion=0.05, mass_sigma=0.5)
    cfg, _ = params(theta=0.7, dt=5e-5)
    sim = nbody_sim.Simulation(cfg)

    def box_of_step_after(setup):
        setup()
        _, p, _ = sim.getBodies()
        sim.step(1)
        sim.synchronize()
        return sim.debugBuffer("bbox")[0], O.bbox(p)

    def steps(k, graph):
        sim.useGraph(graph)
        sim.step(k)
        sim.synchronize()

    sim.setBodies(*bodies)
    for setup in (lambda: None, lambda: steps(4, False), lambda: steps(5, True),
                  lambda: sim.setBodies(*nbody_sim.generate("uniform-disk", 20000, seed=5)),
                  lambda: steps(3, True)):
        got, want = box_of_step_after(setup)
        assert np.array_equal(got, np.asarray(want, np.float32)), (got, want)
    sim.useGraph(False)
    sim.adjointForward(4, 2)
    sim.adjointBackward(np.ones((20000, 2), np.float32), None)
    got, want = box_of_step_after(lambda: None)
    assert np.array_equal(got, np.asarray(want, np.float32))
    sim.close()


def test_wave_bodies_identical():
    """The pair walks with 64, 32, 16 or 8 bodies per wavefront (and the
    automatic choice) take the same decisions and give the same bits: each
    lane's terms come in its own fixed order."""
    bodies = nbody_sim.generate("uniform-disk", 16384, seed=31)
    cfg, _ = params(theta=0.5)
    cfg = cfg.with_(mathMode="fast")
    ref, ref_counts = None, None
    for variant in (2, 4):
        for wb in (64, 32, 16, 8, 0):
            sim = nbody_sim.Simulation(cfg)
            sim.setTuning(walk_variant=variant, wave_bodies=wb)
            for stats in (False, True):
                sim.setBodies(*bodies)
                sim.setDebug(accel=True, stats=stats)
                sim.walkStats(reset=True)
                sim.step(2)
                sim.synchronize()
                got = [sim.debugBuffer("accel"), sim.getBodies()[1]]
                if ref is None:
                    ref = got
                for a, b in zip(ref, got):
                    assert np.array_equal(a, b), (variant, wb, stats)
                if stats:
                    st = sim.walkStats()
                    counts = (st["accepted"], st["opened"], sim.debugBuffer("walk_counts").copy())
                    if ref_counts is None:
                        ref_counts = counts
                    assert counts[:2] == ref_counts[:2], (variant, wb)
                    assert np.array_equal(counts[2], ref_counts[2]), (variant, wb)
            sim.close()
