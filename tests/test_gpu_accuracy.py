"""Barnes-Hut force error of the GPU walk against the FP64 direct sum, for
every benchmark workload (median, p99 and max relative error over sampled
bodies), plus the library's own FP64 direct sum (nbody_direct_accel_f64,
which the bench uses to state the error) against the oracle's.

The error is the method's (monopole Barnes-Hut with the reference's opening
test length * rsqrt(d^2 + eps^2) < theta, barnes_hut.wgsl:89-94), not the
arithmetic's: the oracle's binary32 walk gives the same numbers.  The bounds
are the measured values with a 10 % margin (the inputs are seeded, so the
values only move if the method changes).  At theta 0.5-0.7 the reference
criterion leaves a p99 of 0.17-0.55 on these 2D disks, far above 1e-2."""
import numpy as np
import pytest

import nbody_sim
from nbody_sim import workloads as W
from oracle import oracle as O

pytestmark = pytest.mark.gpu

# workload: (median bound, p99 bound), measured values + 10 %
BOUNDS = {
    "plain_16k": (0.025, 0.37),
    "plain_800k": (0.043, 0.23),
    "collision_4M": (0.044, 0.19),
    "adjoint_65k": (0.036, 0.49),
    "ensemble_256": (0.040, 0.60),
}
NT = 4096  # sampled target bodies


def bh_error(sim, mass, pos, vel, targets, ref):
    sim.setBodies(mass, pos, vel)
    sim.setDebug(accel=True)
    sim.step(1)
    acc = sim.debugBuffer("accel")[targets]
    e = np.linalg.norm(acc - ref, axis=1) / np.linalg.norm(ref, axis=1)
    return e


def targets_of(n, seed=0):
    return np.sort(np.random.default_rng(seed).choice(n, min(n, NT), replace=False))


@pytest.mark.parametrize("name", ["plain_16k", "plain_800k", "collision_4M", "adjoint_65k"])
def test_force_error_vs_fp64_direct_sum(name):
    w = W.WORKLOADS[name]
    mass, pos, vel = w.bodies()
    sim = nbody_sim.Simulation(w.config("fast"))
    tg = targets_of(w.n)
    P = O.Params(grav_constant=1.0, delta_time=w.dt, bh_theta=w.theta, epsilon=w.eps,
                 softening_mode=O.SOFTEN_CONSTANT)
    ref = O.direct_sum_f64(pos, mass, P, targets=tg)
    sim.setBodies(mass, pos, vel)
    lib_ref = sim.directAccelF64(tg)
    assert np.max(np.abs(lib_ref - ref) / np.linalg.norm(ref, axis=1)[:, None]) < 1e-10
    e = bh_error(sim, mass, pos, vel, tg, ref)
    sim.close()
    med, p99 = np.median(e), np.percentile(e, 99)
    print(f"force error {name}: n={w.n} theta={w.theta} median {med:.3e} p99 {p99:.3e} "
          f"max {e.max():.3e}")
    assert np.all(np.isfinite(e))
    assert med <= BOUNDS[name][0] and p99 <= BOUNDS[name][1]


def test_force_error_ensemble_members():
    """The 256-sim ensemble's error over every 16th member (a plain run of a
    member equals its ensemble slice bit for bit)."""
    sizes, _, parts = W.ensemble_members()
    cfg = W.ensemble_config("fast")
    errs = []
    for s in range(0, len(sizes), 16):
        mass, pos, vel = parts[s]
        sim = nbody_sim.Simulation(cfg)
        tg = targets_of(int(sizes[s]), seed=s)
        sim.setBodies(mass, pos, vel)
        ref = sim.directAccelF64(tg)
        errs.append(bh_error(sim, mass, pos, vel, tg, ref))
        sim.close()
    e = np.concatenate(errs)
    med, p99 = np.median(e), np.percentile(e, 99)
    print(f"force error ensemble_256 (16 members, {e.size} bodies): median {med:.3e} "
          f"p99 {p99:.3e} max {e.max():.3e}")
    assert med <= BOUNDS["ensemble_256"][0] and p99 <= BOUNDS["ensemble_256"][1]


@pytest.mark.parametrize("soft", ["constant", "reference"])
def test_library_direct_sum_matches_oracle(soft):
    """nbody_direct_accel_f64 against the oracle's FP64 direct sum, both
    softening modes, all targets of a small planetary disk (unequal masses)."""
    mass, pos, vel = nbody_sim.generate("planetary-disk", 3000, seed=2)
    cfg = nbody_sim.Config(softeningMode=soft, epsilon=1e-2, deltaTime=1e-3)
    P = O.Params(grav_constant=1.0, delta_time=1e-3, epsilon=1e-2, epsilon_multiplier=1.0,
                 softening_mode=O.SOFTEN_CONSTANT if soft == "constant" else O.SOFTEN_REFERENCE)
    sim = nbody_sim.Simulation(cfg)
    sim.setBodies(mass, pos, vel)
    got = sim.directAccelF64()
    sub = sim.directAccelF64([5, 0, 2999, 5])
    sim.close()
    ref = O.direct_sum_f64(pos, mass, P)
    scale = np.linalg.norm(ref, axis=1)[:, None]
    assert np.max(np.abs(got - ref) / scale) < 1e-10
    assert np.allclose(sub, got[[5, 0, 2999, 5]], rtol=1e-12, atol=0)
