"""Barnes-Hut force error of the GPU walk against the FP64 direct sum, for
the shapes of every configuration (median and p99 relative error over
sampled bodies).  The error is the method's (monopole Barnes-Hut with the
reference's opening test), not the arithmetic's: the same numbers come out
of the oracle's binary32 walk."""
import numpy as np
import pytest

import nbody_sim
from oracle import oracle as O

pytestmark = pytest.mark.gpu

CASES = {
    # name: (scenario, n, theta, eps, dt, ic)
    "plain_small_16k": ("uniform-disk", 16384, 0.5, 1e-3, 1e-4, {}),
    "plain_800k": ("exp-disk", 800_000, 0.7, 1e-3, 5e-5,
                   dict(scale=0.2, dispersion=0.05, mass_sigma=0.5)),
    "collision_4M": ("collision", 4_194_304, 0.7, 5e-4, 2e-5, dict(scale=0.15)),
    "ensemble_member_plummer_8k": ("plummer", 8192, 0.6, 1e-3, 1e-4, {}),
    "adjoint_65k": ("exp-disk", 65536, 0.6, 1e-3, 1e-4,
                    dict(scale=0.2, dispersion=0.05, mass_sigma=0.5)),
}


@pytest.mark.parametrize("name", list(CASES))
def test_force_error_vs_fp64_direct_sum(name):
    scen, n, theta, eps, dt, ic = CASES[name]
    mass, pos, vel = nbody_sim.generate(scen, n, seed=1, **ic)
    cfg = nbody_sim.Config(bhTheta=theta, deltaTime=dt, epsilon=eps, softeningMode="constant",
                           mortonMode="bbox30", mathMode="fast")
    sim = nbody_sim.Simulation(cfg)
    sim.setBodies(mass, pos, vel)
    sim.setDebug(accel=True)
    sim.step(1)
    acc = sim.debugBuffer("accel")
    sim.close()
    nt = min(n, 4096)
    tg = np.sort(np.random.default_rng(0).choice(n, nt, replace=False))
    P = O.Params(grav_constant=1.0, delta_time=dt, bh_theta=theta, epsilon=eps,
                 softening_mode=O.SOFTEN_CONSTANT)
    ref = O.direct_sum_f64(pos, mass, P, targets=tg)
    e = np.linalg.norm(acc[tg] - ref, axis=1) / np.linalg.norm(ref, axis=1)
    print(f"force error {name}: n={n} theta={theta} median {np.median(e):.3e} "
          f"p99 {np.percentile(e, 99):.3e} max {e.max():.3e}")
    # theta = 0 limit and sanity: errors are finite and bounded by the method
    assert np.all(np.isfinite(e))
    assert np.median(e) < 0.05
