"""The five benchmark workloads, seeded (Philox through nbody_generate).

    plain_16k        uniform unit disk, 16,384 bodies, v=0, m=1/N; theta 0.5,
                     eps 1e-3, dt 1e-4, 100 steps (seed 1)
    plain_800k       exponential disk, scale 0.2, circular v + N(0, 0.05 v_c),
                     lognormal(0.5) masses summing to 1; theta 0.7, eps 1e-3,
                     dt 5e-5, 1,000 steps (seed 1)
    collision_4M     two exponential disks of 2,097,152 bodies (scale 0.15) at
                     (+-0.6, 0) with bulk v (0, +-0.3); theta 0.7, eps 5e-4,
                     dt 2e-5, 500 steps (seed 3)
    ensemble_256     256 simulations, N_s log-uniform in [1,024, 65,536],
                     each a uniform disk or a projected Plummer disk (scale
                     0.1), v=0, m=1/N_s; theta 0.6, eps 1e-3, dt 1e-4, 200
                     steps (seed 4)
    adjoint_65k      the exponential disk at 65,536 bodies; theta 0.6, eps
                     1e-3, dt 1e-4, 200 steps with a checkpoint every 20
                     (seed 5)

All run with constant softening and 30-bit Morton keys over the per-step
bounding box.  The reference has no such workloads (its scenarios are
setScenario's three, simulation.ts:243-361); these are the sizes and
distributions the measurements are quoted on.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from .config import Config
from .simulation import generate


@dataclass(frozen=True)
class Workload:
    name: str
    scenario: str
    n: int
    theta: float
    eps: float
    dt: float
    steps: int
    seed: int
    ic: dict = field(default_factory=dict)

    def config(self, mathMode: str = "fast", **kw) -> Config:
        return Config(bhTheta=self.theta, deltaTime=self.dt, epsilon=self.eps,
                      softeningMode="constant", mortonMode="bbox30", mathMode=mathMode, **kw)

    def bodies(self, n: int | None = None):
        """(mass, pos, vel) of the workload (n bodies: the same distribution)."""
        return generate(self.scenario, self.n if n is None else n, seed=self.seed, **self.ic)


EXP_DISK = dict(scale=0.2, dispersion=0.05, mass_sigma=0.5)

WORKLOADS = {
    "plain_16k": Workload("plain_16k", "uniform-disk", 16384, 0.5, 1e-3, 1e-4, 100, 1),
    "plain_800k": Workload("plain_800k", "exp-disk", 800_000, 0.7, 1e-3, 5e-5, 1000, 1, EXP_DISK),
    "collision_4M": Workload("collision_4M", "collision", 4_194_304, 0.7, 5e-4, 2e-5, 500, 3,
                             dict(scale=0.15)),
    "adjoint_65k": Workload("adjoint_65k", "exp-disk", 65536, 0.6, 1e-3, 1e-4, 200, 5, EXP_DISK),
}

ENSEMBLE = dict(num_sims=256, theta=0.6, eps=1e-3, dt=1e-4, steps=200, seed=4)


def ensemble_config(mathMode: str = "fast") -> Config:
    return Config(bhTheta=ENSEMBLE["theta"], deltaTime=ENSEMBLE["dt"], epsilon=ENSEMBLE["eps"],
                  softeningMode="constant", mortonMode="bbox30", mathMode=mathMode)


def ensemble_members(num_sims: int = 256, seed: int = 4):
    """(sizes, kinds, parts): N_s = round(exp(U[ln 1024, ln 65536])), each sim a
    Plummer disk (kind True) or a uniform disk, bodies of sim i seeded 1000 + i."""
    rng = np.random.default_rng(seed)
    sizes = np.round(np.exp(rng.uniform(np.log(1024), np.log(65536), num_sims))).astype(np.int64)
    kinds = rng.random(num_sims) < 0.5
    parts = [generate("plummer" if k else "uniform-disk", int(nn), seed=1000 + i)
             for i, (nn, k) in enumerate(zip(sizes, kinds))]
    return sizes, kinds, parts


def concat(parts):
    """Concatenate per-sim (mass, pos, vel) in sim order (the ensemble layout)."""
    return tuple(np.concatenate([p[i] for p in parts]) for i in range(3))
