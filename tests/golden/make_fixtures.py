"""Write the golden fixtures of tests/golden/ from the oracle.

The reference ships no test vectors, so these are the oracle's own outputs on
small seeded inputs, committed to pin the oracle (and, through the GPU parity
tests, the HIP path) against regressions:
  karras_fig3.npz   the 8 keys of Karras 2012 Fig. 3 and their tree
  step_n64.npz      one full step of 64 bodies, every intermediate
  step_c1_n1024.npz one step of the small plain configuration's inputs
                    (uniform disk, theta 0.5, eps 1e-3, dt 1e-4) at n = 1024
  step_ref_n512.npz one step in the reference modes (fixed-box keys,
                    mass-dependent softening) on the "default" scenario
Run from the repository root: python tests/golden/make_fixtures.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [ROOT, os.path.join(ROOT, "webgpu-fluid-sim_amd")]

import nbody_sim  # noqa: E402
from oracle import oracle as O  # noqa: E402


def step_fixture(name, mass, pos, vel, P):
    d = O.step_debug(mass, pos, vel, P, threads=1)
    np.savez_compressed(os.path.join(HERE, name), mass=mass, pos_in=pos, vel_in=vel,
                        params=np.array([P.grav_constant, P.delta_time, P.epsilon_multiplier,
                                         P.bh_theta, P.epsilon, P.softening_mode,
                                         P.morton_mode], np.float64),
                        **{k: v for k, v in d.items()})


def main():
    keys = np.array([0b00001, 0b00010, 0b00100, 0b00101, 0b10011, 0b11000, 0b11001, 0b11110],
                    np.uint32)
    left, right, parent = O.build_u32(keys)
    np.savez_compressed(os.path.join(HERE, "karras_fig3.npz"), keys=keys, left=left, right=right,
                        parent=parent)
    rng = np.random.default_rng(64)
    pos = rng.uniform(-1, 1, (64, 2)).astype(np.float32)
    mass = rng.uniform(0.5, 1.5, 64).astype(np.float32)
    vel = rng.normal(0, 0.1, (64, 2)).astype(np.float32)
    step_fixture("step_n64.npz", mass, pos, vel,
                 O.Params(1.0, 1e-3, 1.0, 0.5, 1e-2, O.SOFTEN_CONSTANT, O.MORTON_BBOX30))
    mass, pos, vel = nbody_sim.generate("uniform-disk", 1024, seed=1)
    step_fixture("step_c1_n1024.npz", mass, pos, vel,
                 O.Params(1.0, 1e-4, 1.0, 0.5, 1e-3, O.SOFTEN_CONSTANT, O.MORTON_BBOX30))
    mass, pos, vel = nbody_sim.generate("default", 512, seed=2)
    step_fixture("step_ref_n512.npz", mass, pos, vel,
                 O.Params(1.0, 0.1 / 60, 1.0, 0.6, 1e-3, O.SOFTEN_REFERENCE, O.MORTON_REFERENCE))


if __name__ == "__main__":
    main()
