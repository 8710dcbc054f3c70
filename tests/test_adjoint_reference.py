"""CPU pins of the FP64 frozen-decision reference used by the adjoint tests:
its forces reproduce the oracle's binary32 Barnes-Hut forces, and its torch
autograd gradient agrees with central finite differences."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

import nbody_sim  # noqa: E402
from frozen_bh import FrozenTrajectory, direct_trajectory  # noqa: E402
from oracle import oracle as O  # noqa: E402


def setup(n=512, steps=3, theta=0.6, seed=2):
    mass, pos, vel = nbody_sim.generate("exp-disk", n, seed=seed, dispersion=0.05, mass_sigma=0.5)
    P = O.Params(grav_constant=1.0, delta_time=1e-3, bh_theta=theta, epsilon=1e-2,
                 softening_mode=O.SOFTEN_CONSTANT, morton_mode=O.MORTON_BBOX30)
    return mass, pos, vel, P, FrozenTrajectory(mass, pos, vel, P, steps)


def test_frozen_forces_reproduce_oracle_forces():
    mass, pos, vel, P, F = setup()
    a = F.forces(0, torch.from_numpy(pos.astype(np.float64)),
                 torch.from_numpy(mass.astype(np.float64))).numpy()
    a32 = F.steps[0]["accel32"]
    e = np.linalg.norm(a - a32, axis=1) / np.linalg.norm(a32, axis=1)
    assert np.median(e) < 1e-6 and e.max() < 1e-4


def test_frozen_trajectory_tracks_oracle():
    mass, pos, vel, P, F = setup(steps=4)
    x, v = F.run(torch.from_numpy(pos.astype(np.float64)), torch.from_numpy(vel.astype(np.float64)),
                 torch.from_numpy(mass.astype(np.float64)))
    assert np.allclose(x.numpy(), F.final32[0], atol=1e-5)


def test_frozen_gradient_matches_central_differences():
    mass, pos, vel, P, F = setup(n=256, steps=3)
    x0 = torch.from_numpy(pos.astype(np.float64)).requires_grad_()
    v0 = torch.from_numpy(vel.astype(np.float64)).requires_grad_()
    m0 = torch.from_numpy(mass.astype(np.float64)).requires_grad_()
    w = torch.from_numpy(np.random.default_rng(0).normal(size=(256, 2)))

    def loss(x, v, m):
        xt, vt = F.run(x, v, m)
        return (w * xt).sum() + 0.3 * (w * vt).sum()

    L = loss(x0, v0, m0)
    gx, gv, gm = torch.autograd.grad(L, (x0, v0, m0))
    rng = np.random.default_rng(1)
    for which, base, g in (("x", x0, gx), ("v", v0, gv), ("m", m0, gm)):
        flat = base.detach().reshape(-1)
        for k in rng.choice(flat.numel(), 8, replace=False):
            # h = 1e-4 for positions and velocities; masses (~1/n) get 1e-4 of their value
            h = 1e-4 if which != "m" else 1e-4 * float(flat[k])
            d = torch.zeros_like(flat)
            d[k] = h
            args = [x0.detach(), v0.detach(), m0.detach()]
            i = "xvm".index(which)
            args[i] = (flat + d).reshape(base.shape)
            lp = loss(*args).item()
            args[i] = (flat - d).reshape(base.shape)
            lm = loss(*args).item()
            fd = (lp - lm) / (2 * h)
            # central differences carry an O(h^2) truncation error (~1e-6 here)
            assert abs(fd - g.reshape(-1)[k].item()) <= 2e-5 * max(1.0, abs(fd)), (which, k)


def test_direct_trajectory_matches_oracle_theta_zero():
    mass, pos, vel = nbody_sim.generate("uniform-disk", 300, seed=4)
    P = O.Params(grav_constant=1.0, delta_time=1e-3, bh_theta=0.0, epsilon=1e-2,
                 softening_mode=O.SOFTEN_CONSTANT, morton_mode=O.MORTON_BBOX30)
    p32, _ = O.step(mass, pos, vel, P, nsteps=3)
    x, _ = direct_trajectory(torch.from_numpy(pos.astype(np.float64)),
                             torch.from_numpy(vel.astype(np.float64)),
                             torch.from_numpy(mass.astype(np.float64)), 1.0,
                             float(np.float32(1e-2)), float(np.float32(1e-3)), 3)
    assert np.allclose(x.numpy(), p32, atol=1e-6)
