"""The allocation-free step (nbody_step_static) against the handle step
(nbody_create + nbody_set_bodies + nbody_step): same config, same bodies,
bit-identical positions and velocities, over two calls in a row on one GPU,
the second with fewer bodies (the fixed scratch keeps the first call's
contents, which the second must not see).  Device buffers for the bodies are
the caller's (torch tensors here); the library allocates nothing."""
import ctypes

import numpy as np
import pytest
import torch

import nbody_sim
from nbody_sim import _lib

CASES = [
    # (scenario, n, steps, theta, mode, seed)
    ("exp-disk", 60000, 3, 0.7, "fast", 11),
    ("uniform-disk", 5000, 2, 0.5, "fast", 12),
]


def handle_run(cfg, mass, pos, vel, steps):
    sim = nbody_sim.Simulation(cfg)
    try:
        sim.setBodies(mass, pos, vel)
        sim.step(steps)
        sim.synchronize()
        _, p, v = sim.getBodies()
        return p, v
    finally:
        sim.close()


def static_run(cfg, mass, pos, vel, steps):
    dev = torch.device("cuda:0")
    m = torch.from_numpy(mass).to(dev)
    p = torch.from_numpy(np.ascontiguousarray(pos)).to(dev)
    v = torch.from_numpy(np.ascontiguousarray(vel)).to(dev)
    torch.cuda.synchronize()
    c = cfg.to_c()
    stream = torch.cuda.current_stream().cuda_stream
    rc = _lib.lib().nbody_step_static(ctypes.byref(c), mass.shape[0], m.data_ptr(), p.data_ptr(),
                                       v.data_ptr(), steps, ctypes.c_void_p(stream))
    _lib.check(rc)
    torch.cuda.synchronize()
    return p.cpu().numpy(), v.cpu().numpy()


@pytest.mark.gpu
def test_static_step_matches_handle_step_twice():
    for scen, n, steps, theta, mode, seed in CASES:
        ic = dict(scale=0.2, dispersion=0.05, mass_sigma=0.5) if scen == "exp-disk" else {}
        mass, pos, vel = nbody_sim.generate(scen, n, seed=seed, **ic)
        cfg = nbody_sim.Config(bhTheta=theta, deltaTime=1e-4, epsilon=1e-3,
                               softeningMode="constant", mortonMode="bbox30", mathMode=mode)
        print(f"[no-alloc] {scen} n={n} steps={steps}", flush=True)
        want_p, want_v = handle_run(cfg, mass, pos, vel, steps)
        got_p, got_v = static_run(cfg, mass, pos, vel, steps)
        assert np.array_equal(got_p, want_p), f"{scen} n={n}: positions differ"
        assert np.array_equal(got_v, want_v), f"{scen} n={n}: velocities differ"
        assert not np.array_equal(got_p, pos)  # the bodies moved


@pytest.mark.gpu
def test_static_step_rejects_too_many_bodies():
    c = nbody_sim.Config().to_c()
    n = _lib.STATIC_MAX_BODIES + 1
    rc = _lib.lib().nbody_step_static(ctypes.byref(c), n, None, None, None, 1, None)
    assert rc == _lib.ERR_CAPACITY
    assert "65536" in _lib.last_error(None)
