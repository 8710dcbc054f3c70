"""The hand-written adjoint (through the C ABI) against FP64 references:
torch autograd of the frozen-decision Barnes-Hut trajectory, FP64 central
finite differences of it (h = 1e-4 on positions and velocities), and
autograd of an O(N^2) direct-sum trajectory at N = 4096 (theta = 0, where the
Barnes-Hut walk is the direct sum)."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

import nbody_sim  # noqa: E402
from frozen_bh import FrozenTrajectory, direct_trajectory  # noqa: E402
from helpers import params  # noqa: E402
from nbody_sim.adjoint import trajectory  # noqa: E402

pytestmark = pytest.mark.gpu


def rel(a, b):
    a = np.asarray(a, np.float64).reshape(-1)
    b = np.asarray(b, np.float64).reshape(-1)
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def gpu_grads(cfg, mass, pos, vel, steps, every, wx, wv, lane_bodies=None):
    sim = nbody_sim.Simulation(cfg)
    sim.setBodies(mass, pos, vel)
    if lane_bodies is not None:
        sim.setTuning(lane_bodies=lane_bodies)
    sim.adjointForward(steps, every)
    _, pT, vT = sim.getBodies()
    g = sim.adjointBackward(wx, wv)
    _, pT2, _ = sim.getBodies()
    sim.close()
    assert np.array_equal(pT, pT2)  # final state restored
    return g, pT, vT


@pytest.mark.parametrize("math", ["exact", "fast"])
def test_adjoint_matches_frozen_fp64_autograd(math):
    n, steps, every = 2048, 10, 4
    mass, pos, vel = nbody_sim.generate("exp-disk", n, seed=6, dispersion=0.05, mass_sigma=0.5)
    cfg, P = params(theta=0.6, dt=1e-4, eps=1e-3)
    cfg = cfg.with_(mathMode=math)
    rng = np.random.default_rng(0)
    wx = rng.normal(size=(n, 2)).astype(np.float32)
    wv = (0.3 * rng.normal(size=(n, 2))).astype(np.float32)
    (gx, gv, gm), pT, _ = gpu_grads(cfg, mass, pos, vel, steps, every, wx, wv)

    F = FrozenTrajectory(mass, pos, vel, P, steps)
    if math == "exact":
        assert np.array_equal(pT, F.final32[0])  # same decisions as the oracle
    x0 = torch.from_numpy(pos.astype(np.float64)).requires_grad_()
    v0 = torch.from_numpy(vel.astype(np.float64)).requires_grad_()
    m0 = torch.from_numpy(mass.astype(np.float64)).requires_grad_()
    xt, vt = F.run(x0, v0, m0)
    L = (torch.from_numpy(wx.astype(np.float64)) * xt).sum() + \
        (torch.from_numpy(wv.astype(np.float64)) * vt).sum()
    rx, rv, rm = (t.numpy() for t in torch.autograd.grad(L, (x0, v0, m0)))
    ex, ev, em = rel(gx, rx), rel(gv, rv), rel(gm, rm)
    print(f"adjoint vs FP64 autograd ({math}): pos {ex:.2e} vel {ev:.2e} mass {em:.2e}")
    tol = 1e-4 if math == "exact" else 1e-3
    assert ex < tol and ev < tol and em < tol


def test_adjoint_lane_bodies_agree():
    """EXACT mode (the register-stack VJP walk): bodies per lane change only
    the order of the node-cotangent atomics."""
    n, steps, every = 5000, 6, 3
    mass, pos, vel = nbody_sim.generate("plummer", n, seed=9)
    cfg, _ = params(theta=0.6, dt=1e-4, eps=1e-3)
    cfg = cfg.with_(mathMode="exact")
    rng = np.random.default_rng(1)
    wx = rng.normal(size=(n, 2)).astype(np.float32)
    wv = rng.normal(size=(n, 2)).astype(np.float32)
    ref, pT, _ = gpu_grads(cfg, mass, pos, vel, steps, every, wx, wv, lane_bodies=1)
    for lb in (2, 4):
        got, pT2, _ = gpu_grads(cfg, mass, pos, vel, steps, every, wx, wv, lane_bodies=lb)
        assert np.array_equal(pT, pT2)
        for a, b in zip(got, ref):
            assert rel(a, b) < 1e-5, lb


def test_adjoint_fast_matches_exact_many_nodes():
    """FAST mode (the pair-walk VJP with per-workgroup cotangent tables) at a
    small theta, where a workgroup's wavefronts accept more distinct nodes
    than its table holds (the rest go straight to memory): the gradients
    agree with EXACT mode's register-stack VJP walk up to the two modes'
    arithmetic and their few theta-boundary decisions."""
    n, steps, every = 20000, 2, 2
    mass, pos, vel = nbody_sim.generate("exp-disk", n, seed=11, dispersion=0.05, mass_sigma=0.5)
    cfg, _ = params(theta=0.3, dt=1e-4, eps=1e-3)
    rng = np.random.default_rng(2)
    wx = rng.normal(size=(n, 2)).astype(np.float32)
    wv = rng.normal(size=(n, 2)).astype(np.float32)
    fast, _, _ = gpu_grads(cfg.with_(mathMode="fast"), mass, pos, vel, steps, every, wx, wv)
    exact, _, _ = gpu_grads(cfg.with_(mathMode="exact"), mass, pos, vel, steps, every, wx, wv)
    errs = [rel(a, b) for a, b in zip(fast, exact)]
    print("fast vs exact adjoint (rel L2) x, v, m:", errs)
    assert all(np.all(np.isfinite(a)) for a in fast)
    assert errs[0] < 1e-3 and errs[1] < 1e-5 and errs[2] < 1e-3


def test_adjoint_central_differences_64_coordinates():
    """FP64 central differences (h = 1e-4; masses: 1e-4 of their value) of the
    frozen trajectory at 64 random coordinates, and its torch autograd; the
    loss is the mean squared distance to target positions from a run with
    initial velocities perturbed by N(0, 0.01)."""
    n, steps, every = 1024, 20, 6
    mass, pos, vel = nbody_sim.generate("exp-disk", n, seed=8, dispersion=0.05, mass_sigma=0.5)
    cfg, P = params(theta=0.6, dt=1e-4, eps=1e-3)
    cfg = cfg.with_(mathMode="exact")
    rng = np.random.default_rng(3)
    vel_t = (vel + rng.normal(0, 0.01, vel.shape)).astype(np.float32)
    sim = nbody_sim.Simulation(cfg)
    sim.setBodies(mass, pos, vel_t)
    sim.step(steps)
    _, target, _ = sim.getBodies()
    sim.close()

    F = FrozenTrajectory(mass, pos, vel, P, steps)
    tgt = torch.from_numpy(target.astype(np.float64))
    base = [torch.from_numpy(a.astype(np.float64)) for a in (pos, vel, mass)]

    def loss(x, v, m):
        xt, _ = F.run(x, v, m)
        return (((xt - tgt) ** 2).sum(1)).mean()

    leaves = [b.clone().requires_grad_() for b in base]
    L = loss(*leaves)
    auto = [g.reshape(-1).numpy() for g in torch.autograd.grad(L, leaves)]
    # The seed dL/dx_T comes from the FP64 trajectory the differences are taken
    # of: x_T - target is ~1e-4, so binary32 rounding of x_T would perturb it.
    with torch.no_grad():
        xT64, _ = F.run(*base)
    gx_T = ((2.0 / n) * (xT64 - tgt)).numpy()
    (gx, gv, gm), _, _ = gpu_grads(cfg, mass, pos, vel, steps, every, gx_T.astype(np.float32), None)
    adj = [gx.reshape(-1), gv.reshape(-1), gm]

    picks = [(0, k) for k in rng.choice(2 * n, 24, replace=False)] + \
            [(1, k) for k in rng.choice(2 * n, 24, replace=False)] + \
            [(2, k) for k in rng.choice(n, 16, replace=False)]
    err_fd = [0.0, 0.0, 0.0]    # plain central differences, h as stated
    err_rich = [0.0, 0.0, 0.0]  # Richardson-extrapolated differences
    err_adj = [0.0, 0.0, 0.0]
    for i, k in picks:
        flat = base[i].reshape(-1)

        def fd(h):
            args = [b.clone() for b in base]
            args[i].reshape(-1)[k] += h
            lp = loss(*args).item()
            args[i].reshape(-1)[k] -= 2 * h
            return (lp - loss(*args).item()) / (2 * h)

        h = 1e-4 if i < 2 else 1e-4 * float(flat[k])
        f1, f2 = fd(h), fd(h / 2)
        rich = (4 * f2 - f1) / 3  # Richardson: removes the O(h^2) term
        sc = np.abs(auto[i]).max()
        err_fd[i] = max(err_fd[i], abs(f1 - auto[i][k]) / sc)
        err_rich[i] = max(err_rich[i], abs(rich - auto[i][k]) / sc)
        err_adj[i] = max(err_adj[i], abs(adj[i][k] - rich) / sc)
    rel_all = [rel(adj[i], auto[i]) for i in range(3)]
    # With softening 1e-3, h = 1e-4 on a position is a tenth of the interaction
    # scale, so the plain difference carries an O(h^2) truncation error of a
    # few percent on some coordinates; the Richardson value removes it.
    print(f"plain FD (h=1e-4) vs autograd (worst/max|g|) x,v,m: {err_fd}")
    print(f"Richardson FD vs autograd (worst/max|g|) x,v,m: {err_rich}")
    print(f"adjoint vs Richardson FD (worst/max|g|) x,v,m: {err_adj}")
    print(f"adjoint vs autograd (rel L2) x,v,m: {rel_all}")
    assert max(err_rich) < 1e-4          # the FD reference itself is pinned
    assert max(err_adj) < 1e-3           # the adjoint against FP64 differences
    assert max(rel_all) < 1e-3           # and against FP64 autograd


def test_adjoint_theta0_matches_direct_sum_autograd_n4096():
    n, steps = 4096, 3
    mass, pos, vel = nbody_sim.generate("uniform-disk", n, seed=9)
    vel = (0.1 * np.random.default_rng(1).normal(size=(n, 2))).astype(np.float32)
    cfg, _ = params(theta=0.0, dt=1e-4, eps=1e-3)
    rng = np.random.default_rng(2)
    wx = rng.normal(size=(n, 2)).astype(np.float32)
    (gx, gv, gm), _, _ = gpu_grads(cfg.with_(mathMode="fast"), mass, pos, vel, steps, 2, wx, None)
    x0 = torch.from_numpy(pos.astype(np.float64)).requires_grad_()
    v0 = torch.from_numpy(vel.astype(np.float64)).requires_grad_()
    m0 = torch.from_numpy(mass.astype(np.float64)).requires_grad_()
    xt, _ = direct_trajectory(x0, v0, m0, 1.0, float(np.float32(1e-3)), float(np.float32(1e-4)),
                              steps)
    L = (torch.from_numpy(wx.astype(np.float64)) * xt).sum()
    rx, rv, rm = (t.numpy() for t in torch.autograd.grad(L, (x0, v0, m0)))
    ex, ev, em = rel(gx, rx), rel(gv, rv), rel(gm, rm)
    print(f"adjoint (theta=0) vs direct-sum autograd: pos {ex:.2e} vel {ev:.2e} mass {em:.2e}")
    assert ex < 1e-3 and ev < 1e-3 and em < 1e-3


def test_autograd_function_wrapper():
    n = 512
    mass, pos, vel = nbody_sim.generate("plummer", n, seed=4)
    cfg, _ = params(theta=0.6, dt=1e-4, eps=1e-2)
    sim = nbody_sim.Simulation(cfg)
    x0 = torch.from_numpy(pos).requires_grad_()
    v0 = torch.from_numpy(vel).requires_grad_()
    m0 = torch.from_numpy(mass).requires_grad_()
    xT, vT = trajectory(x0, v0, m0, sim, 5, 2)
    L = (xT ** 2).sum() + (vT ** 2).sum()
    L.backward()
    gx, gv, gm = sim.adjointBackward(2 * xT.detach().numpy(), 2 * vT.detach().numpy())
    sim.close()
    # node cotangents are summed with float atomics: equal up to summation order
    assert np.allclose(x0.grad.numpy(), gx, rtol=1e-4, atol=1e-6 * np.abs(gx).max())
    assert np.allclose(m0.grad.numpy(), gm, rtol=1e-4, atol=1e-6 * np.abs(gm).max())
    assert np.all(np.isfinite(gv))


def test_adjoint_full_size_c5_segment():
    """The C5 configuration at its full size (65,536-body exponential disk,
    theta 0.6, fast mode, the bench's bodies) over a 5-step segment with
    checkpoints every 2 steps, against torch autograd of the FP64
    frozen-decision trajectory."""
    from nbody_sim import workloads as W
    from oracle import oracle as O

    w = W.WORKLOADS["adjoint_65k"]
    mass, pos, vel = w.bodies()
    steps, every = 5, 2
    cfg = w.config("fast")
    P = O.Params(grav_constant=1.0, delta_time=w.dt, bh_theta=w.theta, epsilon=w.eps,
                 softening_mode=O.SOFTEN_CONSTANT, morton_mode=O.MORTON_BBOX30)
    rng = np.random.default_rng(7)
    wx = rng.normal(size=(w.n, 2)).astype(np.float32)
    wv = (0.3 * rng.normal(size=(w.n, 2))).astype(np.float32)
    (gx, gv, gm), _, _ = gpu_grads(cfg, mass, pos, vel, steps, every, wx, wv)
    F = FrozenTrajectory(mass, pos, vel, P, steps)
    x0 = torch.from_numpy(pos.astype(np.float64)).requires_grad_()
    v0 = torch.from_numpy(vel.astype(np.float64)).requires_grad_()
    m0 = torch.from_numpy(mass.astype(np.float64)).requires_grad_()
    xt, vt = F.run(x0, v0, m0)
    L = (torch.from_numpy(wx.astype(np.float64)) * xt).sum() + \
        (torch.from_numpy(wv.astype(np.float64)) * vt).sum()
    rx, rv, rm = (t.numpy() for t in torch.autograd.grad(L, (x0, v0, m0)))
    ex, ev, em = rel(gx, rx), rel(gv, rv), rel(gm, rm)
    print(f"adjoint 65k x 5 steps vs FP64 autograd: pos {ex:.2e} vel {ev:.2e} mass {em:.2e}")
    assert ex < 1e-3 and ev < 1e-3 and em < 1e-3


def test_adjoint_200_steps_checkpoint_spacing():
    """The bench's 200-step, 65,536-body run: gradients with a checkpoint every
    20 steps equal those with a checkpoint every step, within the spread of
    two runs with the same spacing (the node cotangents are summed with float
    atomics, whose order varies)."""
    from nbody_sim import workloads as W

    w = W.WORKLOADS["adjoint_65k"]
    mass, pos, vel = w.bodies()
    cfg = w.config("fast")
    g = np.random.default_rng(0).normal(size=(w.n, 2)).astype(np.float32)
    runs = [gpu_grads(cfg, mass, pos, vel, 200, every, g, None)[0] for every in (20, 20, 1)]
    for k in range(3):
        assert np.all(np.isfinite(runs[0][k]))
        floor = rel(runs[1][k], runs[0][k])
        diff = rel(runs[2][k], runs[0][k])
        print(f"gradient {('pos', 'vel', 'mass')[k]}: every 20 twice {floor:.2e}, "
              f"every 20 vs every 1 {diff:.2e}")
        assert diff < 1e-4 and diff <= 10 * floor + 1e-7
