"""Parity of the HIP step (through the C ABI) with the CPU oracle.

Bit-exact: sorted Morton keys, the sorted permutation (ties by body index),
the LBVH child/parent arrays, node mass/COM/AABB/length, and -- in the exact
arithmetic mode -- accelerations and the kicked/drifted state.  The fast mode
(hardware rsqrt + FMA, a square-root-free opening test) is held to 1e-5
relative and identical walk counters
per body, except bodies whose opening test lies within rounding of theta,
which must equal the oracle walk at a nudged theta instead.
"""
import numpy as np
import pytest

import nbody_sim
from helpers import gpu_step, oracle_step, params, rel_err

pytestmark = pytest.mark.gpu


def _compare_tree(g, r):
    n = g["keys"].shape[0]
    assert np.array_equal(g["keys"], r["keys"]), "sorted keys"
    assert np.array_equal(g["body_indices"], r["idx"]), "permutation"
    assert np.array_equal(g["children"][:, 0], r["left"]), "left children"
    assert np.array_equal(g["children"][:, 1], r["right"]), "right children"
    assert np.array_equal(g["parent"][1:], r["parent"][1:]), "parents"
    if n > 1:
        assert g["parent"][0] == 0xFFFFFFFF
    assert np.array_equal(g["node_cm"], r["node_cm"]), "node com/mass/length"
    assert np.array_equal(g["node_aabb"], r["node_aabb"]), "node aabb"


@pytest.mark.parametrize("scenario,n,seed", [
    ("uniform-disk", 16384, 1),   # the small plain config's inputs
    ("exp-disk", 20000, 2),
    ("plummer", 5000, 3),
    ("uniform-disk", 1000, 4),
])
def test_step_bit_exact_bbox30(scenario, n, seed):
    mass, pos, vel = nbody_sim.generate(scenario, n, seed=seed, dispersion=0.05, mass_sigma=0.5)
    cfg, P = params(theta=0.5, dt=1e-4, eps=1e-3)
    g = gpu_step(cfg, mass, pos, vel)
    r = oracle_step(P, mass, pos, vel)
    _compare_tree(g, r)
    assert np.array_equal(g["bbox"][0], r["bbox"])
    assert np.array_equal(g["accel"], r["accel"]), "accelerations (exact mode)"
    assert np.array_equal(g["vel"], r["vel"]) and np.array_equal(g["pos"], r["pos"])


@pytest.mark.parametrize("scenario", ["default", "two-clusters", "planetary-disk"])
def test_reference_mode_scenarios(scenario):
    """Reference keys (fixed box) and reference softening on the reference's scenarios."""
    n = 8192
    mass, pos, vel = nbody_sim.generate(scenario, n, seed=7)
    cfg, P = params(theta=0.6, dt=0.1 / 60.0, soft="reference", morton="reference")
    g = gpu_step(cfg, mass, pos, vel)
    r = oracle_step(P, mass, pos, vel)
    _compare_tree(g, r)
    assert np.array_equal(g["accel"], r["accel"])
    assert np.array_equal(g["pos"], r["pos"]) and np.array_equal(g["vel"], r["vel"])


def test_reference_keys_heavy_duplicates():
    """Unit-scale bodies under the fixed [-1000,1000] box collapse onto few
    cells: the sort's stability and the Karras index tie-break decide the tree."""
    mass, pos, vel = nbody_sim.generate("uniform-disk", 5000, seed=11)
    cfg, P = params(theta=0.5, soft="constant", morton="reference")
    g = gpu_step(cfg, mass, pos, vel)
    r = oracle_step(P, mass, pos, vel)
    assert len(np.unique(r["keys"])) < 0.6 * len(r["keys"])
    _compare_tree(g, r)
    assert np.array_equal(g["accel"], r["accel"])


def test_clamped_and_out_of_box_keys():
    rng = np.random.default_rng(5)
    n = 3000
    pos = rng.uniform(-1500, 1500, (n, 2)).astype(np.float32)
    mass = np.full(n, 1.0 / n, np.float32)
    vel = np.zeros_like(pos)
    cfg, P = params(morton="reference", theta=0.7)
    g = gpu_step(cfg, mass, pos, vel)
    r = oracle_step(P, mass, pos, vel)
    _compare_tree(g, r)


def test_multistep_trajectory_bit_exact():
    mass, pos, vel = nbody_sim.generate("exp-disk", 4096, seed=3, dispersion=0.05, mass_sigma=0.5)
    cfg, P = params(theta=0.7, dt=5e-5, eps=1e-3)
    g = gpu_step(cfg, mass, pos, vel, nsteps=10)
    r = oracle_step(P, mass, pos, vel, nsteps=10)
    _compare_tree(g, r)
    assert np.array_equal(g["pos"], r["pos"]) and np.array_equal(g["vel"], r["vel"])


def test_graph_replay_matches_eager():
    mass, pos, vel = nbody_sim.generate("uniform-disk", 10000, seed=9)
    cfg, _ = params(theta=0.5)
    a = gpu_step(cfg, mass, pos, vel, nsteps=5, math="fast", graph=False)
    b = gpu_step(cfg, mass, pos, vel, nsteps=5, math="fast", graph=True)
    for k in ("pos", "vel", "keys", "body_indices", "node_cm"):
        assert np.array_equal(a[k], b[k]), k


# The fast mode's opening test (length |length| / theta^2 against the squared
# softened distance, both rounded in binary32) moves the decision boundary by
# a few ULPs from the reference's length * inv < theta; an opening test closer
# than this to theta (relative, about 30 ULPs) may take the other branch than
# the oracle's.
FLIP_MARGIN = 2e-6
NUDGE = 1e-6  # relative theta nudge that flips exactly such near-boundary tests
# Rounding of the fast sums (FMA, hardware rsqrt, G applied once) relative to
# the sum of the terms' lengths; a body whose terms nearly cancel can exceed
# 1e-5 relative to its own small acceleration through rounding alone.
ROUND_BOUND = 1e-6


def test_fast_mode_pinned_against_oracle():
    """The benchmarked path itself -- fast arithmetic, pair walk, the 800k-body
    exponential disk at theta 0.7 -- against the oracle, body by body:
      * a body whose accepted/opened counts differ from the oracle's took
        another opening decision: it must have an opening test within
        FLIP_MARGIN of theta, and agree (1e-5 and both counts) with the oracle
        walk at theta nudged by -1e-6 or +1e-6 relative;
      * every other body took the oracle's decisions: its acceleration is
        within 1e-5 relative, or within ROUND_BOUND of the sum of its terms'
        lengths (cancelling sums).
    The counters come from the stats-mode walk, whose accelerations must equal
    the benchmarked kernel's bit for bit."""
    from nbody_sim import workloads as W
    from oracle import oracle as O

    w = W.WORKLOADS["plain_800k"]
    mass, pos, vel = w.bodies()
    cfg, P = params(theta=w.theta, dt=w.dt, eps=w.eps)
    sim = nbody_sim.Simulation(cfg.with_(mathMode="fast"))
    sim.setBodies(mass, pos, vel)
    sim.setDebug(accel=True)
    sim.step(1)
    acc = sim.debugBuffer("accel")
    sim.setBodies(mass, pos, vel)
    sim.setDebug(accel=True, stats=True)
    sim.step(1)
    acc_stats = sim.debugBuffer("accel")
    counts = sim.debugBuffer("walk_counts").astype(np.int64)
    sim.close()
    assert np.array_equal(acc, acc_stats), "stats-mode walk changed the sums"

    r = O.step_debug(mass, pos, vel, P)
    tree = (r["idx"], r["node_cm"], r["left"], r["right"], mass, pos)
    _, _, margin, tnorm = O.walk(*tree, P, margin=True)
    e = rel_err(acc, r["accel"])
    flip = np.any(counts != r["stats"][:, :2], axis=1)
    dabs = np.linalg.norm(acc.astype(np.float64) - r["accel"], axis=1)
    loose = ~flip & (e > 1e-5)
    print(f"fast mode 800k: median rel {np.median(e):.2e}; {flip.sum()} bodies with other "
          f"decisions (margins <= {margin[flip].max() if flip.any() else 0:.2e}); "
          f"{loose.sum()} same-decision bodies above 1e-5, |da|/sum|term| <= "
          f"{(dabs[loose] / tnorm[loose]).max() if loose.any() else 0:.2e}")
    assert np.median(e) < 1e-6
    assert flip.mean() < 1e-3 and loose.mean() < 1e-3
    assert np.all(dabs[~flip] <= np.maximum(1e-5 * np.linalg.norm(r["accel"][~flip], axis=1),
                                            ROUND_BOUND * tnorm[~flip]))
    assert np.all(margin[flip] < FLIP_MARGIN), "a body changed decisions far from theta"
    if flip.any():
        ok = np.zeros(flip.sum(), bool)
        for sign in (-1.0, 1.0):
            Pn = O.Params(**{**P.__dict__, "bh_theta": float(np.float32(w.theta) *
                                                          (1.0 + sign * NUDGE))})
            acc_n, st_n = O.walk(*tree, Pn)
            d = np.linalg.norm(acc[flip].astype(np.float64) - acc_n[flip], axis=1)
            ok |= ((d <= np.maximum(1e-5 * np.linalg.norm(acc_n[flip], axis=1),
                                    ROUND_BOUND * tnorm[flip]))
                   & np.all(counts[flip] == st_n[flip][:, :2], axis=1))
        assert np.all(ok), f"{(~ok).sum()} flipped bodies match neither nudged oracle walk"


@pytest.mark.parametrize("math", ["exact", "fast"])
@pytest.mark.parametrize("theta", [0.0, -0.5, float("nan")], ids=["zero", "negative", "nan"])
def test_theta_without_approximation_opens_every_node(theta, math):
    """theta <= 0 or NaN: the reference's test length * inv < theta
    (barnes_hut.wgsl:89) fails for every internal node, so each body opens all
    n - 1 of them and accepts the leaves, a direct sum in tree order: the
    oracle's theta = 0 walk, counter for counter; its accelerations bit for
    bit in the exact arithmetic, and in the fast one (opening keys; the pair
    walk's own summation order) within 1e-5 or ROUND_BOUND of the sum of the
    terms' lengths."""
    from oracle import oracle as O

    mass, pos, vel = nbody_sim.generate("plummer", 700, seed=9)
    cfg, _ = params(theta=theta, eps=1e-2)
    _, P0 = params(theta=0.0, eps=1e-2)
    sim = nbody_sim.Simulation(cfg.with_(mathMode=math))
    try:
        sim.setBodies(mass, pos, vel)
        sim.setDebug(accel=True, stats=True)
        sim.step(1)
        sim.synchronize()
        acc = sim.debugBuffer("accel")
        totals = sim.walkStats()
        # per-body counters: the fast arithmetic's pair walk
        counts = sim.debugBuffer("walk_counts").astype(np.int64) if math == "fast" else None
    finally:
        sim.close()
    r = oracle_step(P0, mass, pos, vel)
    assert np.all(r["stats"][:, 1] == len(mass) - 1)
    assert totals["accepted"] == int(r["stats"][:, 0].sum())
    assert totals["opened"] == int(r["stats"][:, 1].sum())
    if counts is not None:
        assert np.array_equal(counts, r["stats"][:, :2])
    if math == "exact":
        assert np.array_equal(acc, r["accel"])
    else:
        tree = (r["idx"], r["node_cm"], r["left"], r["right"], mass, pos)
        _, _, _, tnorm = O.walk(*tree, P0, margin=True)
        dabs = np.linalg.norm(acc.astype(np.float64) - r["accel"], axis=1)
        assert np.all(dabs <= np.maximum(1e-5 * np.linalg.norm(r["accel"], axis=1),
                                         ROUND_BOUND * tnorm))


@pytest.mark.parametrize("n", [1, 2, 3, 63, 64, 65, 257])
def test_small_sizes(n):
    rng = np.random.default_rng(n)
    pos = rng.uniform(-1, 1, (n, 2)).astype(np.float32)
    mass = rng.uniform(0.5, 1.5, n).astype(np.float32)
    vel = rng.normal(0, 0.1, (n, 2)).astype(np.float32)
    cfg, P = params(theta=0.5, dt=1e-3, eps=1e-2)
    g = gpu_step(cfg, mass, pos, vel, nsteps=2)
    r = oracle_step(P, mass, pos, vel, nsteps=2)
    assert np.array_equal(g["pos"], r["pos"]) and np.array_equal(g["vel"], r["vel"])
    assert np.array_equal(g["accel"], r["accel"])


@pytest.mark.parametrize("climb", [False, True])
def test_fill_upper_levels(climb):
    """Five super-blocks of 32768 leaves: the fill's upper levels, resolved
    from the build's node lists in workgroup memory (default) and by the
    global-counter climb, equal the oracle bit for bit."""
    mass, pos, vel = nbody_sim.generate("exp-disk", 150000, seed=7, dispersion=0.05,
                                        mass_sigma=0.5)
    cfg, P = params(theta=0.6, dt=1e-4, eps=1e-3)
    g = gpu_step(cfg, mass, pos, vel, nsteps=2, fill_climb=climb)
    r = oracle_step(P, mass, pos, vel, nsteps=2)
    _compare_tree(g, r)
    assert np.array_equal(g["accel"], r["accel"])
    assert np.array_equal(g["pos"], r["pos"]) and np.array_equal(g["vel"], r["vel"])


def test_step_bit_exact_800k():
    """The bench's 800k-body exponential disk (25 super-blocks, a 27-level
    top chain resolved beside the phase-2 workgroups): keys, permutation,
    tree, accelerations and the new state equal the oracle bit for bit on the
    second step (the first step's fill reset the counters and flags)."""
    mass, pos, vel = nbody_sim.generate("exp-disk", 800_000, seed=1, scale=0.2, dispersion=0.05,
                                        mass_sigma=0.5)
    cfg, P = params(theta=0.7, dt=5e-5, eps=1e-3)
    g = gpu_step(cfg, mass, pos, vel, nsteps=2)
    r = oracle_step(P, mass, pos, vel, nsteps=2)
    _compare_tree(g, r)
    assert np.array_equal(g["accel"], r["accel"])
    assert np.array_equal(g["pos"], r["pos"]) and np.array_equal(g["vel"], r["vel"])


def test_step_bit_exact_4M_collision():
    """The 4M-body collision (128 super-blocks of 32768 leaves, the top list
    resolved beside 128 phase-2 workgroups): tree, accelerations and state
    equal the oracle bit for bit on the second step."""
    from nbody_sim import workloads as W

    w = W.WORKLOADS["collision_4M"]
    mass, pos, vel = w.bodies()
    cfg, P = params(theta=w.theta, dt=w.dt, eps=w.eps)
    g = gpu_step(cfg, mass, pos, vel, nsteps=2)
    r = oracle_step(P, mass, pos, vel, nsteps=2)
    _compare_tree(g, r)
    assert np.array_equal(g["accel"], r["accel"])
    assert np.array_equal(g["pos"], r["pos"]) and np.array_equal(g["vel"], r["vel"])


def test_own_leaf_term_with_tiny_softening():
    """Constant softening so small that eps^-3 overflows binary32 (eps = 1e-14)
    with masses small enough that m * eps^-3 would not: the fast walk must
    leave the own leaf out instead of adding m * inf * 0 = NaN."""
    mass, pos, vel = nbody_sim.generate("uniform-disk", 3000, seed=12)
    mass = np.full_like(mass, 1e-13)
    cfg, P = params(theta=0.5, dt=1e-4, eps=1e-14)
    g = gpu_step(cfg, mass, pos, vel, math="fast")
    r = oracle_step(P, mass, pos, vel)
    assert np.all(np.isfinite(g["accel"])) and np.all(np.isfinite(g["pos"]))
    assert np.median(rel_err(g["accel"], r["accel"])) < 1e-6


def test_coincident_bodies():
    n = 500
    pos = np.zeros((n, 2), np.float32)
    pos[:250] = 0.25
    mass = np.full(n, 1.0 / n, np.float32)
    vel = np.zeros_like(pos)
    cfg, P = params(theta=0.5, eps=1e-2)
    g = gpu_step(cfg, mass, pos, vel)
    r = oracle_step(P, mass, pos, vel)
    _compare_tree(g, r)
    assert np.array_equal(g["accel"], r["accel"])


def test_deep_tree():
    """A tree 46 levels deep (plain trees of up to 2^26 bodies stay below
    57): 30 bodies whose bbox30 keys are single bits, each peeling one level
    off the path to the origin's key cell, a body at the box's far corner,
    and 2^16 bodies inside the origin's key cell at distinct positions (equal
    keys: a sub-tree of 16 levels on the sorted position),
    so that the walks open nodes all the way down.  Exact mode: tree,
    accelerations and state bit for bit.  Fast mode (the pair walk and its
    fixed stack): no overflow, the benchmarked walk's sums equal the
    stats-mode walk's, and the accelerations within 1e-5 of the oracle's."""
    m = 1 << 16
    rng = np.random.default_rng(11)
    cluster = rng.random((m, 2)) * 1.4e-5  # every coordinate rounds to key cell 0
    # equal keys order by body index: index the cell's bodies along a finer
    # Morton curve, so that its sub-tree is spatially coherent
    q = (cluster / 1.4e-5 * 65535).astype(np.uint64)
    code = np.zeros(m, np.uint64)
    for b in range(16):
        code |= ((q[:, 0] >> np.uint64(b)) & np.uint64(1)) << np.uint64(2 * b + 1)
        code |= ((q[:, 1] >> np.uint64(b)) & np.uint64(1)) << np.uint64(2 * b)
    cluster = cluster[np.argsort(code, kind="stable")]
    single = ([(2.0 ** b / 32767, 0.0) for b in range(15)] +
              [(0.0, 2.0 ** b / 32767) for b in range(15)])
    pos = np.vstack([cluster, single, [(1.0, 1.0)]]).astype(np.float32)
    n = len(pos)
    mass = np.full(n, 1.0 / n, np.float32)
    vel = np.zeros_like(pos)
    cfg, P = params(theta=0.5, eps=1e-9)
    g = gpu_step(cfg, mass, pos, vel)
    r = oracle_step(P, mass, pos, vel)
    _compare_tree(g, r)
    assert len(np.unique(g["keys"])) == 32  # the cell's key, 30 single bits, the corner
    depth, frontier = 0, np.array([0])
    while frontier.size:
        ch = g["children"][frontier].ravel()
        frontier = ch[ch < n - 1]
        depth += 1
    assert depth >= 46, depth
    assert np.array_equal(g["accel"], r["accel"])
    assert np.array_equal(g["pos"], r["pos"]) and np.array_equal(g["vel"], r["vel"])
    f = gpu_step(cfg, mass, pos, vel, math="fast")
    fs = gpu_step(cfg, mass, pos, vel, math="fast", stats=True)
    assert fs["stats"]["overflowed"] == 0
    assert np.array_equal(f["accel"], fs["accel"])
    e = rel_err(f["accel"], r["accel"])
    assert np.median(e) < 1e-6 and np.mean(e > 1e-5) < 1e-3


def test_stats_match_oracle():
    mass, pos, vel = nbody_sim.generate("uniform-disk", 8192, seed=2)
    cfg, P = params(theta=0.5)
    g = gpu_step(cfg, mass, pos, vel, stats=True)
    r = oracle_step(P, mass, pos, vel)
    assert g["stats"]["accepted"] == int(r["stats"][:, 0].sum())
    assert g["stats"]["opened"] == int(r["stats"][:, 1].sum())
    assert g["stats"]["overflowed"] == 0


GOLDEN = sorted(__import__("glob").glob(__import__("os").path.join(
    __import__("os").path.dirname(__file__), "golden", "step_*.npz")))


@pytest.mark.parametrize("path", GOLDEN, ids=[p.rsplit("/", 1)[-1] for p in GOLDEN])
def test_gpu_matches_golden_fixture(path):
    f = np.load(path)
    g, dt, em, th, eps, sm, mm = f["params"]
    cfg = nbody_sim.Config(gravConstant=float(g), deltaTime=float(dt), epsilonMultiplier=float(em),
                           bhTheta=float(th), epsilon=float(eps),
                           softeningMode="constant" if int(sm) == 1 else "reference",
                           mortonMode="bbox30" if int(mm) == 1 else "reference")
    out = gpu_step(cfg, f["mass"], f["pos_in"], f["vel_in"])
    ref = {k: f[k] for k in f.files}
    _compare_tree(out, {"keys": ref["keys"], "idx": ref["idx"], "left": ref["left"],
                        "right": ref["right"], "parent": ref["parent"],
                        "node_cm": ref["node_cm"], "node_aabb": ref["node_aabb"]})
    assert np.array_equal(out["accel"], ref["accel"])
    assert np.array_equal(out["pos"], ref["pos"]) and np.array_equal(out["vel"], ref["vel"])
