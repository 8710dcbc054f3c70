"""Pins of the CPU oracle (CPU only).

The reference ships no tests, fixtures or golden vectors, so each oracle
function is checked against what the mathematics fixes: closed-form known
answers, independent implementations (numpy stable sort, a recursive
brute-force radix tree, an FP64 direct sum), invariants, and the Karras 2012
Fig. 3 example.
"""
import numpy as np
import pytest

from oracle import oracle as O


def interleave_magic(x, y):
    """Independent interleave: x -> even bits, y -> odd bits (lbvh.wgsl:59-65)."""
    def part(v):
        v = v.astype(np.uint64) & 0xFFFF
        v = (v | (v << 8)) & 0x00FF00FF
        v = (v | (v << 4)) & 0x0F0F0F0F
        v = (v | (v << 2)) & 0x33333333
        v = (v | (v << 1)) & 0x55555555
        return v
    return (part(x) | (part(y) << 1)).astype(np.uint32)


# ------------------------------------------------------------------ Morton

def test_morton_reference_known_answers():
    pos = np.array([[-1000, -1000], [1000, 1000], [0, 0], [-5000, 7000], [1000, -1000],
                    [np.nan, np.nan]], np.float32)
    keys, idx = O.morton_ref(pos)
    assert keys[0] == 0x00000000
    assert keys[1] == 0xFFFFFFFF
    assert keys[2] == 0xC0000000          # q = 32768 on both axes: bit 15 -> bits 30, 31
    assert keys[3] == 0xAAAAAAAA          # clamped: x -> 0, y -> 65535 (odd bits)
    assert keys[4] == 0x55555555          # x -> 65535 (even bits), y -> 0
    assert keys[5] == 0                   # NaN clamps to 0
    assert np.array_equal(idx, np.arange(6))


def test_interleave_matches_magic_bits():
    rng = np.random.default_rng(0)
    x = rng.integers(0, 65536, 2000)
    y = rng.integers(0, 65536, 2000)
    want = interleave_magic(x, y)
    got = np.array([O.interleave16(int(a), int(b)) for a, b in zip(x, y)], np.uint32)
    assert np.array_equal(got, want)


def _deinterleave(k):
    k = k.astype(np.uint64)
    x = np.zeros_like(k)
    y = np.zeros_like(k)
    for b in range(16):
        x |= ((k >> (2 * b)) & 1) << b
        y |= ((k >> (2 * b + 1)) & 1) << b
    return x, y


def test_morton_reference_quantiser_monotone():
    rng = np.random.default_rng(1)
    pos = rng.uniform(-1200, 1200, (5000, 2)).astype(np.float32)
    keys, _ = O.morton_ref(pos)
    qx, qy = _deinterleave(keys)
    # the decoded coordinates are non-decreasing in each position component
    o = np.argsort(pos[:, 0], kind="stable")
    assert np.all(np.diff(qx[o].astype(np.int64)) >= 0)
    o = np.argsort(pos[:, 1], kind="stable")
    assert np.all(np.diff(qy[o].astype(np.int64)) >= 0)
    # round-half-up of (p+1000)/2000*65535 (float64 check away from ties)
    u = np.clip((pos.astype(np.float64) + 1000) / 2000, 0, 1) * 65535
    frac = u - np.floor(u)
    ok = np.abs(frac - 0.5) > 2e-2   # float32 rounding of the scaled coordinate is < 0.01
    assert np.array_equal(qx[ok[:, 0]], np.floor(u[ok[:, 0], 0] + 0.5).astype(np.uint64))


def test_bbox_and_bbox30_corners():
    rng = np.random.default_rng(2)
    pos = rng.uniform(-3, 5, (4000, 2)).astype(np.float32)
    pos[0] = pos[:, 0].min(), pos[:, 1].min()
    box = O.bbox(pos)
    assert np.array_equal(box, [pos[:, 0].min(), pos[:, 1].min(), pos[:, 0].max(), pos[:, 1].max()])
    keys, _ = O.morton_bbox30(pos, box)
    assert keys.max() < (1 << 30)
    assert keys[0] == 0                                  # the box minimum
    qx, qy = _deinterleave(keys)
    side = max(box[2] - box[0], box[3] - box[1])
    assert qx.max() <= 32767 and qy.max() <= 32767
    # the longer axis spans the whole 15-bit range
    big = 0 if box[2] - box[0] >= box[3] - box[1] else 1
    assert (qx if big == 0 else qy).max() == 32767
    o = np.argsort(pos[:, 0], kind="stable")
    assert np.all(np.diff(qx[o].astype(np.int64)) >= 0)
    assert side > 0


def test_bbox30_degenerate_box():
    pos = np.full((10, 2), 3.5, np.float32)
    keys, _ = O.morton_bbox30(pos)
    assert np.all(keys == 0)


# ------------------------------------------------------------------- sort

@pytest.mark.parametrize("ncells", [4, 300, 1 << 30])
def test_sort_is_numpy_stable_sort(ncells):
    rng = np.random.default_rng(ncells % 97)
    keys = rng.integers(0, ncells, 20000).astype(np.uint32)
    vals = np.arange(20000, dtype=np.uint32)
    k, v = O.sort_pairs(keys, vals, 32)
    o = np.argsort(keys, kind="stable")
    assert np.array_equal(v, o) and np.array_equal(k, keys[o])


def test_sort_u64():
    rng = np.random.default_rng(3)
    keys = (rng.integers(0, 256, 5000).astype(np.uint64) << np.uint64(30)) | \
        rng.integers(0, 1 << 20, 5000).astype(np.uint64)
    k, v = O.sort_pairs(keys, np.arange(5000), 40)
    o = np.argsort(keys, kind="stable")
    assert np.array_equal(v, o) and np.array_equal(k, keys[o])


# ------------------------------------------------------------------ build

def brute_radix_tree(keys, bits):
    """Recursive binary radix tree over augmented keys (key, index): returns
    {(lo, hi): split} for every internal node (split = last leaf of the left half)."""
    aug = [(int(k) << 32) | i for i, k in enumerate(keys)]
    out = {}

    def rec(lo, hi):
        if lo == hi:
            return
        a, b = aug[lo], aug[hi]
        top = (a ^ b).bit_length() - 1              # highest differing bit
        split = lo
        while split + 1 <= hi and not ((aug[split + 1] >> top) & 1):
            split += 1
        out[(lo, hi)] = split
        rec(lo, split)
        rec(split + 1, hi)

    rec(0, len(aug) - 1)
    return out


def karras_ranges(left, right, n):
    """Ranges of every internal node by walking the Karras tree from the root."""
    rng = {}

    def cover(node):
        if node >= n - 1:
            k = node - (n - 1)
            return k, k
        lo, _ = cover(int(left[node]))
        _, hi = cover(int(right[node]))
        rng[node] = (lo, hi)
        return lo, hi

    import sys
    sys.setrecursionlimit(10000)
    cover(0)
    return rng


@pytest.mark.parametrize("n,ncells,seed", [(2, 5, 0), (3, 2, 1), (17, 4, 2), (200, 50, 3),
                                           (500, 1 << 30, 4), (333, 1, 5)])
def test_build_matches_brute_force_radix_tree(n, ncells, seed):
    rng = np.random.default_rng(seed)
    keys = np.sort(rng.integers(0, ncells, n).astype(np.uint32))
    left, right, parent = O.build_u32(keys)
    want = brute_radix_tree(keys, 32)
    got = karras_ranges(left, right, n)
    assert len(got) == n - 1
    for node, (lo, hi) in got.items():
        assert (lo, hi) in want
        assert node in (lo, hi)                     # Karras: a node index is a range end
        lc, rc = int(left[node]), int(right[node])
        lch = (lc - (n - 1),) * 2 if lc >= n - 1 else got[lc]
        assert lch[1] == want[(lo, hi)]             # split position
        assert parent[lc] == node and parent[rc] == node
    assert parent[0] == 0xFFFFFFFF
    assert got[0] == (0, n - 1)


def test_build_karras_fig3_example():
    """Karras 2012 (HPG), Fig. 3: eight 5-bit keys."""
    keys = np.array([0b00001, 0b00010, 0b00100, 0b00101, 0b10011, 0b11000, 0b11001, 0b11110],
                    np.uint32)
    left, right, parent = O.build_u32(keys)
    L = 7  # leaves are nodes 7..14
    want = {0: (3, 4), 1: (L + 0, L + 1), 2: (L + 2, L + 3), 3: (1, 2), 4: (L + 4, 5),
            5: (6, L + 7), 6: (L + 5, L + 6)}
    for i, (a, b) in want.items():
        assert (left[i], right[i]) == (a, b), i


def test_build_u64_ensemble_subtrees_equal_plain_trees():
    """Sim-prefixed keys with sim-local tie-breaks: every sim's sub-tree is
    the plain tree of that sim (same shape and split positions)."""
    rng = np.random.default_rng(7)
    sizes = [37, 1, 64, 100, 5, 2, 80]
    off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    plain = []
    allk = []
    for s, ns in enumerate(sizes):
        k = np.sort(rng.integers(0, 40 if s % 2 else 1 << 30, ns).astype(np.uint32))
        plain.append(k)
        allk.append((np.uint64(s) << np.uint64(30)) | k.astype(np.uint64))
    keys = np.concatenate(allk)
    n = len(keys)
    left, right, parent, roots = O.build_u64(keys, off)
    got = karras_ranges(left, right, n)
    for s, ns in enumerate(sizes):
        lo = off[s]
        r = int(roots[s])
        if ns == 1:
            assert r == n - 1 + lo
            continue
        assert got[r] == (lo, lo + ns - 1)
        pl, pr, _ = O.build_u32(plain[s])
        pg = karras_ranges(pl, pr, ns)
        for node, (a, b) in pg.items():
            # plain node k (k >= 1) is ensemble node lo + k; the root maps to roots[s]
            en = r if node == 0 else lo + node
            assert got[en] == (a + lo, b + lo)


# ------------------------------------------------------------------- fill

def _tree(n, seed, scen="uniform"):
    rng = np.random.default_rng(seed)
    pos = rng.uniform(-1, 1, (n, 2)).astype(np.float32)
    mass = rng.uniform(0.1, 2.0, n).astype(np.float32)
    keys, idx = O.morton_bbox30(pos)
    keys, idx = O.sort_pairs(keys, idx, 32)
    left, right, parent = O.build_u32(keys)
    return pos, mass, idx, left, right, parent


def test_fill_matches_fp64_range_sums():
    n = 3000
    pos, mass, idx, left, right, parent = _tree(n, 11)
    cm, bb = O.fill(idx, mass, pos, left, right, parent)
    rngs = karras_ranges(left, right, n)
    m64 = mass.astype(np.float64)[idx]
    p64 = pos.astype(np.float64)[idx]
    cs = np.concatenate([[0], np.cumsum(m64)])
    cx = np.concatenate([[0], np.cumsum(m64 * p64[:, 0])])
    cy = np.concatenate([[0], np.cumsum(m64 * p64[:, 1])])
    for node, (lo, hi) in rngs.items():
        M = cs[hi + 1] - cs[lo]
        assert abs(cm[node, 2] - M) <= 1e-5 * M
        assert abs(cm[node, 0] - (cx[hi + 1] - cx[lo]) / M) <= 1e-5
        assert abs(cm[node, 1] - (cy[hi + 1] - cy[lo]) / M) <= 1e-5
        seg = pos[idx[lo:hi + 1]]
        assert np.array_equal(bb[node], [seg[:, 0].min(), seg[:, 1].min(), seg[:, 0].max(),
                                         seg[:, 1].max()])
        assert cm[node, 3] == np.float32(max(bb[node, 2] - bb[node, 0], bb[node, 3] - bb[node, 1]))
    leaves = np.arange(n) + n - 1
    assert np.array_equal(cm[leaves, :2], pos[idx]) and np.array_equal(cm[leaves, 2], mass[idx])
    assert np.all(cm[leaves, 3] == 0)


def test_fill_is_independent_of_invocation_order():
    n = 2000
    pos, mass, idx, left, right, parent = _tree(n, 12)
    a = O.fill(idx, mass, pos, left, right, parent, offset=0, order_seed=0)
    b = O.fill(idx, mass, pos, left, right, parent, offset=777, order_seed=0)
    c = O.fill(idx, mass, pos, left, right, parent, offset=0, order_seed=12345)
    for x in (b, c):
        assert np.array_equal(a[0], x[0]) and np.array_equal(a[1], x[1])


# ------------------------------------------------------------------- walk

def test_walk_theta_zero_is_direct_sum():
    n = 1500
    pos, mass, idx, left, right, parent = _tree(n, 13)
    cm, _ = O.fill(idx, mass, pos, left, right, parent)
    P = O.Params(grav_constant=1.0, delta_time=1e-3, bh_theta=0.0, epsilon=1e-2,
                 softening_mode=O.SOFTEN_CONSTANT)
    acc, stats = O.walk(idx, cm, left, right, mass, pos, P)
    ref = O.direct_sum_f64(pos, mass, P)
    e = np.linalg.norm(acc - ref, axis=1) / np.linalg.norm(ref, axis=1)
    assert e.max() < 1e-4
    assert np.all(stats[:, 0] == n - 1)   # every other leaf accepted once


def test_direct_sum_f64_matches_numpy():
    rng = np.random.default_rng(14)
    n = 300
    pos = rng.uniform(-1, 1, (n, 2)).astype(np.float32)
    mass = rng.uniform(0.1, 1, n).astype(np.float32)
    for soft in (O.SOFTEN_CONSTANT, O.SOFTEN_REFERENCE):
        P = O.Params(grav_constant=0.7, delta_time=1e-2, bh_theta=0.5, epsilon=3e-2,
                     epsilon_multiplier=2.0, softening_mode=soft)
        # the oracle receives its parameters as float32
        G, dt, eps = (float(np.float32(v)) for v in (0.7, 1e-2, 3e-2))
        p = pos.astype(np.float64)
        m = mass.astype(np.float64)
        r = p[None, :, :] - p[:, None, :]
        if soft == O.SOFTEN_CONSTANT:
            eps2 = eps ** 2 * np.ones((n, n))
        else:
            eps2 = (2.0 * np.sqrt(G * np.maximum(m[:, None], m[None, :]) * dt)) ** 2
        s = (r ** 2).sum(-1) + eps2
        w = G * m[None, :] * s ** -1.5
        np.fill_diagonal(w, 0)
        want = (w[:, :, None] * r).sum(1)
        got = O.direct_sum_f64(pos, mass, P)
        assert np.allclose(got, want, rtol=1e-9, atol=1e-9)


def test_two_body_known_answer():
    pos = np.array([[0, 0], [1, 0]], np.float32)
    mass = np.array([1, 3], np.float32)
    vel = np.zeros_like(pos)
    P = O.Params(grav_constant=2.0, delta_time=1e-2, bh_theta=0.5, epsilon=0.1,
                 softening_mode=O.SOFTEN_CONSTANT, morton_mode=O.MORTON_BBOX30)
    d = O.step_debug(mass, pos, vel, P)
    f = 2.0 / (1.0 + 0.01) ** 1.5
    assert np.allclose(d["accel"][0], [3 * f, 0], rtol=1e-6)
    assert np.allclose(d["accel"][1], [-1 * f, 0], rtol=1e-6)
    # kick then drift (barnes_hut.wgsl:110, 120)
    assert np.allclose(d["vel"][0], [3 * f * 1e-2, 0], rtol=1e-6)
    assert np.allclose(d["pos"][0], [3 * f * 1e-4, 0], rtol=1e-5)


def test_reference_softening_known_answer():
    pos = np.array([[0, 0], [0, 2]], np.float32)
    mass = np.array([4, 1], np.float32)
    P = O.Params(grav_constant=1.0, delta_time=0.25, bh_theta=0.5, epsilon_multiplier=0.5,
                 softening_mode=O.SOFTEN_REFERENCE, morton_mode=O.MORTON_BBOX30)
    d = O.step_debug(mass, pos, np.zeros_like(pos), P)
    eps = 0.5 * np.sqrt(1.0 * 4 * 0.25)   # max(m1, m2) = 4 for both bodies
    f = 1.0 / (4 + eps ** 2) ** 1.5
    assert np.allclose(d["accel"][0], [0, 2 * 1 * f], rtol=1e-6)
    assert np.allclose(d["accel"][1], [0, -2 * 4 * f], rtol=1e-6)


def test_large_theta_accepts_whole_root_including_self():
    """With a huge theta the root is accepted as one body for every body,
    self included (barnes_hut.wgsl:94 tests only the leaf for self)."""
    pos = np.array([[0, 0], [1, 0], [0, 1]], np.float32)
    mass = np.array([1, 2, 3], np.float32)
    P = O.Params(grav_constant=1.0, delta_time=1e-3, bh_theta=1e6, epsilon=0.1,
                 softening_mode=O.SOFTEN_CONSTANT, morton_mode=O.MORTON_BBOX30)
    d = O.step_debug(mass, pos, np.zeros_like(pos), P)
    M = 6.0
    com = (mass[:, None].astype(np.float64) * pos).sum(0) / M
    r = com[None, :] - pos.astype(np.float64)
    want = M * r / (((r ** 2).sum(1) + 0.01) ** 1.5)[:, None]
    assert np.allclose(d["accel"], want, rtol=1e-5)
    assert np.all(d["stats"][:, 0] == 1)


def test_single_body_free_drift():
    pos = np.array([[0.5, -0.25]], np.float32)
    vel = np.array([[2.0, 4.0]], np.float32)
    P = O.Params(grav_constant=1.0, delta_time=0.125, bh_theta=0.5, epsilon=0.1,
                 softening_mode=O.SOFTEN_CONSTANT, morton_mode=O.MORTON_BBOX30)
    p, v = O.step(np.ones(1, np.float32), pos, vel, P, nsteps=4)
    assert np.array_equal(v, vel)
    assert np.allclose(p, pos + 4 * 0.125 * vel)


def test_force_error_shrinks_with_theta():
    n = 4000
    pos, mass, idx, left, right, parent = _tree(n, 15)
    cm, _ = O.fill(idx, mass, pos, left, right, parent)
    ref = O.direct_sum_f64(pos, mass, O.Params(epsilon=1e-3, softening_mode=O.SOFTEN_CONSTANT))
    med = []
    for th in (1.0, 0.7, 0.5, 0.3):
        P = O.Params(bh_theta=th, epsilon=1e-3, softening_mode=O.SOFTEN_CONSTANT)
        acc, _ = O.walk(idx, cm, left, right, mass, pos, P)
        e = np.linalg.norm(acc - ref, axis=1) / np.linalg.norm(ref, axis=1)
        med.append(np.median(e))
    assert all(a > b for a, b in zip(med, med[1:]))


# ---------------------------------------------------------------- renderer

def test_render_exp_against_exp():
    """orc_render_exp: within 1e-6 relative of exp on the splat's [-10, 0],
    exactly 1 at 0, and monotone."""
    x = np.linspace(-10.0, 0.0, 20001).astype(np.float32)
    got = O.render_exp(x).astype(np.float64)
    want = np.exp(x.astype(np.float64))
    assert np.max(np.abs(got - want) / want) < 1e-6
    assert O.render_exp([0.0])[0] == 1.0
    assert np.all(np.diff(got) >= 0)


def test_fixed_to_f16_matches_numpy_rounding():
    """Fixed point (2^-40 units) -> fp16 bits, round to nearest even, against
    numpy's float64 -> float16 conversion (exact values below 2^53):
    subnormals, the normal range, ties and overflow to inf."""
    rng = np.random.default_rng(3)
    v = np.concatenate([
        np.array([0, 1, 2 ** 15, 2 ** 16, 2 ** 16 + 2 ** 15, 3 * 2 ** 15, 2 ** 26, 2 ** 40,
                  65504 * 2 ** 40, 65519 * 2 ** 40, 65520 * 2 ** 40, 2 ** 52], np.uint64),
        rng.integers(0, 2 ** 30, 2000, dtype=np.uint64),
        rng.integers(0, 2 ** 52, 2000, dtype=np.uint64),
        (rng.integers(1, 2 ** 11, 500, dtype=np.uint64) << np.uint64(30)) + np.uint64(2 ** 29),
    ])
    got = O.fixed40_to_f16(v)
    with np.errstate(over="ignore"):
        want = (v.astype(np.float64) * 2.0 ** -40).astype(np.float16).view(np.uint16)
    assert np.array_equal(got, want)


def _cam(w, h, cx=0.0, cy=0.0, half_y=10.0):
    return np.array([cx, cy, half_y * w / h, half_y], np.float32)


def test_splat_single_body_footprint():
    """A body at a pixel centre: R = 2 px covers the 13 pixel centres with
    dx^2 + dy^2 <= 4; the centre holds base * zoom_scale exactly, the
    footprint is symmetric, and every pixel holds base * exp(-10 r^2) *
    zoom_scale to within the exp's 1e-6 (density.wgsl:77-96)."""
    w = h = 64
    cam = _cam(w, h)
    # pixel (20, 40) centre in world coordinates (texture rows: y_ndc = 1 - (row + .5) 2/h)
    xn = (20 + 0.5) * 2 / w - 1
    yn = 1 - (40 + 0.5) * 2 / h
    pos = np.array([[xn * cam[2], yn * cam[3]]], np.float32)
    zoom_scale = np.float32(1) / (np.float32(1) + np.float32(0.05) * np.float32(max(cam[2], cam[3])))
    for rm, base in ((1.0, 0.02), (3.0, 0.1)):
        fx = O.splat(pos, [rm], cam, w, h).astype(np.float64) * 2.0 ** -40
        R = 2 * rm
        ys, xs = np.nonzero(fx)
        dx, dy = xs - 20, ys - 40
        assert np.all(dx * dx + dy * dy <= R * R)
        assert len(xs) == np.sum([(a * a + b * b <= R * R) for a in range(-8, 9) for b in range(-8, 9)])
        assert fx[40, 20] == np.float64(np.float32(base) * zoom_scale)
        sym = fx[40 - 8:40 + 9, 20 - 8:20 + 9]
        assert np.array_equal(sym, sym[::-1, :]) and np.array_equal(sym, sym[:, ::-1])
        r2 = (dx * dx + dy * dy) / R ** 2
        want = base * np.exp(-10 * r2) * float(zoom_scale)
        assert np.max(np.abs(fx[ys, xs] - want) / want) < 2e-6


def test_splat_is_additive_and_rows_flip():
    """Fixed-point sums: splat(A + B) = splat(A) + splat(B) exactly (order of
    bodies irrelevant); in the RGBA frame a body above the camera centre
    (world y > 0, texture rows near the top) appears in the lower half."""
    rng = np.random.default_rng(5)
    w, h = 96, 48
    cam = _cam(w, h)
    pa = rng.uniform(-9, 9, (300, 2)).astype(np.float32)
    pb = rng.uniform(-9, 9, (200, 2)).astype(np.float32)
    ra, rb = np.ones(300, np.float32), rng.uniform(0.5, 4, 200).astype(np.float32)
    sa, sb = O.splat(pa, ra, cam, w, h), O.splat(pb, rb, cam, w, h)
    both = O.splat(np.concatenate([pb, pa]), np.concatenate([rb, ra]), cam, w, h)
    assert np.array_equal(both, sa + sb)
    top = np.array([[0.0, 6.0]], np.float32)
    d16, rgba = O.render(top, [2.0], cam, w, h)
    rows = np.nonzero(d16.any(axis=1))[0]
    assert rows.max() < h // 2
    lit = np.nonzero(rgba[:, :, :3].any(axis=(1, 2)))[0]
    assert lit.min() >= h // 2 and np.array_equal(np.sort(h - 1 - rows), lit)


def test_tone_table_known_answers():
    """tone_map.wgsl:66-84: black at 0, white at the largest densities and
    inf, opaque, every channel non-decreasing in the density (the ramp's
    stops are channel-wise ordered), and the ramp's stops where the gamma-
    mapped value crosses 0.33 and 0.66."""
    t = O.tone_table()
    rgba = np.stack([(t >> s) & 0xFF for s in (0, 8, 16, 24)], axis=1)
    assert tuple(rgba[0]) == (0, 0, 0, 255)
    assert tuple(rgba[0x7BFF]) == (255, 255, 255, 255)  # 65504
    assert tuple(rgba[0x7C00]) == (255, 255, 255, 255)  # +inf
    assert np.all(rgba[:, 3] == 255)
    pos = np.arange(0, 0x7C01)  # 0 .. +inf in increasing order of value
    assert np.all(np.diff(rgba[pos, :3].astype(int), axis=0) >= 0)
    vals = np.arange(0x7C00, dtype=np.uint16).view(np.float16).astype(np.float64)
    x = 8.0 * vals
    g = (x / (1 + x)) ** float(np.float32(1 / 2.2))
    c1 = np.rint(255 * np.float32([0.1, 0.2, 0.9]).astype(np.float64))
    c2 = np.rint(255 * np.float32([0.2, 0.9, 0.9]).astype(np.float64))
    k1 = np.argmax(g >= float(np.float32(0.33)))
    k2 = np.argmax(g >= float(np.float32(0.66)))
    assert np.all(np.abs(rgba[k1, :3] - c1) <= 1) and np.all(np.abs(rgba[k2, :3] - c2) <= 1)


# ---------------------------------------------------------------- quadrupole

import nbody_sim  # noqa: E402  (host-side generator only: no device call)

def _quad_tree(mass, pos, P):
    r = O.step_debug(mass, pos, np.zeros_like(pos), P)
    return r


def test_node_quad_matches_brute_force():
    """Second moments of every node = sum over its leaf range of
    m (x - c)(x - c)^T about the range's centre of mass (numpy, float64)."""
    rng = np.random.default_rng(21)
    n = 300
    pos = rng.normal(0, 1, (n, 2)).astype(np.float32)
    mass = rng.uniform(0.2, 2.0, n).astype(np.float32)
    P = O.Params(softening_mode=O.SOFTEN_CONSTANT, morton_mode=O.MORTON_BBOX30, bh_theta=0.5)
    r = _quad_tree(mass, pos, P)
    q = O.node_quad(r["idx"], mass, pos, r["left"], r["right"])
    from frozen_bh import node_ranges

    lo, hi = node_ranges(r["left"], r["right"], n)
    for v in range(2 * n - 1):
        b = r["idx"][lo[v]:hi[v] + 1]
        m = mass[b].astype(np.float64)
        x = pos[b].astype(np.float64)
        c = (m[:, None] * x).sum(0) / m.sum()
        d = x - c
        want = np.array([(m * d[:, 0] ** 2).sum(), (m * d[:, 0] * d[:, 1]).sum(),
                         (m * d[:, 1] ** 2).sum()])
        assert np.allclose(q[v], want, rtol=1e-6, atol=1e-12 * m.sum()), v
    assert np.all(q[n - 1:] == 0)


def test_quad_walk_theta_zero_is_the_monopole_walk():
    """At theta = 0 every accepted node is a leaf, whose moments are zero: the
    quadrupole walk gives the monopole walk's bits."""
    mass, pos, _ = nbody_sim.generate("plummer", 2000, seed=3)
    P = O.Params(softening_mode=O.SOFTEN_CONSTANT, morton_mode=O.MORTON_BBOX30, bh_theta=0.0)
    r = _quad_tree(mass, pos, P)
    q = O.node_quad(r["idx"], mass, pos, r["left"], r["right"])
    a0, _ = O.walk(r["idx"], r["node_cm"], r["left"], r["right"], mass, pos, P)
    a2, _ = O.walk(r["idx"], r["node_cm"], r["left"], r["right"], mass, pos, P, quad=q)
    assert np.array_equal(a0, a2)


def test_quad_far_field_order():
    """A two-body node (masses 1 at (-a, 0) and (a, 0)) seen from distance d
    along both axes: the monopole error falls as (a/d)^2 and the quadrupole
    error as (a/d)^4 (halving a/d: errors / 4 and / 16), and the quadrupole
    error is far smaller -- a wrong sign or factor in the term breaks both."""
    a = np.float32(0.01)
    P = O.Params(softening_mode=O.SOFTEN_CONSTANT, morton_mode=O.MORTON_BBOX30, bh_theta=0.99,
                 epsilon=1e-6)
    errs = {}
    for axis in (0, 1):
        for d in (0.05, 0.1, 0.2):
            tgt = np.zeros(2, np.float32)
            tgt[axis] = d
            pos = np.array([[-a, 0], [a, 0], tgt], np.float32)
            mass = np.array([1, 1, 1e-30], np.float32)
            r = _quad_tree(mass, pos, P)
            q = O.node_quad(r["idx"], mass, pos, r["left"], r["right"])
            exact = O.direct_sum_f64(pos, mass, P, targets=[2])[0]
            out = []
            for quad in (None, q):
                acc, st = O.walk(r["idx"], r["node_cm"], r["left"], r["right"], mass, pos, P,
                                 quad=quad)
                assert st[2, 0] == 1  # the pair node accepted as one term
                out.append(np.linalg.norm(acc[2] - exact) / np.linalg.norm(exact))
            errs[(axis, d)] = out
    for axis in (0, 1):
        for d in (0.05, 0.1):
            m_ratio = errs[(axis, d)][0] / errs[(axis, 2 * d)][0]
            q_ratio = errs[(axis, d)][1] / errs[(axis, 2 * d)][1]
            assert 3.0 < m_ratio < 5.0, (axis, d, m_ratio)
            assert 11.0 < q_ratio < 21.0, (axis, d, q_ratio)
            assert errs[(axis, d)][1] < 0.2 * errs[(axis, d)][0]


def test_quad_walk_is_more_accurate():
    """The quadrupole expansion at theta 0.5 on a 16k uniform disk: median and
    p99 force error against the FP64 direct sum several times smaller."""
    mass, pos, _ = nbody_sim.generate("uniform-disk", 16384, seed=1)
    P = O.Params(softening_mode=O.SOFTEN_CONSTANT, morton_mode=O.MORTON_BBOX30, bh_theta=0.5,
                 epsilon=1e-3)
    r = _quad_tree(mass, pos, P)
    q = O.node_quad(r["idx"], mass, pos, r["left"], r["right"])
    tg = np.arange(0, 16384, 8)
    ref = O.direct_sum_f64(pos, mass, P, targets=tg)
    e = []
    for quad in (None, q):
        acc, _ = O.walk(r["idx"], r["node_cm"], r["left"], r["right"], mass, pos, P, quad=quad)
        e.append(np.linalg.norm(acc[tg] - ref, axis=1) / np.linalg.norm(ref, axis=1))
    assert np.median(e[1]) < np.median(e[0]) / 4
    assert np.percentile(e[1], 99) < np.percentile(e[0], 99) / 3
