"""ctypes + numpy front end of the CPU oracle (``libbh_oracle.so``).

TEST INFRASTRUCTURE ONLY: the checker for the HIP path.  Each wrapper calls
the C restatement in ``bh_oracle.c``, which cites the reference lines it
follows (src/shaders/compute/lbvh.wgsl, src/shaders/compute/barnes_hut.wgsl,
src/simulation.ts:393-421).  Arrays use the reference's host layouts:
``pos``/``vel`` are float32 (n, 2) like the vec2f buffers (simulation.ts:237-238),
``mass`` float32 (n,).
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libbh_oracle.so")

SOFTEN_REFERENCE = 0
SOFTEN_CONSTANT = 1
MORTON_REFERENCE = 0
MORTON_BBOX30 = 1


class _Params(ctypes.Structure):
    _fields_ = [
        ("grav_constant", ctypes.c_float),
        ("delta_time", ctypes.c_float),
        ("epsilon_multiplier", ctypes.c_float),
        ("bh_theta", ctypes.c_float),
        ("epsilon", ctypes.c_float),
        ("softening_mode", ctypes.c_int32),
        ("morton_mode", ctypes.c_int32),
    ]


@dataclass
class Params:
    """Physics parameters: Metadata fields of lbvh.wgsl:3-10 plus the modes."""

    grav_constant: float = 1.0
    delta_time: float = 0.1 / 60.0
    epsilon_multiplier: float = 1.0
    bh_theta: float = 0.6
    epsilon: float = 1e-3
    softening_mode: int = SOFTEN_REFERENCE
    morton_mode: int = MORTON_REFERENCE

    def c(self) -> _Params:
        return _Params(self.grav_constant, self.delta_time, self.epsilon_multiplier,
                       self.bh_theta, self.epsilon, self.softening_mode, self.morton_mode)


_lib = None


def _threads():
    """OpenMP threads for the oracle: the host cores, at most 16."""
    return max(1, min(os.cpu_count() or 1, 16))


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle library not built: {LIB_PATH} (run make -C oracle)")
        _lib = ctypes.CDLL(LIB_PATH)
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _f32(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.float32)
    if shape is not None:
        a = a.reshape(shape)
    return a


def interleave16(x: int, y: int) -> int:
    f = lib().orc_interleave16
    f.restype = ctypes.c_uint32
    return int(f(ctypes.c_uint32(x), ctypes.c_uint32(y)))


def morton_ref(pos):
    pos = _f32(pos, (-1, 2))
    n = pos.shape[0]
    keys = np.empty(n, np.uint32)
    idx = np.empty(n, np.uint32)
    lib().orc_morton_ref(ctypes.c_int64(n), _p(pos), _p(keys), _p(idx))
    return keys, idx


def bbox(pos):
    pos = _f32(pos, (-1, 2))
    out = np.empty(4, np.float32)
    lib().orc_bbox(ctypes.c_int64(pos.shape[0]), _p(pos), _p(out))
    return out


def morton_bbox30(pos, box=None):
    pos = _f32(pos, (-1, 2))
    if box is None:
        box = bbox(pos)
    box = _f32(box)
    n = pos.shape[0]
    keys = np.empty(n, np.uint32)
    idx = np.empty(n, np.uint32)
    lib().orc_morton_bbox30(ctypes.c_int64(n), _p(pos), _p(box), _p(keys), _p(idx))
    return keys, idx


def sort_pairs(keys, vals, bits=32):
    keys = np.array(keys, copy=True)
    vals = np.ascontiguousarray(np.array(vals, dtype=np.uint32, copy=True))
    n = keys.shape[0]
    if keys.dtype == np.uint64:
        lib().orc_sort_pairs_u64(ctypes.c_int64(n), _p(keys), _p(vals), ctypes.c_int(bits))
    else:
        keys = np.ascontiguousarray(keys, dtype=np.uint32)
        lib().orc_sort_pairs_u32(ctypes.c_int64(n), _p(keys), _p(vals), ctypes.c_int(bits))
    return keys, vals


def build_u32(keys):
    keys = np.ascontiguousarray(keys, dtype=np.uint32)
    n = keys.shape[0]
    left = np.zeros(max(n - 1, 0), np.uint32)
    right = np.zeros(max(n - 1, 0), np.uint32)
    parent = np.zeros(max(2 * n - 1, 0), np.uint32)
    if n > 0:
        lib().orc_build_u32(ctypes.c_int64(n), _p(keys), _p(left), _p(right), _p(parent))
    return left, right, parent


def build_u64(keys, sim_offset):
    keys = np.ascontiguousarray(keys, dtype=np.uint64)
    sim_offset = np.ascontiguousarray(sim_offset, dtype=np.int64)
    n = keys.shape[0]
    s = sim_offset.shape[0] - 1
    left = np.zeros(max(n - 1, 0), np.uint32)
    right = np.zeros(max(n - 1, 0), np.uint32)
    parent = np.zeros(max(2 * n - 1, 0), np.uint32)
    roots = np.zeros(max(s, 1), np.uint32)
    lib().orc_build_u64(ctypes.c_int64(n), _p(keys), _p(sim_offset), ctypes.c_int32(s),
                        _p(left), _p(right), _p(parent), _p(roots))
    return left, right, parent, roots[:s]


def fill(idx, mass, pos, left, right, parent, offset=0, order_seed=0):
    idx = np.ascontiguousarray(idx, np.uint32)
    mass = _f32(mass)
    pos = _f32(pos, (-1, 2))
    left = np.ascontiguousarray(left, np.uint32)
    right = np.ascontiguousarray(right, np.uint32)
    parent = np.ascontiguousarray(parent, np.uint32)
    n = idx.shape[0]
    cm = np.zeros((2 * n - 1, 4), np.float32)
    bb = np.zeros((2 * n - 1, 4), np.float32)
    lib().orc_fill(ctypes.c_int64(n), _p(idx), _p(mass), _p(pos), _p(left), _p(right),
                   _p(parent), ctypes.c_uint32(offset), ctypes.c_uint64(order_seed),
                   _p(cm), _p(bb))
    return cm, bb


def walk(idx, cm, left, right, mass, pos, params: Params, roots=None, stack_size=256,
         threads=None, margin=False, quad=None):
    """orc_walk: (accel (n,2), stats (n,3) {accepted, opened, dropped} by body)
    and, with margin=True, each body's smallest relative distance
    |length*inv - theta| / theta of an opening test from theta and the sum of
    its accepted terms' lengths |G m r inv^3|: (accel, stats, margin, term_norm).
    quad: per-node second moments (node_quad) for the quadrupole expansion."""
    idx = np.ascontiguousarray(idx, np.uint32)
    cm = _f32(cm, (-1, 4))
    left = np.ascontiguousarray(left, np.uint32)
    right = np.ascontiguousarray(right, np.uint32)
    mass = _f32(mass)
    pos = _f32(pos, (-1, 2))
    n = idx.shape[0]
    acc = np.zeros((n, 2), np.float32)
    stats = np.zeros((n, 3), np.int64)
    pr = params.c()
    rp = None
    if roots is not None:
        roots = np.ascontiguousarray(roots, np.uint32)
        rp = _p(roots)
    mg = np.zeros(n, np.float32) if margin else None
    tn = np.zeros(n, np.float32) if margin else None
    qp = None
    if quad is not None:
        quad = _f32(quad, (-1, 3))
        qp = _p(quad)
    lib().orc_walk(ctypes.c_int64(n), _p(idx), _p(cm), _p(left), _p(right), rp, _p(mass),
                   _p(pos), ctypes.byref(pr), ctypes.c_int(stack_size), _p(acc), _p(stats), qp,
                   _p(mg) if margin else None, _p(tn) if margin else None,
                   ctypes.c_int(threads or _threads()))
    return (acc, stats, mg, tn) if margin else (acc, stats)


def kick_drift(acc, dt, pos, vel):
    pos = _f32(pos, (-1, 2)).copy()
    vel = _f32(vel, (-1, 2)).copy()
    acc = _f32(acc, (-1, 2))
    lib().orc_kick_drift(ctypes.c_int64(pos.shape[0]), _p(acc), ctypes.c_float(dt), _p(pos),
                         _p(vel))
    return pos, vel


def step(mass, pos, vel, params: Params, nsteps=1, threads=None):
    """nsteps full steps (simulation.ts:393-421); returns new (pos, vel)."""
    mass = _f32(mass)
    pos = _f32(pos, (-1, 2)).copy()
    vel = _f32(vel, (-1, 2)).copy()
    pr = params.c()
    rc = lib().orc_step(ctypes.c_int64(pos.shape[0]), ctypes.byref(pr), _p(mass), _p(pos),
                        _p(vel), ctypes.c_int(nsteps), ctypes.c_int(threads or _threads()))
    if rc != 0:
        raise RuntimeError("orc_step failed")
    return pos, vel


def step_debug(mass, pos, vel, params: Params, threads=None):
    """One step with every intermediate returned as a dict."""
    mass = _f32(mass)
    pos = _f32(pos, (-1, 2))
    vel = _f32(vel, (-1, 2))
    n = pos.shape[0]
    out = dict(
        keys=np.zeros(n, np.uint32), idx=np.zeros(n, np.uint32),
        left=np.zeros(max(n - 1, 1), np.uint32), right=np.zeros(max(n - 1, 1), np.uint32),
        parent=np.zeros(max(2 * n - 1, 1), np.uint32),
        node_cm=np.zeros((max(2 * n - 1, 1), 4), np.float32),
        node_aabb=np.zeros((max(2 * n - 1, 1), 4), np.float32),
        accel=np.zeros((n, 2), np.float32), pos=np.zeros((n, 2), np.float32),
        vel=np.zeros((n, 2), np.float32), stats=np.zeros((n, 3), np.int64),
        bbox=np.zeros(4, np.float32),
    )
    pr = params.c()
    rc = lib().orc_step_debug(
        ctypes.c_int64(n), ctypes.byref(pr), _p(mass), _p(pos), _p(vel), _p(out["keys"]),
        _p(out["idx"]), _p(out["left"]), _p(out["right"]), _p(out["parent"]),
        _p(out["node_cm"]), _p(out["node_aabb"]), _p(out["accel"]), _p(out["pos"]),
        _p(out["vel"]), _p(out["stats"]), _p(out["bbox"]),
        ctypes.c_int(threads or _threads()))
    if rc != 0:
        raise RuntimeError("orc_step_debug failed")
    out["left"] = out["left"][: max(n - 1, 0)]
    out["right"] = out["right"][: max(n - 1, 0)]
    return out


def direct_sum_f64(pos, mass, params: Params, targets=None, threads=None):
    pos = _f32(pos, (-1, 2))
    mass = _f32(mass)
    n = pos.shape[0]
    tp = None
    nt = n
    if targets is not None:
        targets = np.ascontiguousarray(targets, np.int64)
        tp = _p(targets)
        nt = targets.shape[0]
    acc = np.zeros((nt, 2), np.float64)
    pr = params.c()
    lib().orc_direct_sum_f64(ctypes.c_int64(n), _p(pos), _p(mass), ctypes.byref(pr), tp,
                             ctypes.c_int64(nt), _p(acc),
                             ctypes.c_int(threads or _threads()))
    return acc


def walk_lists(idx, cm, left, right, mass, pos, params: Params):
    """Accepted nodes of each sorted position's walk (CSR offsets, nodes)."""
    idx = np.ascontiguousarray(idx, np.uint32)
    cm = _f32(cm, (-1, 4))
    left = np.ascontiguousarray(left, np.uint32)
    right = np.ascontiguousarray(right, np.uint32)
    mass = _f32(mass)
    pos = _f32(pos, (-1, 2))
    n = idx.shape[0]
    off = np.zeros(n + 1, np.int64)
    pr = params.c()
    f = lib().orc_walk_lists
    f.restype = ctypes.c_int64
    cap = 0
    nodes = np.zeros(1, np.uint32)
    while True:
        total = f(ctypes.c_int64(n), _p(idx), _p(cm), _p(left), _p(right), _p(mass), _p(pos),
                  ctypes.byref(pr), _p(off), _p(nodes), ctypes.c_int64(cap))
        if total <= cap:
            return off, nodes[:total]
        cap = int(total)
        nodes = np.zeros(cap, np.uint32)


# ---------------------------------------------------------------- renderer

def render_exp(x):
    """orc_render_exp over an array of float32 values."""
    f = lib().orc_render_exp
    f.restype = ctypes.c_float
    f.argtypes = [ctypes.c_float]
    x = np.asarray(x, np.float32).ravel()
    return np.array([f(float(v)) for v in x], np.float32)


def fixed40_to_f16(v):
    """fp16 bits of v * 2^-40 (round to nearest even), for an array of uint64."""
    f = lib().orc_fixed40_to_f16
    f.restype = ctypes.c_uint16
    f.argtypes = [ctypes.c_uint64]
    return np.array([f(int(u)) for u in np.asarray(v, np.uint64).ravel()], np.uint16)


def splat(pos, radius, cam4, width, height):
    pos = _f32(pos, (-1, 2))
    radius = _f32(radius)
    cam4 = _f32(cam4)
    out = np.zeros((height, width), np.uint64)
    lib().orc_splat(ctypes.c_int64(pos.shape[0]), _p(pos), _p(radius), _p(cam4),
                    ctypes.c_int32(width), ctypes.c_int32(height), _p(out))
    return out


def tone_table():
    t = np.zeros(65536, np.uint32)
    lib().orc_tone_table(_p(t))
    return t


def render(pos, radius, cam4, width, height):
    """(fp16 density bits [h][w] in texture rows, RGBA uint8 [h][w][4])."""
    pos = _f32(pos, (-1, 2))
    radius = _f32(radius)
    cam4 = _f32(cam4)
    d16 = np.zeros((height, width), np.uint16)
    rgba = np.zeros((height, width, 4), np.uint8)
    lib().orc_render(ctypes.c_int64(pos.shape[0]), _p(pos), _p(radius), _p(cam4),
                     ctypes.c_int32(width), ctypes.c_int32(height), _p(d16), _p(rgba))
    return d16, rgba


def node_quad(idx, mass, pos, left, right):
    """Central second mass moments {xx, xy, yy} of every node (2n-1, 3)."""
    idx = np.ascontiguousarray(idx, np.uint32)
    mass = _f32(mass)
    pos = _f32(pos, (-1, 2))
    left = np.ascontiguousarray(left, np.uint32)
    right = np.ascontiguousarray(right, np.uint32)
    n = idx.shape[0]
    q = np.zeros((max(2 * n - 1, 1), 3), np.float32)
    lib().orc_node_quad(ctypes.c_int64(n), _p(idx), _p(mass), _p(pos), _p(left), _p(right), _p(q))
    return q
