"""FP64 reference of a Barnes-Hut trajectory with frozen opening decisions.

Along the oracle's binary32 trajectory, each step's sorted order, node leaf
ranges and accepted-node lists are recorded (oracle.walk_lists).  The FP64
function then recomputes every node's mass and centre of mass from the
current leaf positions (prefix sums over the node's leaf range) and sums the
softened pulls of exactly the recorded nodes, followed by kick and drift.
This is the function the hand-written adjoint differentiates; torch autograd
and central finite differences of it are the gradient references.
"""
from __future__ import annotations

import numpy as np
import torch

from oracle import oracle as O


def node_ranges(left, right, n):
    """(lo, hi) sorted-leaf range of every node id (internal then leaves)."""
    lo = np.zeros(2 * n - 1, np.int64)
    hi = np.zeros(2 * n - 1, np.int64)
    lo[n - 1:] = np.arange(n)
    hi[n - 1:] = np.arange(n)
    if n < 2:
        return lo, hi
    # post-order without recursion
    order = []
    stack = [0]
    while stack:
        v = stack.pop()
        order.append(v)
        for c in (int(left[v]), int(right[v])):
            if c < n - 1:
                stack.append(c)
    for v in reversed(order):
        lo[v] = lo[int(left[v])]
        hi[v] = hi[int(right[v])]
    return lo, hi


class FrozenTrajectory:
    """Record the frozen structure of `nsteps` steps from (pos, vel, mass)."""

    def __init__(self, mass, pos, vel, P: O.Params, nsteps: int):
        self.P = P
        self.n = len(mass)
        self.steps = []
        p, v = pos, vel
        for _ in range(nsteps):
            d = O.step_debug(mass, p, v, P)
            off, nodes = O.walk_lists(d["idx"], d["node_cm"], d["left"], d["right"], mass, p, P)
            lo, hi = node_ranges(d["left"], d["right"], self.n)
            t_of = np.repeat(np.arange(self.n), np.diff(off))
            self.steps.append(dict(
                idx=torch.from_numpy(d["idx"].astype(np.int64)),
                t=torch.from_numpy(t_of),
                lo=torch.from_numpy(lo[nodes.astype(np.int64)]),
                hi=torch.from_numpy(hi[nodes.astype(np.int64)]),
                leaf=torch.from_numpy(nodes.astype(np.int64) >= self.n - 1),
                accel32=d["accel"]))
            p, v = d["pos"], d["vel"]
        self.final32 = (p, v)

    def forces(self, k, x, m):
        """FP64 accelerations (by body) of step k at positions x, masses m."""
        S = self.steps[k]
        P = self.P
        idx = S["idx"]
        xs, ms = x[idx], m[idx]
        zero = torch.zeros(1, dtype=x.dtype)
        cm = torch.cat([zero, torch.cumsum(ms, 0)])
        cmx = torch.cat([torch.zeros(1, 2, dtype=x.dtype), torch.cumsum(ms[:, None] * xs, 0)])
        lo, hi, leaf = S["lo"], S["hi"], S["leaf"]
        M = torch.where(leaf, ms[lo], cm[hi + 1] - cm[lo])
        C = torch.where(leaf[:, None], xs[lo], (cmx[hi + 1] - cmx[lo]) / M[:, None])
        r = C - xs[S["t"]]
        s = (r * r).sum(1) + float(np.float32(P.epsilon)) ** 2
        f = float(np.float32(P.grav_constant)) * M * s ** -1.5
        acc_s = torch.zeros_like(xs).index_add(0, S["t"], f[:, None] * r)
        return torch.zeros_like(x).index_copy(0, idx, acc_s)

    def run(self, x, v, m):
        dt = float(np.float32(self.P.delta_time))
        for k in range(len(self.steps)):
            a = self.forces(k, x, m)
            v = v + a * dt
            x = x + v * dt
        return x, v


def direct_trajectory(x, v, m, G, eps, dt, nsteps):
    """FP64 O(N^2) softened direct-sum kick-drift trajectory (autograd-able)."""
    for _ in range(nsteps):
        r = x[None, :, :] - x[:, None, :]
        s = (r * r).sum(-1) + eps * eps
        w = G * m[None, :] * s ** -1.5
        w = w * (1.0 - torch.eye(len(m), dtype=x.dtype))
        a = (w[:, :, None] * r).sum(1)
        v = v + a * dt
        x = x + v * dt
    return x, v
