"""The oracle reproduces the committed golden fixtures bit for bit, and the
fixtures' trees satisfy the structural invariants (CPU only)."""
import glob
import os

import numpy as np
import pytest

from oracle import oracle as O

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
STEPS = sorted(glob.glob(os.path.join(HERE, "step_*.npz")))


def test_karras_fixture():
    f = np.load(os.path.join(HERE, "karras_fig3.npz"))
    left, right, parent = O.build_u32(f["keys"])
    assert np.array_equal(left, f["left"]) and np.array_equal(right, f["right"])
    assert np.array_equal(parent, f["parent"])


@pytest.mark.parametrize("path", STEPS, ids=[os.path.basename(p) for p in STEPS])
def test_oracle_reproduces_step_fixture(path):
    f = np.load(path)
    g, dt, em, th, eps, sm, mm = f["params"]
    P = O.Params(float(g), float(dt), float(em), float(th), float(eps), int(sm), int(mm))
    d = O.step_debug(f["mass"], f["pos_in"], f["vel_in"], P, threads=2)
    for k, v in d.items():
        assert np.array_equal(v, f[k]), k


@pytest.mark.parametrize("path", STEPS, ids=[os.path.basename(p) for p in STEPS])
def test_fixture_tree_invariants(path):
    f = np.load(path)
    n = f["keys"].shape[0]
    left, right, parent = f["left"], f["right"], f["parent"]
    assert np.all(np.diff(f["keys"].astype(np.int64)) >= 0)          # sorted
    assert sorted(f["idx"].tolist()) == list(range(n))               # a permutation
    children = np.concatenate([left, right])
    assert sorted(children.tolist()) == list(range(1, 2 * n - 1))    # every non-root once
    assert np.all(parent[left] == np.arange(n - 1)) and np.all(parent[right] == np.arange(n - 1))
    # node mass of the root = total mass; COM = mass-weighted mean (FP64)
    m = f["mass"].astype(np.float64)
    assert abs(f["node_cm"][0, 2] - m.sum()) <= 1e-5 * m.sum()
    com = (m[:, None] * f["pos_in"].astype(np.float64)).sum(0) / m.sum()
    assert np.allclose(f["node_cm"][0, :2], com, atol=1e-5)
