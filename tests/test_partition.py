"""Partition parity: the engine's host clustering + block tree against the
oracle restatement and the hand-derived fixtures (bit-identical permutation
and preorder block list)."""
import json
import math
import os

import numpy as np
import pytest

from oracle import hmat_oracle as ho

hmat_amd = pytest.importorskip("hmat_amd")

GOLD = os.path.join(os.path.dirname(__file__), "golden", "partition_fixtures.json")
KIND = {0: "F", 1: "R", 2: "H"}


def engine_partition(points, leaf, eta=2.0, sym=True):
    t = hmat_amd.ClusterTree(points, leaf)
    h = hmat_amd.HMatrix(t, eta=eta, lower_symmetric=sym)
    part = [(int(a), int(b), int(c), int(d), KIND[int(k)]) for a, b, c, d, k in h.partition()]
    idx = t.indices
    h.close()
    return idx, part


def oracle_partition(points, leaf, eta=2.0, sym=True):
    root, idx = ho.build_cluster_tree(points, leaf)
    bt = ho.build_block_tree(root, root, eta, sym)
    return idx, ho.partition_signature(bt)


@pytest.mark.parametrize("case", json.load(open(GOLD))["cases"], ids=lambda c: c["name"])
def test_hand_fixtures(case):
    pts = np.array(case["points"], dtype=float).reshape(len(case["points"]), case["dim"])
    idx, part = engine_partition(pts, case["leaf"], case["eta"])
    assert list(idx) == case["indices"]
    oidx, opart = oracle_partition(pts, case["leaf"], case["eta"])
    assert list(oidx) == case["indices"]
    if "partition" in case:
        exp = [tuple(p) for p in case["partition"]]
        assert part == exp
        assert opart == exp


@pytest.mark.parametrize("n,height,seed,leaf", [
    (300, 1.0, 1, 16), (2000, 0.8, 7, 64), (10000, 4.0, 12345, 128), (5000, 2.0, 99, 128)])
def test_engine_matches_oracle(n, height, seed, leaf):
    pts = ho.cylinder_points(n, 1.0, height, seed)
    idx, part = engine_partition(pts, leaf)
    oidx, opart = oracle_partition(pts, leaf)
    assert np.array_equal(idx, oidx)
    assert part == opart


def test_nonsymmetric_partition_matches():
    pts = ho.cylinder_points(1500, 1.0, 1.0, 5)
    idx, part = engine_partition(pts, 64, sym=False)
    oidx, opart = oracle_partition(pts, 64, sym=False)
    assert part == opart


def test_invariants_config0():
    """Permutation is a bijection, leaves <= leaf size, children tile their
    parent, and a leaf is Rk exactly when the eta rule holds on bounding boxes
    recomputed here from the points."""
    n, leaf, eta = 10000, 128, 2.0
    pts = ho.cylinder_points(n, 1.0, 4.0, 12345)
    idx, part = engine_partition(pts, leaf, eta)
    assert sorted(idx) == list(range(n))
    P = pts[idx]

    def box(o, s):
        q = P[o:o + s]
        return q.min(0), q.max(0)

    def diam(b):
        return math.sqrt(float(((b[1] - b[0]) ** 2).sum()))

    def dist(a, b):
        d = np.maximum(0, a[0] - b[1]) ** 2 + np.maximum(0, b[0] - a[1]) ** 2
        return math.sqrt(float(d.sum()))

    covered = np.zeros((n // 50 + 1,), dtype=np.int64)
    area = 0
    for r0, rs, c0, cs, k in part:
        if k == "H":
            continue
        assert min(rs, cs) <= leaf or k == "R"
        area += rs * cs if r0 != c0 else rs * cs
        br, bc = box(r0, rs), box(c0, cs)
        low = min(diam(br), diam(bc)) > 0 and min(diam(br), diam(bc)) <= eta * dist(br, bc)
        assert (k == "R") == low
    # lower triangle incl. diagonal blocks is covered exactly once
    diag = sum(rs * cs for r0, rs, c0, cs, k in part if k != "H" and r0 == c0)
    off = sum(rs * cs for r0, rs, c0, cs, k in part if k != "H" and r0 != c0)
    assert diag + 2 * off == n * n


@pytest.mark.parametrize("n,height,seed", [(200000, 16.0, 2)])
def test_full_size_partition_matches(n, height, seed):
    """Precision workload geometry (N=200k, radius 1, height 16), leaf 128."""
    pts = ho.cylinder_points(n, 1.0, height, seed)
    idx, part = engine_partition(pts, 128)
    oidx, opart = oracle_partition(pts, 128)
    assert np.array_equal(idx, oidx)
    assert part == opart


@pytest.mark.parametrize("n,leaf", [(1500, 64), (2050, 128)])
def test_complex_partition_matches(n, leaf):
    """A complex matrix keeps the complex partition: its blocks over the 2 x 2
    real representation cover exactly twice the rows and columns."""
    pts = ho.cylinder_points(n, 1.0, 1.0, 5)
    t = hmat_amd.ClusterTree(pts, leaf)
    h = hmat_amd.HMatrix(t, lower_symmetric=False, value_type=hmat_amd.HMAT_DOUBLE_COMPLEX)
    part = [(int(a), int(b), int(c), int(d), KIND[int(k)]) for a, b, c, d, k in h.partition()]
    oidx, opart = oracle_partition(pts, leaf, sym=False)
    assert np.array_equal(t.indices, oidx)
    assert part == opart
    info = h.info()
    assert info["uncompressed_size"] == n * n
    h.close()
    with pytest.raises(hmat_amd.EngineError, match="stored whole"):
        hmat_amd.HMatrix(t, lower_symmetric=True, value_type=hmat_amd.HMAT_DOUBLE_COMPLEX)
