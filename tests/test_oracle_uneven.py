"""Pins the oracle on partitions whose leaves sit at more than one depth.

When the cluster tree is uneven (N=2050 with leaf 128, N=1030 with leaf 64),
the H-LLt / H-LU recursion reaches branches that uniform trees never do:
an H right-hand side on a leaf diagonal (evaluate, solve, clear, add back:
h_matrix.cpp:2386-2397), Full x H products (multiplyFullH h_matrix.cpp:1623),
fullHHGemm into a full leaf (h_matrix.cpp:1449-1472) and the H += Full axpy
(h_matrix.cpp:1034-1073).  Each of those oracle functions is checked here in
situ against dense arithmetic on its own operands, and the factorisations
that use them against the dense matrix (reconstruction and residual scale
with eps; logdet equals half the dense log-determinant)."""
import collections

import numpy as np
import pytest

from oracle import hmat_oracle as ho

UNEVEN = [(2050, 1.0, 5, 128), (1030, 0.6, 5, 64)]


def _tree(n, height, seed, leaf, sym=True, kernel="exp", eps=1e-4):
    pts = ho.cylinder_points(n, 1.0, height, seed)
    root, idx = ho.build_cluster_tree(pts, leaf)
    bt = ho.build_block_tree(root, root, 2.0, sym)
    if kernel == "exp":
        k = ho.ExpKernel(pts, 0.5, 1e-2)
    else:
        k = ho.ShiftedExpKernel(pts, 0.5, 0.2, 0.3)
    ho.assemble(bt, idx, k, eps=eps, lowrank_eps=eps)
    return pts, root, idx, bt, k


def _leaf_depths(root):
    return {c.depth for c in ho.cluster_leaves(root)}


@pytest.mark.parametrize("n,height,seed,leaf", UNEVEN)
def test_partition_is_uneven(n, height, seed, leaf):
    pts, root, idx, bt, k = _tree(n, height, seed, leaf)
    assert len(_leaf_depths(root)) > 1
    # a leaf diagonal block with a non-leaf block below it exists
    found = []

    def rec(b):
        if b.is_leaf():
            return
        for i in range(b.nrow if b.rows is b.cols else 0):
            d = b.get(i, i)
            for r in range(i + 1, b.nrow):
                below = b.get(r, i)
                if d is not None and d.is_leaf() and below is not None and not below.is_leaf():
                    found.append((d, below))
        for c in b.children:
            if c is not None:
                rec(c)

    rec(bt)
    assert found


class _Spy:
    """Wraps oracle functions, checks each call against dense arithmetic on
    its own operands, and counts the calls."""

    def __init__(self, monkeypatch, eps):
        self.calls = collections.Counter()
        self.worst = collections.defaultdict(float)
        self.eps = eps
        self._wrap(monkeypatch, "multiply_full", self._multiply_full)
        self._wrap(monkeypatch, "axpy_full", self._axpy_full)
        self._wrap(monkeypatch, "full_hh_gemm", self._full_hh_gemm)
        self._wrap(monkeypatch, "_clear_and_set", self._clear_and_set)
        self._wrap(monkeypatch, "_eval", self._eval)

    def _wrap(self, mp, name, checker):
        orig = getattr(ho, name)

        def f(*a, **kw):
            self.calls[name] += 1
            return checker(orig, *a, **kw)

        mp.setattr(ho, name, f)
        self._orig = getattr(self, "_orig", {})
        self._orig[name] = orig

    def _dense(self, h):
        # dense value by summing leaves into an array (independent of _eval's walk order)
        d = np.zeros((h.m, h.n))
        for lf in ho.leaves(h):
            r0, c0 = lf.rows.offset - h.rows.offset, lf.cols.offset - h.cols.offset
            if lf.kind == ho.FULL and lf.full is not None:
                d[r0:r0 + lf.m, c0:c0 + lf.n] += lf.full
            elif lf.kind == ho.RK and lf.rk is not None and lf.rk.rank:
                d[r0:r0 + lf.m, c0:c0 + lf.n] += lf.rk.a @ lf.rk.b.T
        return d

    @staticmethod
    def _op(d, t):
        return d if t == "N" else d.T

    def _note(self, name, err):
        self.worst[name] = max(self.worst[name], err)

    def _eval(self, orig, h):
        out = orig(h)
        ref = self._dense(h)
        self._note("_eval", np.abs(out - ref).max() / max(np.abs(ref).max(), 1e-300))
        return out

    def _multiply_full(self, orig, ta, tb, a, b):
        out = orig(ta, tb, a, b)
        if out is not None and (not a.is_leaf() or not b.is_leaf()):
            ref = self._op(self._dense(a), ta) @ self._op(self._dense(b), tb)
            self._note("multiply_full", np.linalg.norm(out - ref) / max(np.linalg.norm(ref), 1e-300))
        return out

    def _axpy_full(self, orig, h, alpha, rows, cols, f, eps):
        before = self._dense(h)
        r0, c0 = h.rows.offset - rows.offset, h.cols.offset - cols.offset
        ref = before + alpha * f[r0:r0 + h.m, c0:c0 + h.n]
        orig(h, alpha, rows, cols, f, eps)
        self._note("axpy_full", np.linalg.norm(self._dense(h) - ref) / max(np.linalg.norm(ref), 1e-300))

    def _full_hh_gemm(self, orig, c, ta, tb, alpha, a, b, eps):
        before = self._dense(c)
        prod = self._op(self._dense(a), ta) @ self._op(self._dense(b), tb)
        orig(c, ta, tb, alpha, a, b, eps)
        ref = before + alpha * prod
        self._note("full_hh_gemm", np.linalg.norm(self._dense(c) - ref) / max(np.linalg.norm(ref), 1e-300))

    def _clear_and_set(self, orig, b, d, eps):
        orig(b, d, eps)
        for lf in ho.leaves(b):
            r0, c0 = lf.rows.offset - b.rows.offset, lf.cols.offset - b.cols.offset
            sub = d[r0:r0 + lf.m, c0:c0 + lf.n]
            if lf.kind == ho.FULL:
                assert np.array_equal(lf.full, sub)
            else:
                got = lf.rk.a @ lf.rk.b.T if lf.rk.rank else np.zeros_like(sub)
                # truncated SVD at eps: spectral error at most eps * sigma_0
                s0 = np.linalg.norm(sub, 2)
                self._note("_clear_and_set", np.linalg.norm(sub - got, 2) / max(s0, 1e-300))


@pytest.mark.parametrize("n,height,seed,leaf", UNEVEN)
def test_llt_uneven_partition(monkeypatch, n, height, seed, leaf):
    eps = 1e-4
    pts, root, idx, bt, k = _tree(n, height, seed, leaf, eps=eps)
    H = ho.to_dense(bt, n, n)
    spy = _Spy(monkeypatch, eps)
    ho.llt(bt, eps)
    # every uneven-depth branch is reached
    for name in ("full_hh_gemm", "axpy_full", "_eval", "_clear_and_set", "multiply_full"):
        assert spy.calls[name] > 0, name
    assert spy.worst["_eval"] == 0.0
    assert spy.worst["multiply_full"] <= 1e-13
    assert spy.worst["_clear_and_set"] <= 1.0001 * eps
    # Rk leaves of the target are recompressed at eps
    assert spy.worst["axpy_full"] <= 2 * eps
    assert spy.worst["full_hh_gemm"] <= 2 * eps
    L = ho.factor_dense(bt, n)
    assert np.allclose(np.triu(L, 1), 0)
    rec = np.linalg.norm(L @ L.T - H) / np.linalg.norm(H)
    assert rec <= 3e-5
    assert np.isclose(ho.logdet(bt), 0.5 * np.linalg.slogdet(H)[1], rtol=1e-4)
    b = np.random.default_rng(1).standard_normal((n, 2))
    x = b.copy()
    ho.solve_llt(bt, x)
    assert np.linalg.norm(L @ (L.T @ x) - b) <= 1e-11 * np.linalg.norm(b)


@pytest.mark.parametrize("eps,bound", [(1e-6, 1e-6), (1e-10, 1e-10)])
def test_llt_uneven_reconstruction_tracks_eps(eps, bound):
    n, height, seed, leaf = UNEVEN[1]
    pts, root, idx, bt, k = _tree(n, height, seed, leaf, eps=eps)
    H = ho.to_dense(bt, n, n)
    ho.llt(bt, eps)
    L = ho.factor_dense(bt, n)
    assert np.linalg.norm(L @ L.T - H) <= bound * np.linalg.norm(H)


@pytest.mark.parametrize("n,height,seed,leaf", UNEVEN)
@pytest.mark.parametrize("eps,bound", [(1e-8, 1e-5), (1e-12, 1e-9)])
def test_lu_uneven_partition(monkeypatch, n, height, seed, leaf, eps, bound):
    pts, root, idx, bt, k = _tree(n, height, seed, leaf, sym=False, kernel="shifted", eps=eps)
    H = ho.to_dense(bt, n, n, symmetric=False)
    spy = _Spy(monkeypatch, eps)
    ho.lu(bt, eps)
    for name in ("full_hh_gemm", "axpy_full", "_eval", "_clear_and_set", "multiply_full"):
        assert spy.calls[name] > 0, name
    assert spy.worst["multiply_full"] <= 1e-13
    assert spy.worst["_clear_and_set"] <= 1.0001 * eps
    b = np.random.default_rng(3).standard_normal((n, 2))
    x = b.copy()
    ho.solve_lu(bt, x)
    res = np.linalg.norm(H @ x - b) / np.linalg.norm(b)
    assert res <= bound
    if eps == 1e-12:
        xe = np.linalg.solve(H, b)
        assert np.linalg.norm(x - xe) <= 1e-7 * np.linalg.norm(xe)
