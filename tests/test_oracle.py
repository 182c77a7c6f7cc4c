"""Pins the oracle restatement with checks the mathematics fixes: closed-form
truncation errors, exact-rank recovery of ACA on separable blocks, dense
reconstruction of the H-matrix and of its LLt factor, exactness of the
triangular solves, and the determinant identity of logdet.  (The reference
ships no golden vectors for this path.)"""
import numpy as np
import pytest

from oracle import hmat_oracle as ho


def orthonormal(rng, m, k):
    q, _ = np.linalg.qr(rng.standard_normal((m, k)))
    return q


def test_splitmix64_known_answer():
    # first outputs of splitmix64 seeded with 0 (published test vector)
    s, z1 = ho.splitmix64(0)
    s, z2 = ho.splitmix64(s)
    assert z1 == 0xE220A8397B1DCDAF
    assert z2 == 0x6E789E6AA1B965F4


def test_find_k_rule():
    sig = np.array([1.0, 0.5, 1e-3, 1e-4, 2e-5])
    assert ho.find_k(sig, 1e-4) == 3          # sigma_3 == eps * sigma_0 is cut (<=)
    assert ho.find_k(sig, 1e-6) == 5
    assert ho.find_k(np.zeros(4), 1e-4) == 0


def test_truncated_svd_closed_form():
    rng = np.random.default_rng(1)
    s = np.array([3.0, 1.0, 2e-2, 5e-4, 1e-4, 1e-7])
    u, v = orthonormal(rng, 40, 6), orthonormal(rng, 30, 6)
    M = (u * s) @ v.T
    rk = ho.truncated_svd(M, 1e-4)
    assert rk.rank == 4
    # best rank-4 approximation: spectral error = sigma_4
    assert np.isclose(np.linalg.norm(M - rk.a @ rk.b.T, 2), 1e-4, rtol=1e-6)


def test_rk_truncate_recovers_rank_and_error():
    rng = np.random.default_rng(2)
    s = np.array([1.0, 0.3, 1e-2, 3e-5, 1e-9])
    qa, qb = orthonormal(rng, 50, 5), orthonormal(rng, 60, 5)
    # non-orthogonal panels of the same product
    g = rng.standard_normal((5, 5)) + 5 * np.eye(5)
    A = (qa * np.sqrt(s)) @ g
    B = (qb * np.sqrt(s)) @ np.linalg.inv(g).T
    M = A @ B.T
    out = ho.rk_truncate(ho.Rk(A, B), 50, 60, 1e-4)
    assert out.rank == 3
    assert np.isclose(np.linalg.norm(M - out.a @ out.b.T, 2), 3e-5, rtol=1e-5)


def test_rk_truncate_of_duplicated_panels():
    rng = np.random.default_rng(3)
    A1, B1 = rng.standard_normal((40, 3)), rng.standard_normal((35, 3))
    A = np.hstack([A1, A1])
    B = np.hstack([B1, -0.5 * B1])
    out = ho.rk_truncate(ho.Rk(A, B), 40, 35, 1e-10)
    assert out.rank == 3
    assert np.allclose(out.a @ out.b.T, 0.5 * A1 @ B1.T, atol=1e-12)


class Separable:
    """K(i, j) = sum_l g_l(x_i) h_l(y_j): exact rank r."""

    def __init__(self, x, y, r):
        self.x, self.y, self.r = x, y, r

    def block(self, ri, ci):
        G = np.stack([np.cos((l + 1) * self.x[ri]) for l in range(self.r)], 1)
        H = np.stack([np.exp(-(l + 1) * self.y[ci]) for l in range(self.r)], 1)
        return G @ H.T


@pytest.mark.parametrize("algo", ["aca_plus", "aca_partial", "aca_full"])
@pytest.mark.parametrize("r", [1, 3, 6])
def test_aca_recovers_exact_rank(algo, r):
    rng = np.random.default_rng(r)
    x, y = rng.uniform(0, 1, 300), rng.uniform(0, 1, 250)
    k = Separable(x, y, r)
    ri, ci = np.arange(300), np.arange(250)
    M = k.block(ri, ci)
    blk = ho._BlockFn(k, ri, ci)
    if algo == "aca_plus":
        rk = ho.aca_plus(blk, 300, 250, 1e-12)
    elif algo == "aca_partial":
        rk = ho.aca_partial(blk, 300, 250, 1e-12)
    else:
        rk = ho.aca_full(M, 1e-12)
    assert np.linalg.norm(M - rk.a @ rk.b.T) <= 1e-9 * np.linalg.norm(M)
    t = ho.rk_truncate(rk, 300, 250, 1e-10)
    assert t.rank == r


def _problem(n=2000, h=0.8, seed=12, leaf=128):
    pts = ho.cylinder_points(n, 1.0, h, seed)
    root, idx = ho.build_cluster_tree(pts, leaf)
    bt = ho.build_block_tree(root, root, 2.0, True)
    k = ho.ExpKernel(pts, 0.5, 1e-2)
    ho.assemble(bt, idx, k)
    return pts, root, idx, bt, k


def test_assembly_blocks_within_eps():
    pts, root, idx, bt, k = _problem()
    for b in ho.leaves(bt):
        exact = k.block(idx[b.rows.offset:b.rows.offset + b.m], idx[b.cols.offset:b.cols.offset + b.n])
        got = b.full if b.kind == ho.FULL else b.rk.eval(b.m, b.n, float)
        assert np.linalg.norm(got - exact) <= 3e-4 * np.linalg.norm(exact)


def test_hgemv_and_llt_against_dense():
    pts, root, idx, bt, k = _problem()
    n = len(idx)
    H = ho.to_dense(bt, n, n)
    assert np.allclose(H, H.T)
    x = np.random.default_rng(0).standard_normal((n, 2))
    y = np.ones_like(x)
    ho.gemv(bt, "N", 2.0, x, -1.0, y)
    assert np.allclose(y, 2.0 * H @ x - 1.0, rtol=0, atol=1e-12 * np.abs(H @ x).max())
    ho.llt(bt, 1e-4)
    L = ho.factor_dense(bt, n)
    assert np.allclose(np.triu(L, 1), 0)
    assert np.linalg.norm(L @ L.T - H) <= 2e-4 * np.linalg.norm(H)
    b = np.random.default_rng(1).standard_normal((n, 1))
    xs = b.copy()
    ho.solve_llt(bt, xs)
    assert np.linalg.norm(L @ (L.T @ xs) - b) <= 1e-10 * np.linalg.norm(b) * np.linalg.cond(L) ** 0
    # logdet returns log det L = 0.5 log det A
    assert np.isclose(ho.logdet(bt), np.log(np.diag(L)).sum(), rtol=1e-12)
    assert np.isclose(ho.logdet(bt), 0.5 * np.linalg.slogdet(H)[1], rtol=1e-4)


def test_info_counts_match_partition():
    pts, root, idx, bt, k = _problem(1000, 0.4, 3, 64)
    inf = ho.info(bt)
    lv = ho.leaves(bt)
    assert inf["full_count"] + inf["rk_count"] == len(lv)
    n = len(idx)
    # uncompressed size of the stored lower triangle incl. diagonal blocks
    diag = sum(b.m * b.n for b in lv if b.rows is b.cols)
    assert diag + 2 * (inf["uncompressed_size"] - diag) == n * n
