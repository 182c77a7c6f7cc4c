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
    # the gemv of the factor applies L and L^T (not the mirrored matrix)
    v = np.random.default_rng(2).standard_normal((n, 2))
    for tr, op in (("N", L), ("T", L.T)):
        w = np.zeros_like(v)
        ho.factor_gemv(bt, tr, v, w)
        assert np.linalg.norm(w - op @ v) <= 1e-13 * np.linalg.norm(op @ v)
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


def _lu_problem(n=1200, eps=1e-4, leaf=64, shift=0.3, nugget=0.2, seed=4):
    pts = ho.cylinder_points(n, 1.0, 0.5, seed)
    root, idx = ho.build_cluster_tree(pts, leaf)
    bt = ho.build_block_tree(root, root, 2.0, False)
    k = ho.ShiftedExpKernel(pts, 0.5, nugget, shift)
    ho.assemble(bt, idx, k, eps=eps, lowrank_eps=eps)
    H = ho.to_dense(bt, n, n, symmetric=False)
    return bt, idx, k, H


def test_shifted_kernel_is_nonsymmetric_translation_invariant():
    pts = ho.cylinder_points(50, 1.0, 0.5, 1)
    k = ho.ShiftedExpKernel(pts, 0.5, 0.0, 0.3)
    i = np.arange(50)
    K = k.block(i, i)
    assert not np.allclose(K, K.T)
    # K(x, y) = exp(-|x - y - s e_z| / l): moving y by +s e_z maps to the symmetric kernel
    d = pts[:, None, :] - pts[None, :, :]
    d[:, :, 2] -= 0.3
    assert np.allclose(K, np.exp(-np.sqrt((d * d).sum(axis=2)) / 0.5), rtol=0, atol=1e-15)


def test_lu_single_leaf_is_getrf():
    """A matrix that is one leaf: the H-LU is LAPACK getrf and the solve is
    getrs (interchanges applied in getrf order, unit L, then U)."""
    import scipy.linalg as sla
    bt, idx, k, H = _lu_problem(n=60, leaf=64)
    assert bt.is_leaf()
    ho.lu(bt, 1e-4)
    lu_, piv = sla.lu_factor(H)
    assert np.array_equal(bt.full, lu_) and np.array_equal(bt.piv, piv)
    assert (piv != np.arange(len(piv))).any()          # pivoting really happens
    b = np.random.default_rng(2).standard_normal((60, 3))
    x = b.copy()
    ho.solve_lu(bt, x)
    assert np.allclose(x, np.linalg.solve(H, b), rtol=0, atol=1e-10 * np.abs(x).max())


@pytest.mark.parametrize("eps,bound", [(1e-4, 5e-3), (1e-8, 1e-6), (1e-12, 1e-10)])
def test_hlu_solve_residual_tracks_eps(eps, bound):
    """H-LU on non-symmetric storage of a non-symmetric kernel: the residual of
    the solve against the assembled matrix scales with the compression eps,
    and at eps=1e-12 the solution is the dense solve's."""
    bt, idx, k, H = _lu_problem(eps=eps)
    ho.lu(bt, eps)
    pivoted = [lf for lf in ho.leaves(bt) if lf.piv is not None and (lf.piv != np.arange(len(lf.piv))).any()]
    assert pivoted                                       # leaf interchanges are exercised
    b = np.random.default_rng(3).standard_normal((H.shape[0], 2))
    x = b.copy()
    ho.solve_lu(bt, x)
    res = np.linalg.norm(H @ x - b) / np.linalg.norm(b)
    assert res <= bound
    if eps == 1e-12:
        xe = np.linalg.solve(H, b)
        assert np.linalg.norm(x - xe) <= 1e-8 * np.linalg.norm(xe)


def test_hlu_upper_solves_against_dense():
    """The U-side solves of the recursion (x U^{-1}, U^{-T} x, U^{-1} x) on a
    factorised tree agree with dense triangular solves of its U part."""
    import scipy.linalg as sla
    bt, idx, k, H = _lu_problem(n=700, eps=1e-12, leaf=64)
    ho.lu(bt, 1e-12)
    n = H.shape[0]
    F = ho.to_dense(bt, n, n, symmetric=False)
    # U part of the factor: upper triangles of the diagonal leaves and every
    # block above them (block-upper storage of U)
    U = np.triu(F)
    rng = np.random.default_rng(5)
    x = rng.standard_normal((n, 2))
    y = x.copy()
    ho.lu_solve_upper_left_dense(bt, y)
    assert np.allclose(U @ y, x, rtol=0, atol=1e-9 * np.abs(x).max())
    y = x.copy()
    ho.lu_solve_upper_t_left_dense(bt, y)
    assert np.allclose(U.T @ y, x, rtol=0, atol=1e-9 * np.abs(x).max())
    z = rng.standard_normal((3, n))
    w = z.copy()
    ho.lu_solve_upper_right_dense(bt, w)
    assert np.allclose(w @ U, z, rtol=0, atol=1e-9 * np.abs(z).max())


def test_ldlt_leaf_identities():
    rng = np.random.default_rng(17)
    n = 40
    # SPD: L D L^T = A, L unit lower, and the Cholesky identities
    # D = diag(C)^2, L = C diag(C)^-1 for A = C C^T
    g = rng.standard_normal((n, n))
    a = g @ g.T + n * np.eye(n)
    L, d = ho.ldlt_leaf(a)
    assert np.allclose(np.diag(L), 1.0) and np.allclose(np.triu(L, 1), 0.0)
    assert np.linalg.norm(L @ np.diag(d) @ L.T - a) <= 1e-13 * np.linalg.norm(a)
    c = np.linalg.cholesky(a)
    assert np.allclose(d, np.diag(c) ** 2, rtol=1e-12)
    assert np.allclose(L, c / np.diag(c)[None, :], rtol=1e-11, atol=1e-13)
    # symmetric indefinite (leading minors nonzero): reconstruction, and the
    # signs of D follow Sylvester's law of inertia
    s = a - 1.3 * n * np.eye(n)
    L, d = ho.ldlt_leaf(s)
    assert np.linalg.norm(L @ np.diag(d) @ L.T - s) <= 1e-11 * np.linalg.norm(s)
    ev = np.linalg.eigvalsh(s)
    assert (d < 0).sum() == (ev < 0).sum() and 0 < (ev < 0).sum() < n
    with pytest.raises(ZeroDivisionError):
        ho.ldlt_leaf(np.array([[0.0, 1.0], [1.0, 2.0]]))


def _hmatrix(n, leaf, nugget, seed=5, height=1.5):
    pts = ho.cylinder_points(n, 1.0, height, seed)
    root, idx = ho.build_cluster_tree(pts, leaf)
    bt = ho.build_block_tree(root, root, 2.0, True)
    ho.assemble(bt, idx, ho.ExpKernel(pts, 0.5, nugget))
    return bt, ho.to_dense(bt, n, n)


def test_hldlt_against_llt_and_dense():
    n, leaf, eps = 1500, 64, 1e-4
    bt, H = _hmatrix(n, leaf, 1e-2)
    d = np.zeros(n)
    ho.ldlt(bt, eps, d)
    L = ho.factor_dense(bt, n)
    assert np.allclose(np.diag(L), 1.0)
    assert np.linalg.norm(L @ np.diag(d) @ L.T - H) <= 5e-5 * np.linalg.norm(H)
    b = np.random.default_rng(1).standard_normal((n, 2))
    x = b.copy()
    ho.solve_ldlt(bt, d, x)
    # the H-LLt factor of the same matrix: D = diag(C)^2 and the same solution
    # (measured 2e-14: the recompressions see the same blocks up to a diagonal
    # scaling, so they keep the same ranks)
    bc, _ = _hmatrix(n, leaf, 1e-2)
    ho.llt(bc, eps)
    C = ho.factor_dense(bc, n)
    xc = b.copy()
    ho.solve_llt(bc, xc)
    assert np.abs(d / np.diag(C) ** 2 - 1).max() <= 1e-12
    assert np.linalg.norm(x - xc) <= 1e-12 * np.linalg.norm(xc)
    assert np.linalg.norm(H @ x - b) <= 2e-3 * np.linalg.norm(b)


def test_hldlt_indefinite_inertia_and_residual():
    # exp kernel minus 87.44 I: three positive eigenvalues (the shift sits in
    # the gap between the third and fourth largest), condition number ~4
    n, leaf, eps = 1500, 64, 1e-4
    bt, H = _hmatrix(n, leaf, -87.44226910401406)
    ev = np.linalg.eigvalsh(H)
    assert (ev > 0).sum() == 3
    d = np.zeros(n)
    ho.ldlt(bt, eps, d)
    L = ho.factor_dense(bt, n)
    assert (d > 0).sum() == 3          # Sylvester's law of inertia
    assert np.linalg.norm(L @ np.diag(d) @ L.T - H) <= 2e-4 * np.linalg.norm(H)
    b = np.random.default_rng(2).standard_normal((n, 3))
    x = b.copy()
    ho.solve_ldlt(bt, d, x)
    assert np.linalg.norm(H @ x - b) <= 2e-4 * np.linalg.norm(b)


def _hodlr(n, leaf, height=1.0, seed=3):
    pts = ho.cylinder_points(n, 1.0, height, seed)
    root, idx = ho.build_cluster_tree(pts, leaf)
    bt = ho.build_block_tree(root, root, 2.0, True, hodlr=True)
    ho.assemble(bt, idx, ho.ExpKernel(pts, 0.5, 1e-2))
    return bt, ho.to_dense(bt, n, n)


def test_hodlr_partition_rule():
    # every inner block a lower-symmetric 2 x 2 node with an Rk (1, 0) child
    bt, _ = _hodlr(1000, 64)

    def walk(b):
        if b.is_leaf():
            assert b.kind == ho.FULL and b.rows is b.cols
            return
        assert b.sym and b.nrow == b.ncol == 2 and b.get(0, 1) is None
        assert b.get(1, 0).kind == ho.RK
        walk(b.get(0, 0))
        walk(b.get(1, 1))
    walk(bt)


def test_hodlr_nonsym_is_exact_factorisation():
    # the non-symmetric HODLR factorisation of the same lower-stored matrix
    # (Woodbury with an LU-factorised 2r x 2r core) solves exactly too
    n = 2000
    bt, H = _hodlr(n, 64)
    node = ho.hodlr_factor_ns(bt)
    b = np.random.default_rng(8).standard_normal((n, 3))
    x = b.copy()
    ho.hodlr_solve_ns(bt, x, node)
    xe = np.linalg.solve(H, b)
    assert np.linalg.norm(x - xe) <= 1e-10 * np.linalg.norm(xe)


def test_hodlr_sym_is_exact_factorisation():
    # the HODLR factorisation truncates nothing: its solve is the exact solve
    # of the assembled matrix, its logdet half the log-determinant, and
    # y = L^-1 b has |y|^2 = b^T A^-1 b
    n = 2000
    bt, H = _hodlr(n, 64)
    node = ho.hodlr_factor_sym(bt)
    b = np.random.default_rng(5).standard_normal((n, 3))
    x = b.copy()
    ho.hodlr_lower(bt, x, node)
    y = x.copy()
    ho.hodlr_upper(bt, x, node)
    xe = np.linalg.solve(H, b)
    assert np.linalg.norm(x - xe) <= 1e-10 * np.linalg.norm(xe)
    sign, ld = np.linalg.slogdet(H)
    assert sign > 0 and abs(ho.hodlr_logdet(bt, node) - 0.5 * ld) <= 1e-10 * abs(ld)
    assert np.allclose((y * y).sum(axis=0), (b * xe).sum(axis=0), rtol=1e-10)
    # the gemv of the factor: (L W) (L W)^T = A exactly, and (L W)^{-1} is the
    # lower solve (hodlr.cpp:325-378 against 113-156)
    v = np.random.default_rng(6).standard_normal((n, 2))
    t = np.zeros_like(v)
    ho.hodlr_gemv(bt, "T", v, t, node)
    w = np.zeros_like(v)
    ho.hodlr_gemv(bt, "N", t, w, node)
    assert np.linalg.norm(w - H @ v) <= 1e-12 * np.linalg.norm(H @ v)
    z = np.zeros_like(v)
    ho.hodlr_gemv(bt, "N", v, z, node)
    ho.hodlr_lower(bt, z, node)
    assert np.linalg.norm(z - v) <= 1e-12 * np.linalg.norm(v)
