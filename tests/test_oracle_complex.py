"""The oracle's complex (Z) path pinned by mathematics: the truncated SVD's
A B^T reproduces a complex low-rank matrix, partial ACA on a Helmholtz block
reaches the compression tolerance with the SVD's rank, the H-gemv equals the
dense product of the stored leaves, and the complex H-LU solve agrees with a
dense LAPACK solve of the same H-matrix and of the exact kernel matrix.
numpy's ComplexWarning (a complex value stored into a real array) is an error
here, so a real temporary left in the arithmetic fails the tests."""
import warnings

import numpy as np
import pytest

from oracle import hmat_oracle as ho


@pytest.fixture(autouse=True)
def complex_warnings_are_errors():
    with warnings.catch_warnings():
        warnings.simplefilter("error", np.exceptions.ComplexWarning)
        yield


def helmholtz_problem(n, height, seed=3):
    """Points on a cylinder, mean nearest-neighbour spacing h, ten points per
    wavelength (k = 2 pi / (10 h))."""
    from scipy.spatial import cKDTree
    pts = ho.cylinder_points(n, 1.0, height, seed)
    d, _ = cKDTree(pts).query(pts, 2)
    h = float(d[:, 1].mean())
    return pts, h, 2 * np.pi / (10 * h)


def test_truncated_svd_complex_is_a_b_transpose():
    rng = np.random.default_rng(0)
    a = rng.standard_normal((40, 5)) + 1j * rng.standard_normal((40, 5))
    b = rng.standard_normal((30, 5)) + 1j * rng.standard_normal((30, 5))
    m = a @ b.T
    rk = ho.truncated_svd(m, 1e-12)
    assert rk.rank == 5
    assert np.linalg.norm(rk.a @ rk.b.T - m) <= 1e-13 * np.linalg.norm(m)
    t = ho.rk_truncate(ho.Rk(np.hstack([a, a]), np.hstack([b, -0.5 * b])), 40, 30, 1e-12)
    assert t.rank == 5
    assert np.linalg.norm(t.a @ t.b.T - 0.5 * m) <= 1e-13 * np.linalg.norm(m)


def test_aca_partial_complex_helmholtz_block():
    pts, h, k = helmholtz_problem(4000, 2.0)
    ker = ho.HelmholtzKernel(pts, k, h, 1.0)
    # two well separated groups of 300 points (bottom and top of the cylinder)
    ri = np.argsort(pts[:, 2])[:300]
    ci = np.argsort(pts[:, 2])[-300:]
    full = ker.block(ri, ci)
    rk = ho.aca_partial(ho._BlockFn(ker, ri, ci), 300, 300, 1e-4)
    err = np.linalg.norm(rk.a @ rk.b.T - full) / np.linalg.norm(full)
    s = np.linalg.svd(full, compute_uv=False)
    k_svd = ho.find_k(s, 1e-4)
    t = ho.rk_truncate(rk, 300, 300, 1e-4)
    print(f"partial ACA rank {rk.rank}, error {err:.2e}; truncated {t.rank}, SVD rank {k_svd}")
    assert err <= 1e-3
    assert rk.rank >= k_svd
    assert abs(t.rank - k_svd) <= 2
    # the kernel itself: the diagonal value and the symmetric off-diagonal value
    d = ker.block(np.array([5, 7]), np.array([5, 7]))
    assert d[0, 0] == 1.0 / (4 * np.pi * h) + 1.0
    r = np.sqrt(((pts[5] - pts[7]) ** 2).sum() + h * h)
    assert abs(d[0, 1] - np.exp(1j * k * r) / (4 * np.pi * r)) <= 1e-15 and d[0, 1] == d[1, 0]


@pytest.mark.parametrize("n,leaf,height", [(1200, 64, 1.0), (2050, 128, 1.0)])
def test_complex_h_lu_solve(n, leaf, height):
    pts, h, k = helmholtz_problem(n, height)
    root, idx = ho.build_cluster_tree(pts, leaf)
    bt = ho.build_block_tree(root, root, 2.0, False)
    ker = ho.HelmholtzKernel(pts, k, h, 1.0)
    with ho.scalar_type(np.complex128):
        ho.assemble(bt, idx, ker, method="aca_partial", eps=1e-4)
        H = ho.to_dense(bt, n, n, dtype=np.complex128, symmetric=False)
        exact = ker.block(idx, idx)
        assert np.linalg.norm(H - exact) <= 1e-4 * np.linalg.norm(exact)
        rng = np.random.default_rng(1)
        x = rng.standard_normal((n, 2)) + 1j * rng.standard_normal((n, 2))
        y = np.zeros((n, 2), dtype=np.complex128)
        ho.gemv(bt, "N", 1.0, x, 0.0, y)
        assert np.linalg.norm(y - H @ x) <= 1e-13 * np.linalg.norm(y)
        b = rng.standard_normal((n, 3)) + 1j * rng.standard_normal((n, 3))
        ho.lu(bt, 1e-4)
        xs = b.copy()
        ho.solve_lu(bt, xs)
    xe = np.linalg.solve(H, b)
    rel = np.linalg.norm(xs - xe) / np.linalg.norm(xe)
    res = np.linalg.norm(H @ xs - b) / np.linalg.norm(b)
    print(f"n={n}: H-LU solution vs dense solve of H {rel:.2e}, residual {res:.2e}")
    assert xs.dtype == np.complex128
    assert rel <= 1e-3 and res <= 1e-3
