"""GPU tests of the whole-matrix slots of hmat_interface_t: copy, copy_struct,
norm, scale, add_identity, extract_diagonal, truncate, axpy and
solve_lower_triangular(_dense), each checked against dense linear algebra on
the stored leaves (to_dense) or on the exact kernel."""
import sys

import numpy as np
import pytest

from oracle import hmat_oracle as ho

hmat_amd = pytest.importorskip("hmat_amd")

EPS = 1e-4


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _matrix(n=3000, height=1.2, seed=3, leaf=128, length=0.5, lowrank_eps=EPS, vt=None):
    pts = ho.cylinder_points(n, 1.0, height, seed)
    t = hmat_amd.ClusterTree(pts, leaf)
    kw = {} if vt is None else {"value_type": vt}
    h = hmat_amd.HMatrix(t, lowrank_eps=lowrank_eps, **kw)
    h.assemble(length=length, nugget=1e-2, eps=EPS)
    return pts, t, h


@pytest.mark.gpu
def test_copy_scale_add_identity_norm():
    say("copy / copy_struct / scale / add_identity / norm n=3000")
    pts, t, h = _matrix()
    D = h.to_dense()
    c = h.copy()
    assert np.array_equal(c.to_dense(), D)
    assert c.info()["compressed_size"] == h.info()["compressed_size"]
    z = h.copy(structure_only=True)
    assert not np.any(z.to_dense())
    assert z.info()["rk_count"] == h.info()["rk_count"]
    # norm: Frobenius norm of the whole (mirrored) matrix
    nrm = h.norm()
    ref = np.linalg.norm(D)
    say("norm", nrm, "dense", ref)
    assert abs(nrm - ref) <= 1e-12 * ref
    # scale the copy: the copy changes, the original does not
    c.scale(-0.75)
    assert np.allclose(c.to_dense(), -0.75 * D, rtol=0, atol=1e-15 * np.abs(D).max())
    assert np.array_equal(h.to_dense(), D)
    assert abs(c.norm() - 0.75 * ref) <= 1e-12 * ref
    # add_identity
    c.add_identity(2.5)
    E = c.to_dense() - (-0.75 * D)
    assert np.allclose(E, 2.5 * np.eye(h.n), rtol=0, atol=1e-14)
    # extract_diagonal, user numbering
    d = h.extract_diagonal()
    idx = t.indices
    du = np.empty(h.n)
    du[idx] = np.diag(D)
    assert np.array_equal(d, du)
    k = ho.ExpKernel(pts, 0.5, 1e-2)
    assert np.allclose(d, 1.0 + 1e-2, rtol=0, atol=1e-15)   # exp(0) + nugget
    for m in (h, c, z):
        m.close()


@pytest.mark.gpu
def test_truncate_matches_svd_cut():
    say("truncate at a larger eps n=3000")
    pts, t, h = _matrix()
    before = list(h.leaves())
    D = h.to_dense()
    new_eps = 1e-2
    h.iface.set_low_rank_epsilon(h.handle, new_eps)
    h.truncate()
    after = list(h.leaves())
    worst = 0.0
    mism = 0
    for (m, n, r0, c0, kind, rank, a, b), (_, _, _, _, _, rank2, a2, b2) in zip(before, after):
        if kind == 0 or rank == 0:
            continue
        blk = a @ b.T
        s = np.linalg.svd(blk, compute_uv=False)
        # findK (scalar_array.cpp:98-116): keep sigma_i > eps * sigma_0
        want = int(np.sum(s > new_eps * s[0]))
        if rank2 != want:
            mism += 1
        # the QR + core-SVD truncation is the best rank-`want` approximation:
        # its Frobenius error is the norm of the dropped singular values
        new = a2 @ b2.T if rank2 > 0 else np.zeros((m, n))
        opt = np.sqrt(np.sum(s[rank2:] ** 2))
        got = np.linalg.norm(new - blk)
        worst = max(worst, abs(got - opt) / np.linalg.norm(blk))
    say("worst |error - optimal| / ||leaf||", worst, "rank mismatches", mism)
    assert mism == 0
    assert worst <= 1e-10
    assert sum(r[5] for r in after) < sum(r[5] for r in before)
    h.close()


@pytest.mark.gpu
def test_axpy_same_partition():
    say("axpy n=3000")
    pts, t, x = _matrix(length=0.5)
    y = hmat_amd.HMatrix(t)
    y.assemble(length=0.3, nugget=1e-2, eps=EPS)
    X, Y = x.to_dense(), y.to_dense()
    y.axpy(0.5, x)
    R = y.to_dense()
    err = np.linalg.norm(R - (Y + 0.5 * X)) / np.linalg.norm(Y + 0.5 * X)
    say("axpy rel err", err)
    assert err <= EPS
    # x + x: the same panels twice recompress to the same block
    c = x.copy()
    c.axpy(1.0, x)
    err2 = np.linalg.norm(c.to_dense() - 2 * X) / np.linalg.norm(2 * X)
    say("x + x rel err", err2)
    assert err2 <= 1e-12
    for m in (x, y, c):
        m.close()


@pytest.mark.gpu
def test_solve_lower_triangular():
    say("solve_lower_triangular n=3000")
    pts, t, h = _matrix()
    h.factorize(hmat_amd.FACT_LLT)
    L = np.tril(h.to_dense(mirror=False))
    b = np.random.default_rng(4).standard_normal((h.n, 3))
    idx = t.indices
    bh = b[idx]
    y = h.solve_lower(bh, transpose=False, permute=False)
    assert np.linalg.norm(L @ y - bh) <= 1e-12 * np.linalg.norm(bh)
    x = h.solve_lower(y, transpose=True, permute=False)
    assert np.linalg.norm(L.T @ x - y) <= 1e-12 * np.linalg.norm(y)
    # the two sweeps together are the full solve
    xs = h.solve(bh, permute=False)
    assert np.linalg.norm(xs - x) <= 1e-13 * np.linalg.norm(x)
    # user numbering
    yu = h.solve_lower(b, transpose=False, permute=True)
    yh = np.empty_like(yu)
    yh[:] = yu[idx]
    assert np.linalg.norm(yh - y) <= 1e-13 * np.linalg.norm(y)
    h.close()


@pytest.mark.gpu
def test_get_block_and_get_values():
    """get_block / get_values (c_wrapping.hpp:832-890): blocks in hmat
    numbering, entries at 1-based user indices, symmetric storage mirrored,
    a factor's stored blocks as they are."""
    n = 3000
    pts = ho.cylinder_points(n, 1.0, 1.2, 31)
    t = hmat_amd.ClusterTree(pts, 128)
    h = hmat_amd.HMatrix(t)
    h.assemble(length=0.5, nugget=1e-2, eps=1e-4)
    D = h.to_dense()                               # hmat numbering
    idx = t.indices
    B = h.get_block(0, n, 0, n)
    e1 = np.abs(B - D).max()
    S = h.get_block(700, 900, 1500, 1100)
    e2 = np.abs(S - D[700:1600, 1500:2600]).max()
    R = h.get_block(0, n, 100, 50, renumber_rows=True)
    U = np.empty_like(D)
    U[idx] = D
    e3 = np.abs(R - U[:, 100:150]).max()
    rng = np.random.default_rng(5)
    rows, cols = rng.choice(n, 37, replace=False), rng.choice(n, 23, replace=False)
    V = h.get_values(rows, cols)
    Du = np.empty_like(D)
    Du[np.ix_(idx, idx)] = D                       # user numbering
    e4 = np.abs(V - Du[np.ix_(rows, cols)]).max()
    h.factorize(hmat_amd.FACT_LLT)
    F = h.get_block(0, n, 0, n)
    e5 = np.abs(F - h.to_dense(mirror=False)).max()
    print(f"get_block full {e1:.1e}, sub {e2:.1e}, renumbered {e3:.1e}; get_values {e4:.1e}; factor {e5:.1e}")
    assert max(e1, e2, e3, e4) <= 1e-14 * np.abs(D).max() and e5 <= 1e-14 * np.abs(F).max()
    h.close()


@pytest.mark.gpu
def test_transpose():
    """hmat_interface_t.transpose (HMatrix::transpose): the stored matrix of a
    non-symmetric real matrix becomes its exact transpose (leaves exchanged
    with their partners, data moved, nothing recomputed) and its gemv the
    transposed product; a complex matrix becomes its transpose (not the
    conjugate transpose); symmetric storage is unchanged; factors are refused."""
    n, leaf = 3000, 64
    say(f"transpose n={n}")
    pts = ho.cylinder_points(n, 1.0, 1.2, 6)
    t = hmat_amd.ClusterTree(pts, leaf)
    h = hmat_amd.HMatrix(t, lower_symmetric=False)
    h.assemble(kind=hmat_amd.KERNEL_SHIFTED_EXPONENTIAL, length=0.5, nugget=1e-2, shift=0.3,
               eps=1e-4)
    D = h.to_dense()
    h.transpose()
    Dt = h.to_dense()
    assert np.array_equal(Dt, D.T)
    x = np.random.default_rng(1).standard_normal((n, 3))
    idx = t.indices
    ref = np.empty_like(x)
    ref[idx] = D.T @ x[idx]
    e = np.linalg.norm(h.gemv(x) - ref) / np.linalg.norm(ref)
    h.transpose()
    assert np.array_equal(h.to_dense(), D)
    from scipy.spatial import cKDTree
    dd, _ = cKDTree(pts).query(pts, 2)
    hreg = float(dd[:, 1].mean())
    z = hmat_amd.HMatrix(t, lower_symmetric=False, value_type=hmat_amd.HMAT_DOUBLE_COMPLEX)
    z.assemble(kind=hmat_amd.KERNEL_HELMHOLTZ, wavenumber=2 * np.pi / (10 * hreg), reg=hreg,
               nugget=1.0, eps=1e-4, method="aca_partial")
    Z = z.to_dense()
    z.transpose()
    ez = np.abs(z.to_dense() - Z.T).max() / np.abs(Z).max()
    say(f"gemv of the transpose {e:.2e}; complex transpose {ez:.2e}")
    assert e <= 1e-13 and ez <= 1e-15
    s = hmat_amd.HMatrix(t)
    s.assemble(length=0.5, nugget=1e-2, eps=1e-4)
    Ds = s.to_dense()
    s.transpose()
    assert np.array_equal(s.to_dense(), Ds)
    s.factorize(hmat_amd.FACT_LLT)
    h.factorize(hmat_amd.FACT_LU)
    with pytest.raises(hmat_amd.EngineError):
        h.transpose()
    for m in (h, z, s):
        m.close()
