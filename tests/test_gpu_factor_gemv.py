"""GPU test of the gemv of a factorised matrix (the gemv / gemm_dense slots
after factorize_generic): with an LLt factor they apply L or L^T (HMatrix::gemv
of the triangular factor, h_matrix.cpp:771-857), with a symmetric HODLR
factor L W or (L W)^T (HODLR<T>::gemv, hodlr.cpp:325-378) -- the products the
reference's examples/hodlrvsllt.c chains to rebuild A x from a factor."""
import sys

import numpy as np
import pytest

from oracle import hmat_oracle as ho

hmat_amd = pytest.importorskip("hmat_amd")

EPS = 1e-4


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def user(M, idx):
    """A dense matrix in hmat numbering moved to user numbering."""
    U = np.empty_like(M)
    U[np.ix_(idx, idx)] = M
    return U


@pytest.mark.gpu
def test_llt_factor_gemv():
    n, leaf = 3000, 128
    say(f"gemv of an LLt factor n={n}")
    pts = ho.cylinder_points(n, 1.0, 1.2, 4)
    t = hmat_amd.ClusterTree(pts, leaf)
    idx = t.indices
    h = hmat_amd.HMatrix(t)
    h.assemble(length=0.5, nugget=1e-2, eps=EPS)
    H = user(h.to_dense(), idx)
    h.factorize(hmat_amd.FACT_LLT)
    L = user(np.tril(h.to_dense(mirror=False)), idx)   # the stored factor
    x = np.random.default_rng(1).standard_normal((n, 2))
    yn = h.gemv(x)
    yt = h.gemv(x, trans="T")
    en = np.linalg.norm(yn - L @ x) / np.linalg.norm(L @ x)
    et = np.linalg.norm(yt - L.T @ x) / np.linalg.norm(L.T @ x)
    # hodlrvsllt.c: gemm_dense('T') then gemm_dense('N') gives A x to the factor's accuracy
    y2 = h.gemv(h.gemv(x, trans="T"), alpha=2.0, beta=-1.0, y=x.copy())
    rec = np.linalg.norm(y2 - (2 * H @ x - x)) / np.linalg.norm(2 * H @ x - x)
    say(f"L x vs dense {en:.2e}, L^T x {et:.2e}; 2 L L^T x - x vs 2 A x - x {rec:.2e}")
    assert en <= 1e-13 and et <= 1e-13      # measured 5e-16, 6e-16
    assert rec <= 10 * EPS
    h.close()
    # an LU factor's product is not provided
    g = hmat_amd.HMatrix(t, lower_symmetric=False)
    g.assemble(length=0.5, nugget=1e-2, eps=EPS)
    g.factorize(hmat_amd.FACT_LU)
    with pytest.raises(hmat_amd.EngineError, match="LLt and HODLR"):
        g.gemv(x)
    g.close()


@pytest.mark.gpu
def test_hodlr_factor_gemv():
    n, leaf = 2000, 64
    say(f"gemv of a HODLR factor n={n}")
    pts = ho.cylinder_points(n, 1.0, 1.0, 3)
    t = hmat_amd.ClusterTree(pts, leaf)
    idx = t.indices
    h = hmat_amd.HMatrix(t, hodlr=True)
    h.assemble(length=0.5, nugget=1e-2, eps=EPS)
    H = user(h.to_dense(), idx)
    h.factorize(hmat_amd.FACT_HODLRSYM)
    x = np.random.default_rng(2).standard_normal((n, 3))
    # (L W) (L W)^T = A exactly (the HODLR factorisation truncates nothing)
    w = h.gemv(h.gemv(x, trans="T"))
    rec = np.linalg.norm(w - H @ x) / np.linalg.norm(H @ x)
    # (L W)^{-1} is the lower solve
    back = h.solve_lower(h.gemv(x))
    inv = np.linalg.norm(back - x) / np.linalg.norm(x)
    # the oracle's restatement on its own factor of the same blocks
    root, oidx = ho.build_cluster_tree(pts, leaf)
    bt = ho.build_block_tree(root, root, 2.0, True, hodlr=True)
    ho.assemble(bt, oidx, ho.ExpKernel(pts, 0.5, 1e-2))
    node = ho.hodlr_factor_sym(bt)
    worst = 0.0
    for tr in ("N", "T"):
        yo = np.zeros((n, 3))
        ho.hodlr_gemv(bt, tr, x[idx], yo, node)
        ye = h.gemv(x, trans=tr)[idx]
        worst = max(worst, np.linalg.norm(ye - yo) / np.linalg.norm(yo))
    say(f"(LW)(LW)^T x vs A x {rec:.2e}; L^-1 (L W) x vs x {inv:.2e}; vs oracle {worst:.2e}")
    assert rec <= 1e-12 and inv <= 1e-12 and worst <= 1e-11   # measured 1.6e-15, 1.2e-15, 6.2e-14
    h.close()


@pytest.mark.gpu
def test_factorisation_progress():
    """hmat_progress_t of the assembly context (leaves assembled) and of the
    factorisation context: max = N, current = the end row of the last diagonal
    leaf factorised, non-decreasing, ending at N."""
    n = 20000
    say(f"factorisation progress n={n}")
    pts = ho.cylinder_points(n, 1.0, 1.6, 2)
    t = hmat_amd.ClusterTree(pts, 128)
    h = hmat_amd.HMatrix(t)
    asm = []
    h.assemble(length=0.5, nugget=1e-2, eps=EPS, progress=lambda cur, mx: asm.append((cur, mx)))
    # the assembly's reporter: max = the number of leaves, current = leaves
    # assembled, non-decreasing, ending at max
    nleaf = len(h.leaf_table())
    say(f"assembly progress calls: {asm}, {nleaf} leaves")
    assert asm and all(mx == nleaf for _, mx in asm)
    assert asm[-1][0] == nleaf and all(a[0] <= b[0] for a, b in zip(asm, asm[1:]))
    seen = []
    h.factorize(hmat_amd.FACT_LLT, progress=lambda cur, mx: seen.append((cur, mx)))
    say(f"progress calls: {seen}")
    assert seen and all(mx == n for _, mx in seen)
    assert seen[-1][0] == n and all(a[0] <= b[0] for a, b in zip(seen, seen[1:]))
    h.close()


@pytest.mark.gpu
def test_solve_autotune():
    """hmx_autotune_solve times the solve with each size of inverted diagonal
    blocks and keeps the fastest; the solution does not depend on it beyond
    rounding."""
    n = 20000
    say(f"solve autotune n={n}")
    pts = ho.cylinder_points(n, 1.0, 1.6, 2)
    t = hmat_amd.ClusterTree(pts, 128)
    h = hmat_amd.HMatrix(t)
    h.assemble(length=0.5, nugget=1e-2, eps=EPS)
    h.factorize(hmat_amd.FACT_LLT)
    b = np.random.default_rng(3).standard_normal((n, 8))
    x0 = h.solve(b)
    rows, ms = h.autotune_solve(8)
    x1 = h.solve(b)
    h.set_solve_block_rows(512)
    x2 = h.solve(b)
    d1 = np.linalg.norm(x1 - x0) / np.linalg.norm(x0)
    d2 = np.linalg.norm(x2 - x0) / np.linalg.norm(x0)
    say(f"autotune picked {rows} rows ({ms:.2f} ms per 8-RHS solve); solution changes {d1:.2e}, {d2:.2e}")
    assert rows in (512, 1024, 2048, 4096) and ms > 0
    assert d1 <= 1e-11 and d2 <= 1e-11
    h.close()
