"""GPU test of the HODLR factorisations (hmat_create_admissibility_hodlr,
factorize_generic with hmat_factorization_hodlrsym, hodlr.cpp:113-166,
221-312, and hmat_factorization_hodlr, hodlr.cpp:45-116): partition, factor,
solve, lower / upper solves and logdet against the oracle's restatement on the same assembled blocks and against the exact
dense algebra of the assembled matrix (the HODLR factorisation truncates
nothing), beside the H-LLt of the same matrix (the reference's
examples/hodlrvsllt.c comparison)."""
import sys

import numpy as np
import pytest

from oracle import hmat_oracle as ho

hmat_amd = pytest.importorskip("hmat_amd")

EPS = 1e-4
KIND = {0: "F", 1: "R", 2: "H"}


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.gpu
@pytest.mark.parametrize("n,leaf,height", [(2000, 64, 1.0), (4000, 128, 4.0)])
def test_hodlr_sym_matches_oracle(n, leaf, height):
    say(f"hodlrsym n={n} leaf={leaf} height={height}")
    pts = ho.cylinder_points(n, 1.0, height, 3)
    root, idx = ho.build_cluster_tree(pts, leaf)
    t = hmat_amd.ClusterTree(pts, leaf)
    assert np.array_equal(t.indices, idx)
    h = hmat_amd.HMatrix(t, hodlr=True)
    part = [(int(a), int(b), int(c), int(d), KIND[int(k)]) for a, b, c, d, k in h.partition()]
    bt = ho.build_block_tree(root, root, 2.0, True, hodlr=True)
    assert part == ho.partition_signature(bt)
    h.assemble(length=0.5, nugget=1e-2, eps=EPS)
    H = h.to_dense()
    h.factorize(hmat_amd.FACT_HODLRSYM)
    rng = np.random.default_rng(n)
    b = rng.standard_normal((n, 3))
    x = h.solve(b)
    ld = h.logdet()
    bh = b[idx]
    y = h.solve_lower(bh, permute=False)
    # the oracle on its own assembly of the same partition
    ho.assemble(bt, idx, ho.ExpKernel(pts, 0.5, 1e-2))
    node = ho.hodlr_factor_sym(bt)
    xo = bh.copy()
    ho.hodlr_lower(bt, xo, node)
    ho.hodlr_upper(bt, xo, node)
    xou = np.empty_like(xo)
    xou[idx] = xo
    ld_o = ho.hodlr_logdet(bt, node)
    # exact algebra of the engine's own matrix
    xe = np.linalg.solve(H, bh)
    sign, lde = np.linalg.slogdet(H)
    rel_o = np.linalg.norm(x - xou) / np.linalg.norm(xou)
    rel_e = np.linalg.norm(x[idx] - xe) / np.linalg.norm(xe)
    quad = np.abs((y * y).sum(axis=0) / (bh * xe).sum(axis=0) - 1).max()
    say(f"solution vs oracle {rel_o:.2e}, vs dense solve of H {rel_e:.2e}; logdet {ld:.12g} "
        f"oracle {ld_o:.12g} half slogdet {0.5 * lde:.12g}; |L^-1 b|^2 vs b^T H^-1 b {quad:.2e}")
    assert rel_o <= 1e-11               # measured 5.8e-14 / 4.9e-14
    assert rel_e <= 1e-11               # measured 6.6e-14 / 3.8e-14
    assert abs(ld - ld_o) <= 1e-12 * abs(ld_o)
    assert sign > 0 and abs(ld - 0.5 * lde) <= 1e-12 * abs(lde)
    assert quad <= 1e-12                # measured 1.1e-15 / 8.9e-16
    # the H-LLt of the same matrix (recompressions at eps): the same logdet to eps
    g = hmat_amd.HMatrix(t, hodlr=True)
    g.assemble(length=0.5, nugget=1e-2, eps=EPS)
    g.factorize(hmat_amd.FACT_LLT)
    ld_llt = g.logdet()
    xl = g.solve(b)
    say(f"H-LLt on the HODLR partition: logdet {ld_llt:.12g} (rel {abs(ld_llt - ld) / abs(ld):.2e}), "
        f"solution rel diff {np.linalg.norm(xl - x) / np.linalg.norm(x):.2e}")
    assert abs(ld_llt - ld) <= 1e-5 * abs(ld)   # measured 7.0e-7 / 3.4e-7
    g.close()
    h.close()


@pytest.mark.gpu
def test_hodlr_nonsym_matches_oracle():
    """hmat_factorization_hodlr (HODLR<T>::factorize / solve, hodlr.cpp:45-116):
    the same lower-stored matrix through the non-symmetric Woodbury form."""
    n, leaf = 3000, 64
    say(f"hodlr (non-symmetric) n={n} leaf={leaf}")
    pts = ho.cylinder_points(n, 1.0, 1.5, 3)
    root, idx = ho.build_cluster_tree(pts, leaf)
    t = hmat_amd.ClusterTree(pts, leaf)
    h = hmat_amd.HMatrix(t, hodlr=True)
    h.assemble(length=0.5, nugget=1e-2, eps=EPS)
    H = h.to_dense()
    h.factorize(hmat_amd.FACT_HODLR)
    b = np.random.default_rng(4).standard_normal((n, 4))
    x = h.solve(b)
    bt = ho.build_block_tree(root, root, 2.0, True, hodlr=True)
    ho.assemble(bt, idx, ho.ExpKernel(pts, 0.5, 1e-2))
    node = ho.hodlr_factor_ns(bt)
    xo = b[idx].copy()
    ho.hodlr_solve_ns(bt, xo, node)
    xou = np.empty_like(xo)
    xou[idx] = xo
    xe = np.linalg.solve(H, b[idx])
    rel_o = np.linalg.norm(x - xou) / np.linalg.norm(xou)
    rel_e = np.linalg.norm(x[idx] - xe) / np.linalg.norm(xe)
    say(f"solution vs oracle {rel_o:.2e}, vs dense solve of H {rel_e:.2e}")
    assert rel_o <= 1e-11 and rel_e <= 1e-11
    # what the reference's non-symmetric HODLR factor does not provide either
    with pytest.raises(hmat_amd.EngineError, match="logdet"):
        h.logdet()
    with pytest.raises(hmat_amd.EngineError, match="solve_lower_triangular"):
        h.solve_lower(b)
    h.close()


@pytest.mark.gpu
def test_hodlr_mixed_precision_and_errors():
    import torch
    n, leaf = 2000, 64
    say(f"hodlrsym FP32 factor + refinement n={n}")
    pts = ho.cylinder_points(n, 1.0, 1.0, 3)
    t = hmat_amd.ClusterTree(pts, leaf)
    h = hmat_amd.HMatrix(t, hodlr=True)
    h.assemble(length=0.5, nugget=1e-2, eps=EPS)
    H = h.to_dense()
    h.factorize(hmat_amd.FACT_HODLRSYM, mixed=True)
    b = np.random.default_rng(2).standard_normal((n, 2))
    bd = torch.tensor(b.T.copy(), dtype=torch.float64, device="cuda").contiguous()
    it, rel = h.solve_refined_device(bd.data_ptr(), 2, tol=1e-10, max_iter=40)
    torch.cuda.synchronize()
    x = bd.cpu().numpy().T
    res = np.linalg.norm(H @ x - b, axis=0) / np.linalg.norm(b, axis=0)
    say(f"{it} sweeps, residual {res.max():.2e}")
    assert res.max() <= 1e-10
    h.close()
    # a standard-admissibility partition is not HODLR
    s = hmat_amd.HMatrix(t)
    s.assemble(length=0.5, nugget=1e-2, eps=EPS)
    with pytest.raises(hmat_amd.EngineError, match="Not HODLR"):
        s.factorize(hmat_amd.FACT_HODLRSYM)
    with pytest.raises(hmat_amd.EngineError, match="HODLR"):
        s.factorize(hmat_amd.FACT_HODLR)
    s.close()
