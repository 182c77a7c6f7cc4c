"""GPU parity of the H-LLt factorisation, the H-triangular solves and logdet
against the oracle (same seeded inputs, same partition)."""
import sys

import numpy as np
import pytest

from oracle import hmat_oracle as ho

hmat_amd = pytest.importorskip("hmat_amd")

EPS = 1e-4


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def dense_lower_factor(h, precision=0):
    return np.tril(h.to_dense(precision=precision, mirror=False))


@pytest.mark.gpu
@pytest.mark.parametrize("n,height,seed,leaf", [(1500, 0.6, 3, 64), (4000, 1.6, 9, 128)])
def test_llt_fp64_matches_oracle(n, height, seed, leaf):
    say(f"llt fp64 n={n} leaf={leaf} seed={seed}")
    pts = ho.cylinder_points(n, 1.0, height, seed)
    t = hmat_amd.ClusterTree(pts, leaf)
    h = hmat_amd.HMatrix(t)
    h.assemble(length=0.5, nugget=1e-2, eps=EPS)
    H = h.to_dense()
    h.factorize(hmat_amd.FACT_LLT)
    st = h.stats()
    say("stats", st)
    L = dense_lower_factor(h)
    rec = np.linalg.norm(L @ L.T - H) / np.linalg.norm(H)
    # oracle on the same inputs
    root, idx = ho.build_cluster_tree(pts, leaf)
    bt = ho.build_block_tree(root, root, 2.0, True)
    k = ho.ExpKernel(pts, 0.5, 1e-2)
    ho.assemble(bt, idx, k)
    Ho = ho.to_dense(bt, n, n)
    ho.llt(bt, EPS)
    Lo = ho.factor_dense(bt, n)
    rec_o = np.linalg.norm(Lo @ Lo.T - Ho) / np.linalg.norm(Ho)
    say("||LL^T - H|| / ||H|| gpu", rec, "oracle", rec_o)
    # the same recompressions at the same eps: the factorisation error is the
    # oracle's to a few digits (and about eps)
    assert abs(rec - rec_o) <= 1e-2 * rec_o and rec < 10 * EPS
    ld_o = ho.logdet(bt)
    ld_g = h.logdet()
    say("logdet gpu", ld_g, "oracle", ld_o, "rel", abs(ld_g - ld_o) / abs(ld_o))
    assert abs(ld_g - ld_o) <= 1e-8 * abs(ld_o)       # measured 1.5e-9 / 2.2e-9
    sign, ld_dense = np.linalg.slogdet(H)
    assert abs(ld_g - 0.5 * ld_dense) <= 1e-4 * abs(ld_dense)
    # solve: user numbering through solve_systems
    rng = np.random.default_rng(seed)
    b = rng.standard_normal((n, 2))
    x = h.solve(b)
    A = k.block(np.arange(n), np.arange(n))
    res = np.linalg.norm(A @ x - b) / np.linalg.norm(b)
    xo = b[idx].copy()
    ho.solve_llt(bt, xo)
    xou = np.empty_like(xo)
    xou[idx] = xo
    res_o = np.linalg.norm(A @ xou - b) / np.linalg.norm(b)
    rel = np.linalg.norm(x - xou) / np.linalg.norm(xou)
    say("residual gpu", res, "oracle", res_o, "solution rel diff", rel)
    assert abs(res - res_o) <= 1e-3 * res_o                # agree to 5 digits measured
    assert rel <= 1e-5                                     # measured 2.8e-8 / 1.3e-6
    # solve against the H-matrix itself: H x = b holds to the factor accuracy,
    # the same as the oracle's factor against its own H
    xh = h.solve(b[idx], permute=False)
    res_h = np.linalg.norm(H @ xh - b[idx]) / np.linalg.norm(b)
    xoh = b[idx].copy()
    ho.solve_llt(bt, xoh)
    res_oh = np.linalg.norm(Ho @ xoh - b[idx]) / np.linalg.norm(b)
    say("residual vs H gpu", res_h, "oracle", res_oh)
    assert abs(res_h - res_oh) <= 1e-2 * res_oh
    h.close()


@pytest.mark.gpu
def test_llt_fp32_type():
    say("llt fp32 n=2500")
    n = 2500
    pts = ho.cylinder_points(n, 1.0, 1.0, 17)
    t = hmat_amd.ClusterTree(pts, 128)
    h = hmat_amd.HMatrix(t, value_type=hmat_amd.HMAT_SIMPLE_PRECISION)
    h.assemble(length=0.5, nugget=1e-2, eps=EPS)
    H = h.to_dense(precision=1)
    h.factorize(hmat_amd.FACT_LLT)
    L = dense_lower_factor(h, precision=1)
    rec = np.linalg.norm(L @ L.T - H) / np.linalg.norm(H)
    say("fp32 ||LL^T - H|| / ||H||", rec)
    assert rec < 10 * EPS                                   # measured 2.0e-5
    b = np.random.default_rng(0).standard_normal((n, 1)).astype(np.float32)
    x = h.solve(b)
    k = ho.ExpKernel(pts, 0.5, 1e-2)
    A = k.block(np.arange(n), np.arange(n))
    res = np.linalg.norm(A @ x.astype(np.float64) - b) / np.linalg.norm(b)
    # the oracle's FP64 H-LLt solve on the same inputs: FP32 storage changes
    # the solution by far less than the compression does
    root, idx = ho.build_cluster_tree(pts, 128)
    bt = ho.build_block_tree(root, root, 2.0, True)
    ho.assemble(bt, idx, k)
    ho.llt(bt, EPS)
    xo = b[idx].astype(np.float64)
    ho.solve_llt(bt, xo)
    xou = np.empty_like(xo)
    xou[idx] = xo
    res_o = np.linalg.norm(A @ xou - b) / np.linalg.norm(b)
    rel = np.linalg.norm(x - xou) / np.linalg.norm(xou)
    say("fp32 residual", res, "oracle fp64", res_o, "solution rel diff", rel)
    assert res <= 1.1 * res_o
    assert rel <= 1e-4
    h.close()


@pytest.mark.gpu
def test_not_spd_reports_error():
    """A negative nugget large enough to break positive definiteness makes the
    factorisation fail with an error (as potrf's info > 0 in the reference)."""
    say("non-SPD input")
    pts = ho.cylinder_points(800, 1.0, 0.4, 1)
    t = hmat_amd.ClusterTree(pts, 64)
    h = hmat_amd.HMatrix(t)
    h.assemble(length=0.5, nugget=-5.0, eps=EPS)
    with pytest.raises(hmat_amd.EngineError, match="positive definite"):
        h.factorize(hmat_amd.FACT_LLT)
    h.close()


@pytest.mark.gpu
def test_mixed_precision_refinement_reaches_1e10():
    """FP32 H-LLt + FP64 iterative refinement: the residual against the FP64
    H-matrix reaches 1e-10 and the solution agrees with the plain FP64 path."""
    import torch
    say("mixed precision refinement n=3000")
    n = 3000
    pts = ho.cylinder_points(n, 1.0, 1.2, 23)
    t = hmat_amd.ClusterTree(pts, 128)
    h = hmat_amd.HMatrix(t)
    h.assemble(length=0.5, nugget=1e-2, eps=EPS)
    H = h.to_dense()
    h.factorize(hmat_amd.FACT_LLT, mixed=True)
    rng = np.random.default_rng(5)
    b = rng.standard_normal((n, 3))
    bd = torch.tensor(b.T.copy(), dtype=torch.float64, device="cuda").contiguous()
    it, rel = h.solve_refined_device(bd.data_ptr(), 3, tol=1e-10, max_iter=40)
    torch.cuda.synchronize()
    x = bd.cpu().numpy().T
    res = np.linalg.norm(H @ x - b, axis=0) / np.linalg.norm(b, axis=0)
    say("IR sweeps", it, "reported", rel, "checked", res.max())
    assert res.max() <= 1e-10
    xe = np.linalg.solve(H, b)
    assert np.linalg.norm(x - xe) / np.linalg.norm(xe) < 1e-8
    # plain FP64 factor on a second matrix: both solutions approximate H^{-1} b
    h2 = hmat_amd.HMatrix(t)
    h2.assemble(length=0.5, nugget=1e-2, eps=EPS)
    h2.factorize(hmat_amd.FACT_LLT)
    x64 = h2.solve(b, permute=False)
    say("fp64 direct vs refined", np.linalg.norm(x64 - x) / np.linalg.norm(x))
    h.close()
    h2.close()


@pytest.mark.gpu
def test_solve_graph_replays_match_eager():
    """The first solve for a given nrhs runs its recorded task list eagerly;
    later solves replay it as a captured HIP graph.  Every replay, with other
    eager work and a re-capture (after another plan moved the flag buffer) in
    between, gives bitwise the eager result."""
    say("solve graph replays n=3000")
    n = 3000
    pts = ho.cylinder_points(n, 1.0, 1.2, 31)
    t = hmat_amd.ClusterTree(pts, 128)
    h = hmat_amd.HMatrix(t)
    h.assemble(length=0.5, nugget=1e-2, eps=EPS)
    h.factorize(hmat_amd.FACT_LLT)
    rng = np.random.default_rng(8)
    b = rng.standard_normal((n, 2))
    c = rng.standard_normal((n, 3))
    x_eager = h.solve(b, permute=False)            # records the nrhs=2 plan, eager
    for rep in range(4):
        x = h.solve(b, permute=False)              # graph capture / replay
        say("replay", rep, "max diff", float(np.abs(x - x_eager).max()))
        assert np.array_equal(x, x_eager)
        if rep == 1:
            y_eager = h.solve(c, permute=False)    # another plan, recorded eagerly
            y = h.solve(c, permute=False)
            assert np.array_equal(y, y_eager)
        if rep == 2:
            _ = h.to_dense()                       # unrelated eager work
    h.close()


def _shifted_problem(n, height, seed, leaf, nugget=0.2, shift=0.3):
    pts = ho.cylinder_points(n, 1.0, height, seed)
    t = hmat_amd.ClusterTree(pts, leaf)
    h = hmat_amd.HMatrix(t, lower_symmetric=False)
    h.assemble(kind=hmat_amd.KERNEL_SHIFTED_EXPONENTIAL, length=0.5, nugget=nugget, eps=EPS,
               shift=shift)
    return pts, t, h


@pytest.mark.gpu
@pytest.mark.parametrize("n,height,seed,leaf,nugget,shift",
                         [(1200, 0.5, 4, 64, 0.2, 0.3), (3000, 1.2, 6, 128, 0.4, 0.4)])
def test_lu_fp64_matches_oracle(n, height, seed, leaf, nugget, shift):
    """H-LU (non-symmetric storage, non-symmetric kernel, leaf interchanges
    present) against the oracle's H-LU on the same inputs: solution and residual."""
    say(f"lu fp64 n={n} leaf={leaf} seed={seed}")
    pts, t, h = _shifted_problem(n, height, seed, leaf, nugget, shift)
    H = h.to_dense(mirror=False)
    h.factorize(hmat_amd.FACT_LU)
    say("stats", h.stats())
    root, idx = ho.build_cluster_tree(pts, leaf)
    bt = ho.build_block_tree(root, root, 2.0, False)
    k = ho.ShiftedExpKernel(pts, 0.5, nugget, shift)
    ho.assemble(bt, idx, k)
    ho.lu(bt, EPS)
    rng = np.random.default_rng(seed)
    b = rng.standard_normal((n, 2))
    x = h.solve(b)                                   # user numbering
    A = k.block(np.arange(n), np.arange(n))
    res = np.linalg.norm(A @ x - b) / np.linalg.norm(b)
    xo = b[idx].copy()
    ho.solve_lu(bt, xo)
    xou = np.empty_like(xo)
    xou[idx] = xo
    res_o = np.linalg.norm(A @ xou - b) / np.linalg.norm(b)
    rel = np.linalg.norm(x - xou) / np.linalg.norm(xou)
    say("residual gpu", res, "oracle", res_o, "solution rel diff", rel)
    assert abs(res - res_o) <= 1e-2 * res_o               # measured 2e-5 relative
    assert rel <= 5e-5                                     # measured 3.2e-6 / 8.0e-6
    xh = h.solve(b[idx], permute=False)
    res_h = np.linalg.norm(H @ xh - b[idx]) / np.linalg.norm(b)
    say("residual vs H", res_h)
    assert res_h <= res_o                                  # H is closer to itself than A
    h.close()


@pytest.mark.gpu
def test_lu_mixed_precision_refinement_reaches_1e10():
    import torch
    say("lu mixed precision refinement n=2500")
    n = 2500
    pts, t, h = _shifted_problem(n, 1.0, 12, 128)
    H = h.to_dense(mirror=False)
    h.factorize(hmat_amd.FACT_LU, mixed=True)
    b = np.random.default_rng(9).standard_normal((n, 2))
    bd = torch.tensor(b.T.copy(), dtype=torch.float64, device="cuda").contiguous()
    it, rel = h.solve_refined_device(bd.data_ptr(), 2, tol=1e-10, max_iter=40)
    torch.cuda.synchronize()
    x = bd.cpu().numpy().T
    res = np.linalg.norm(H @ x - b, axis=0) / np.linalg.norm(b, axis=0)
    say("LU IR sweeps", it, "reported", rel, "checked", res.max())
    assert res.max() <= 1e-10
    h.close()


@pytest.mark.gpu
def test_lu_singular_reports_error():
    """Two identical points and no nugget give two identical rows: getrf ends
    on an exact zero pivot and the factorisation reports an error."""
    say("singular LU")
    pts = ho.cylinder_points(40, 1.0, 0.3, 2)
    pts[7] = pts[3]
    t = hmat_amd.ClusterTree(pts, 64)
    h = hmat_amd.HMatrix(t, lower_symmetric=False)
    h.assemble(length=0.5, nugget=0.0, eps=EPS)
    with pytest.raises(hmat_amd.EngineError, match="singular"):
        h.factorize(hmat_amd.FACT_LU)
    h.close()
