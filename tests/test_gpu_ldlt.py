"""GPU test of the H-LDLt factorisation (factorize_generic with
hmat_factorization_ldlt; HMatrix::ldltDecomposition h_matrix.cpp:2747-2769,
recursiveLdltDecomposition recursion.cpp:37-83, solveLdlt :2990-3002).

* SPD exponential kernels (uniform and uneven leaf depths): the engine's
  factor and solution against the oracle's H-LDLt, and against the engine's
  own H-LLt of the same matrix (D = diag(C)^2, the same solution).
* A symmetric indefinite matrix (the kernel minus 87.44 I, three positive
  eigenvalues, condition ~4): solution against the oracle, residual against
  the dense H-matrix, and the inertia of D (Sylvester's law).
* The lower-triangular solve slot with the unit L, the FP32 factor with FP64
  refinement, and a serialized LDLt factor read back."""
import struct
import sys

import numpy as np
import pytest

from oracle import hmat_oracle as ho

hmat_amd = pytest.importorskip("hmat_amd")

EPS = 1e-4
SHIFT = 87.44226910401406     # between the 3rd and 4th largest eigenvalues (N=1500 case)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _case(n, height, seed, leaf, nugget):
    pts = ho.cylinder_points(n, 1.0, height, seed)
    root, idx = ho.build_cluster_tree(pts, leaf)
    t = hmat_amd.ClusterTree(pts, leaf)
    assert np.array_equal(t.indices, idx)
    bt = ho.build_block_tree(root, root, 2.0, True)
    ho.assemble(bt, idx, ho.ExpKernel(pts, 0.5, nugget))
    return pts, t, idx, bt


def _to_user(xo, idx):
    x = np.empty_like(xo)
    x[idx] = xo
    return x


@pytest.mark.gpu
@pytest.mark.parametrize("n,height,seed,leaf", [(4000, 1.6, 3, 128), (2050, 1.0, 5, 128)])
def test_ldlt_spd_matches_oracle_and_llt(n, height, seed, leaf):
    say(f"ldlt spd n={n} leaf={leaf}")
    pts, t, idx, bt = _case(n, height, seed, leaf, 1e-2)
    h = hmat_amd.HMatrix(t)
    h.assemble(length=0.5, nugget=1e-2, eps=EPS)
    H = h.to_dense()
    h.factorize(hmat_amd.FACT_LDLT)
    L = np.tril(h.to_dense(mirror=False))
    b = np.random.default_rng(seed).standard_normal((n, 2))
    x = h.solve(b)
    d = np.zeros(n)
    ho.ldlt(bt, EPS, d)
    Lo = ho.factor_dense(bt, n)
    xo = b[idx].copy()
    ho.solve_ldlt(bt, d, xo)
    xou = _to_user(xo, idx)
    rel = np.linalg.norm(x - xou) / np.linalg.norm(xou)
    lrel = np.linalg.norm(L - Lo) / np.linalg.norm(Lo)
    # the engine's own LLt of the same matrix
    g = hmat_amd.HMatrix(t)
    g.assemble(length=0.5, nugget=1e-2, eps=EPS)
    g.factorize(hmat_amd.FACT_LLT)
    C = np.tril(g.to_dense(mirror=False))
    xc = g.solve(b)
    g.close()
    res = np.linalg.norm(H @ x[idx] - b[idx]) / np.linalg.norm(b)
    rel_c = np.linalg.norm(x - xc) / np.linalg.norm(xc)
    lc = np.linalg.norm(L - C / np.diag(C)[None, :]) / np.linalg.norm(L)
    dl = h.extract_diagonal()[idx]
    dc = np.abs(dl / np.diag(C) ** 2 - 1).max()
    do = np.abs(dl / d - 1).max()
    say(f"solution vs oracle {rel:.2e}, L vs oracle {lrel:.2e}, D vs oracle {do:.2e}; vs own LLt: "
        f"solution {rel_c:.2e}, L vs C diag(C)^-1 {lc:.2e}, D vs diag(C)^2 {dc:.2e}; "
        f"residual vs H {res:.2e}")
    assert np.allclose(np.diag(L), 1.0)
    assert rel <= 1e-5                  # measured 2.9e-6 / 2.3e-8
    assert lrel <= 1e-5                 # measured 2.3e-6 / 2.7e-8
    assert rel_c <= 1e-12               # measured 4.8e-14 / 4.4e-14
    assert lc <= 1e-12                  # measured 1.3e-14 / 1.1e-14
    assert dc <= 1e-12
    assert do <= 5e-5                   # measured 5.8e-6 / 6.8e-7
    assert res <= 2e-3                  # measured 1.4e-3 / 5.9e-4
    h.close()


@pytest.mark.gpu
def test_ldlt_indefinite():
    n, leaf = 1500, 64
    say(f"ldlt indefinite n={n} shift {SHIFT}")
    pts, t, idx, bt = _case(n, 1.5, 5, leaf, -SHIFT)
    h = hmat_amd.HMatrix(t)
    h.assemble(length=0.5, nugget=-SHIFT, eps=EPS)
    H = h.to_dense()
    ev = np.linalg.eigvalsh(H)
    h.factorize(hmat_amd.FACT_LDLT)
    dd = h.extract_diagonal()[idx]     # D of the factor (extract_diagonal), hmat order
    b = np.random.default_rng(7).standard_normal((n, 3))
    x = h.solve(b)
    d = np.zeros(n)
    ho.ldlt(bt, EPS, d)
    xo = b[idx].copy()
    ho.solve_ldlt(bt, d, xo)
    xou = _to_user(xo, idx)
    rel = np.linalg.norm(x - xou) / np.linalg.norm(xou)
    res = np.linalg.norm(H @ x[idx] - b[idx]) / np.linalg.norm(b)
    say(f"positive eigenvalues {(ev > 0).sum()}, positive D {(dd > 0).sum()}; "
        f"D vs oracle {np.abs(dd / d - 1).max():.2e}; "
        f"solution vs oracle {rel:.2e}; residual vs H {res:.2e}")
    assert (ev > 0).sum() == 3 and (dd > 0).sum() == 3
    assert np.abs(dd / d - 1).max() <= 1e-8      # measured 1.2e-10
    assert rel <= 1e-10                 # measured 8.4e-14
    assert res <= 2e-4                  # measured 6.3e-05
    h.close()


@pytest.mark.gpu
def test_ldlt_lower_solve_refinement_and_serialization():
    import torch
    n, leaf = 3000, 128
    say(f"ldlt lower solve, mixed refinement, serialization n={n}")
    pts, t, idx, bt = _case(n, 1.2, 23, leaf, 1e-2)
    h = hmat_amd.HMatrix(t)
    h.assemble(length=0.5, nugget=1e-2, eps=EPS)
    H = h.to_dense()
    h.factorize(hmat_amd.FACT_LDLT)
    L = np.tril(h.to_dense(mirror=False))
    b = np.random.default_rng(4).standard_normal((n, 2))
    # lower-triangular slot: L y = b and L^T y = b with the unit L (hmat order)
    for tr in (False, True):
        y = h.solve_lower(b, transpose=tr, permute=False)
        r = np.linalg.norm((L.T if tr else L) @ y - b) / np.linalg.norm(b)
        say(f"solve_lower transpose={tr}: residual {r:.2e}")
        assert r <= 1e-12
    # serialized factor read back: same bytes, same solve
    s, dta = h.write_struct(), h.write_data()
    assert struct.unpack_from("<ii", s) == (hmat_amd.HMAT_DOUBLE_PRECISION, 1)
    g = hmat_amd.HMatrix.read_struct(s)
    g.read_data(dta)
    assert g.write_data() == dta
    assert np.array_equal(g.solve(b), h.solve(b))
    g.close()
    h.close()
    # FP32 LDLt + FP64 refinement sweeps against the FP64 H-matrix
    m = hmat_amd.HMatrix(t)
    m.assemble(length=0.5, nugget=1e-2, eps=EPS)
    m.factorize(hmat_amd.FACT_LDLT, mixed=True)
    bd = torch.tensor(b.T.copy(), dtype=torch.float64, device="cuda").contiguous()
    it, rel = m.solve_refined_device(bd.data_ptr(), 2, tol=1e-10, max_iter=40)
    torch.cuda.synchronize()
    x = bd.cpu().numpy().T
    res = np.linalg.norm(H @ x - b, axis=0) / np.linalg.norm(b, axis=0)
    say(f"FP32 LDLt + refinement: {it} sweeps, residual {res.max():.2e}")
    assert res.max() <= 1e-10
    m.close()
