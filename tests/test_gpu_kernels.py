"""GPU tests of the individual task kernels, one batched launch each through
the hmx_run_tasks test hook, against LAPACK / numpy and the oracle at leaf
sizes 16 to 256 (the dense leaf kernels: POTRF, GETRF, TRSM in its modes,
GEMM tiles, LDLt, diagonal scaling) and panel sizes up to 50,000 rows and
ranks up to ~120 (the recompression of a formatted sum on each of its paths:
LDS panels, global panels, the dense path and the single-precision Gram path;
the Rk product core; acaFull; the HODLR node blocks).  Several tasks of one
type go in one launch, as in the executor's waves."""
import math
import sys

import numpy as np
import pytest
import scipy.linalg as sl

from oracle import hmat_oracle as ho

hmat_amd = pytest.importorskip("hmat_amd")
torch = pytest.importorskip("torch")
import task_records as tr  # noqa: E402

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def spd(n, rng):
    g = rng.standard_normal((n, n))
    return g @ g.T / n + np.eye(n)


TOL = {False: 1e-12, True: 2e-5}


@pytest.mark.parametrize("fp32", [False, True])
def test_potrf_and_ldlt_leaves(fp32):
    rng = np.random.default_rng(1)
    sizes = [16, 64, 98, 128, 200, 256]          # above 128 the kernels work in global memory
    mats = [spd(n, rng) for n in sizes]
    A = [tr.dev(a, fp32) for a in mats]
    tasks = []
    for a, t_ in zip(mats, A):
        t = tr.FTask()
        t.type, t.m.s, t.A, t.lda = tr.T_POTRF, a.shape[0], t_.data_ptr(), a.shape[0]
        tasks.append(t)
    assert tr.run(tasks, tr.T_POTRF, fp32) == [0, 0]
    worst = max(rel(tr.host(t_), np.linalg.cholesky(a)) for a, t_ in zip(mats, A))
    say(f"POTRF fp32={fp32}: worst rel diff vs LAPACK {worst:.2e}")
    assert worst <= TOL[fp32]
    # LDLt of the same leaves: unit lower L and D against the oracle's restatement
    A = [tr.dev(a, fp32) for a in mats]
    D = [tr.dev(np.zeros(a.shape[0]), fp32) for a in mats]
    tasks = []
    for a, t_, d_ in zip(mats, A, D):
        t = tr.FTask()
        t.type, t.m.s, t.A, t.lda, t.B = tr.T_LDLT, a.shape[0], t_.data_ptr(), a.shape[0], d_.data_ptr()
        tasks.append(t)
    assert tr.run(tasks, tr.T_LDLT, fp32) == [0, 0]
    worst = 0.0
    for a, t_, d_ in zip(mats, A, D):
        L, d = ho.ldlt_leaf(a)
        worst = max(worst, rel(tr.host(t_), L), rel(tr.host(d_)[:, 0], d))
    say(f"LDLt fp32={fp32}: worst rel diff vs oracle {worst:.2e}")
    assert worst <= TOL[fp32]


def test_potrf_reports_non_spd():
    a = np.eye(32)
    a[5, 5] = -1.0
    t_ = tr.dev(a)
    t = tr.FTask()
    t.type, t.m.s, t.A, t.lda = tr.T_POTRF, 32, t_.data_ptr(), 32
    assert tr.run([t], tr.T_POTRF)[1] == 1


@pytest.mark.parametrize("fp32", [False, True])
def test_getrf_pivots_equal_lapack(fp32):
    rng = np.random.default_rng(2)
    sizes = [16, 98, 128, 200]
    mats = [rng.standard_normal((n, n)) + 0.1 * np.eye(n) for n in sizes]
    A = [tr.dev(a, fp32) for a in mats]
    P = [tr.ints(np.zeros(n)) for n in sizes]
    tasks = []
    for a, t_, p_ in zip(mats, A, P):
        t = tr.FTask()
        t.type, t.m.s, t.A, t.lda, t.B = tr.T_GETRF, a.shape[0], t_.data_ptr(), a.shape[0], p_.data_ptr()
        tasks.append(t)
    assert tr.run(tasks, tr.T_GETRF, fp32) == [0, 0]
    worst, same = 0.0, True
    for a, t_, p_ in zip(mats, A, P):
        lu, piv = sl.lu_factor(a.astype(np.float32) if fp32 else a)
        worst = max(worst, rel(tr.host(t_), lu))
        same &= np.array_equal(p_.cpu().numpy(), piv)
    say(f"GETRF fp32={fp32}: worst rel diff vs LAPACK {worst:.2e}, pivots equal {same}")
    assert same and worst <= (1e-4 if fp32 else 1e-11)


@pytest.mark.parametrize("fp32", [False, True])
def test_trsm_modes(fp32):
    rng = np.random.default_rng(3)
    cases = []  # (mode, n, other)
    for mode in (0, 1, 2, 1 | tr.TRSM_UNIT, 2 | tr.TRSM_TRANS_A, 1 | tr.TRSM_TRANS_A | tr.TRSM_UNIT):
        for n, other in ((16, 5), (98, 8), (128, 98), (200, 37), (98, 1562)):
            cases.append((mode, n, other))
    tasks, keep = [], []
    for mode, n, other in cases:
        Lf = np.linalg.cholesky(spd(n, rng))
        stored = Lf.T.copy() if mode & tr.TRSM_TRANS_A else Lf      # upper U when TransA
        X = rng.standard_normal((other, n) if (mode & 3) == 0 else (n, other))
        a_, x_ = tr.dev(stored, fp32), tr.dev(X, fp32)
        t = tr.FTask()
        t.type, t.mode, t.n.s, t.m.s = tr.T_TRSM, mode, n, other
        t.A, t.lda, t.C, t.ldc = a_.data_ptr(), n, x_.data_ptr(), X.shape[0]
        tasks.append(t)
        keep.append((mode, Lf, X, a_, x_))
    assert tr.run(tasks, tr.T_TRSM, fp32) == [0, 0]
    worst = 0.0
    for mode, Lf, X, a_, x_ in keep:
        # the lower operand is Lf (stored as is, or as its transpose with
        # TransA); the unit flag reads ones instead of its diagonal
        L = Lf.copy()
        if mode & tr.TRSM_UNIT:
            np.fill_diagonal(L, 1.0)
        base = mode & 3
        if base == 0:
            want = sl.solve_triangular(L, X.T, lower=True).T          # X L^-T
        elif base == 1:
            want = sl.solve_triangular(L, X, lower=True)              # L^-1 X
        else:
            want = sl.solve_triangular(L, X, lower=True, trans=1)     # L^-T X
        worst = max(worst, rel(tr.host(x_), want))
    say(f"TRSM fp32={fp32}: {len(cases)} cases, worst rel diff vs LAPACK {worst:.2e}")
    assert worst <= (5e-5 if fp32 else 1e-11)


@pytest.mark.parametrize("fp32", [False, True])
def test_gemm_tiles(fp32):
    rng = np.random.default_rng(4)
    cases = [(98, 98, 98, 0, 1, 0), (130, 70, 33, 1, 0, 1), (300, 8, 1562, 0, 0, 1), (1, 1, 1, 1, 1, 0),
             (64, 64, 4096, 1, 0, 0), (257, 129, 65, 0, 1, 1), (16, 200, 5, 1, 1, 1)]
    tasks, keep = [], []
    for m, n, p, ta, tb, acc in cases:
        A = rng.standard_normal((p, m) if ta else (m, p))
        B = rng.standard_normal((n, p) if tb else (p, n))
        Cm = rng.standard_normal((m, n))
        a_, b_, c_ = tr.dev(A, fp32), tr.dev(B, fp32), tr.dev(Cm, fp32)
        t = tr.FTask()
        t.type, t.mode, t.ta, t.tb, t.alpha = tr.T_GEMM, acc, ta, tb, -0.75
        t.m.s, t.n.s, t.p.s = m, n, p
        t.A, t.B, t.C = a_.data_ptr(), b_.data_ptr(), c_.data_ptr()
        t.lda, t.ldb, t.ldc = A.shape[0], B.shape[0], m
        tasks.append(t)
        keep.append((A, B, Cm, a_, b_, acc, c_))   # the operands stay alive until the launch
    assert tr.run(tasks, tr.T_GEMM, fp32) == [0, 0]
    worst = 0.0
    for (m, n, p, ta, tb, acc), (A, B, Cm, _, _, _, c_) in zip(cases, keep):  # noqa: B007
        want = acc * Cm - 0.75 * ((A.T if ta else A) @ (B.T if tb else B))
        e = rel(tr.host(c_), want)
        say(f"  GEMM {m}x{n}x{p} ta={ta} tb={tb} acc={acc}: {e:.2e}")
        worst = max(worst, e)
    say(f"GEMM fp32={fp32}: worst rel diff vs numpy {worst:.2e}")
    assert worst <= (2e-5 if fp32 else 1e-13)


def _rk(rng, m, n, k, decay=0.5):
    """A smooth-spectrum Rk block: A (m x k), B (n x k), singular values decaying."""
    qa, _ = np.linalg.qr(rng.standard_normal((m, k)))
    qb, _ = np.linalg.qr(rng.standard_normal((n, k)))
    s = decay ** np.arange(k)
    return qa * s, qb


@pytest.mark.parametrize("fp32", [False, True])
def test_fadd_recompression_paths(fp32):
    """RkMatrix::formattedAddParts + truncate: target + alpha_i parts at row /
    column offsets, against the truncated SVD of the dense sum at eps."""
    rng = np.random.default_rng(5)
    eps = 1e-4
    # (m, n, target rank, [(part rows, part cols, rank, roff, coff, alpha)]):
    # LDS panels, K >= min(m, n) (dense path), large global panels, long panels
    cases = [
        (98, 98, 6, [(98, 98, 5, 0, 0, -1.0), (49, 98, 3, 49, 0, 0.5)]),
        (40, 36, 20, [(40, 36, 20, 0, 0, -1.0)]),
        (390, 195, 10, [(195, 195, 8, 0, 0, -1.0), (195, 195, 8, 195, 0, 2.0), (390, 98, 4, 0, 97, 1.0)]),
        (1562, 1562, 12, [(1562, 1562, 12, 0, 0, -1.0), (781, 781, 9, 781, 781, 0.25)]),
        (20000, 300, 8, [(20000, 300, 10, 0, 0, -0.5), (5000, 300, 6, 15000, 0, 1.0)]),
        (50000, 98, 10, [(50000, 98, 12, 0, 0, -1.0)]),
        (1562, 781, 30, [(1562, 781, 30, 0, 0, -1.0), (1562, 781, 30, 0, 0, 0.5), (781, 781, 28, 0, 0, 1.0)]),
    ]
    tasks, keep = [], []
    for m, n, kt, parts in cases:
        K = kt + sum(p[2] for p in parts)
        cap = min(m, n, K)
        ta, tb = _rk(rng, m, n, kt)
        S = ta @ tb.T
        TA, TB = np.zeros((m, cap)), np.zeros((n, cap))
        TA[:, :kt], TB[:, :kt] = ta, tb
        a_, b_, r_ = tr.dev(TA, fp32), tr.dev(TB, fp32), tr.ints([kt])
        t = tr.FTask()
        t.type, t.eps, t.nparts = tr.T_FADD, eps, len(parts)
        t.tgt = tr.rkref(a_, b_, m, n, cap, r_)
        hold = [a_, b_, r_]
        for q, (pm, pn, pk, ro, co, al) in enumerate(parts):
            pa, pb = _rk(rng, pm, pn, pk, decay=0.6)
            S[ro:ro + pm, co:co + pn] += al * (pa @ pb.T)
            pa_, pb_, pr_ = tr.dev(pa, fp32), tr.dev(pb, fp32), tr.ints([pk])
            hold += [pa_, pb_, pr_]
            t.parts[q].rk = tr.rkref(pa_, pb_, pm, pn, pk, pr_)
            t.parts[q].roff, t.parts[q].coff, t.parts[q].alpha = ro, co, al
        tasks.append(t)
        keep.append((m, n, S, a_, b_, r_, hold))
    assert tr.run(tasks, tr.T_FADD, fp32) == [0, 0]
    for (m, n, S, a_, b_, r_, _), (_, _, kt, parts) in zip(keep, cases):
        k = int(r_.cpu().numpy()[0])
        u, s, vt = np.linalg.svd(S, full_matrices=False)
        kw = tr.find_k(s, eps)
        Sk = (u[:, :kw] * s[:kw]) @ vt[:kw]
        R = tr.host(a_)[:, :k] @ tr.host(b_)[:, :k].T
        err_opt = np.linalg.norm(S - Sk) / np.linalg.norm(S)
        err = np.linalg.norm(S - R) / np.linalg.norm(S)
        say(f"FADD {m}x{n} K={kt + sum(p[2] for p in parts)} fp32={fp32}: rank {k} (SVD rule {kw}), "
            f"error {err:.2e} (optimal {err_opt:.2e}), vs truncated SVD {rel(R, Sk):.2e}")
        if fp32:
            assert abs(k - kw) <= 1 and err <= 2 * err_opt + 1e-5
        else:
            assert k == kw and rel(R, Sk) <= 1e-10


@pytest.mark.parametrize("fp32", [False, True])
def test_rkrk_and_acafull(fp32):
    rng = np.random.default_rng(6)
    eps = 1e-4
    # RKRK: (A1 B1^T)(A2 B2^T), the k1 x k2 core B1^T A2 truncated at eps
    tasks, keep = [], []
    for m, p, n, k1, k2 in ((98, 98, 98, 10, 12), (390, 195, 781, 20, 8), (1562, 1562, 98, 30, 30)):
        a1, b1 = _rk(rng, m, p, k1)
        a2, b2 = _rk(rng, p, n, k2)
        cap = min(k1, k2)
        h = [tr.dev(a1, fp32), tr.dev(b1, fp32), tr.dev(a2, fp32), tr.dev(b2, fp32),
             tr.dev(np.zeros((m, cap)), fp32), tr.dev(np.zeros((n, cap)), fp32),
             tr.ints([k1]), tr.ints([k2]), tr.ints([0])]
        t = tr.FTask()
        t.type, t.eps = tr.T_RKRK, eps
        t.parts[0].rk = tr.rkref(h[0], h[1], m, p, k1, h[6])
        t.parts[1].rk = tr.rkref(h[2], h[3], p, n, k2, h[7])
        t.tgt = tr.rkref(h[4], h[5], m, n, cap, h[8])
        tasks.append(t)
        core = b1.T @ a2
        u, s, vt = np.linalg.svd(core)
        kw = tr.find_k(s, eps)
        keep.append((a1 @ (u[:, :kw] * s[:kw]) @ vt[:kw] @ b2.T, kw, h))
    assert tr.run(tasks, tr.T_RKRK, fp32) == [0, 0]
    for want, kw, h in keep:
        k = int(h[8].cpu().numpy()[0])
        got = tr.host(h[4])[:, :k] @ tr.host(h[5])[:, :k].T
        say(f"RKRK fp32={fp32}: rank {k} (rule {kw}), rel diff {rel(got, want):.2e}")
        assert abs(k - kw) <= (1 if fp32 else 0) and rel(got, want) <= (1e-5 if fp32 else 1e-11)
    # ACAFULL: full-pivot ACA of a dense block, against the oracle's acaFull
    tasks, keep = [], []
    for m, n, k in ((98, 98, 15), (128, 70, 30), (60, 128, 8)):
        a, b = _rk(rng, m, n, k, decay=0.4)
        M = a @ b.T + 1e-9 * rng.standard_normal((m, n))
        cap = min(m, n)
        h = [tr.dev(M, fp32), tr.dev(np.zeros((m, cap)), fp32), tr.dev(np.zeros((n, cap)), fp32),
             tr.ints([0])]
        t = tr.FTask()
        t.type, t.eps, t.A, t.lda, t.m.s, t.n.s = tr.T_ACAFULL, eps, h[0].data_ptr(), m, m, n
        t.tgt = tr.rkref(h[1], h[2], m, n, cap, h[3])
        tasks.append(t)
        keep.append((M, h))
    assert tr.run(tasks, tr.T_ACAFULL, fp32) == [0, 0]
    for M, h in keep:
        k = int(h[3].cpu().numpy()[0])
        got = tr.host(h[1])[:, :k] @ tr.host(h[2])[:, :k].T
        ref = ho.aca_full(M, eps)
        say(f"ACAFULL fp32={fp32}: rank {k} (oracle {ref.rank}), error {rel(got, M):.2e}, "
            f"vs oracle {rel(got, ref.eval(*M.shape, np.float64)):.2e}")
        assert abs(k - ref.rank) <= (1 if fp32 else 0)
        assert rel(got, M) <= 10 * eps
        if not fp32:
            assert rel(got, ref.eval(*M.shape, np.float64)) <= 1e-10


def test_hodlr_node_blocks():
    """The node blocks of a symmetric HODLR factor against the oracle's QR form."""
    rng = np.random.default_rng(7)
    tasks, keep = [], []
    for m1, m0, r in ((200, 180, 12), (1000, 1000, 40), (64, 64, 1)):
        a = 0.1 * rng.standard_normal((m1, r)) / math.sqrt(m1)
        b = 0.1 * rng.standard_normal((m0, r)) / math.sqrt(m0)
        G, P = tr.dev(a.T @ a), tr.dev(b.T @ b)
        h = [G, P, tr.dev(np.zeros((r, r))), tr.dev(np.zeros((r, r))), tr.dev(np.zeros(1)), tr.ints([r])]
        t = tr.FTask()
        t.type, t.lda = tr.T_HODLR, r
        t.m.s, t.m.d = r, h[5].data_ptr()
        t.A, t.B, t.C = G.data_ptr(), P.data_ptr(), h[2].data_ptr()
        t.tgt.a, t.tgt.b = h[3].data_ptr(), h[4].data_ptr()
        tasks.append(t)
        ra = np.linalg.qr(a, mode="r")
        x = ra @ (-(b.T @ b)) @ ra.T + np.eye(r)
        m11 = np.linalg.cholesky(x)
        x = sl.solve_triangular(ra, m11 - np.eye(r), lower=False)
        tmp = x @ ra + np.eye(r)
        x11 = sl.solve_triangular(ra, x.T, lower=False).T
        kk = np.linalg.solve(tmp, x11)
        keep.append((x11, kk, float(np.sum(np.log(np.diag(m11)))), h))
    assert tr.run(tasks, tr.T_HODLR) == [0, 0]
    for x11, kk, ld, h in keep:
        e1, e2 = rel(tr.host(h[2]), x11), rel(tr.host(h[3]), kk)
        e3 = abs(float(h[4].cpu()[0, 0]) - ld) / max(abs(ld), 1e-300)
        say(f"HODLR node r={x11.shape[0]}: X11 {e1:.2e}, KK {e2:.2e}, log det M11 {e3:.2e}")
        assert e1 <= 1e-10 and e2 <= 1e-10 and e3 <= 1e-10


def test_dscale_and_sumk():
    rng = np.random.default_rng(8)
    X = rng.standard_normal((300, 7))
    d = rng.uniform(0.5, 2.0, 300)
    x_, d_ = tr.dev(X), tr.dev(d)
    t = tr.FTask()
    t.type, t.mode, t.m.s, t.n.s = tr.T_DSCALE, 2, 300, 7          # rows, by the inverse
    t.A, t.lda, t.B, t.C, t.ldc = x_.data_ptr(), 300, d_.data_ptr(), x_.data_ptr(), 300
    assert tr.run([t], tr.T_DSCALE) == [0, 0]
    assert rel(tr.host(x_), X / d[:, None]) <= 1e-15
    W = rng.standard_normal((40 * 3, 5))                             # three 40 x 5 slices
    w_, c_ = tr.dev(W.reshape(3, 40, 5).transpose(1, 0, 2).reshape(40, 15)), tr.dev(np.ones((40, 5)))
    t = tr.FTask()
    t.type, t.mode, t.m.s, t.n.s, t.p.s = tr.T_SUMK, 1, 40, 5, 3
    t.A, t.lda, t.ldb, t.C, t.ldc = w_.data_ptr(), 40, 200, c_.data_ptr(), 40
    assert tr.run([t], tr.T_SUMK) == [0, 0]
    assert rel(tr.host(c_), 1.0 + W.reshape(3, 40, 5).sum(axis=0)) <= 1e-15
