"""GPU parity: assembly (full-leaf evaluation, SVD compression, ACA+ and Rk
truncation) and the leaf-level H-gemv against the oracle on the same seeded
inputs."""
import sys

import numpy as np
import pytest

from oracle import hmat_oracle as ho

hmat_amd = pytest.importorskip("hmat_amd")

EPS = 1e-4


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def build(n, height, seed, leaf=128, kind=hmat_amd.KERNEL_EXPONENTIAL, length=0.5, nugget=1e-2,
          method="aca_plus", vt=hmat_amd.HMAT_DOUBLE_PRECISION, eps=EPS, lowrank_eps=EPS):
    pts = ho.cylinder_points(n, 1.0, height, seed)
    t = hmat_amd.ClusterTree(pts, leaf)
    h = hmat_amd.HMatrix(t, value_type=vt, lowrank_eps=lowrank_eps)
    h.assemble(kind=kind, length=length, nugget=nugget, eps=eps, method=method)
    return pts, t, h


def oracle_build(pts, leaf=128, length=0.5, nugget=1e-2, kernel=ho.ExpKernel, eps=EPS,
                 lowrank_eps=EPS):
    root, idx = ho.build_cluster_tree(pts, leaf)
    bt = ho.build_block_tree(root, root, 2.0, True)
    k = kernel(pts, length, nugget)
    ho.assemble(bt, idx, k, eps=eps, lowrank_eps=lowrank_eps)
    return bt, idx, k


def _near_cut(block, rank_a, rank_b, cut):
    """True when a singular value between the two ranks lies within 1e-8 of
    the cut (a tie that rounding may decide either way)."""
    s = np.linalg.svd(block, compute_uv=False)
    lo, hi = min(rank_a, rank_b), max(rank_a, rank_b)
    return all(abs(s[i] / s[0] - cut) <= 1e-8 * cut for i in range(lo, hi))


@pytest.mark.gpu
@pytest.mark.parametrize("n,height,seed,leaf,eps,lowrank_eps",
                         [(3000, 1.2, 3, 128, EPS, EPS), (6000, 2.4, 11, 64, EPS, EPS),
                          (3000, 1.2, 3, 128, 1e-6, 1e-3)])
def test_assembly_leaves_match_oracle(n, height, seed, leaf, eps, lowrank_eps):
    say(f"assembly parity n={n} leaf={leaf} seed={seed} eps={eps} lowrank_eps={lowrank_eps}")
    pts, t, h = build(n, height, seed, leaf, eps=eps, lowrank_eps=lowrank_eps)
    bt, idx, k = oracle_build(pts, leaf, eps=eps, lowrank_eps=lowrank_eps)
    assert np.array_equal(t.indices, idx)
    oleaves = ho.leaves(bt)
    gleaves = list(h.leaves())
    assert len(oleaves) == len(gleaves)
    tol = max(eps, lowrank_eps)
    worst_full = 0.0
    worst_rk = 0.0
    worst_vs_oracle = 0.0
    mismatch = {"svd": 0, "aca": 0}
    count = {"svd": 0, "aca": 0}
    for ob, (m, nn, r0, c0, kind, rank, a, b) in zip(oleaves, gleaves):
        assert (ob.rows.offset, ob.m, ob.cols.offset, ob.n) == (r0, m, c0, nn)
        exact = k.block(idx[r0:r0 + m], idx[c0:c0 + nn])
        if kind == 0:
            assert ob.kind == ho.FULL
            worst_full = max(worst_full, np.abs(a - exact).max() / np.abs(exact).max())
        else:
            assert ob.kind == ho.RK
            g = a @ b.T if rank > 0 else np.zeros((m, nn))
            o = ob.rk.eval(m, nn, np.float64)
            nrm = np.linalg.norm(exact)
            # both are tol-approximations of the same block
            worst_rk = max(worst_rk, np.linalg.norm(g - exact) / nrm)
            worst_vs_oracle = max(worst_vs_oracle, np.linalg.norm(g - o) / nrm)
            cls = "svd" if max(m, nn) < 100 else "aca"
            count[cls] += 1
            if rank != ob.rk.rank:
                mismatch[cls] += 1
                if cls == "svd":
                    # SVD leaves: the same singular values cut by the same rule
                    assert _near_cut(exact, rank, ob.rk.rank, tol), (r0, c0, rank, ob.rk.rank)
                else:
                    assert abs(rank - ob.rk.rank) <= 1, (r0, c0, rank, ob.rk.rank)
    say("worst full", worst_full, "worst rk", worst_rk, "vs oracle", worst_vs_oracle,
        "rank mismatches", mismatch, "of", count)
    assert worst_full < 1e-14
    assert worst_rk < 3 * tol
    # same pivots, same cuts: the leaves equal the oracle's to rounding
    # (measured 3e-14), and every rank is the oracle's
    assert worst_vs_oracle < 1e-10
    assert mismatch["aca"] == 0
    inf = h.info()
    oinf = ho.info(bt)
    assert inf["full_count"] == oinf["full_count"] and inf["rk_count"] == oinf["rk_count"]
    assert inf["nr_block_clusters"] == oinf["nr_block_clusters"]
    if mismatch["svd"] == 0:
        assert inf["compressed_size"] == oinf["compressed_size"]
    h.close()


@pytest.mark.gpu
def test_aca_partial_and_svd_methods():
    say("aca partial / svd methods n=2500")
    pts = ho.cylinder_points(2500, 1.0, 1.0, 21)
    for method in ("aca_partial",):
        t = hmat_amd.ClusterTree(pts, 128)
        h = hmat_amd.HMatrix(t)
        h.assemble(eps=EPS, method=method)
        k = ho.ExpKernel(pts, 0.5, 1e-2)
        idx = t.indices
        A = k.block(idx, idx)
        D = h.to_dense()
        assert np.linalg.norm(D - A) / np.linalg.norm(A) < 4 * EPS
        h.close()


@pytest.mark.gpu
def test_svd_compression_every_size():
    """hmat_create_compression_svd (truncatedSvd, compression.cpp:224-239) on
    blocks of every size: below 100 rows by k_svd_compress, above by a dense
    recompression task; leaves equal the oracle's truncated SVDs."""
    n = 3000
    say(f"svd compression n={n}")
    pts = ho.cylinder_points(n, 1.0, 1.2, 23)
    t = hmat_amd.ClusterTree(pts, 128)
    h = hmat_amd.HMatrix(t)
    h.assemble(eps=EPS, method="svd")
    root, idx = ho.build_cluster_tree(pts, 128)
    bt = ho.build_block_tree(root, root, 2.0, True)
    ho.assemble(bt, idx, ho.ExpKernel(pts, 0.5, 1e-2), method="svd")
    worst, mism, big = 0.0, 0, 0
    for ob, (m, nn, r0, c0, kind, rank, a, b) in zip(ho.leaves(bt), h.leaves()):
        if kind == 0:
            continue
        big += max(m, nn) >= 100
        g = a @ b.T if rank > 0 else np.zeros((m, nn))
        o = ob.rk.eval(m, nn, np.float64)
        worst = max(worst, np.linalg.norm(g - o) / max(np.linalg.norm(o), 1e-300))
        mism += rank != ob.rk.rank
    say(f"svd leaves vs oracle: worst {worst:.2e}, rank mismatches {mism}, {big} blocks of 100 rows or more")
    assert big > 0 and mism == 0 and worst <= 1e-10
    h.close()


# N=6000 on a cylinder of height 8 with leaf 64 has Rk panels of up to 1500
# rows (oracle partition, eta 2), so the one-RHS z pass sees chunks in every
# row class (<= 64, 128, 256, 512 and 1024 rows), including full 1024-row
# chunks and the 476-row remainders of the 1500-row panels
@pytest.mark.gpu
@pytest.mark.parametrize("nrhs,n,height,leaf", [(1, 4000, 1.6, 128), (3, 4000, 1.6, 128),
                                                (9, 4000, 1.6, 128), (1, 6000, 8.0, 64),
                                                (1, 4000, 1.6, 400), (9, 4000, 1.6, 400)])
def test_gemv_device_matches_dense(nrhs, n, height, leaf):
    import torch
    say(f"gemv device nrhs={nrhs} n={n} leaf={leaf}")
    pts, t, h = build(n, height, 5, leaf=leaf)
    D = h.to_dense()
    rng = np.random.default_rng(nrhs)
    x = rng.standard_normal((h.n, nrhs))
    y0 = rng.standard_normal((h.n, nrhs))
    xd = torch.tensor(x.T.copy(), dtype=torch.float64, device="cuda").contiguous()  # column-major
    yd = torch.tensor(y0.T.copy(), dtype=torch.float64, device="cuda").contiguous()
    h.gemv_device(xd.data_ptr(), yd.data_ptr(), nrhs, alpha=0.7, beta=-0.3)
    torch.cuda.synchronize()
    y = yd.cpu().numpy().T
    ref = 0.7 * (D @ x) - 0.3 * y0
    err = np.linalg.norm(y - ref) / np.linalg.norm(ref)
    say("gemv rel err", err)
    assert err < 1e-13
    h.close()


@pytest.mark.gpu
def test_gemv_host_api_user_numbering():
    say("gemv through hmat_interface_t.gemv (user numbering)")
    pts, t, h = build(2500, 1.0, 8)
    k = ho.ExpKernel(pts, 0.5, 1e-2)
    A = k.block(np.arange(h.n), np.arange(h.n))
    x = np.random.default_rng(2).standard_normal((h.n, 2))
    y = h.gemv(x)
    err = np.linalg.norm(y - A @ x) / np.linalg.norm(A @ x)
    say("gemv vs dense kernel", err)
    assert err < 4 * EPS
    # oracle H-gemv on the oracle's own assembly agrees to the same tolerance
    bt, idx, _ = oracle_build(pts)
    xo = x[idx]
    yo = np.zeros_like(xo)
    ho.gemv(bt, "N", 1.0, xo, 0.0, yo)
    yo_user = np.empty_like(yo)
    yo_user[idx] = yo
    assert np.linalg.norm(y - yo_user) / np.linalg.norm(yo_user) < 4 * EPS
    h.close()


@pytest.mark.gpu
def test_gemv_fp32_storage():
    import torch
    say("gemv fp32 matrix")
    pts, t, h = build(3000, 1.2, 4, vt=hmat_amd.HMAT_SIMPLE_PRECISION)
    D = h.to_dense(precision=1)
    x = np.random.default_rng(3).standard_normal((h.n, 1)).astype(np.float32)
    xd = torch.tensor(x[:, 0], device="cuda")
    yd = torch.zeros_like(xd)
    h.gemv_device(xd.data_ptr(), yd.data_ptr(), 1, precision=1)
    torch.cuda.synchronize()
    ref = D @ x.astype(np.float64)
    err = np.linalg.norm(yd.cpu().numpy()[:, None] - ref) / np.linalg.norm(ref)
    say("fp32 gemv rel err", err)
    assert err < 1e-6
    h.close()


@pytest.mark.gpu
def test_rank_capacity_grows_during_assembly():
    """A panel capacity below the ranks (hmx_set_rank_capacity(4)): the
    assembly doubles the capacity limit and runs again until the ranks fit;
    the leaves equal those of an assembly at the default capacity."""
    n = 3000
    say(f"rank capacity growth n={n}")
    pts = ho.cylinder_points(n, 1.0, 1.2, 29)
    t = hmat_amd.ClusterTree(pts, 128)
    lib = hmat_amd.load()
    lib.hmx_set_rank_capacity(4)
    try:
        small = hmat_amd.HMatrix(t)
        small.assemble(eps=EPS)
    finally:
        lib.hmx_set_rank_capacity(0)    # back to the default
    ref = hmat_amd.HMatrix(t)
    ref.assemble(eps=EPS)
    ranks = ref.leaf_table()[:, 5]
    diff = np.linalg.norm(small.to_dense() - ref.to_dense()) / np.linalg.norm(ref.to_dense())
    same = np.array_equal(small.leaf_table()[:, 5], ranks)
    say(f"largest rank {ranks.max()}, ranks equal {same}, dense diff {diff:.2e}")
    assert ranks.max() > 4 and same and diff <= 1e-14
    small.close()
    ref.close()


@pytest.mark.gpu
def test_gemv_leaf_clusters_up_to_512_rows():
    """Leaf clusters above 256 rows take the y pass with 512-row destination
    items: non-symmetric storage with leaf 512 and a complex matrix with leaf
    256 (512 rows of its real representation), against the dense product of
    the stored matrix; 625-row clusters are refused."""
    say("gemv with leaf clusters up to 512 rows")
    pts = ho.cylinder_points(5000, 1.0, 2.0, 3)
    t = hmat_amd.ClusterTree(pts, 512)
    h = hmat_amd.HMatrix(t, lower_symmetric=False)
    h.assemble(length=0.5, nugget=1e-2, eps=EPS)
    assert max(int(r[0]) for r in h.leaf_table() if r[4] == 0) > 256
    D = h.to_dense()
    rng = np.random.default_rng(7)
    x = rng.standard_normal((h.n, 9))
    idx = t.indices
    ref = np.empty_like(x)
    ref[idx] = D @ x[idx]
    e1 = np.linalg.norm(h.gemv(x) - ref) / np.linalg.norm(ref)
    e2 = np.linalg.norm(h.gemv(x[:, :1]) - ref[:, :1]) / np.linalg.norm(ref[:, :1])
    h.close()
    from scipy.spatial import cKDTree
    dd, _ = cKDTree(pts).query(pts, 2)
    hreg = float(dd[:, 1].mean())
    tz = hmat_amd.ClusterTree(pts, 256)
    hz = hmat_amd.HMatrix(tz, lower_symmetric=False, value_type=hmat_amd.HMAT_DOUBLE_COMPLEX)
    hz.assemble(kind=hmat_amd.KERNEL_HELMHOLTZ, wavenumber=2 * np.pi / (10 * hreg), reg=hreg,
                nugget=1.0, eps=EPS, method="aca_partial")
    Dz = hz.to_dense(embedded=True)
    xz = rng.standard_normal((hz.n, 2)) + 1j * rng.standard_normal((hz.n, 2))
    e = np.empty(2 * hz.n, dtype=np.int64)
    e[0::2], e[1::2] = 2 * tz.indices, 2 * tz.indices + 1
    xr = np.empty((2 * hz.n, 2))
    xr[0::2], xr[1::2] = xz.real, xz.imag
    yr = np.empty_like(xr)
    yr[e] = Dz @ xr[e]
    refz = yr[0::2] + 1j * yr[1::2]
    e3 = np.linalg.norm(hz.gemv(xz) - refz) / np.linalg.norm(refz)
    hz.close()
    say(f"leaf 512 non-symmetric: 9 RHS {e1:.2e}, 1 RHS {e2:.2e}; complex leaf 256: {e3:.2e}")
    assert e1 <= 1e-13 and e2 <= 1e-13 and e3 <= 1e-13
    # 625-row leaf clusters (5000 points halved three times, leaf 700): the
    # matrix assembles, its H-gemv is refused
    t6 = hmat_amd.ClusterTree(pts, 700)
    h6 = hmat_amd.HMatrix(t6, lower_symmetric=False)
    h6.assemble(length=0.5, nugget=1e-2, eps=EPS)
    with pytest.raises(hmat_amd.EngineError, match="512"):
        h6.gemv(x)
    h6.close()
