"""GPU tests of complex (HMAT_DOUBLE_COMPLEX) matrices: the regularised
Helmholtz kernel (the complex interaction of examples/c-cylinder.c:68-78),
partial ACA (compression.cpp:384-533), H-gemv and H-LU factor + solve
(recursion.cpp:238-287, h_matrix.cpp:2567, 2772-2779) against the exact kernel
matrix, the oracle's complex restatement on the same inputs and dense LAPACK
algebra of the engine's own matrix.  The engine holds the 2 x 2 real
representation of every entry; the tests see complex arrays only."""
import ctypes as C
import sys

import numpy as np
import pytest

from oracle import hmat_oracle as ho

hmat_amd = pytest.importorskip("hmat_amd")

EPS = 1e-4


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def problem(n, height, seed):
    from scipy.spatial import cKDTree
    pts = ho.cylinder_points(n, 1.0, height, seed)
    d, _ = cKDTree(pts).query(pts, 2)
    h = float(d[:, 1].mean())
    return pts, h, 2 * np.pi / (10 * h)


def cnormal(rng, shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


def real_view(z):
    """A complex (n, k) array as the real (2n, k) array of its interleaved
    (re, im) pairs: the embedded numbering."""
    r = np.empty((2 * z.shape[0],) + z.shape[1:])
    r[0::2], r[1::2] = z.real, z.imag
    return r


def complex_view(r):
    return r[0::2] + 1j * r[1::2]


def embedded_user(h, idx):
    """The stored real matrix of a complex H-matrix in user numbering."""
    D = h.to_dense(embedded=True)
    e = np.empty(2 * len(idx), dtype=np.int64)
    e[0::2], e[1::2] = 2 * idx, 2 * idx + 1
    Du = np.empty_like(D)
    Du[np.ix_(e, e)] = D
    return D, Du


@pytest.mark.gpu
@pytest.mark.parametrize("n,leaf,height", [(3000, 64, 1.0), (2050, 128, 1.0)])
def test_helmholtz_assembly_gemv_lu(n, leaf, height):
    say(f"complex Helmholtz n={n} leaf={leaf}")
    pts, hreg, k = problem(n, height, 11)
    t = hmat_amd.ClusterTree(pts, leaf)
    h = hmat_amd.HMatrix(t, lower_symmetric=False, value_type=hmat_amd.HMAT_DOUBLE_COMPLEX)
    h.assemble(kind=hmat_amd.KERNEL_HELMHOLTZ, wavenumber=k, reg=hreg, nugget=1.0, eps=EPS,
               method="aca_partial")
    idx = t.indices
    H = h.to_dense()                     # hmat numbering
    ker = ho.HelmholtzKernel(pts, k, hreg, 1.0)
    exact = ker.block(idx, idx)
    err_exact = np.linalg.norm(H - exact) / np.linalg.norm(exact)
    # the oracle's complex assembly of the same partition
    root, oidx = ho.build_cluster_tree(pts, leaf)
    assert np.array_equal(oidx, idx)
    bt = ho.build_block_tree(root, root, 2.0, False)
    with ho.scalar_type(np.complex128):
        ho.assemble(bt, idx, ker, method="aca_partial", eps=EPS)
        Ho = ho.to_dense(bt, n, n, dtype=np.complex128, symmetric=False)
    err_oracle = np.linalg.norm(H - Ho) / np.linalg.norm(Ho)
    inf = h.info()
    oinf = ho.info(bt)
    say(f"assembly: vs exact kernel {err_exact:.2e}, vs oracle {err_oracle:.2e}; compressed "
        f"size {inf['compressed_size']} (oracle {oinf['compressed_size']}), "
        f"full {inf['full_count']} / rk {inf['rk_count']}")
    assert err_exact <= 2 * EPS and err_oracle <= 2 * EPS
    assert (inf["full_count"], inf["rk_count"]) == (oinf["full_count"], oinf["rk_count"])
    assert inf["uncompressed_size"] == n * n
    assert 0.5 <= inf["compressed_size"] / oinf["compressed_size"] <= 2.0
    # the stored real representation keeps the complex structure to the tolerance
    D, Du = embedded_user(h, idx)
    skew = max(np.linalg.norm(D[0::2, 0::2] - D[1::2, 1::2]),
               np.linalg.norm(D[1::2, 0::2] + D[0::2, 1::2])) / np.linalg.norm(D)
    # norm of the complex matrix: that of its real representation / sqrt(2)
    nrm = np.linalg.norm(D) / np.sqrt(2)
    say(f"structure of the real representation kept to {skew:.2e}; norm {h.norm():.12g} vs {nrm:.12g}")
    assert skew <= EPS
    assert abs(h.norm() - nrm) <= 1e-12 * nrm
    # H-gemv, user numbering: real and complex scalars, against the dense
    # product of the stored matrix
    rng = np.random.default_rng(n)
    x = cnormal(rng, (n, 3))
    ax = complex_view(Du @ real_view(x))
    y = h.gemv(x)
    g1 = np.linalg.norm(y - ax) / np.linalg.norm(ax)
    y0 = cnormal(rng, (n, 3))
    al, be = 0.5 + 2.0j, 1.0 - 1.0j
    y2 = h.gemv(x, alpha=al, beta=be, y=y0.copy())
    g2 = np.linalg.norm(y2 - (al * ax + be * y0)) / np.linalg.norm(y2)
    say(f"gemv vs dense product of the stored matrix: {g1:.2e}, complex alpha / beta {g2:.2e}")
    assert g1 <= 1e-13 and g2 <= 1e-13
    # H-LU factor + solve of 16 complex right-hand sides
    h.factorize(hmat_amd.FACT_LU)
    b = cnormal(rng, (n, 16))
    xs = h.solve(b)
    xe = complex_view(np.linalg.solve(Du, real_view(b)))
    rel = np.linalg.norm(xs - xe) / np.linalg.norm(xe)
    res = np.linalg.norm(complex_view(Du @ real_view(xs)) - b) / np.linalg.norm(b)
    res_exact = np.linalg.norm(exact @ xs[idx] - b[idx]) / np.linalg.norm(b)
    with ho.scalar_type(np.complex128):
        ho.lu(bt, EPS)
        xo = b[idx].copy()
        ho.solve_lu(bt, xo)
    xou = np.empty_like(xo)
    xou[idx] = xo
    rel_o = np.linalg.norm(xs - xou) / np.linalg.norm(xou)
    say(f"H-LU solve: vs dense solve of the stored matrix {rel:.2e}, residual {res:.2e}, "
        f"residual vs the exact kernel {res_exact:.2e}, vs the oracle's H-LU solution {rel_o:.2e}")
    # measured 1.2e-4 / 2.6e-4 (dense), 1.1e-4 / 2.1e-4 (oracle)
    assert rel <= 1e-3 and res <= 1e-3 and rel_o <= 1e-3 and res_exact <= 2e-3
    h.close()


@pytest.mark.gpu
def test_helmholtz_large_blocks_on_the_whole_device():
    """Blocks of at least 300 complex rows and columns compressed by the
    randomized SVD over the whole device instead of partial ACA: each such
    block equals the exact kernel block to the cut (eps of its largest
    singular value) with the rank of the exact truncated SVD (numpy), and the
    matrix, its copy and its H-LU solve behave as after an ACA assembly."""
    n, leaf, rows = 3000, 64, 300
    say(f"complex Helmholtz, whole-device compression of blocks >= {rows} rows, n={n}")
    pts, hreg, k = problem(n, 1.0, 11)
    t = hmat_amd.ClusterTree(pts, leaf)
    idx = t.indices
    ker = ho.HelmholtzKernel(pts, k, hreg, 1.0)
    ranks = {}
    dense = {}
    for big in (False, True):
        h = hmat_amd.HMatrix(t, lower_symmetric=False, value_type=hmat_amd.HMAT_DOUBLE_COMPLEX)
        if big:
            h.set_large_block_rows(rows)
        h.assemble(kind=hmat_amd.KERNEL_HELMHOLTZ, wavenumber=k, reg=hreg, nugget=1.0, eps=EPS,
                   method="aca_partial")
        tab = h.leaf_table()
        ranks[big] = tab[:, 5].copy()
        dense[big] = h.to_dense()
        if big:
            hb = h
        else:
            h.close()
    tab = hb.leaf_table()
    sel = [i for i, (m, nn, r0, c0, kind, _) in enumerate(tab)
           if kind == 1 and min(m, nn) >= 2 * rows]
    say(f"{len(sel)} blocks compressed on the whole device")
    assert len(sel) >= 2
    H = dense[True]
    worst = 0.0
    for i in sel:
        m, nn, r0, c0 = (int(v) for v in tab[i, :4])
        rs, cs = slice(r0 // 2, (r0 + m) // 2), slice(c0 // 2, (c0 + nn) // 2)
        ex = ker.block(idx[rs], idx[cs])
        sv = np.linalg.svd(ex, compute_uv=False)
        kx = ho.find_k(sv, EPS)
        err = np.linalg.norm(H[rs, cs] - ex, 2) / sv[0]
        worst = max(worst, err)
        say(f"block {m // 2} x {nn // 2}: complex rank {ranks[True][i] / 2:.1f} (exact truncated SVD {kx}, "
            f"partial ACA {ranks[False][i] / 2:.1f}), spectral error / s0 {err:.2e}")
        # the real representation doubles every singular value's multiplicity
        assert abs(int(ranks[True][i]) - 2 * kx) <= 2
        assert int(ranks[True][i]) <= int(ranks[False][i]) + 2
        assert err <= 2 * EPS
    exact = ker.block(idx, idx)
    err_exact = np.linalg.norm(H - exact) / np.linalg.norm(exact)
    say(f"whole matrix vs exact kernel {err_exact:.2e} (ACA assembly "
        f"{np.linalg.norm(dense[False] - exact) / np.linalg.norm(exact):.2e})")
    assert err_exact <= 2 * EPS
    rng = np.random.default_rng(3)
    x = cnormal(rng, (n, 2))
    D, Du = embedded_user(hb, idx)
    ax = complex_view(Du @ real_view(x))
    c = hb.copy()
    g = np.linalg.norm(hb.gemv(x) - ax) / np.linalg.norm(ax)
    gc = np.linalg.norm(c.gemv(x) - ax) / np.linalg.norm(ax)
    c.close()
    say(f"gemv vs dense product of the stored matrix {g:.2e}, of its copy {gc:.2e}")
    assert g <= 1e-13 and gc <= 1e-13
    hb.factorize(hmat_amd.FACT_LU)
    b = cnormal(rng, (n, 4))
    xs = hb.solve(b)
    xe = complex_view(np.linalg.solve(Du, real_view(b)))
    rel = np.linalg.norm(xs - xe) / np.linalg.norm(xe)
    say(f"H-LU solve vs dense solve of the stored matrix {rel:.2e}")
    assert rel <= 1e-3
    hb.close()


@pytest.mark.gpu
def test_complex_block_callback_and_device_entries():
    import torch
    n, leaf = 2000, 64
    say(f"complex block_compute callback + device entries n={n}")
    pts, hreg, k = problem(n, 0.8, 5)
    ker = ho.HelmholtzKernel(pts, k, hreg, 1.0)
    t = hmat_amd.ClusterTree(pts, leaf)
    blocks = {}
    counter = [0]

    @hmat_amd.PREPARE_FUNC
    def prepare(r0, rc, c0, cc, rows_h2c, rows_c2h, cols_h2c, cols_c2h, ctx, info):
        counter[0] += 1
        blocks[counter[0]] = (np.array(rows_h2c[r0:r0 + rc]), np.array(cols_h2c[c0:c0 + cc]))
        info[0].user_data = counter[0]

    @hmat_amd.COMPUTE_FUNC
    def compute(data, r0, rc, c0, cc, out):
        ri, ci = blocks.pop(data)
        v = ker.block(ri[r0:r0 + rc], ci[c0:c0 + cc])
        np.ctypeslib.as_array(C.cast(out, C.POINTER(C.c_double)), shape=(cc, rc, 2))[:] = \
            np.stack([v.T.real, v.T.imag], axis=2)

    h = hmat_amd.HMatrix(t, lower_symmetric=False, value_type=hmat_amd.HMAT_DOUBLE_COMPLEX)
    h.assemble_functions(block=compute, prepare=prepare, eps=EPS, method="aca_partial")
    assert not blocks
    g = hmat_amd.HMatrix(t, lower_symmetric=False, value_type=hmat_amd.HMAT_DOUBLE_COMPLEX)
    g.assemble(kind=hmat_amd.KERNEL_HELMHOLTZ, wavenumber=k, reg=hreg, nugget=1.0, eps=EPS,
               method="aca_partial")
    Hh, Hg = h.to_dense(), g.to_dense()
    dd = np.linalg.norm(Hh - Hg) / np.linalg.norm(Hg)
    idx = t.indices
    ex = np.linalg.norm(Hh - ker.block(idx, idx)) / np.linalg.norm(Hg)
    say(f"host-function matrix vs device-kernel matrix {dd:.2e}; vs exact kernel {ex:.2e}")
    assert dd <= 1e-12 and ex <= 2 * EPS      # measured 4.4e-15 and 4.0e-5
    # device entries on complex tensors (interleaved re, im): gemv, LU solve
    rng = np.random.default_rng(3)
    x = cnormal(rng, (n, 2))
    xt = torch.tensor(x.T.copy(), dtype=torch.complex128, device="cuda")   # hmat numbering
    yt = torch.zeros_like(xt)
    D = g.to_dense(embedded=True)
    Hx = complex_view(D @ real_view(x))       # the device entries work in hmat numbering
    g.gemv_device(xt.data_ptr(), yt.data_ptr(), 2)
    torch.cuda.synchronize()
    dv = np.linalg.norm(yt.cpu().numpy().T - Hx) / np.linalg.norm(Hx)
    g.factorize(hmat_amd.FACT_LU)
    bt_ = torch.tensor(Hx.T.copy(), dtype=torch.complex128, device="cuda")
    g.solve_device(bt_.data_ptr(), 2)
    torch.cuda.synchronize()
    sv = np.linalg.norm(bt_.cpu().numpy().T - x) / np.linalg.norm(x)
    say(f"device gemv vs dense {dv:.2e}; device LU solve of H x recovers x to {sv:.2e}")
    assert dv <= 1e-13 and sv <= 1e-3
    # what complex matrices do not provide
    with pytest.raises(hmat_amd.EngineError, match="hmat_factorization_lu"):
        h.factorize(hmat_amd.FACT_LLT)
    with pytest.raises(hmat_amd.EngineError, match="logdet"):
        g.logdet()
    with pytest.raises(hmat_amd.EngineError, match="imaginary"):
        h.scale(1.0 + 1.0j)
    n0 = h.norm()
    h.scale(2.0)
    assert abs(h.norm() - 2 * n0) <= 1e-12 * n0
    h.close()
    g.close()


@pytest.mark.gpu
def test_complex_svd_compression():
    n, leaf = 800, 32
    say(f"complex svd compression n={n}")
    pts, hreg, k = problem(n, 0.3, 8)
    t = hmat_amd.ClusterTree(pts, leaf)
    h = hmat_amd.HMatrix(t, lower_symmetric=False, value_type=hmat_amd.HMAT_DOUBLE_COMPLEX)
    h.assemble(kind=hmat_amd.KERNEL_HELMHOLTZ, wavenumber=k, reg=hreg, nugget=1.0, eps=EPS,
               method="svd")
    idx = t.indices
    ker = ho.HelmholtzKernel(pts, k, hreg, 1.0)
    exact = ker.block(idx, idx)
    err = np.linalg.norm(h.to_dense() - exact) / np.linalg.norm(exact)
    root, _ = ho.build_cluster_tree(pts, leaf)
    bt = ho.build_block_tree(root, root, 2.0, False)
    with ho.scalar_type(np.complex128):
        ho.assemble(bt, idx, ker, method="svd", eps=EPS)
    ratio = h.info()["compressed_size"] / ho.info(bt)["compressed_size"]
    say(f"complex svd: vs exact kernel {err:.2e}, compressed size / oracle's {ratio:.4f}")
    assert err <= EPS and abs(ratio - 1) <= 0.02
    h.close()


@pytest.mark.gpu
def test_complex_get_block_and_get_values():
    """get_block (hmat numbering, renumber_rows) and get_values (1-based user
    indices) of a complex matrix equal the complex matrix of the stored real
    representation (to_dense), and get_values equals the exact kernel to the
    compression."""
    n, leaf = 2000, 64
    say(f"complex get_block / get_values n={n}")
    pts, hreg, k = problem(n, 0.8, 9)
    t = hmat_amd.ClusterTree(pts, leaf)
    h = hmat_amd.HMatrix(t, lower_symmetric=False, value_type=hmat_amd.HMAT_DOUBLE_COMPLEX)
    h.assemble(kind=hmat_amd.KERNEL_HELMHOLTZ, wavenumber=k, reg=hreg, nugget=1.0, eps=EPS,
               method="aca_partial")
    H = h.to_dense()            # hmat numbering
    idx = t.indices
    B = h.get_block(300, 500, 1200, 400)
    e1 = np.abs(B - H[300:800, 1200:1600]).max()
    F = h.get_block(0, n, 700, 64, renumber_rows=True)
    Fu = np.empty_like(F)
    Fu[idx] = H[:, 700:764]
    e2 = np.abs(F - Fu).max()
    rng = np.random.default_rng(2)
    ur, uc = rng.choice(n, 37, replace=False), rng.choice(n, 29, replace=False)
    V = h.get_values(ur, uc)
    inv = np.empty(n, dtype=np.int64)
    inv[idx] = np.arange(n)
    e3 = np.abs(V - H[np.ix_(inv[ur], inv[uc])]).max()
    ker = ho.HelmholtzKernel(pts, k, hreg, 1.0)
    ex = ker.block(ur, uc)
    e4 = np.linalg.norm(V - ex) / np.linalg.norm(ex)
    say(f"get_block vs to_dense {e1:.2e}, renumbered {e2:.2e}; get_values vs to_dense {e3:.2e}, "
        f"vs the exact kernel {e4:.2e}")
    scale = np.abs(H).max()
    assert e1 <= 1e-14 * scale and e2 <= 1e-14 * scale and e3 <= 1e-14 * scale
    assert e4 <= 5 * EPS
    h.factorize(hmat_amd.FACT_LU)
    with pytest.raises(hmat_amd.EngineError):
        h.get_block(0, 10, 0, 10)
    h.close()


@pytest.mark.gpu
def test_single_precision_complex():
    """HMAT_SIMPLE_COMPLEX (C): the real representation in FP32 (assembled in
    FP64, stored in FP32), against the exact kernel, the dense algebra of its
    own stored matrix, and the FP64 complex matrix of the same partition;
    get_values and the complex stream format (scalar type 2) included."""
    import struct
    n, leaf = 2000, 64
    say(f"single-precision complex n={n}")
    pts, hreg, k = problem(n, 0.8, 21)
    t = hmat_amd.ClusterTree(pts, leaf)
    hz = hmat_amd.HMatrix(t, lower_symmetric=False, value_type=hmat_amd.HMAT_DOUBLE_COMPLEX)
    hc = hmat_amd.HMatrix(t, lower_symmetric=False, value_type=hmat_amd.HMAT_SIMPLE_COMPLEX)
    for h in (hz, hc):
        h.assemble(kind=hmat_amd.KERNEL_HELMHOLTZ, wavenumber=k, reg=hreg, nugget=1.0, eps=EPS,
                   method="aca_partial")
    idx = t.indices
    Hc, Hz = hc.to_dense(), hz.to_dense()
    ker = ho.HelmholtzKernel(pts, k, hreg, 1.0)
    exact = ker.block(idx, idx)
    e_exact = np.linalg.norm(Hc - exact) / np.linalg.norm(exact)
    e_z = np.linalg.norm(Hc - Hz) / np.linalg.norm(Hz)
    say(f"C matrix vs exact kernel {e_exact:.2e}, vs the Z matrix {e_z:.2e}")
    assert e_exact <= 2 * EPS and e_z <= 1e-6
    assert np.array_equal(hc.partition(), hz.partition())
    rng = np.random.default_rng(8)
    x = (rng.standard_normal((n, 2)) + 1j * rng.standard_normal((n, 2))).astype(np.complex64)
    D, Du = embedded_user(hc, idx)
    ax = complex_view(Du @ real_view(x.astype(np.complex128)))
    y = hc.gemv(x)
    assert y.dtype == np.complex64
    g = np.linalg.norm(y - ax) / np.linalg.norm(ax)
    y2 = hc.gemv(x, alpha=0.5 - 1.5j)
    g2 = np.linalg.norm(y2 - (0.5 - 1.5j) * ax) / np.linalg.norm(ax)
    ur, uc = rng.choice(n, 23, replace=False), rng.choice(n, 31, replace=False)
    inv = np.empty(n, dtype=np.int64)
    inv[idx] = np.arange(n)
    V = hc.get_values(ur, uc)
    gv = np.abs(V - Hc[np.ix_(inv[ur], inv[uc])]).max() / np.abs(Hc).max()
    say(f"gemv vs dense product of the stored matrix {g:.2e}, complex alpha {g2:.2e}; "
        f"get_values {gv:.2e}")
    assert g <= 1e-5 and g2 <= 1e-5 and gv <= 1e-6
    # the stream: scalar type 2, complex<float> leaves; read back
    s, d = hc.write_struct(), hc.write_data()
    assert struct.unpack_from("<ii", s) == (2, -1)
    r = hmat_amd.HMatrix.read_struct(s, hmat_amd.HMAT_SIMPLE_COMPLEX)
    assert r.read_data(d) == len(d)
    back = np.linalg.norm(r.to_dense() - Hc) / np.linalg.norm(Hc)
    say(f"read-back C matrix vs the original {back:.2e}")
    # the stream keeps the first column of each 2 x 2 block of the Rk leaves,
    # which the compression keeps equal to the other to eps (to_dense averages)
    assert back <= EPS
    r.close()
    # H-LU in FP32 and its solve
    hc.factorize(hmat_amd.FACT_LU)
    b = (rng.standard_normal((n, 4)) + 1j * rng.standard_normal((n, 4))).astype(np.complex64)
    xs = hc.solve(b)
    xe = complex_view(np.linalg.solve(Du, real_view(b.astype(np.complex128))))
    rel = np.linalg.norm(xs - xe) / np.linalg.norm(xe)
    say(f"FP32 H-LU solve vs dense solve of the stored matrix {rel:.2e}")
    assert rel <= 5e-3
    hc.close()
    hz.close()
