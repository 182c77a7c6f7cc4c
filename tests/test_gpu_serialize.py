"""GPU test of serialization in hmat-oss's stream format (write_struct /
read_struct / write_data / read_data, c_wrapping.hpp:921-946,
serialization.cpp:34-130 for the structure, MatrixDataMarshaller for the
leaves).

* The structure bytes of an assembled matrix equal, byte for byte, the oracle's
  restatement of MatrixStructMarshaller::write on the same partition; those of
  an LLt-factorised matrix equal it with the factor's ranks and the
  lltDecomposition flags (isTriLower on diagonal blocks, h_matrix.cpp:2562).
* A written matrix read back (read_struct, then read_data) reproduces every
  byte it was written from, its partition, its dense image, its gemv, and, for
  LLt and LU factors, its solve and log-determinant, bit for bit.
* Malformed streams (another scalar type, an Rk rank beyond its block) raise
  errors."""
import struct
import sys

import numpy as np
import pytest

from oracle import hmat_oracle as ho

hmat_amd = pytest.importorskip("hmat_amd")

EPS = 1e-4


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _matrix(n, leaf, seed, value_type=hmat_amd.HMAT_DOUBLE_PRECISION, sym=True):
    pts = ho.cylinder_points(n, 1.0, 1.5, seed)
    t = hmat_amd.ClusterTree(pts, leaf)
    h = hmat_amd.HMatrix(t, lower_symmetric=sym, value_type=value_type)
    h.assemble(length=0.5, nugget=1e-2, eps=EPS)
    return pts, t, h


def _round_trip(h):
    s, d = h.write_struct(), h.write_data()
    g = hmat_amd.HMatrix.read_struct(s, h.value_type)
    assert g._consumed == len(s)
    assert g.read_data(d) == len(d)
    assert g.write_struct() == s
    assert g.write_data() == d
    assert np.array_equal(g.partition(), h.partition())
    assert np.array_equal(g.tree.indices, h.tree.indices)
    return s, d, g


@pytest.mark.gpu
def test_structure_bytes_match_oracle():
    n, leaf = 1500, 64
    say(f"structure bytes vs oracle n={n} leaf={leaf}")
    pts, t, h = _matrix(n, leaf, 5)
    root, idx = ho.build_cluster_tree(pts, leaf)
    assert np.array_equal(idx, t.indices)
    bt = ho.build_block_tree(root, root, 2.0, True)
    ho.assemble(bt, idx, ho.ExpKernel(pts, 0.5, 1e-2))
    s = h.write_struct()
    ref = ho.write_struct(bt, root, root, idx, idx, pts, EPS)
    say(f"structure {len(s)} bytes, oracle {len(ref)} bytes")
    assert struct.unpack_from("<ii", s) == (hmat_amd.HMAT_DOUBLE_PRECISION, -1)
    assert s == ref
    # the LLt factor: same partition, the factor's ranks, isTriLower flags
    h.factorize(hmat_amd.FACT_LLT)
    ranks = [int(r[5]) for r in h.leaf_table() if r[4] != 0]
    ho.llt(bt, EPS)
    ref = ho.write_struct(bt, root, root, idx, idx, pts, EPS, factorization=2, ranks=ranks)
    s = h.write_struct()
    assert struct.unpack_from("<ii", s) == (hmat_amd.HMAT_DOUBLE_PRECISION, 2)
    assert s == ref
    h.close()


@pytest.mark.gpu
@pytest.mark.parametrize("vt", [hmat_amd.HMAT_DOUBLE_PRECISION, hmat_amd.HMAT_SIMPLE_PRECISION])
def test_round_trip_assembled(vt):
    n, leaf = 3000, 128
    say(f"round trip of an assembled matrix n={n} type={vt}")
    pts, t, h = _matrix(n, leaf, 7, value_type=vt)
    s, d, g = _round_trip(h)
    say(f"structure {len(s)} bytes, data {len(d)} bytes")
    prec = 1 if vt == hmat_amd.HMAT_SIMPLE_PRECISION else 0   # the store of the matrix's type
    assert np.array_equal(g.to_dense(prec), h.to_dense(prec))
    x = np.random.default_rng(1).standard_normal((n, 3))
    assert np.array_equal(g.gemv(x), h.gemv(x))
    # the read matrix factorises and solves like the original
    for m in (h, g):
        m.factorize(hmat_amd.FACT_LLT)
    b = np.random.default_rng(2).standard_normal((n, 2))
    assert np.array_equal(g.solve(b), h.solve(b))
    g.close()
    h.close()


@pytest.mark.gpu
@pytest.mark.parametrize("algo,sym", [(hmat_amd.FACT_LLT, True), (hmat_amd.FACT_LU, False)])
def test_round_trip_factor(algo, sym):
    n, leaf = 4000, 128
    say(f"round trip of a factor n={n} algo={algo} lower_symmetric={sym}")
    pts, t, h = _matrix(n, leaf, 11, sym=sym)
    h.factorize(algo)
    s, d, g = _round_trip(h)
    assert struct.unpack_from("<ii", s)[1] == (2 if algo == hmat_amd.FACT_LLT else 0)
    say(f"structure {len(s)} bytes, data {len(d)} bytes")
    b = np.random.default_rng(3).standard_normal((n, 4))
    xh, xg = h.solve(b), g.solve(b)
    assert np.array_equal(xg, xh)
    if algo == hmat_amd.FACT_LLT:
        assert g.logdet() == h.logdet()
    # and the solve is that of the matrix: residual against a fresh assembly
    _, _, a = _matrix(n, leaf, 11, sym=sym)
    r = np.linalg.norm(a.gemv(xg) - b) / np.linalg.norm(b)
    say(f"residual of the read factor's solve {r:.2e}")
    # (the factor at eps 1e-4 leaves about 1e-3 on this system, nugget 1e-2)
    assert r <= 5e-3
    for m in (a, g, h):
        m.close()


@pytest.mark.gpu
def test_malformed_streams_raise():
    say("malformed streams")
    _, _, h = _matrix(1200, 64, 13)
    s, d = h.write_struct(), h.write_data()
    # a single-precision reader refuses a double-precision stream
    with pytest.raises(hmat_amd.EngineError):
        hmat_amd.HMatrix.read_struct(s, hmat_amd.HMAT_SIMPLE_PRECISION)
    # an Rk leaf whose rank exceeds its block: the data stream holds the leaves
    # in preorder, a full leaf as flag + m n values, an Rk leaf as rank +
    # (m + n) k values + two flags
    pos = 0
    for m, n, _, _, kind, rank in h.leaf_table():
        if kind != 0:
            break
        pos += 4 + 8 * int(m) * int(n)
    bad = bytearray(d)
    bad[pos:pos + 4] = struct.pack("<i", 1 << 20)
    g = hmat_amd.HMatrix.read_struct(s)
    with pytest.raises(hmat_amd.EngineError, match="rank"):
        g.read_data(bytes(bad))
    g.close()
    h.close()


def _complex_leaves(data, table):
    """The complex matrix (hmat numbering) of a complex data stream: leaves in
    preorder, a full leaf as flag + m n complex values, an Rk leaf as rank k +
    (m + n) k complex values + two flags (MatrixDataMarshaller with Z_t);
    `table` gives each leaf's embedded shape and offsets."""
    nr = int(max(r[2] + r[0] for r in table)) // 2
    nc = int(max(r[3] + r[1] for r in table)) // 2
    Z = np.zeros((nr, nc), dtype=np.complex128)
    pos = 0
    for m2, n2, r2, c2, kind, _ in table:
        m, n, r0, c0 = int(m2) // 2, int(n2) // 2, int(r2) // 2, int(c2) // 2
        head = struct.unpack_from("<i", data, pos)[0]
        pos += 4
        if kind == 0:
            v = np.frombuffer(data, "<c16", m * n, pos).reshape(n, m).T
            pos += 16 * m * n
            Z[r0:r0 + m, c0:c0 + n] = v
        elif head > 0:
            a = np.frombuffer(data, "<c16", m * head, pos).reshape(head, m).T
            pos += 16 * m * head
            b = np.frombuffer(data, "<c16", n * head, pos).reshape(head, n).T
            pos += 16 * n * head + 8
            Z[r0:r0 + m, c0:c0 + n] = a @ b.T
    assert pos == len(data)
    return Z


@pytest.mark.gpu
def test_complex_round_trip():
    """A complex Helmholtz matrix in the complex stream format: the structure
    bytes equal the oracle's MatrixStructMarshaller::write of the complex
    partition (scalar type 3, the caller's trees, each Rk leaf at its rank);
    the data stream holds the complex leaves (equal to the stored matrix to the
    compression); read back into the real representation, the matrix is that
    of the stream to rounding and its gemv is the original's."""
    from scipy.spatial import cKDTree
    n, leaf = 2000, 64
    say(f"complex round trip n={n} leaf={leaf}")
    pts = ho.cylinder_points(n, 1.0, 0.8, 5)
    dd, _ = cKDTree(pts).query(pts, 2)
    hreg = float(dd[:, 1].mean())
    k = 2 * np.pi / (10 * hreg)
    t = hmat_amd.ClusterTree(pts, leaf)
    h = hmat_amd.HMatrix(t, lower_symmetric=False, value_type=hmat_amd.HMAT_DOUBLE_COMPLEX)
    h.assemble(kind=hmat_amd.KERNEL_HELMHOLTZ, wavenumber=k, reg=hreg, nugget=1.0, eps=EPS,
               method="aca_partial")
    s, d = h.write_struct(), h.write_data()
    assert struct.unpack_from("<ii", s) == (3, -1)
    tab = h.leaf_table()
    root, idx = ho.build_cluster_tree(pts, leaf)
    assert np.array_equal(idx, t.indices)
    bt = ho.build_block_tree(root, root, 2.0, False)
    ranks = [int(r[5]) for r in tab if r[4] != 0]
    ref = ho.write_struct(bt, root, root, idx, idx, pts, EPS, value_type=3, lower_symmetric=False,
                          ranks=ranks)
    say(f"structure {len(s)} bytes (oracle {len(ref)}), data {len(d)} bytes")
    assert s == ref
    Zs = _complex_leaves(d, tab)
    H = h.to_dense()
    dev = np.linalg.norm(Zs - H) / np.linalg.norm(H)
    say(f"stream's complex matrix vs the stored matrix {dev:.2e}")
    assert dev <= EPS
    g = hmat_amd.HMatrix.read_struct(s, hmat_amd.HMAT_DOUBLE_COMPLEX)
    assert g._consumed == len(s)
    assert g.read_data(d) == len(d)
    G = g.to_dense()
    back = np.linalg.norm(G - Zs) / np.linalg.norm(Zs)
    x = np.random.default_rng(4).standard_normal((n, 2)) + 1j * np.random.default_rng(5).standard_normal((n, 2))
    gx, hx = g.gemv(x), h.gemv(x)
    gv = np.linalg.norm(gx - hx) / np.linalg.norm(hx)
    say(f"read matrix vs the stream's {back:.2e}; gemv vs the original's {gv:.2e}")
    assert back <= 1e-14 and gv <= EPS
    # factorised complex matrices are not written; a real reader refuses the stream
    with pytest.raises(hmat_amd.EngineError):
        hmat_amd.HMatrix.read_struct(s, hmat_amd.HMAT_DOUBLE_PRECISION)
    h.factorize(hmat_amd.FACT_LU)
    with pytest.raises(hmat_amd.EngineError):
        h.write_struct()
    g.close()
    h.close()
