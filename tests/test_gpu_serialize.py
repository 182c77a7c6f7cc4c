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
