"""CPU pins of the oracle's structure serializer (MatrixStructMarshaller::write,
serialization.cpp:34-130), which the GPU serialization test compares the
engine's bytes with.

* line8: the bytes written out by hand for the hand-derived 8-point partition
  of the golden fixtures, from the reference's rules: int type and
  factorisation; per cluster tree the size, the dimension, the coordinates
  (user numbering, doubles), a zero group flag, the permutation and the nodes
  in preorder as (offset, size, number of children); per block node a flag
  byte, approximate rank, rank (-1 full, -2 inner), epsilon and number of
  children, with 0x80 for a null child.
* A stream of a larger partition, read back by a parser written from the same
  rules, gives the oracle's preorder partition, coordinates and permutation,
  and its length follows from the node counts."""
import json
import os
import struct

import numpy as np

from oracle import hmat_oracle as ho

GOLD = os.path.join(os.path.dirname(__file__), "golden", "partition_fixtures.json")


def _line8():
    case = next(c for c in json.load(open(GOLD))["cases"] if c["name"] == "line8")
    pts = np.array(case["points"], dtype=float).reshape(8, 1)
    root, idx = ho.build_cluster_tree(pts, case["leaf"])
    assert list(idx) == case["indices"]
    bt = ho.build_block_tree(root, root, case["eta"], True)
    return pts, root, idx, bt


def test_line8_known_bytes():
    pts, root, idx, bt = _line8()
    eps = 1e-4
    ranks = [3, 5, 7, 11, 13, 17]          # stand-ins for the six Rk leaves, in preorder
    got = ho.write_struct(bt, root, root, idx, idx, pts, eps, ranks=ranks)

    P = struct.pack
    tree = P("<ii", 8, 1) + P("<8d", 5, 2, 7, 0, 3, 6, 1, 4) + P("<i", 0) + \
        P("<8i", 3, 6, 1, 4, 7, 0, 5, 2) + \
        b"".join(P("<iii", *t) for t in [(0, 8, 2), (0, 4, 2), (0, 2, 0), (2, 2, 0),
                                         (4, 4, 2), (4, 2, 0), (6, 2, 0)])
    NULL = P("<B", 0x80)
    SYM, SAME = 0x02, 0x30           # isLower; keepSameRows | keepSameCols (leaves)

    def node(bits, approx, rank, nchild):
        return P("<Biidi", bits, approx, rank, eps, nchild)

    full = node(SAME, 25, -1, 0)

    def rk(r):
        return node(SAME, 25, r, 0)
    blocks = (
        node(SYM, -3, -2, 4) +
        node(SYM, -3, -2, 4) + full + rk(3) + NULL + full +          # [0,4)x[0,4)
        node(0, -3, -2, 4) + rk(5) + rk(7) + rk(11) + rk(13) +       # [4,8)x[0,4)
        NULL +                                                       # upper block
        node(SYM, -3, -2, 4) + full + rk(17) + NULL + full           # [4,8)x[4,8)
    )
    want = P("<ii", 1, -1) + tree + tree + blocks
    assert got == want

    # the LLt factor's flags: isTriLower on diagonal blocks, no isLower
    got = ho.write_struct(bt, root, root, idx, idx, pts, eps, factorization=2, ranks=ranks)
    TRI = 0x08
    blocks = (
        node(TRI, -3, -2, 4) +
        node(TRI, -3, -2, 4) + node(TRI | SAME, 25, -1, 0) + rk(3) + NULL +
        node(TRI | SAME, 25, -1, 0) +
        node(0, -3, -2, 4) + rk(5) + rk(7) + rk(11) + rk(13) + NULL +
        node(TRI, -3, -2, 4) + node(TRI | SAME, 25, -1, 0) + rk(17) + NULL +
        node(TRI | SAME, 25, -1, 0)
    )
    assert got == P("<ii", 1, 2) + tree + tree + blocks


class _Reader:
    def __init__(self, data):
        self.d, self.p = data, 0

    def get(self, fmt):
        v = struct.unpack_from(fmt, self.d, self.p)
        self.p += struct.calcsize(fmt)
        return v


def _parse(data):
    """Reads a structure stream back: (header, [trees], preorder partition)."""
    r = _Reader(data)
    head = r.get("<ii")
    trees = []
    for _ in range(2):
        n, dim = r.get("<ii")
        coords = np.array(r.get(f"<{n * dim}d")).reshape(n, dim)
        assert r.get("<i") == (0,)
        perm = np.array(r.get(f"<{n}i"))

        def cnode():
            off, size, nc = r.get("<iii")
            return (off, size, [cnode() for _ in range(nc)])
        trees.append((coords, perm, cnode()))
    part = []

    def bnode(rows, cols):
        (bits,) = r.get("<B")
        if bits & 0x80:
            return
        approx, rank, eps, nc = r.get("<iidi")
        kind = "H" if rank == -2 else ("F" if rank == -1 else "R")
        part.append((rows[0], rows[1], cols[0], cols[1], kind))
        rkids = [rows] if bits & 0x10 else rows[2]
        ckids = [cols] if bits & 0x20 else cols[2]
        for k in range(nc):
            bnode(rkids[k % len(rkids)], ckids[k // len(rkids)])
    bnode(trees[0][2], trees[1][2])
    assert r.p == len(data)
    return head, trees, part


def test_parse_back_matches_oracle_partition():
    n, leaf = 3000, 64
    pts = ho.cylinder_points(n, 1.0, 2.0, 21)
    root, idx = ho.build_cluster_tree(pts, leaf)
    for sym in (True, False):
        bt = ho.build_block_tree(root, root, 2.0, sym)
        nrk = sum(1 for b in ho.leaves(bt) if b.kind == ho.RK)
        data = ho.write_struct(bt, root, root, idx, idx, pts, 1e-4, lower_symmetric=sym,
                               ranks=[1] * nrk)
        head, trees, part = _parse(data)
        assert head == (1, -1)
        for coords, perm, _ in trees:
            assert np.array_equal(coords, pts)
            assert np.array_equal(perm, idx)
        assert part == ho.partition_signature(bt)
        # length: header, two trees, 21 bytes per block node, 1 per null child
        ncl = sum(1 for _ in _walk(root))
        nnull = sum(1 for b in _blocks(bt) if b is None)
        tree_bytes = 8 + n * 3 * 8 + 4 + 4 * n + 12 * ncl
        assert len(data) == 8 + 2 * tree_bytes + 21 * len(part) + nnull


def _walk(c):
    yield c
    for k in c.children:
        yield from _walk(k)


def _blocks(b):
    """Every block node and every written null child, preorder."""
    yield b
    if b is not None and not b.is_leaf():
        last = max(k for k, ch in enumerate(b.children) if ch is not None)
        for ch in b.children[:last + 1]:
            yield from _blocks(ch)
