"""CPU restatement of hmat-oss's hot path, used ONLY as test infrastructure.

This module is the oracle: tests, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` call it to check (or time beside) the HIP
engine.  Nothing in the shipped library imports it, and the library never falls
back to it.

It restates, with numpy/LAPACK, the reference algorithms at the call sites named
below (paths are relative to the hmat-oss source tree):

* median clustering          clustering.cpp:109-114, 136-152, 258-301, 489-565
* bounding boxes             cluster_tree.cpp:212-289
* Hackbusch admissibility    admissibility.cpp:39-117
* block tree                 h_matrix.cpp:144-221
* assembly + compression     assembly.cpp:49-71, compression.cpp:82-239, 384-773
* Rk arithmetic              rk_matrix.cpp:156-297, 490-585, 761-1010, 1013-1415
* H-gemv                     h_matrix.cpp:772-857
* H-gemm / axpy              h_matrix.cpp:870-1073, 1202-1620, 1874-1950
* H-LLt and triangular solves recursion.cpp:86-134, 191-233, 349-444,
                              h_matrix.cpp:2297-2564, 3010-3018
* HODLR                      hodlr.cpp:37-166, 221-312, 325-378, 365-404,
                              admissibility.cpp:231-241
* H-LDLt                     recursion.cpp:37-83, h_matrix.cpp:1953-1998,
                              2585-2690, 2747-2769, 2969-3002,
                              scalar_array.cpp:737-779
* H-LU                       recursion.cpp:238-287, h_matrix.cpp:2567, 2772-2779,
                              scalar_array.cpp:694-705, 852-874
* complex (Z) scalars        the same routines with complex128 temporaries
                              (``scalar_type``), conjugated dot products in the
                              ACA norm estimates (blas_overloads.hpp:446-463)
* logdet                     h_matrix.cpp:2925-2936
* info counters              h_matrix.cpp:464-502
* structure serialization    serialization.cpp:34-130 (MatrixStructMarshaller)

Every function here is pinned by the property tests under ``tests/`` (dense
reconstruction bounds, exact-rank recovery, LAPACK identities, partition
invariants and the hand-derived partition fixtures in ``tests/golden``).
"""
from __future__ import annotations

import math
import numpy as np

# ---------------------------------------------------------------------------
# settings (hmat_cpp_interface.hpp:88-95, h_matrix.cpp:151, c_default_interface.cpp:195-200)
# ---------------------------------------------------------------------------
COMPRESSION_MIN_LEAF_SIZE = 100

# Scalar type of the temporaries the H-arithmetic creates (float64, or
# complex128 for the complex (Z) type: the reference templates every routine on
# its scalar type, data_types.hpp:32-35).
_DTYPE = [np.float64]


class scalar_type:
    """``with scalar_type(np.complex128): ...`` runs assembly and H-arithmetic
    with complex temporaries."""

    def __init__(self, dtype):
        self.dtype = dtype

    def __enter__(self):
        self.old = _DTYPE[0]
        _DTYPE[0] = self.dtype
        return self

    def __exit__(self, *a):
        _DTYPE[0] = self.old


def _zeros(shape):
    return np.zeros(shape, dtype=_DTYPE[0])


# ---------------------------------------------------------------------------
# cluster tree
# ---------------------------------------------------------------------------
class Cluster:
    __slots__ = ("offset", "size", "children", "depth", "bbmin", "bbmax", "id")

    def __init__(self, offset, size, depth):
        self.offset = offset
        self.size = size
        self.depth = depth
        self.children = []
        self.bbmin = None
        self.bbmax = None
        self.id = -1

    def is_leaf(self):
        return not self.children

    def diameter(self):
        # AxisAlignedBoundingBox::diameterSqr cluster_tree.cpp:236-246
        r = 0.0
        for i in range(len(self.bbmin)):
            d = self.bbmin[i] - self.bbmax[i]
            r += d * d
        return math.sqrt(r)

    def distance_to(self, other):
        # AxisAlignedBoundingBox::distanceToSqr cluster_tree.cpp:271-285
        r = 0.0
        for i in range(len(self.bbmin)):
            d = max(0.0, self.bbmin[i] - other.bbmax[i])
            r += d * d
            d = max(0.0, other.bbmin[i] - self.bbmax[i])
            r += d * d
        return math.sqrt(r)


def _bbox(coords, idx):
    pts = coords[idx]
    return [float(v) for v in pts.min(axis=0)], [float(v) for v in pts.max(axis=0)]


def _largest_dimension(bbmin, bbmax, to_avoid, avoid_ratio=1.2):
    # AxisAlignClusteringAlgorithm::largestDimension clustering.cpp:136-152
    dim = len(bbmin)
    size_dim = sorted((bbmax[i] - bbmin[i], i) for i in range(dim))
    if (to_avoid < 0 or dim < 2 or size_dim[-1][1] != to_avoid
            or size_dim[-1][0] > avoid_ratio * size_dim[-2][0]):
        return size_dim[-1][1]
    return size_dim[-2][1]


def build_cluster_tree(coords, max_leaf_size):
    """Median bisection (ClusterTreeBuilder::build clustering.cpp:489-565).

    Returns (root, indices) where indices[i] is the user index of the i-th DoF in
    hmat numbering (hmat_cluster_get_indices)."""
    coords = np.ascontiguousarray(coords, dtype=np.float64)
    n = coords.shape[0]
    indices = np.arange(n, dtype=np.int64)
    root = Cluster(0, n, 0)

    def divide(node, current_axis):
        if node.size <= max_leaf_size:
            return
        idx = indices[node.offset:node.offset + node.size]
        bbmin, bbmax = _bbox(coords, idx)
        d = _largest_dimension(bbmin, bbmax, current_axis)
        order = np.argsort(coords[idx, d], kind="stable")
        indices[node.offset:node.offset + node.size] = idx[order]
        mid = node.size * 1 // 2
        kids = []
        if mid > 0:
            kids.append(Cluster(node.offset, mid, node.depth + 1))
        kids.append(Cluster(node.offset + mid, node.size - mid, node.depth + 1))
        node.children = kids
        for c in kids:
            divide(c, d)

    import sys
    old = sys.getrecursionlimit()
    sys.setrecursionlimit(max(old, 10000))
    divide(root, -1)
    sys.setrecursionlimit(old)

    # admissibility prepare: AABB per node from scratch (admissibility.cpp:56-66)
    cid = [0]

    def boxes(node):
        node.bbmin, node.bbmax = _bbox(coords, indices[node.offset:node.offset + node.size])
        node.id = cid[0]
        cid[0] += 1
        for c in node.children:
            boxes(c)

    boxes(root)
    return root, indices


def cluster_leaves(root):
    out = []

    def rec(c):
        if c.is_leaf():
            out.append(c)
        for k in c.children:
            rec(k)

    rec(root)
    return out


# ---------------------------------------------------------------------------
# block tree
# ---------------------------------------------------------------------------
FULL, RK, NODE = "F", "R", "H"


class Rk:
    """A @ B.T with A (m,k), B (n,k); rank 0 means a null block."""
    __slots__ = ("a", "b")

    def __init__(self, a, b):
        self.a = a
        self.b = b

    @property
    def rank(self):
        return 0 if self.a is None else self.a.shape[1]

    def eval(self, m, n, dtype):
        if self.rank == 0:
            return np.zeros((m, n), dtype=dtype)
        return self.a @ self.b.T


class Block:
    __slots__ = ("rows", "cols", "kind", "children", "nrow", "ncol", "sym",
                 "full", "rk", "tri_lower", "depth", "piv")

    def __init__(self, rows, cols, depth):
        self.rows = rows
        self.cols = cols
        self.depth = depth
        self.kind = None
        self.children = None   # flat list, index i + j * nrow (h_matrix.hpp:634-650)
        self.nrow = 0
        self.ncol = 0
        self.sym = False       # isLower
        self.full = None
        self.rk = None
        self.tri_lower = False
        self.piv = None        # getrf row interchanges of an LU-factorised diagonal leaf

    def is_leaf(self):
        return self.kind != NODE

    def get(self, i, j):
        return self.children[i + j * self.nrow]

    @property
    def m(self):
        return self.rows.size

    @property
    def n(self):
        return self.cols.size


def _is_low_rank(r, c, eta):
    # StandardAdmissibilityCondition::isLowRank admissibility.cpp:110-117
    min_d = min(r.diameter(), c.diameter())
    return min_d > 0.0 and min_d <= eta * r.distance_to(c)


def build_block_tree(row_root, col_root, eta=2.0, lower_symmetric=True, hodlr=False):
    """HMatrix constructor + split (h_matrix.cpp:144-221) with the standard
    admissibility (ratio 0, max_width unbounded), or with hodlr=True the HODLR
    one (HODLRAdmissibilityCondition admissibility.cpp:231-241: a block between
    different clusters is low rank; no size limits)."""

    def make(r, c, sym, depth):
        b = Block(r, c, depth)
        if hodlr:
            low = not (r.offset == c.offset and r.size == c.size)
            stop = False
        else:
            low = _is_low_rank(r, c, eta)
            stop = r.size < 2 or c.size < 2
        if (r.is_leaf() and c.is_leaf()) or stop or low:
            force_full = not hodlr and (r.size < 2 or c.size < 2)
            b.kind = RK if (low and not force_full) else FULL
            return b
        split_r, split_c = (not r.is_leaf()), (not c.is_leaf())
        rkids = r.children if split_r else [r]
        ckids = c.children if split_c else [c]
        b.kind = NODE
        b.nrow, b.ncol = len(rkids), len(ckids)
        b.sym = bool(sym)
        b.children = [None] * (b.nrow * b.ncol)
        for i in range(b.nrow):
            for j in range(b.ncol):
                if (not sym) or i >= j:
                    b.children[i + j * b.nrow] = make(rkids[i], ckids[j], sym and i == j, depth + 1)
        return b

    return make(row_root, col_root, lower_symmetric, 0)


def leaves(block):
    """Leaves in preorder (children in flat index order)."""
    out = []

    def rec(b):
        if b is None:
            return
        if b.is_leaf():
            out.append(b)
        else:
            for ch in b.children:
                rec(ch)

    rec(block)
    return out


def partition_signature(block):
    """Preorder list of (row_offset, row_size, col_offset, col_size, kind)."""
    out = []

    def rec(b):
        if b is None:
            return
        out.append((b.rows.offset, b.rows.size, b.cols.offset, b.cols.size, b.kind))
        if not b.is_leaf():
            for ch in b.children:
                rec(ch)

    rec(block)
    return out


# ---------------------------------------------------------------------------
# built-in kernels (the problem definitions of the workloads)
# ---------------------------------------------------------------------------
class ExpKernel:
    """K_ij = exp(-|x_i - x_j| / l) + nugget * delta_ij (user numbering)."""

    def __init__(self, points, length, nugget):
        self.p = np.ascontiguousarray(points, dtype=np.float64)
        self.l = float(length)
        self.nugget = float(nugget)
        self.dtype = np.float64

    def block(self, ri, ci):
        d = self.p[ri][:, None, :] - self.p[ci][None, :, :]
        r = np.sqrt((d * d).sum(axis=2))
        k = np.exp(-r / self.l)
        if self.nugget != 0.0:
            k = k + self.nugget * (ri[:, None] == ci[None, :])
        return k


class ShiftedExpKernel(ExpKernel):
    """K_ij = exp(-|x_i - x_j - s e_z| / l) + nugget * delta_ij (non-symmetric)."""

    def __init__(self, points, length, nugget, shift):
        super().__init__(points, length, nugget)
        self.shift = float(shift)

    def block(self, ri, ci):
        d = self.p[ri][:, None, :] - self.p[ci][None, :, :]
        d[:, :, 2] -= self.shift
        r = np.sqrt((d * d).sum(axis=2))
        k = np.exp(-r / self.l)
        if self.nugget != 0.0:
            k = k + self.nugget * (ri[:, None] == ci[None, :])
        return k


class GaussKernel(ExpKernel):
    """K_ij = exp(-|x_i - x_j|^2 / l^2) + nugget * delta_ij."""

    def block(self, ri, ci):
        d = self.p[ri][:, None, :] - self.p[ci][None, :, :]
        r2 = (d * d).sum(axis=2)
        k = np.exp(-r2 / (self.l * self.l))
        if self.nugget != 0.0:
            k = k + self.nugget * (ri[:, None] == ci[None, :])
        return k


class HelmholtzKernel:
    """Regularised Helmholtz kernel (complex): K_ij = exp(i k r) / (4 pi r) with
    r = sqrt(|x_i - x_j|^2 + h^2) off the diagonal, 1 / (4 pi h) + nugget on it
    (user indices equal); the complex interaction of examples/c-cylinder.c:68-78
    with the distance regularised by h."""

    def __init__(self, points, wavenumber, reg, nugget):
        self.p = np.ascontiguousarray(points, dtype=np.float64)
        self.k = float(wavenumber)
        self.h = float(reg)
        self.nugget = float(nugget)
        self.dtype = np.complex128

    def block(self, ri, ci):
        d = self.p[ri][:, None, :] - self.p[ci][None, :, :]
        r = np.sqrt((d * d).sum(axis=2) + self.h * self.h)
        v = np.exp(1j * self.k * r) / (4.0 * np.pi * r)
        same = ri[:, None] == ci[None, :]
        return np.where(same, 1.0 / (4.0 * np.pi * self.h) + self.nugget, v)


# ---------------------------------------------------------------------------
# dense helpers
# ---------------------------------------------------------------------------
def find_k(sigma, eps):
    # findK scalar_array.cpp:98-116 (sigma sorted descending)
    if sigma.size == 0:
        return 0
    thr = sigma[0] * eps
    k = 0
    while k < sigma.size and not (sigma[k] <= thr):
        k += 1
    return k


def truncated_svd(m, eps):
    """truncatedSvd compression.cpp:224-239 -> Rk with U sqrt(S), V sqrt(S)."""
    if not np.any(m):
        return Rk(None, None)
    u, s, vt = np.linalg.svd(m, full_matrices=False)
    k = find_k(s, eps)
    if k == 0:
        return Rk(None, None)
    sq = np.sqrt(s[:k])
    # the reference transposes LAPACK's V^H without conjugating it
    # (scalar_array.cpp:1010): A B^T = U S V^H
    return Rk(u[:, :k] * sq, vt[:k].T * sq)


def rk_truncate(rk, m, n, eps):
    """RkMatrix::truncate rk_matrix.cpp:490-585 (QR of both panels, SVD of Ra Rb^T)."""
    k = rk.rank
    if k == 0:
        return rk
    if k > min(m, n):
        return truncated_svd(rk.a @ rk.b.T, eps)
    qa, ra = np.linalg.qr(rk.a)
    qb, rb = np.linalg.qr(rk.b)
    r = ra @ rb.T
    u, s, vt = np.linalg.svd(r)
    nk = find_k(s, eps)
    if nk == 0:
        return Rk(None, None)
    sq = np.sqrt(s[:nk])
    return Rk(qa @ (u[:, :nk] * sq), qb @ (vt[:nk].T * sq))


def aca_full(m, eps):
    """acaFull compression.cpp:270-355 on a dense block (copy is modified)."""
    m = np.array(m, copy=True)
    rows, cols = m.shape
    maxk = min(rows, cols)
    acols, bcols = [], []
    est = 0.0
    for nu in range(maxk):
        flat = int(np.argmax(np.abs(m.reshape(-1, order="F"))))
        i, j = flat % rows, flat // rows
        delta = m[i, j]
        if delta == 0:
            break
        va = m[:, j].copy()
        vb = m[i, :] / delta
        m -= np.outer(va, vb)
        new = 0.0
        for l in range(nu - 1):   # (sic) the reference sums l < nu-1
            new += float(np.real(np.vdot(va, acols[l]) * np.vdot(vb, bcols[l])))
        est += 2.0 * new
        ab = float(np.real(np.vdot(va, va)) * np.real(np.vdot(vb, vb)))
        est += ab
        if ab < eps * eps * est:
            break                 # the pair that meets the criterion is dropped (resize(nu))
        acols.append(va)
        bcols.append(vb)
    if not acols:
        return Rk(None, None)
    return Rk(np.stack(acols, axis=1), np.stack(bcols, axis=1))


# ---------------------------------------------------------------------------
# ACA+ / ACA partial (compression.cpp:384-724), on a block-entry oracle
# ---------------------------------------------------------------------------
class _BlockFn:
    def __init__(self, kernel, ri, ci):
        self.k, self.ri, self.ci = kernel, ri, ci
        self.dtype = getattr(kernel, "dtype", np.float64)

    def row(self, i):
        return self.k.block(self.ri[i:i + 1], self.ci)[0]

    def col(self, j):
        return self.k.block(self.ri, self.ci[j:j + 1])[:, 0]

    def full(self):
        return self.k.block(self.ri, self.ci)


def _update(vec, idx, rows, cols):
    # updateRow compression.cpp:82-89: vec -= sum_l cols[l][idx] * rows[l]
    for l in range(len(rows)):
        vec -= cols[l][idx] * rows[l]


def _amax(v):
    return int(np.argmax(np.abs(v)))


def aca_partial(blk, m, n, eps):
    """doCompressionAcaPartial compression.cpp:384-533 (no random pivots)."""
    maxk = min(m, n)
    row_free = np.ones(m, bool)
    col_free = np.ones(n, bool)
    acols, bcols = [], []
    est = 0.0
    row_index, J, k, count = 0, 0, 0, 0
    while True:
        b = blk.row(row_index).astype(blk.dtype)
        _update(b, row_index, bcols, acols)
        row_free[row_index] = False
        mx = 0.0
        for j in range(n):
            v = abs(b[j]) ** 2
            if col_free[j] and v > mx:
                mx, J = v, j
        if b[J] == 0:
            row_index = 0
            while row_index < m and not row_free[row_index]:
                row_index += 1
        else:
            b = b * (1.0 / b[J])
            bcols.append(b)
            a = blk.col(J).astype(b.dtype)
            _update(a, J, acols, bcols)
            col_free[J] = False
            acols.append(a)
            mx = 0.0
            for i in range(m):
                v = abs(a[i]) ** 2
                if row_free[i] and v > mx:
                    mx, row_index = v, i
            new = 0.0
            for l in range(k):
                new += float(np.real(np.vdot(a, acols[l]) * np.vdot(b, bcols[l])))
            est += 2.0 * new
            ab = float(np.real(np.vdot(a, a)) * np.real(np.vdot(b, b)))
            est += ab
            k += 1
            if ab < eps * eps * est:
                break
        count += 1
        if not (count < maxk and row_index < m):
            break
    if k == 0:
        return Rk(None, None)
    return Rk(np.stack(acols, 1), np.stack(bcols, 1))


def aca_plus(blk, m, n, eps):
    """doCompressionAcaPlus compression.cpp:555-724."""
    if n * 100 < m:
        return aca_partial(blk, m, n, eps)
    maxk = min(m, n)
    row_free = np.ones(m, bool)
    col_free = np.ones(n, bool)
    acols, bcols = [], []
    est = 0.0

    def find_col():
        for j in range(n):
            if col_free[j]:
                c = blk.col(j).astype(blk.dtype)
                col_free[j] = False
                if np.any(c != 0):
                    return j, c
        return -1, None

    def find_min_row(a_ref):
        while True:
            i_ref, mn = -1, np.inf
            for i in range(m):
                if row_free[i]:
                    v = abs(a_ref[i]) ** 2
                    if v < mn:
                        i_ref, mn = i, v
            if i_ref == -1:
                return -1, None
            r = blk.row(i_ref).astype(blk.dtype)
            _update(r, i_ref, bcols, acols)
            row_free[i_ref] = False
            if np.any(r != 0):
                return i_ref, r

    def find_min_col(b_ref):
        while True:
            j_ref, mn = -1, np.inf
            for j in range(n):
                if col_free[j]:
                    v = abs(b_ref[j]) ** 2
                    if v < mn:
                        j_ref, mn = j, v
            if j_ref == -1:
                return -1, None
            c = blk.col(j_ref).astype(blk.dtype)
            _update(c, j_ref, acols, bcols)
            col_free[j_ref] = False
            if np.any(c != 0):
                return j_ref, c

    j_ref, a_ref = find_col()
    if j_ref == -1:
        return Rk(None, None)
    i_ref, b_ref = find_min_row(a_ref)
    if b_ref is None:
        b_ref = np.zeros(n, dtype=blk.dtype)
    k = 0
    while True:
        i_star = _amax(a_ref)
        j_star = _amax(b_ref)
        if abs(a_ref[i_star]) ** 2 > abs(b_ref[j_star]) ** 2:
            b = blk.row(i_star).astype(blk.dtype)
            _update(b, i_star, bcols, acols)
            j_star = _amax(b)
            piv = b[j_star]
            a = blk.col(j_star).astype(blk.dtype)
            _update(a, j_star, acols, bcols)
            if piv != 0:
                a = a * (1.0 / piv)
        else:
            a = blk.col(j_star).astype(blk.dtype)
            _update(a, j_star, acols, bcols)
            i_star = _amax(a)
            piv = a[i_star]
            b = blk.row(i_star).astype(blk.dtype)
            _update(b, i_star, bcols, acols)
            if piv != 0:
                b = b * (1.0 / piv)
        row_free[i_star] = False
        col_free[j_star] = False
        acols.append(a)
        bcols.append(b)
        new = 0.0
        for l in range(k):
            new += float(np.real(np.vdot(a, acols[l]) * np.vdot(b, bcols[l])))
        est += 2.0 * new
        ab = float(np.real(np.vdot(a, a)) * np.real(np.vdot(b, b)))
        est += ab
        k += 1
        if ab < eps * eps * est:
            break
        a_ref = a_ref - bcols[k - 1][j_ref] * acols[k - 1]
        b_ref = b_ref - acols[k - 1][i_ref] * bcols[k - 1]
        need_a = (j_star == j_ref) or not np.any(a_ref != 0)
        need_b = (i_star == i_ref) or not np.any(b_ref != 0)
        if need_a and need_b:
            found = False
            while not found:
                j_ref, a_ref = find_col()
                if j_ref == -1:
                    break
                _update(a_ref, j_ref, acols, bcols)
                found = np.any(a_ref != 0)
            if not found:
                break
            i_ref, b_ref = find_min_row(a_ref)
            if i_ref == -1:
                break
        elif need_b:
            i_ref, b_ref = find_min_row(a_ref)
            if i_ref == -1:
                break
        elif need_a:
            j_ref, a_ref = find_min_col(b_ref)
            if j_ref == -1:
                break
        if not k < maxk:
            break
    return Rk(np.stack(acols, 1), np.stack(bcols, 1))


# ---------------------------------------------------------------------------
# assembly (assembly.cpp:49-71, compression.cpp:746-773, h_matrix.cpp:359-447)
# ---------------------------------------------------------------------------
def assemble(root, indices, kernel, method="aca_plus", eps=1e-4, lowrank_eps=1e-4):
    for b in leaves(root):
        ri = indices[b.rows.offset:b.rows.offset + b.m]
        ci = indices[b.cols.offset:b.cols.offset + b.n]
        if b.kind == RK:
            blk = _BlockFn(kernel, ri, ci)
            if max(b.m, b.n) < COMPRESSION_MIN_LEAF_SIZE or method == "svd":
                rk = truncated_svd(blk.full(), eps)
            elif method == "aca_partial":
                rk = aca_partial(blk, b.m, b.n, eps)
            else:
                rk = aca_plus(blk, b.m, b.n, eps)
            b.rk = rk_truncate(rk, b.m, b.n, lowrank_eps)
        else:
            b.full = kernel.block(ri, ci)
    return root


def info(root):
    """HMatrix::info h_matrix.cpp:464-502 (subset of the counters)."""
    res = dict(full_count=0, rk_count=0, compressed_size=0, uncompressed_size=0,
               full_size=0, nr_block_clusters=0)

    def rec(b):
        if b is None:
            return
        res["nr_block_clusters"] += 1
        if b.is_leaf():
            s = b.m * b.n
            res["uncompressed_size"] += s
            if b.kind == RK:
                res["compressed_size"] += (b.rk.rank if b.rk else 0) * (b.m + b.n)
                res["rk_count"] += 1
            else:
                res["compressed_size"] += s
                res["full_count"] += 1
                res["full_size"] += s
        else:
            for ch in b.children:
                rec(ch)

    rec(root)
    return res


def to_dense(root, n_rows, n_cols, dtype=np.float64, symmetric=True):
    """Dense matrix (hmat numbering) of the stored H-matrix; symmetric lower
    storage mirrors each off-diagonal leaf."""
    d = np.zeros((n_rows, n_cols), dtype=dtype)
    for b in leaves(root):
        r0, c0 = b.rows.offset, b.cols.offset
        blk = b.full if b.kind == FULL else b.rk.eval(b.m, b.n, dtype)
        if blk is None:
            continue
        d[r0:r0 + b.m, c0:c0 + b.n] = blk
        if symmetric and b.rows is not b.cols:
            d[c0:c0 + b.n, r0:r0 + b.m] = blk.T
    return d


# ---------------------------------------------------------------------------
# H-gemv (h_matrix.cpp:772-857, rk_matrix.cpp:156-206)
# ---------------------------------------------------------------------------
def _child_for_gemm(h, t, i, j):
    # getChildForGEMM h_matrix.cpp:772-788
    if t != "N":
        i, j = j, i
    if h.sym and j > i:
        return h.get(j, i), ("T" if t == "N" else "N")
    return h.get(i, j), t


def gemv(h, trans, alpha, x, beta, y):
    """y <- alpha op(h) x + beta y on 2-D arrays (rows = h numbering, local)."""
    if beta != 1:
        y *= beta
    if not h.is_leaf():
        ni = h.nrow if trans == "N" else h.ncol
        nj = h.ncol if trans == "N" else h.nrow
        for i in range(ni):
            for j in range(nj):
                ch, t = _child_for_gemm(h, trans, i, j)
                if ch is None:
                    continue
                co = ch.cols.offset - h.cols.offset
                cs = ch.n
                ro = ch.rows.offset - h.rows.offset
                rs = ch.m
                if t != "N":
                    co, ro = ro, co
                    cs, rs = rs, cs
                gemv(ch, t, alpha, x[co:co + cs], 1, y[ro:ro + rs])
        return
    if h.kind == FULL:
        if h.full is None:
            return
        mat = h.full if trans == "N" else h.full.T
        y += alpha * (mat @ x)
    else:
        rk = h.rk
        if rk is None or rk.rank == 0:
            return
        if trans == "N":
            y += alpha * (rk.a @ (rk.b.T @ x))
        else:
            y += alpha * (rk.b @ (rk.a.T @ x))


# ---------------------------------------------------------------------------
# H-arithmetic for the LLt recursion
# ---------------------------------------------------------------------------
def _eval(h):
    """Dense value of a (sub)block with its own stored children only."""
    d = _zeros((h.m, h.n))
    if h.is_leaf():
        if h.kind == FULL:
            if h.full is not None:
                d[:] = h.full
        elif h.rk is not None and h.rk.rank:
            d[:] = h.rk.a @ h.rk.b.T
        return d
    for ch in h.children:
        if ch is None:
            continue
        r0 = ch.rows.offset - h.rows.offset
        c0 = ch.cols.offset - h.cols.offset
        d[r0:r0 + ch.m, c0:c0 + ch.n] = _eval(ch)
    return d


def _is_null(h):
    if h.kind == FULL:
        return h.full is None
    if h.kind == RK:
        return h.rk is None or h.rk.rank == 0
    return False


def _recursively_null(h):
    if h is None:
        return True
    if h.is_leaf():
        return _is_null(h)
    return all(_recursively_null(c) for c in h.children)


def _op_dims(h, t):
    return (h.rows, h.cols) if t == "N" else (h.cols, h.rows)


def _op_rk(rk, t):
    """Rk of op(R): swap panels for 'T'."""
    return rk if t == "N" else Rk(rk.b, rk.a)


def _op_full(f, t):
    return f if t == "N" else f.T


def formatted_add_parts(target, parts, eps):
    """RkMatrix::formattedAddParts (Rk parts) rk_matrix.cpp:846-981.

    target: Block (RK kind) or a ('rows','cols',rk) holder; parts: list of
    (alpha, rows_cluster, cols_cluster, Rk)."""
    m, n = target.m, target.n
    used = []
    if target.rk is not None and target.rk.rank:
        used.append((1.0, target.rows, target.cols, target.rk))
    for (al, rr, cc, rk) in parts:
        if rk is None or rk.rank == 0 or rr.size == 0 or cc.size == 0 or al == 0:
            continue
        used.append((al, rr, cc, rk))
    if not used:
        return
    total = sum(u[3].rank for u in used)
    if total >= min(m, n):
        full = _eval_rk_block(target)
        for (al, rr, cc, rk) in used[(1 if (target.rk is not None and target.rk.rank) else 0):]:
            r0 = rr.offset - target.rows.offset
            c0 = cc.offset - target.cols.offset
            full[r0:r0 + rr.size, c0:c0 + cc.size] += al * (rk.a @ rk.b.T)
        target.rk = truncated_svd(full, abs(eps))
        return
    A = _zeros((m, total))
    B = _zeros((n, total))
    off = 0
    for (al, rr, cc, rk) in used:
        r0 = rr.offset - target.rows.offset
        c0 = cc.offset - target.cols.offset
        k = rk.rank
        A[r0:r0 + rr.size, off:off + k] = al * rk.a
        B[c0:c0 + cc.size, off:off + k] = rk.b
        off += k
    target.rk = Rk(A, B)
    if len(used) > 1 and eps >= 0:
        target.rk = rk_truncate(target.rk, m, n, eps)


def _eval_rk_block(b):
    if b.rk is None or b.rk.rank == 0:
        return _zeros((b.m, b.n))
    return b.rk.a @ b.rk.b.T


def formatted_add_full(target, parts, eps):
    """RkMatrix::formattedAddParts (Full parts) rk_matrix.cpp:984-1010."""
    me = _eval_rk_block(target)
    for (al, rr, cc, f) in parts:
        r0 = rr.offset - target.rows.offset
        c0 = cc.offset - target.cols.offset
        me[r0:r0 + rr.size, c0:c0 + cc.size] += al * f
    target.rk = truncated_svd(me, eps)


class _Tmp:
    """Temporary Rk holder with cluster index sets (used by gemmRk subRks)."""
    __slots__ = ("rows", "cols", "rk")

    def __init__(self, rows, cols):
        self.rows, self.cols, self.rk = rows, cols, Rk(None, None)

    @property
    def m(self):
        return self.rows.size

    @property
    def n(self):
        return self.cols.size


def multiply_rk(eps, ta, tb, a, b):
    """HMatrix::multiplyRkMatrix h_matrix.cpp:1874-1917 -> (rows, cols, Rk)."""
    ar, _ = _op_dims(a, ta)
    _, bc = _op_dims(b, tb)
    if a.kind == RK and not b.is_leaf():
        rk = _op_rk(a.rk, ta)
        if rk is None or rk.rank == 0:
            return ar, bc, Rk(None, None)
        # multiplyRkH rk_matrix.cpp:1127-1189: newB = op(H)^T B
        x = rk.b
        y = _zeros((bc.size, x.shape[1]))
        gemv(b, "T" if tb == "N" else "N", 1.0, x, 0.0, y)
        return ar, bc, Rk(rk.a.copy(), y)
    if not a.is_leaf() and b.kind == RK:
        rk = _op_rk(b.rk, tb)
        if rk is None or rk.rank == 0:
            return ar, bc, Rk(None, None)
        # multiplyHRk rk_matrix.cpp:1192-1243: newA = op(H) A
        x = rk.a
        y = _zeros((ar.size, x.shape[1]))
        gemv(a, ta, 1.0, x, 0.0, y)
        return ar, bc, Rk(y, rk.b.copy())
    if a.kind == RK and b.kind == RK:
        r1, r2 = _op_rk(a.rk, ta), _op_rk(b.rk, tb)
        if r1.rank == 0 or r2.rank == 0:
            return ar, bc, Rk(None, None)
        # multiplyRkRk rk_matrix.cpp:1245-1333 (SVD of B1^T A2)
        tmp = r1.b.T @ r2.a
        u, s, vt = np.linalg.svd(tmp, full_matrices=False)
        k = find_k(s, eps)
        if k == 0:
            return ar, bc, Rk(None, None)
        sq = np.sqrt(s[:k])
        return ar, bc, Rk(r1.a @ (u[:, :k] * sq), r2.b @ (vt[:k].T * sq))
    if a.kind == RK and b.kind == FULL:
        rk = _op_rk(a.rk, ta)
        if rk is None or rk.rank == 0 or b.full is None:
            return ar, bc, Rk(None, None)
        # multiplyRkFull rk_matrix.cpp:1013-1077: newB = op(M)^T B
        return ar, bc, Rk(rk.a.copy(), _op_full(b.full, tb).T @ rk.b)
    if a.kind == FULL and b.kind == RK:
        rk = _op_rk(b.rk, tb)
        if rk is None or rk.rank == 0 or a.full is None:
            return ar, bc, Rk(None, None)
        # multiplyFullRk rk_matrix.cpp:1080-1124: newA = op(M) A
        return ar, bc, Rk(_op_full(a.full, ta) @ rk.a, rk.b.copy())
    return ar, bc, Rk(None, None)


def multiply_full(ta, tb, a, b):
    """HMatrix::multiplyFullMatrix h_matrix.cpp:1920-1950 -> dense or None."""
    ar, _ = _op_dims(a, ta)
    _, bc = _op_dims(b, tb)
    if not a.is_leaf() and b.kind == FULL:
        if b.full is None:
            return None
        x = _op_full(b.full, tb)
        y = _zeros((ar.size, bc.size))
        gemv(a, ta, 1.0, x, 0.0, y)
        return y
    if a.kind == FULL and not b.is_leaf():
        if a.full is None:
            return None
        # multiplyFullH: result^T = op(H)^T op(M)^T
        x = _op_full(a.full, ta).T
        y = _zeros((bc.size, ar.size))
        gemv(b, "T" if tb == "N" else "N", 1.0, np.ascontiguousarray(x), 0.0, y)
        return y.T
    if a.kind == FULL and b.kind == FULL:
        if a.full is None or b.full is None:
            return None
        return _op_full(a.full, ta) @ _op_full(b.full, tb)
    return None


def gemm_rk(target, eps, ta, tb, alpha, ha, hb):
    """RkMatrix::gemmRk rk_matrix.cpp:1352-1415 (target: Block or _Tmp)."""
    if not ha.is_leaf() and not hb.is_leaf():
        nbr = ha.nrow if ta == "N" else ha.ncol
        nbc = hb.ncol if tb == "N" else hb.nrow
        nbk = ha.ncol if ta == "N" else ha.nrow
        subs = [None] * (nbr * nbc)
        for i in range(nbr):
            for j in range(nbc):
                p = i + j * nbr
                for k in range(nbk):
                    aik = ha.get(i, k) if ta == "N" else ha.get(k, i)
                    bkj = hb.get(k, j) if tb == "N" else hb.get(j, k)
                    if aik is not None and bkj is not None:
                        if subs[p] is None:
                            sr = aik.rows if ta == "N" else aik.cols
                            sc = bkj.cols if tb == "N" else bkj.rows
                            subs[p] = _Tmp(sr, sc)
                        gemm_rk(subs[p], eps, ta, tb, alpha, aik, bkj)
        parts = [(1.0, s.rows, s.cols, s.rk) for s in subs if s is not None]
        formatted_add_parts(target, parts, eps)
        return
    rk = None
    if _recursively_null(ha) or _recursively_null(hb):
        pass
    elif ha.kind == RK or hb.kind == RK:
        rr, cc, rk = multiply_rk(eps, ta, tb, ha, hb)
    else:
        f = multiply_full(ta, tb, ha, hb)
        if f is not None:
            rk = aca_full(f, eps)
    if rk is not None:
        if target.rk is None or target.rk.rank == 0:
            if rk.rank:
                target.rk = Rk(alpha * rk.a, rk.b)
            else:
                target.rk = rk
        else:
            ar, _ = _op_dims(ha, ta)
            _, bc = _op_dims(hb, tb)
            formatted_add_parts(target, [(alpha, ar, bc, rk)], eps)


def axpy_rk(h, alpha, rows, cols, rk, eps):
    """HMatrix::axpy(RkMatrix) h_matrix.cpp:976-1030; rk spans (rows, cols)."""
    if rk is None or rk.rank == 0 or h.m == 0 or h.n == 0:
        return
    if not h.is_leaf():
        for c in h.children:
            if c is None:
                continue
            ra = rk.a[c.rows.offset - rows.offset:c.rows.offset - rows.offset + c.m]
            rb = rk.b[c.cols.offset - cols.offset:c.cols.offset - cols.offset + c.n]
            sub = Rk(ra, rb)
            if rk.rank < min(c.m, c.n) and rk.rank > 10:
                sub = rk_truncate(Rk(ra.copy(), rb.copy()), c.m, c.n, eps)
            axpy_rk(c, alpha, c.rows, c.cols, sub, eps)
        return
    ra = rk.a[h.rows.offset - rows.offset:h.rows.offset - rows.offset + h.m]
    rb = rk.b[h.cols.offset - cols.offset:h.cols.offset - cols.offset + h.n]
    if h.kind == RK:
        formatted_add_parts(h, [(alpha, h.rows, h.cols, Rk(ra, rb))], eps)
    else:
        d = ra @ rb.T
        if h.full is None:
            h.full = alpha * d
        else:
            h.full += alpha * d


def axpy_full(h, alpha, rows, cols, f, eps):
    """HMatrix::axpy(FullMatrix) h_matrix.cpp:1034-1073."""
    if not h.is_leaf():
        for c in h.children:
            if c is not None:
                axpy_full(c, alpha, rows, cols, f, eps)
        return
    r0 = h.rows.offset - rows.offset
    c0 = h.cols.offset - cols.offset
    sub = f[r0:r0 + h.m, c0:c0 + h.n]
    if h.kind == RK:
        formatted_add_full(h, [(alpha, h.rows, h.cols, sub)], eps)
    else:
        if h.full is None:
            h.full = alpha * sub
        else:
            h.full += alpha * sub


def leaf_gemm(c, ta, tb, alpha, a, b, eps):
    """HMatrix::leafGemm h_matrix.cpp:1474-1564."""
    if not c.is_leaf():
        if a.kind == RK or b.kind == RK:
            if (a.kind == RK and _is_null(a)) or (b.kind == RK and _is_null(b)):
                return
            rr, cc, rk = multiply_rk(eps, ta, tb, a, b)
            axpy_rk(c, alpha, rr, cc, rk, eps)
        else:
            f = multiply_full(ta, tb, a, b)
            if f is not None:
                ar, _ = _op_dims(a, ta)
                _, bc = _op_dims(b, tb)
                axpy_full(c, alpha, ar, bc, f, eps)
        return
    if c.kind == RK:
        gemm_rk(c, eps, ta, tb, alpha, a, b)
        return
    if (not a.is_leaf() and not b.is_leaf()) or _is_null(c):
        full_hh_gemm(c, ta, tb, alpha, a, b, eps)
        return
    if a.kind == RK or b.kind == RK:
        if (a.kind == RK and _is_null(a)) or (b.kind == RK and _is_null(b)):
            return
        _, _, rk = multiply_rk(eps, ta, tb, a, b)
        f = rk.eval(c.m, c.n, _DTYPE[0])
    elif a.is_leaf() and b.is_leaf():
        if a.full is None or b.full is None:
            return
        c.full += alpha * (_op_full(a.full, ta) @ _op_full(b.full, tb))
        return
    else:
        f = multiply_full(ta, tb, a, b)
    if f is not None:
        c.full += alpha * f


def full_hh_gemm(c, ta, tb, alpha, a, b, eps):
    """fullHHGemm h_matrix.cpp:1449-1472."""
    if not a.is_leaf() and not b.is_leaf():
        ni = a.nrow if ta == "N" else a.ncol
        nj = b.ncol if tb == "N" else b.nrow
        nk = a.ncol if ta == "N" else a.nrow
        for i in range(ni):
            for j in range(nj):
                for k in range(nk):
                    ca, tA = _child_for_gemm(a, ta, i, k)
                    cb, tB = _child_for_gemm(b, tb, k, j)
                    if ca is not None and cb is not None:
                        full_hh_gemm(c, tA, tB, alpha, ca, cb, eps)
    elif not _recursively_null(a) and not _recursively_null(b):
        if c.full is None:
            c.full = _zeros((c.m, c.n))
        hgemm(c, ta, tb, alpha, a, b, eps)


def hgemm(c, ta, tb, alpha, a, b, eps):
    """HMatrix::gemm with beta=1 (h_matrix.cpp:1570-1620) + recursiveGemm."""
    if (a.is_leaf() and _is_null(a)) or (b.is_leaf() and _is_null(b)):
        return
    if not c.is_leaf() and not a.is_leaf() and not b.is_leaf():
        row_a = a.nrow if ta == "N" else a.ncol
        col_a = a.ncol if ta == "N" else a.nrow
        row_b = b.nrow if tb == "N" else b.ncol
        col_b = b.ncol if tb == "N" else b.nrow
        for i in range(c.nrow):
            for j in range(c.ncol):
                ch = c.get(i, j)
                if ch is None:
                    continue
                for ia in range(row_a):
                    for jb in range(col_b):
                        for k in range(col_a):
                            ca, tA = _child_for_gemm(a, ta, ia, k)
                            if ca is None:
                                continue
                            ar, ac = _op_dims(ca, tA)
                            if not _intersects(ar, ch.rows):
                                continue
                            for l in range(row_b):
                                cb, tB = _child_for_gemm(b, tb, l, jb)
                                if cb is None:
                                    continue
                                br, bcc = _op_dims(cb, tB)
                                if not _intersects(bcc, ch.cols) or not _intersects(ac, br):
                                    continue
                                hgemm(ch, tA, tB, alpha, ca, cb, eps)
        return
    # uncompatibleGemm with compatible index sets (tree-aligned operands)
    ar, ac = _op_dims(a, ta)
    br, bc = _op_dims(b, tb)
    assert (ar.offset, ar.size) == (c.rows.offset, c.m), "non tree-aligned gemm"
    assert (bc.offset, bc.size) == (c.cols.offset, c.n), "non tree-aligned gemm"
    assert (ac.offset, ac.size) == (br.offset, br.size), "non tree-aligned gemm"
    if c.is_leaf() and c.kind == FULL and c.full is None:
        full_hh_gemm(c, ta, tb, alpha, a, b, eps)
        return
    leaf_gemm(c, ta, tb, alpha, a, b, eps)


def _intersects(x, y):
    return x.offset < y.offset + y.size and y.offset < x.offset + x.size


def solve_lower_left_dense(L, x):
    """HMatrix::solveLowerTriangularLeft(ScalarArray) h_matrix.cpp:2338-2372,
    Uplo LOWER, NONUNIT: x <- L^{-1} x (x rows aligned with L)."""
    if L.is_leaf():
        import scipy.linalg as sla
        x[:] = sla.solve_triangular(L.full, x, lower=True, trans=0)
        return
    off = 0
    subs = []
    for i in range(L.nrow):
        d = L.get(i, i)
        subs.append((off, d.n))
        for j in range(i):
            lij = L.get(i, j)
            if lij is not None:
                o_i, s_i = subs[i]
                o_j, s_j = subs[j]
                gemv(lij, "N", -1.0, x[o_j:o_j + s_j], 1.0, x[o_i:o_i + s_i])
        solve_lower_left_dense(d, x[off:off + d.n])
        off += d.n


def solve_upper_left_dense(L, x):
    """HMatrix::solveUpperTriangularLeft(ScalarArray) h_matrix.cpp:2504-2540
    with Uplo LOWER: x <- L^{-T} x."""
    if L.is_leaf():
        import scipy.linalg as sla
        x[:] = sla.solve_triangular(L.full, x, lower=True, trans=1)
        return
    subs = []
    off = 0
    for i in range(L.nrow):
        d = L.get(i, i)
        subs.append((off, d.n))
        off += d.n
    for i in range(L.nrow - 1, -1, -1):
        o_i, s_i = subs[i]
        solve_upper_left_dense(L.get(i, i), x[o_i:o_i + s_i])
        for j in range(i):
            lij = L.get(i, j)
            if lij is not None:
                o_j, s_j = subs[j]
                gemv(lij, "T", -1.0, x[o_i:o_i + s_i], 1.0, x[o_j:o_j + s_j])


def solve_upper_right_dense(L, x):
    """HMatrix::solveUpperTriangularRight(ScalarArray) h_matrix.cpp:2417-2451,
    Uplo LOWER: x <- x L^{-T} (x columns aligned with L)."""
    if L.is_leaf():
        import scipy.linalg as sla
        # x L^T = b  <=>  L x^T = b^T
        x[:] = sla.solve_triangular(L.full, x.T, lower=True, trans=0).T
        return
    off = 0
    subs = []
    for i in range(L.ncol):
        d = L.get(i, i)
        subs.append((off, d.m))
        for j in range(i):
            lij = L.get(i, j)
            if lij is not None:
                o_i, s_i = subs[i]
                o_j, s_j = subs[j]
                # x_i -= x_j L_ij^T  (gemv 'T' on the right side)
                y = x[:, o_i:o_i + s_i].T
                gemv(lij, "N", -1.0, np.ascontiguousarray(x[:, o_j:o_j + s_j].T), 1.0, y)
                x[:, o_i:o_i + s_i] = y.T
        solve_upper_right_dense(d, x[:, off:off + d.m])
        off += d.m


def solve_upper_right_h(L, b, eps):
    """HMatrix::solveUpperTriangularRight(HMatrix) h_matrix.cpp:2380-2415 +
    recursiveSolveUpperTriangularRight recursion.cpp:86-134 (LLt, LOWER)."""
    if b is None:
        return
    if not L.is_leaf() and not b.is_leaf():
        if L.nrow == b.ncol:
            for k in range(b.nrow):
                for i in range(L.nrow):
                    bki = b.get(k, i)
                    if bki is None:
                        continue
                    for j in range(i):
                        u_ji = L.get(i, j)
                        bkj = b.get(k, j)
                        if bkj is not None and u_ji is not None:
                            hgemm(bki, "N", "T", -1.0, bkj, u_ji, eps)
                    solve_upper_right_h(L.get(i, i), bki, eps)
        elif L.nrow > 1 and b.ncol == 1 and b.nrow > 1:
            for k in range(b.nrow):
                solve_upper_right_h(L, b.get(k, 0), eps)
        else:
            raise NotImplementedError("solveUpperTriangularRight case")
        return
    if not b.is_leaf():
        d = _eval(b)
        solve_upper_right_dense(L, d)
        _clear_and_set(b, d, eps)
        return
    if _is_null(b):
        return
    if b.kind == FULL:
        solve_upper_right_dense(L, b.full)
    else:
        # Xb^T L^T = Bb^T  <=>  L Xb = Bb
        solve_lower_left_dense(L, b.rk.b)


def _clear_and_set(b, d, eps):
    for lf in leaves(b):
        r0 = lf.rows.offset - b.rows.offset
        c0 = lf.cols.offset - b.cols.offset
        sub = d[r0:r0 + lf.m, c0:c0 + lf.n]
        if lf.kind == FULL:
            lf.full = sub.copy()
        else:
            lf.rk = truncated_svd(sub, eps)


def llt(h, eps):
    """HMatrix::lltDecomposition h_matrix.cpp:2547-2564 +
    recursiveLltDecomposition recursion.cpp:349-390."""
    if h.is_leaf():
        c = np.linalg.cholesky(h.full)   # potrf('L'), upper zeroed
        h.full = np.tril(c)
    else:
        n = h.nrow
        for k in range(n):
            llt(h.get(k, k), eps)
            for i in range(k + 1, n):
                if h.get(i, k) is not None:
                    solve_upper_right_h(h.get(k, k), h.get(i, k), eps)
            for i in range(k + 1, n):
                if h.get(i, k) is None:
                    continue
                for j in range(k + 1, i + 1):
                    if h.get(i, j) is not None and h.get(j, k) is not None:
                        hgemm(h.get(i, j), "N", "T", -1.0, h.get(i, k), h.get(j, k), eps)
    h.tri_lower = True
    h.sym = False


def ldlt_leaf(a):
    """ScalarArray::ldltDecomposition scalar_array.cpp:737-779: A = L D L^T
    without pivoting; returns the unit lower L (diagonal 1, upper triangle 0)
    and D.  Raises on a zero pivot above the last row, as the reference does."""
    n = a.shape[0]
    L = np.array(a, dtype=np.float64, copy=True)
    d = np.empty(n)
    for j in range(n):
        v = L[j, :j] * d[:j]
        dj = L[j, j] - L[j, :j] @ v
        d[j] = dj
        if j + 1 < n:
            if dj == 0.0:
                raise ZeroDivisionError("ldltDecomposition: zero pivot")
            L[j + 1:, j] = (L[j + 1:, j] - L[j + 1:, :j] @ v) / dj
    return np.tril(L, -1) + np.eye(n), d


def _copy_block(b):
    """HMatrix::copy of a block subtree (data copied)."""
    if b is None:
        return None
    c = Block(b.rows, b.cols, b.depth)
    c.kind, c.nrow, c.ncol, c.sym = b.kind, b.nrow, b.ncol, b.sym
    if b.is_leaf():
        c.full = None if b.full is None else b.full.copy()
        if b.rk is not None:
            c.rk = Rk(None if b.rk.a is None else b.rk.a.copy(),
                      None if b.rk.b is None else b.rk.b.copy())
    else:
        c.children = [_copy_block(ch) for ch in b.children]
    return c


def multiply_with_diag(b, d, inverse):
    """HMatrix::multiplyWithDiag(D, Side::RIGHT, inverse) h_matrix.cpp:1953-1998:
    each leaf's columns times d (the diagonal in hmat numbering) or 1 / d; an
    Rk leaf scales the rows of B."""
    for lf in leaves(b):
        s = d[lf.cols.offset:lf.cols.offset + lf.n]
        if inverse:
            s = 1.0 / s
        if lf.kind == FULL:
            if lf.full is not None:
                lf.full = lf.full * s[None, :]
        elif lf.rk is not None and lf.rk.rank > 0:
            lf.rk = Rk(lf.rk.a, lf.rk.b * s[:, None])


def ldlt(h, eps, d):
    """HMatrix::ldltDecomposition h_matrix.cpp:2747-2769 +
    recursiveLdltDecomposition recursion.cpp:37-83.  The leaves get the unit
    lower L; D goes into d (hmat numbering).  L_ik = H_ik L_kk^-T D_k^-1, then
    H_ij -= (L_ik D_k) L_jk^T for k < j <= i (mdntProduct / mdmtProduct
    h_matrix.cpp:2585-2690: a scaled copy of L_ik times L_jk^T)."""
    if h.is_leaf():
        h.full, dd = ldlt_leaf(h.full)
        d[h.rows.offset:h.rows.offset + h.m] = dd
    else:
        n = h.nrow
        for k in range(n):
            ldlt(h.get(k, k), eps, d)
            for i in range(k + 1, n):
                if h.get(i, k) is not None:
                    solve_upper_right_h(h.get(k, k), h.get(i, k), eps)
                    multiply_with_diag(h.get(i, k), d, True)
            for i in range(k + 1, n):
                if h.get(i, k) is None:
                    continue
                w = _copy_block(h.get(i, k))
                multiply_with_diag(w, d, False)
                for j in range(k + 1, i + 1):
                    if h.get(i, j) is not None and h.get(j, k) is not None:
                        hgemm(h.get(i, j), "N", "T", -1.0, w, h.get(j, k), eps)
    h.tri_lower = True
    h.sym = False


def solve_ldlt(L, d, b):
    """HMatrix::solveLdlt h_matrix.cpp:2990-3002 on a (N, nrhs) array in hmat
    order: L y = b, y <- D^-1 y (solveDiagonal :2969-2983), L^T x = y."""
    solve_lower_left_dense(L, b)
    b *= (1.0 / d)[:, None]
    solve_upper_left_dense(L, b)


class HodlrNode:
    """HODLRNode hodlr.cpp:365-404: the r x r blocks X11 and KK of a symmetric
    node and det(M11)."""
    __slots__ = ("x11", "kk", "detm11", "child0", "child1")

    def __init__(self):
        self.x11 = self.kk = None
        self.detm11 = 1.0
        self.child0 = self.child1 = None


def _tri_solve(L, x, lower, trans=False):
    import scipy.linalg as sl
    return sl.solve_triangular(L, x, lower=lower, trans=1 if trans else 0)


def hodlr_factor_sym(a):
    """HODLR<T>::factorizeSym hodlr.cpp:287-312 on a HODLR block tree (every
    inner block a lower-symmetric 2 x 2 node with an Rk (1, 0) child); the
    node blocks in the reference's default QR form (computeQRSymX11 /
    computeQRSymKK hodlr.cpp:221-285).  The leaves get their Cholesky factors,
    the Rk blocks the panels a <- L11^-1 a, b <- L00^-1 b.  Returns the node
    tree."""
    if a.is_leaf():
        raise ValueError("Not HODLR matrix")
    a00, a11, a10 = a.get(0, 0), a.get(1, 1), a.get(1, 0)
    if a10 is None or a10.kind != RK or a.get(0, 1) is not None:
        raise ValueError("Not HODLR matrix")
    node = HodlrNode()
    for k, ch in ((0, a00), (1, a11)):
        if ch.is_leaf():
            ch.full = np.linalg.cholesky(ch.full)
        elif k == 0:
            node.child0 = hodlr_factor_sym(ch)
        else:
            node.child1 = hodlr_factor_sym(ch)
    r = a10.rk.rank
    if r == 0:
        node.x11 = node.kk = np.zeros((0, 0))
        return node
    b = a10.rk.b.copy()
    hodlr_lower(a00, b, node.child0)
    ua = a10.rk.a.copy()
    hodlr_lower(a11, ua, node.child1)
    a10.rk = Rk(ua, b)
    ra = np.linalg.qr(ua, mode="r")
    x = ra @ (-(b.T @ b)) @ ra.T + np.eye(r)
    m11 = np.linalg.cholesky(x)
    node.detm11 = float(np.prod(np.diag(m11)))
    x = _tri_solve(ra, m11 - np.eye(r), lower=False)          # Ra^-1 (M11 - I)
    tmp = x @ ra + np.eye(r)                                  # Ra^-1 (M11 - I) Ra + I
    node.x11 = _tri_solve(ra, x.T, lower=False).T             # Ra^-1 (M11 - I) Ra^-T
    node.kk = np.linalg.solve(tmp, node.x11)
    return node


def hodlr_lower(m, x, node):
    """x <- L^-1 x (solveLowerTriangularLeft hodlr.cpp:138-156)."""
    if m.is_leaf():
        x[:] = _tri_solve(m.full, x, lower=True)
        return
    n0 = m.get(0, 0).m
    x0, x1 = x[:n0], x[n0:]
    hodlr_lower(m.get(0, 0), x0, node.child0)
    hodlr_lower(m.get(1, 1), x1, node.child1)
    rk = m.get(1, 0).rk
    if rk.rank == 0:
        return
    x1 -= rk.a @ (rk.b.T @ x0)
    x1 -= rk.a @ (node.kk @ (rk.a.T @ x1))


class HodlrNodeNs:
    """Per 2 x 2 node of a non-symmetric HODLR factor: KK = I + [[0, b01^T a10],
    [b10^T a01, 0]] as an LU factorisation, and the transposed copy (a01, b01)
    of the node's Rk block (hodlr.cpp:37-43, 81-116, 384-428)."""
    __slots__ = ("kk", "a01", "b01", "child0", "child1")

    def __init__(self):
        self.kk = self.a01 = self.b01 = self.child0 = self.child1 = None


def hodlr_solve_ns(m, x, node):
    """x <- M^-1 x with a non-symmetric HODLR factor (solve hodlr.cpp:45-74)."""
    import scipy.linalg as sla
    if m.is_leaf():
        x[:] = _tri_solve(m.full, _tri_solve(m.full, x, lower=True), lower=True, trans=True)
        return
    n0 = m.get(0, 0).m
    x0, x1 = x[:n0], x[n0:]
    hodlr_solve_ns(m.get(0, 0), x0, node.child0)
    hodlr_solve_ns(m.get(1, 1), x1, node.child1)
    rk = m.get(1, 0).rk
    r = rk.rank
    if r == 0:
        return
    tv = np.vstack([node.b01.T @ x1, rk.b.T @ x0])
    tv = sla.lu_solve(node.kk, tv)
    x0 -= node.a01 @ tv[:r]
    x1 -= rk.a @ tv[r:]


def hodlr_factor_ns(a):
    """HODLR<T>::factorize (factorize hodlr.cpp:81-116) of a lower-stored HODLR
    tree in place: leaves by Cholesky, a01 b01^T = (a10 b10^T)^T, a01 <- M00^-1
    a01, a10 <- M11^-1 a10, KK = I + [[0, b01^T a10], [b10^T a01, 0]] by LU."""
    import scipy.linalg as sla
    a00, a11, a10 = a.get(0, 0), a.get(1, 1), a.get(1, 0)
    if a10 is None or a10.kind != RK or a.get(0, 1) is not None:
        raise ValueError("Not HODLR matrix")
    node = HodlrNodeNs()
    for k, ch in ((0, a00), (1, a11)):
        if ch.is_leaf():
            ch.full = np.linalg.cholesky(ch.full)
        elif k == 0:
            node.child0 = hodlr_factor_ns(ch)
        else:
            node.child1 = hodlr_factor_ns(ch)
    r = a10.rk.rank
    if r == 0:
        return node
    node.a01, node.b01 = a10.rk.b.copy(), a10.rk.a.copy()
    hodlr_solve_ns(a00, node.a01, node.child0)
    ua = a10.rk.a.copy()
    hodlr_solve_ns(a11, ua, node.child1)
    a10.rk = Rk(ua, a10.rk.b)
    kk = np.eye(2 * r)
    kk[:r, r:] = node.b01.T @ ua
    kk[r:, :r] = a10.rk.b.T @ node.a01
    node.kk = sla.lu_factor(kk)
    return node


def hodlr_gemv(m, trans, x, y, node):
    """y += op(L W) x for a symmetric HODLR factor (HODLR<T>::gemv
    hodlr.cpp:325-378, alpha 1, beta 1): W = [[I, 0], [a b^T, I + a X11 a^T]]
    per node, L the leaves' Cholesky factors."""
    if m.is_leaf():
        y += (m.full if trans == "N" else m.full.T) @ x
        return
    n0 = m.get(0, 0).m
    x0, x1, y0, y1 = x[:n0], x[n0:], y[:n0], y[n0:]
    rk = m.get(1, 0).rk
    r = rk.rank
    if trans == "N":
        x1b = x1.copy()
        if r:
            x1b += rk.a @ (rk.b.T @ x0 + node.x11 @ (rk.a.T @ x1))
        hodlr_gemv(m.get(1, 1), trans, x1b, y1, node.child1)
        hodlr_gemv(m.get(0, 0), trans, x0, y0, node.child0)
        return
    hodlr_gemv(m.get(0, 0), trans, x0, y0, node.child0)
    l1 = np.zeros_like(y1)
    hodlr_gemv(m.get(1, 1), trans, x1, l1, node.child1)
    if r:
        t = rk.a.T @ l1
        y0 += rk.b @ t
        y1 += rk.a @ (node.x11.T @ t)
    y1 += l1


def factor_gemv(L, trans, x, y):
    """y += op(L) x for an LLt-factorised tree: HMatrix::gemv of the factor
    (isTriLower, no mirrored upper blocks; h_matrix.cpp:771-857); diagonal
    leaves hold L with a zero upper triangle (scalar_array.cpp:822-826)."""
    for b in leaves(L):
        r0, c0 = b.rows.offset - L.rows.offset, b.cols.offset - L.cols.offset
        if b.kind == FULL:
            f = np.tril(b.full) if b.rows is b.cols else b.full
        elif b.rk is not None and b.rk.rank:
            f = b.rk.a @ b.rk.b.T
        else:
            continue
        if trans == "N":
            y[r0:r0 + b.m] += f @ x[c0:c0 + b.n]
        else:
            y[c0:c0 + b.n] += f.T @ x[r0:r0 + b.m]


def hodlr_upper(m, x, node):
    """x <- L^-T x (solveUpperTriangularLeft hodlr.cpp:113-136)."""
    if m.is_leaf():
        x[:] = _tri_solve(m.full, x, lower=True, trans=True)
        return
    n0 = m.get(0, 0).m
    x0, x1 = x[:n0], x[n0:]
    rk = m.get(1, 0).rk
    if rk.rank > 0:
        x1 -= rk.a @ (node.kk.T @ (rk.a.T @ x1))
        x0 -= rk.b @ (rk.a.T @ x1)
    hodlr_upper(m.get(0, 0), x0, node.child0)
    hodlr_upper(m.get(1, 1), x1, node.child1)


def hodlr_logdet(m, node):
    """logdet hodlr.cpp:158-166 (log det L = 1/2 log det A)."""
    if m.is_leaf():
        return float(np.sum(np.log(np.diag(m.full))))
    return (hodlr_logdet(m.get(0, 0), node.child0) + hodlr_logdet(m.get(1, 1), node.child1)
            + math.log(node.detm11))


def solve_llt(L, b):
    """HMatrix::solveLlt h_matrix.cpp:3010-3018 on a (N, nrhs) array in hmat order."""
    solve_lower_left_dense(L, b)
    solve_upper_left_dense(L, b)


def logdet(L):
    """HMatrix::logdet h_matrix.cpp:2925-2936 (= log det L = 0.5 log det A)."""
    if L.is_leaf():
        return float(np.log(np.prod(np.diag(L.full))))
    return sum(logdet(L.get(i, i)) for i in range(L.nrow))


# ---------------------------------------------------------------------------
# H-LU (non-symmetric storage): recursion.cpp:86-134 (UPPER), 191-233, 238-287,
# 393-444; h_matrix.cpp:2297-2415, 2504-2540, 2567, 2772-2779;
# leaves: scalar_array.cpp:694-705 (getrf), 852-874 (laswp + trsm)
# ---------------------------------------------------------------------------
def _laswp(piv, x):
    """LAPACK laswp (forward, incx=1) of the rows of x with getrf pivots."""
    for i, p in enumerate(piv):
        if p != i:
            x[[i, p]] = x[[p, i]]


def _subs_diag(h, along_rows=True):
    out, off = [], 0
    n = h.nrow if along_rows else h.ncol
    for i in range(n):
        d = h.get(i, i)
        size = d.n if along_rows else d.m
        out.append((off, size))
        off += size
    return out


def lu_solve_lower_left_dense(L, x):
    """solveLowerTriangularLeft(ScalarArray) h_matrix.cpp:2338-2372 with LU,
    UNIT, LOWER: at a leaf laswp then unit lower trsm (scalar_array.cpp:852-857)."""
    import scipy.linalg as sla
    if L.is_leaf():
        _laswp(L.piv, x)
        x[:] = sla.solve_triangular(L.full, x, lower=True, unit_diagonal=True)
        return
    subs = _subs_diag(L)
    for i in range(L.nrow):
        o_i, s_i = subs[i]
        for j in range(i):
            lij = L.get(i, j)
            if lij is not None:
                o_j, s_j = subs[j]
                gemv(lij, "N", -1.0, x[o_j:o_j + s_j], 1.0, x[o_i:o_i + s_i])
        lu_solve_lower_left_dense(L.get(i, i), x[o_i:o_i + s_i])


def lu_solve_upper_left_dense(U, x):
    """solveUpperTriangularLeft(ScalarArray) h_matrix.cpp:2504-2540 with
    UPPER, NONUNIT: x <- U^{-1} x (u_ji = get(j, i), gemv 'N')."""
    import scipy.linalg as sla
    if U.is_leaf():
        x[:] = sla.solve_triangular(U.full, x, lower=False, trans=0)
        return
    subs = _subs_diag(U)
    for i in range(U.nrow - 1, -1, -1):
        o_i, s_i = subs[i]
        lu_solve_upper_left_dense(U.get(i, i), x[o_i:o_i + s_i])
        for j in range(i):
            uji = U.get(j, i)
            if uji is not None:
                o_j, s_j = subs[j]
                gemv(uji, "N", -1.0, x[o_i:o_i + s_i], 1.0, x[o_j:o_j + s_j])


def lu_solve_upper_t_left_dense(U, x):
    """solveLowerTriangularLeft(ScalarArray) h_matrix.cpp:2338-2372 with
    Uplo UPPER (no pivots): x <- U^{-T} x (u_ji = get(j, i), gemv 'T')."""
    import scipy.linalg as sla
    if U.is_leaf():
        x[:] = sla.solve_triangular(U.full, x, lower=False, trans=1)
        return
    subs = _subs_diag(U)
    for i in range(U.nrow):
        o_i, s_i = subs[i]
        for j in range(i):
            uji = U.get(j, i)
            if uji is not None:
                o_j, s_j = subs[j]
                gemv(uji, "T", -1.0, x[o_j:o_j + s_j], 1.0, x[o_i:o_i + s_i])
        lu_solve_upper_t_left_dense(U.get(i, i), x[o_i:o_i + s_i])


def lu_solve_upper_right_dense(U, x):
    """solveUpperTriangularRight(ScalarArray) h_matrix.cpp:2417-2451 with
    UPPER: x <- x U^{-1} (trsm RIGHT, UPPER, 'N' at the leaves)."""
    x_t = np.ascontiguousarray(x.T)
    lu_solve_upper_t_left_dense(U, x_t)     # U^T x^T = b^T
    x[:] = x_t.T


def lu_solve_lower_left_h(L, b, eps):
    """HMatrix::solveLowerTriangularLeft(HMatrix) h_matrix.cpp:2297-2336 +
    recursiveSolveLowerTriangularLeft recursion.cpp:191-233 (LU, UNIT, LOWER)."""
    if b is None:
        return
    if not L.is_leaf() and not b.is_leaf():
        if L.ncol == b.nrow:
            for k in range(b.ncol):
                for i in range(L.nrow):
                    bik = b.get(i, k)
                    if bik is None:
                        continue
                    for j in range(i):
                        lij = L.get(i, j)
                        bjk = b.get(j, k)
                        if lij is not None and bjk is not None:
                            hgemm(bik, "N", "N", -1.0, lij, bjk, eps)
                    lu_solve_lower_left_h(L.get(i, i), bik, eps)
        elif L.ncol > 1 and b.nrow == 1 and b.ncol > 1:
            for k in range(b.ncol):
                lu_solve_lower_left_h(L, b.get(0, k), eps)
        else:
            raise NotImplementedError("solveLowerTriangularLeft case")
        return
    if not b.is_leaf():
        if b.nrow == 1:
            for k in range(b.ncol):
                lu_solve_lower_left_h(L, b.get(0, k), eps)
            return
        d = _eval(b)
        lu_solve_lower_left_dense(L, d)
        _clear_and_set(b, d, eps)
        return
    if _is_null(b):
        return
    if b.kind == FULL:
        lu_solve_lower_left_dense(L, b.full)
    else:
        lu_solve_lower_left_dense(L, b.rk.a)


def lu_solve_upper_right_h(U, b, eps):
    """HMatrix::solveUpperTriangularRight(HMatrix) h_matrix.cpp:2380-2415 +
    recursiveSolveUpperTriangularRight recursion.cpp:86-134 (LU, NONUNIT, UPPER)."""
    if b is None:
        return
    if not U.is_leaf() and not b.is_leaf():
        if U.nrow == b.ncol:
            for k in range(b.nrow):
                for i in range(U.nrow):
                    bki = b.get(k, i)
                    if bki is None:
                        continue
                    for j in range(i):
                        uji = U.get(j, i)
                        bkj = b.get(k, j)
                        if bkj is not None and uji is not None:
                            hgemm(bki, "N", "N", -1.0, bkj, uji, eps)
                    lu_solve_upper_right_h(U.get(i, i), bki, eps)
        elif U.nrow > 1 and b.ncol == 1 and b.nrow > 1:
            for k in range(b.nrow):
                lu_solve_upper_right_h(U, b.get(k, 0), eps)
        else:
            raise NotImplementedError("solveUpperTriangularRight case")
        return
    if not b.is_leaf():
        d = _eval(b)
        lu_solve_upper_right_dense(U, d)
        _clear_and_set(b, d, eps)
        return
    if _is_null(b):
        return
    if b.kind == FULL:
        lu_solve_upper_right_dense(U, b.full)
    else:
        # Xb^T U = Bb^T  <=>  U^T Xb = Bb
        lu_solve_upper_t_left_dense(U, b.rk.b)


def lu(h, eps):
    """HMatrix::luDecomposition h_matrix.cpp:2567 + recursiveLuDecomposition
    recursion.cpp:238-287; a leaf is getrf'd in place (L unit lower + U, pivots)."""
    if h.is_leaf():
        import scipy.linalg as sla
        lu_, piv = sla.lu_factor(h.full, check_finite=True)
        h.full = lu_
        h.piv = piv
        return
    n = h.nrow
    for k in range(n):
        hkk = h.get(k, k)
        if hkk is None:
            continue
        lu(hkk, eps)
        for i in range(k + 1, n):
            if h.get(k, i) is not None:
                lu_solve_lower_left_h(hkk, h.get(k, i), eps)
        for i in range(k + 1, n):
            if h.get(i, k) is not None:
                lu_solve_upper_right_h(hkk, h.get(i, k), eps)
        for i in range(k + 1, n):
            if h.get(i, k) is None:
                continue
            for j in range(k + 1, n):
                if h.get(i, j) is not None and h.get(k, j) is not None:
                    hgemm(h.get(i, j), "N", "N", -1.0, h.get(i, k), h.get(k, j), eps)


def solve_lu(LU, b):
    """HMatrix::solve (LU) h_matrix.cpp:2772-2779 on a (N, nrhs) array in hmat order."""
    lu_solve_lower_left_dense(LU, b)
    lu_solve_upper_left_dense(LU, b)


def factor_dense(L, n):
    """Dense lower factor L (hmat numbering) from a factorised tree."""
    return np.tril(to_dense(L, n, n, symmetric=False))


# ---------------------------------------------------------------------------
# workload inputs (seeded; shared with the engine's generators)
# ---------------------------------------------------------------------------
def splitmix64(state):
    """One step of splitmix64; returns (new_state, value)."""
    state = (state + 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF
    z = state
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & 0xFFFFFFFFFFFFFFFF
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & 0xFFFFFFFFFFFFFFFF
    z = z ^ (z >> 31)
    return state, z


def uniform_stream(seed, count):
    """count doubles in [0,1) from splitmix64 (53-bit mantissa)."""
    out = np.empty(count)
    s = seed & 0xFFFFFFFFFFFFFFFF
    for i in range(count):
        s, z = splitmix64(s)
        out[i] = (z >> 11) * (1.0 / 9007199254740992.0)
    return out


def cylinder_points(n, radius, height, seed):
    """Uniform points on a cylinder surface: theta ~ U[0,2pi), z ~ U[0,H)."""
    u = uniform_stream(seed, 2 * n).reshape(n, 2)
    th = 2.0 * math.pi * u[:, 0]
    p = np.empty((n, 3))
    p[:, 0] = radius * np.cos(th)
    p[:, 1] = radius * np.sin(th)
    p[:, 2] = height * u[:, 1]
    return p


# ---------------------------------------------------------------------------
# structure serialization (MatrixStructMarshaller, serialization.cpp:34-130)
# ---------------------------------------------------------------------------
def write_struct(bt, row_root, col_root, idx_rows, idx_cols, points, eps, value_type=1,
                 factorization=-1, lower_symmetric=True, ranks=None):
    """Bytes of MatrixStructMarshaller::write for an assembled (or, with
    factorization=2, LLt-factorised) block tree: scalar type, factorisation,
    both cluster trees (serialization.cpp:34-53: size, dimension, coordinates in
    user numbering, no group index, the permutation, then the nodes as
    offset, size, number of children) and the block tree (writeTreeNode
    serialization.cpp:74-103: flag byte isUpper | isLower << 1 | isTriUpper << 2
    | isTriLower << 3 | keepSameRows << 4 | keepSameCols << 5, or 1 << 7 for a
    null child; approximate rank 25 at leaves, -3 elsewhere (h_matrix.cpp:150,
    165); rank, -1 full, -2 inner; low-rank epsilon; number of children).
    `ranks`, if given, replaces the Rk leaves' ranks in preorder (a factor
    computed elsewhere on the same partition)."""
    import struct as st
    out = bytearray()
    rank_iter = iter(ranks) if ranks is not None else None
    out += st.pack("<ii", value_type, factorization)

    def tree(root, idx):
        n, dim = points.shape
        out.extend(st.pack("<ii", n, dim))
        out.extend(np.ascontiguousarray(points, dtype="<f8").tobytes())
        out.extend(st.pack("<i", 0))
        out.extend(np.ascontiguousarray(idx, dtype="<i4").tobytes())

        def rec(c):
            out.extend(st.pack("<iii", c.offset, c.size, len(c.children)))
            for k in c.children:
                rec(k)
        rec(root)
    tree(row_root, idx_rows)
    tree(col_root, idx_cols)
    tri = factorization == 2

    def node(b):
        if b is None:
            out.extend(st.pack("<B", 1 << 7))
            return
        leaf = b.is_leaf()
        bits = 0
        diag = lower_symmetric and b.rows is b.cols
        if not leaf and b.sym and not tri:
            bits |= 1 << 1
        if diag and tri:
            bits |= 1 << 3
        if leaf or b.nrow == 1:
            bits |= 1 << 4
        if leaf or b.ncol == 1:
            bits |= 1 << 5
        if leaf and b.kind == RK:
            rank = next(rank_iter) if rank_iter is not None else b.rk.rank
        else:
            rank = -1 if leaf else -2
        kids = []
        if not leaf:
            last = max((k for k, ch in enumerate(b.children) if ch is not None), default=-1)
            kids = b.children[:last + 1]
        out.extend(st.pack("<Biidi", bits, 25 if leaf else -3, rank, eps, len(kids)))
        for ch in kids:
            node(ch)
    node(bt)
    return bytes(out)

