"""Thin ctypes binding over libhmat_mi355x.so (the C ABI in include/hmat).

The structures below mirror include/hmat/hmat.h and include/hmat/hmat_mi355x.h
field by field.  ``HMatrix`` drives the same sequence of calls a C application
makes (see examples/c-cholesky.c of hmat-oss): clustering -> cluster tree ->
admissibility -> create_empty_hmatrix_admissibility -> assemble_generic ->
factorize_generic -> solve / gemv.

The binding never computes anything itself: when the shared library is
missing, ``load()`` raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "libhmat_mi355x.so")

HMAT_SIMPLE_PRECISION, HMAT_DOUBLE_PRECISION, HMAT_SIMPLE_COMPLEX, HMAT_DOUBLE_COMPLEX = 0, 1, 2, 3
_NP_TYPE = {HMAT_SIMPLE_PRECISION: np.float32, HMAT_DOUBLE_PRECISION: np.float64,
            HMAT_SIMPLE_COMPLEX: np.complex64, HMAT_DOUBLE_COMPLEX: np.complex128}
FACT_NONE, FACT_LU, FACT_LDLT, FACT_LLT, FACT_HODLR, FACT_HODLRSYM = -1, 0, 1, 2, 3, 4
KERNEL_EXPONENTIAL, KERNEL_GAUSSIAN, KERNEL_SHIFTED_EXPONENTIAL = 0, 1, 2
# derivatives with respect to the correlation length (for the GP likelihood gradient)
KERNEL_EXPONENTIAL_DLENGTH, KERNEL_GAUSSIAN_DLENGTH = 3, 4
# complex (HMAT_DOUBLE_COMPLEX): exp(i k r) / (4 pi r), r = sqrt(|x - y|^2 + h^2)
KERNEL_HELMHOLTZ = 5

_vp = C.c_void_p


class hmat_cluster_info_t(C.Structure):
    _fields_ = [("spatial_dimension", C.c_int), ("dimension", C.c_int),
                ("nr_tree_nodes", C.c_size_t)]


class hmat_info_t(C.Structure):
    _fields_ = [("full_zeros", C.c_int), ("full_size", C.c_size_t), ("rk_size", C.c_size_t),
                ("full_count", C.c_size_t), ("rk_count", C.c_size_t),
                ("compressed_size", C.c_size_t), ("uncompressed_size", C.c_size_t),
                ("nr_block_clusters", C.c_int), ("largest_rk_dim_rows", C.c_int),
                ("largest_rk_dim_cols", C.c_int), ("largest_rk_mem_rows", C.c_int),
                ("largest_rk_mem_cols", C.c_int), ("largest_rk_mem_rank", C.c_int)]


class hmat_assemble_context_t(C.Structure):
    _fields_ = [("assembly", _vp), ("simple_compute", _vp), ("block_compute", _vp),
                ("advanced_compute", _vp), ("user_context", _vp), ("prepare", _vp),
                ("compression", _vp), ("lower_symmetric", C.c_int), ("factorization", C.c_int),
                ("progress", _vp)]


class hmat_block_info_t(C.Structure):
    _fields_ = [("block_type", C.c_int), ("user_data", _vp), ("release_user_data", _vp),
                ("is_guaranteed_null_row", _vp), ("is_guaranteed_null_col", _vp),
                ("needed_memory", C.c_size_t), ("number_of_strata", C.c_int)]


# host-function types of hmat_assemble_context_t (hmat.h: hmat_interaction_func_t,
# hmat_prepare_func_t, hmat_compute_func_t)
INTERACTION_FUNC = C.CFUNCTYPE(None, _vp, C.c_int, C.c_int, _vp)
PREPARE_FUNC = C.CFUNCTYPE(None, C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_int),
                           C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int), _vp,
                           C.POINTER(hmat_block_info_t))
COMPUTE_FUNC = C.CFUNCTYPE(None, _vp, C.c_int, C.c_int, C.c_int, C.c_int, _vp)


class hmat_get_values_context_t(C.Structure):
    _fields_ = [("matrix", _vp), ("values", _vp), ("row_offset", C.c_int), ("col_offset", C.c_int),
                ("row_size", C.c_int), ("col_size", C.c_int),
                ("row_indices", C.POINTER(C.c_int)), ("col_indices", C.POINTER(C.c_int)),
                ("renumber_rows", C.c_int, 1)]


class hmat_progress_t(C.Structure):
    pass


PROGRESS_FUNC = C.CFUNCTYPE(None, C.POINTER(hmat_progress_t))
hmat_progress_t._fields_ = [("max", C.c_int), ("current", C.c_int), ("update", PROGRESS_FUNC),
                            ("user_data", _vp)]


class hmat_factorization_context_t(C.Structure):
    _fields_ = [("factorization", C.c_int), ("progress", _vp)]


class hmat_solve_context_t(C.Structure):
    _fields_ = [("values", _vp), ("nr_rhs", C.c_int), ("no_permutation", C.c_int),
                ("lower", C.c_int), ("upper", C.c_int), ("progress", _vp)]


class hmat_mi355x_kernel_t(C.Structure):
    _fields_ = [("kind", C.c_int), ("dimension", C.c_int), ("n", C.c_int),
                ("points", C.POINTER(C.c_double)), ("length", C.c_double),
                ("nugget", C.c_double), ("shift", C.c_double), ("wavenumber", C.c_double),
                ("regularization", C.c_double)]


_CREATE = C.CFUNCTYPE(_vp, _vp, _vp, C.c_int, _vp)
_P_INT = C.CFUNCTYPE(C.c_int, _vp)
_ASSEMBLE = C.CFUNCTYPE(C.c_int, _vp, C.POINTER(hmat_assemble_context_t))
_FACTOR = C.CFUNCTYPE(C.c_int, _vp, C.POINTER(hmat_factorization_context_t))
_LOGDET = C.CFUNCTYPE(C.c_int, _vp, C.POINTER(C.c_double))
_SOLVE = C.CFUNCTYPE(C.c_int, _vp, _vp, C.c_int)
_SOLVE_GEN = C.CFUNCTYPE(C.c_int, _vp, C.POINTER(hmat_solve_context_t))
_REORDER = C.CFUNCTYPE(C.c_int, _vp, _vp, C.c_int, _vp, C.c_int)
_GEMV = C.CFUNCTYPE(C.c_int, C.c_char, _vp, _vp, _vp, _vp, _vp, C.c_int)
_GEMM_DENSE = C.CFUNCTYPE(C.c_int, C.c_char, C.c_char, C.c_char, _vp, _vp, _vp, _vp, _vp, C.c_int)
_INFO = C.CFUNCTYPE(C.c_int, _vp, C.POINTER(hmat_info_t))
_VOID_EPS = C.CFUNCTYPE(None, _vp, C.c_double)
_OWN = C.CFUNCTYPE(None, _vp, C.c_int, C.c_int)
_INIT = C.CFUNCTYPE(C.c_int)
_COPY = C.CFUNCTYPE(_vp, _vp)
_NORM = C.CFUNCTYPE(C.c_double, _vp)
_SCALAR_H = C.CFUNCTYPE(C.c_int, _vp, _vp)
_AXPY = C.CFUNCTYPE(C.c_int, _vp, _vp, _vp)
_GEMM_H = C.CFUNCTYPE(C.c_int, C.c_char, C.c_char, _vp, _vp, _vp, _vp, _vp)
_SOLVE_MAT = C.CFUNCTYPE(C.c_int, _vp, _vp)
_DIAG = C.CFUNCTYPE(C.c_int, _vp, _vp, C.c_int)
_SOLVE_TRI = C.CFUNCTYPE(C.c_int, _vp, C.c_int, _vp, C.c_int)
IOSTREAM = C.CFUNCTYPE(None, _vp, C.c_size_t, _vp)   # hmat_iostream
_READ_STRUCT = C.CFUNCTYPE(_vp, IOSTREAM, _vp)
_WRITE_IO = C.CFUNCTYPE(None, _vp, IOSTREAM, _vp)
_GET_TREES = C.CFUNCTYPE(C.c_int, _vp, C.POINTER(_vp), C.POINTER(_vp))
_GET_VALUES = C.CFUNCTYPE(C.c_int, C.POINTER(hmat_get_values_context_t))

# slot order of hmat_interface_t (hmat.h:717-1103); unused slots are plain pointers
_SLOTS = [
    ("create_empty_hmatrix_admissibility", _CREATE), ("own_cluster_trees", _OWN),
    ("set_low_rank_epsilon", _VOID_EPS), ("assemble_generic", _ASSEMBLE), ("copy", _COPY),
    ("copy_struct", _COPY), ("norm", _NORM), ("logdet", _LOGDET), ("inverse", _vp),
    ("factorize_generic", _FACTOR), ("destroy", _P_INT), ("get_child", _vp),
    ("destroy_child", _vp), ("solve_generic", _SOLVE_GEN), ("solve_mat", _SOLVE_MAT),
    ("solve_systems", _SOLVE), ("solve_dense", _SOLVE), ("transpose", _P_INT), ("scale", _SCALAR_H),
    ("truncate", _P_INT), ("vector_reorder", _REORDER), ("vector_restore", _REORDER),
    ("gemm", _GEMM_H), ("axpy", _AXPY), ("trsm", _vp), ("gemv", _GEMV), ("gemm_scalar", _GEMV),
    ("full_gemm", _vp), ("gemm_dense", _GEMM_DENSE), ("add_identity", _SCALAR_H), ("init", _INIT),
    ("finalize", _INIT), ("get_info", _INFO), ("get_profile", _vp), ("get_ratio", _vp),
    ("FPcompress", _vp), ("FPdecompress", _vp), ("GetFPCompressionSettings", _vp),
    ("SetFPCompressionSettings", _vp), ("SetFPCompressionSettingsParams", _vp),
    ("dump_info", _vp), ("get_cluster_trees", _GET_TREES), ("set_cluster_trees", _vp),
    ("extract_diagonal", _DIAG), ("solve_lower_triangular", _SOLVE_TRI),
    ("solve_lower_triangular_dense", _SOLVE_TRI), ("get_block", _GET_VALUES),
    ("get_values", _GET_VALUES),
    ("walk", _vp), ("apply_on_leaf", _vp), ("value_type", C.c_int), ("internal", _vp),
    ("read_struct", _READ_STRUCT), ("write_struct", _WRITE_IO), ("read_data", _WRITE_IO), ("write_data", _WRITE_IO),
    ("set_progressbar", _vp), ("extract_diagonal_block", _vp),
]


class hmat_interface_t(C.Structure):
    _fields_ = _SLOTS


# exported symbols the binding uses, with their signatures
_EXPORTS = {
    "hmat_init_mi355x_interface": (None, [C.POINTER(hmat_interface_t), C.c_int]),
    "hmat_create_clustering_median": (_vp, []),
    "hmat_create_clustering_max_dof": (_vp, [_vp, C.c_int]),
    "hmat_delete_clustering": (None, [_vp]),
    "hmat_create_cluster_tree": (_vp, [_vp, C.c_int, C.c_int, _vp]),
    "hmat_delete_cluster_tree": (None, [_vp]),
    "hmat_cluster_get_indices": (C.POINTER(C.c_int), [_vp]),
    "hmat_tree_nodes_count": (C.c_int, [_vp]),
    "hmat_cluster_get_info": (C.c_int, [_vp, _vp]),
    "hmat_create_admissibility_standard": (_vp, [C.c_double]),
    "hmat_create_admissibility_hodlr": (_vp, []),
    "hmat_delete_admissibility": (None, [_vp]),
    "hmat_create_compression_aca_plus": (_vp, [C.c_double]),
    "hmat_create_compression_aca_partial": (_vp, [C.c_double]),
    "hmat_create_compression_svd": (_vp, [C.c_double]),
    "hmat_delete_compression": (None, [_vp]),
    "hmat_assemble_context_init": (None, [C.POINTER(hmat_assemble_context_t)]),
    "hmat_factorization_context_init": (None, [C.POINTER(hmat_factorization_context_t)]),
    "hmat_solve_context_init": (None, [C.POINTER(hmat_solve_context_t)]),
    "hmat_get_version": (C.c_char_p, []),
    "hmat_mi355x_kernel_compute": (None, [_vp, C.c_int, C.c_int, _vp]),
    "hmx_last_error": (C.c_char_p, []),
    "hmx_gemv_device": (C.c_int, [_vp, C.c_int, C.c_double, _vp, C.c_double, _vp, C.c_int]),
    "hmx_gemv_device_noalloc": (C.c_int, [_vp, C.c_int, C.c_double, _vp, C.c_double, _vp,
                                          C.c_int]),
    "hmx_run_batch": (C.c_int, [_vp, _vp, C.c_int, C.c_int, C.c_int, C.c_double, _vp, _vp,
                                C.c_int, _vp, _vp]),
    "hmx_solve_device": (C.c_int, [_vp, C.c_int, _vp, C.c_int]),
    "hmx_factorize_mixed": (C.c_int, [_vp, C.c_int]),
    "hmx_solve_refined_device": (C.c_int, [_vp, _vp, C.c_int, C.c_double, C.c_int,
                                           C.POINTER(C.c_double)]),
    "hmx_partition": (C.c_int, [_vp, _vp]),
    "hmx_leaf_count": (C.c_int, [_vp]),
    "hmx_leaf_info": (C.c_int, [_vp, C.c_int, C.c_int, C.POINTER(C.c_int)]),
    "hmx_leaf_data": (C.c_int, [_vp, C.c_int, C.c_int, _vp, _vp]),
    "hmx_leaf_table": (C.c_int, [_vp, C.c_int, _vp]),
    "hmx_stats": (C.c_int, [_vp, C.POINTER(C.c_double)]),
    "hmx_set_rank_capacity": (None, [C.c_int]),
    "hmx_autotune_solve": (C.c_int, [_vp, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_double)]),
    "hmx_set_solve_block_rows": (C.c_int, [_vp, C.c_int]),
    "hmx_set_large_block_rows": (C.c_int, [_vp, C.c_int]),
    "hmx_set_stream": (C.c_int, [_vp, _vp]),
}

_lib = None


def load(path: str = LIB_PATH):
    """Loads the engine library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(path):
            raise RuntimeError(f"{path} is missing: run __graft_entry__.build() first")
        lib = C.CDLL(path)
        for name, (res, args) in _EXPORTS.items():
            f = getattr(lib, name)
            f.restype = res
            f.argtypes = args
        _lib = lib
    return _lib


class EngineError(RuntimeError):
    pass


def _check(rc, what):
    if rc != 0:
        raise EngineError(f"{what} failed: {load().hmx_last_error().decode()}")


def interface(value_type=HMAT_DOUBLE_PRECISION):
    lib = load()
    it = hmat_interface_t()
    lib.hmat_init_mi355x_interface(C.byref(it), value_type)
    return it


class ClusterTree:
    def __init__(self, points, max_leaf=128):
        lib = load()
        self.points = np.ascontiguousarray(points, dtype=np.float64)
        self.n, self.dim = self.points.shape
        algo = lib.hmat_create_clustering_median()
        algo2 = lib.hmat_create_clustering_max_dof(algo, int(max_leaf))
        self.handle = lib.hmat_create_cluster_tree(self.points.ctypes.data, self.dim, self.n, algo2)
        lib.hmat_delete_clustering(algo2)
        lib.hmat_delete_clustering(algo)
        if not self.handle:
            raise EngineError(lib.hmx_last_error().decode())

    @property
    def indices(self):
        p = load().hmat_cluster_get_indices(self.handle)
        return np.ctypeslib.as_array(p, shape=(self.n,)).copy()

    def __del__(self):
        if getattr(self, "handle", None) and _lib is not None:
            _lib.hmat_delete_cluster_tree(self.handle)
            self.handle = None


def _source(data, pos):
    """An hmat_iostream reading `data` from pos[0]; past the end it yields zeros."""
    data = bytes(data)

    @IOSTREAM
    def source(buf, n, ud):
        if n:
            chunk = data[pos[0]:pos[0] + n]
            C.memmove(buf, chunk, len(chunk))
            if len(chunk) < n:
                C.memset(buf + len(chunk), 0, n - len(chunk))
            pos[0] += n
    return source


class _BorrowedTree:
    """A cluster tree owned by a matrix (read from a stream): size and permutation."""

    def __init__(self, handle):
        self.handle = handle
        info = hmat_cluster_info_t()
        load().hmat_cluster_get_info(handle, C.byref(info))
        self.n, self.dim = int(info.dimension), int(info.spatial_dimension)

    @property
    def indices(self):
        p = load().hmat_cluster_get_indices(self.handle)
        return np.ctypeslib.as_array(p, shape=(self.n,)).copy()


class HMatrix:
    """An H-matrix created and driven through hmat_interface_t."""

    def __init__(self, tree: ClusterTree, eta=2.0, lower_symmetric=True,
                 value_type=HMAT_DOUBLE_PRECISION, lowrank_eps=1e-4, hodlr=False):
        lib = load()
        self.tree = tree
        self.iface = interface(value_type)
        self.value_type = value_type
        adm = (lib.hmat_create_admissibility_hodlr() if hodlr
               else lib.hmat_create_admissibility_standard(float(eta)))
        self.handle = self.iface.create_empty_hmatrix_admissibility(
            tree.handle, tree.handle, 1 if lower_symmetric else 0, adm)
        lib.hmat_delete_admissibility(adm)
        if not self.handle:
            raise EngineError(lib.hmx_last_error().decode())
        self.lower_symmetric = lower_symmetric
        self.iface.set_low_rank_epsilon(self.handle, float(lowrank_eps))
        self._kernel = None

    @property
    def n(self):
        return self.tree.n

    @property
    def is_complex(self):
        return self.value_type in (HMAT_SIMPLE_COMPLEX, HMAT_DOUBLE_COMPLEX)

    @property
    def dtype(self):
        return _NP_TYPE[self.value_type]

    def assemble_context(self, kind=KERNEL_EXPONENTIAL, length=0.5, nugget=1e-2, eps=1e-4,
                         method="aca_plus", factorization=FACT_NONE, shift=0.0, wavenumber=0.0,
                         reg=0.0):
        """An hmat_assemble_context_t for this matrix (the kernel as data, see
        hmat_mi355x_kernel_t), plus the compression object to delete after use."""
        lib = load()
        k = hmat_mi355x_kernel_t()
        k.kind, k.dimension, k.n = kind, self.tree.dim, self.tree.n
        k.points = self.tree.points.ctypes.data_as(C.POINTER(C.c_double))
        k.length, k.nugget, k.shift = float(length), float(nugget), float(shift)
        k.wavenumber, k.regularization = float(wavenumber), float(reg)
        self._kernel = k
        ctx = hmat_assemble_context_t()
        lib.hmat_assemble_context_init(C.byref(ctx))
        comp = {"aca_plus": lib.hmat_create_compression_aca_plus,
                "aca_partial": lib.hmat_create_compression_aca_partial,
                "svd": lib.hmat_create_compression_svd}[method](float(eps))
        ctx.compression = comp
        ctx.simple_compute = C.cast(lib.hmat_mi355x_kernel_compute, _vp).value
        ctx.user_context = C.cast(C.pointer(k), _vp).value
        ctx.lower_symmetric = 1 if self.lower_symmetric else 0
        ctx.factorization = factorization
        return ctx, comp

    def _progress_reporter(self, progress):
        """An hmat_progress_t calling progress(current, max), kept alive with
        the matrix."""
        @PROGRESS_FUNC
        def update(p):
            progress(p[0].current, p[0].max)
        prog = hmat_progress_t(0, 0, update, None)
        self._progress = (prog, update)
        return C.cast(C.pointer(prog), _vp)

    def assemble(self, kind=KERNEL_EXPONENTIAL, length=0.5, nugget=1e-2, eps=1e-4,
                 method="aca_plus", factorization=FACT_NONE, shift=0.0, wavenumber=0.0, reg=0.0,
                 progress=None):
        """assemble_generic with the device kernel; progress(current, max) is
        called with the number of leaves assembled as the assembly advances."""
        ctx, comp = self.assemble_context(kind, length, nugget, eps, method, factorization, shift,
                                          wavenumber, reg)
        if progress is not None:
            ctx.progress = self._progress_reporter(progress)
        rc = self.iface.assemble_generic(self.handle, C.byref(ctx))
        load().hmat_delete_compression(comp)
        _check(rc, "assemble_generic")

    def assemble_functions(self, simple=None, block=None, prepare=None, eps=1e-4,
                           method="aca_plus"):
        """assemble_generic with the caller's host functions (ctypes callbacks
        of INTERACTION_FUNC, or COMPUTE_FUNC + PREPARE_FUNC): every block is
        evaluated on the host, then compressed on the device."""
        lib = load()
        ctx = hmat_assemble_context_t()
        lib.hmat_assemble_context_init(C.byref(ctx))
        comp = {"aca_plus": lib.hmat_create_compression_aca_plus,
                "aca_partial": lib.hmat_create_compression_aca_partial,
                "svd": lib.hmat_create_compression_svd}[method](float(eps))
        ctx.compression = comp
        if simple is not None:
            ctx.simple_compute = C.cast(simple, _vp).value
        if block is not None:
            ctx.block_compute = C.cast(block, _vp).value
            ctx.prepare = C.cast(prepare, _vp).value
        ctx.lower_symmetric = 1 if self.lower_symmetric else 0
        self._callbacks = (simple, block, prepare)
        rc = self.iface.assemble_generic(self.handle, C.byref(ctx))
        lib.hmat_delete_compression(comp)
        _check(rc, "assemble_generic")

    def factorize(self, algo=FACT_LLT, mixed=False, progress=None):
        """factorize_generic; progress(current, max) is called as the
        factorisation advances (hmat_progress_t of the context)."""
        if mixed:
            _check(load().hmx_factorize_mixed(self.handle, algo), "hmx_factorize_mixed")
            return
        ctx = hmat_factorization_context_t()
        load().hmat_factorization_context_init(C.byref(ctx))
        ctx.factorization = algo
        if progress is not None:
            ctx.progress = self._progress_reporter(progress)
        _check(self.iface.factorize_generic(self.handle, C.byref(ctx)), "factorize_generic")

    def info(self):
        inf = hmat_info_t()
        _check(self.iface.get_info(self.handle, C.byref(inf)), "get_info")
        return {k: getattr(inf, k) for k, _ in hmat_info_t._fields_}

    def partition(self):
        lib = load()
        cnt = lib.hmx_partition(self.handle, None)
        out = np.zeros(cnt * 5, dtype=np.int32)
        lib.hmx_partition(self.handle, out.ctypes.data)
        out = out.reshape(cnt, 5)
        if self.is_complex:
            # the engine's blocks cover the 2 x 2 real representation
            out[:, :4] //= 2
        return out

    def leaf_table(self, precision=0):
        """(count, 6) int array: m, n, row offset, col offset, kind, rank."""
        lib = load()
        cnt = lib.hmx_leaf_count(self.handle)
        out = np.zeros(cnt * 6, dtype=np.int32)
        _check(lib.hmx_leaf_table(self.handle, precision, out.ctypes.data), "hmx_leaf_table")
        return out.reshape(cnt, 6)

    def leaf_shapes(self):
        for m, n, r0, c0, kind, _ in self.leaf_table():
            yield int(m), int(n), int(r0), int(c0), int(kind)

    def leaves(self, precision=0):
        """Yields (m, n, r0, c0, kind, rank, a, b) with host copies of the data."""
        lib = load()
        for leaf, row in enumerate(self.leaf_table(precision)):
            m, n, r0, c0, kind, rank = (int(v) for v in row)
            if kind == 0:
                a = np.zeros(m * n)
                _check(lib.hmx_leaf_data(self.handle, leaf, precision, a.ctypes.data, None), "leaf")
                yield m, n, r0, c0, kind, rank, a.reshape(n, m).T, None
            else:
                k = max(rank, 0)
                a = np.zeros(m * max(k, 1))
                b = np.zeros(n * max(k, 1))
                _check(lib.hmx_leaf_data(self.handle, leaf, precision, a.ctypes.data, b.ctypes.data), "leaf")
                yield m, n, r0, c0, kind, rank, a[:m * k].reshape(k, m).T, b[:n * k].reshape(k, n).T

    def to_dense(self, precision=0, mirror=True, embedded=False):
        """Dense matrix (hmat numbering) from the stored leaves.  A complex
        matrix is stored as the 2 x 2 real representation [[Re, -Im], [Im, Re]]
        of its entries, which compression keeps to the tolerance only: with
        embedded=True the stored real 2N x 2N matrix (hmat row / column 2p + c
        is component c of hmat DoF p), otherwise its nearest complex matrix
        (Re and Im averaged over the two places that hold them)."""
        if self.is_complex:
            if self.value_type == HMAT_SIMPLE_COMPLEX:
                precision = 1    # the single-precision store holds the matrix
            d = np.zeros((2 * self.n, 2 * self.n))
            for m, n, r0, c0, kind, rank, a, b in self.leaves(precision):
                d[r0:r0 + m, c0:c0 + n] = a if kind == 0 else (a @ b.T if rank > 0 else 0.0)
            if embedded:
                return d
            return 0.5 * (d[0::2, 0::2] + d[1::2, 1::2]) + 0.5j * (d[1::2, 0::2] - d[0::2, 1::2])
        d = np.zeros((self.n, self.n))
        for m, n, r0, c0, kind, rank, a, b in self.leaves(precision):
            blk = a if kind == 0 else (a @ b.T if rank > 0 else np.zeros((m, n)))
            d[r0:r0 + m, c0:c0 + n] = blk
            if mirror and self.lower_symmetric and (r0 != c0 or m != n):
                d[c0:c0 + n, r0:r0 + m] = blk.T
        return d

    # host-pointer API entries (user numbering unless stated)
    def gemv(self, x, alpha=1.0, beta=0.0, y=None, permute=True, trans="N"):
        """y = alpha op(A) x + beta y; on an LLt / HODLR-factorised matrix op
        applies to the factor (L, or L W), as the reference's gemv does."""
        dt = self.dtype
        x = np.asfortranarray(x, dtype=dt)
        nrhs = 1 if x.ndim == 1 else x.shape[1]
        if x.shape[0] != self.n:
            raise ValueError(f"gemv: x has {x.shape[0]} rows, the matrix {self.n} columns")
        if y is None:
            y = np.zeros_like(x)
        elif y.shape != x.shape:
            raise ValueError(f"gemv: y has shape {y.shape}, expected {x.shape}")
        out = np.asfortranarray(y, dtype=dt)   # a converted copy when y's dtype/layout differs
        a, b = self._scalar(alpha), self._scalar(beta)
        fn = self.iface.gemv if permute else self.iface.gemm_scalar
        _check(fn(trans.encode(), C.cast(C.byref(a), _vp), self.handle, x.ctypes.data,
                  C.cast(C.byref(b), _vp), out.ctypes.data, nrhs), "gemv")
        if out is not y and isinstance(y, np.ndarray):
            y[...] = out
        return out

    def solve(self, b, permute=True):
        x = np.array(b, dtype=self.dtype, order="F", copy=True)
        nrhs = 1 if x.ndim == 1 else x.shape[1]
        fn = self.iface.solve_systems if permute else self.iface.solve_dense
        _check(fn(self.handle, x.ctypes.data, nrhs), "solve")
        return x

    # ---- whole-matrix slots ----------------------------------------------------
    def _scalar(self, v):
        if self.is_complex:
            ct = C.c_double if self.value_type == HMAT_DOUBLE_COMPLEX else C.c_float
            return (ct * 2)(complex(v).real, complex(v).imag)
        return (C.c_double if self.value_type else C.c_float)(v)

    def copy(self, structure_only=False):
        """hmat_interface_t.copy / copy_struct: a new matrix sharing this one's
        cluster trees (a null matrix of the same structure when structure_only)."""
        fn = self.iface.copy_struct if structure_only else self.iface.copy
        h = fn(self.handle)
        if not h:
            raise EngineError(f"copy failed: {load().hmx_last_error().decode()}")
        c = object.__new__(HMatrix)
        c.tree, c.iface, c.value_type = self.tree, self.iface, self.value_type
        c.handle, c.lower_symmetric, c._kernel = h, self.lower_symmetric, self._kernel
        return c

    def transpose(self):
        """hmat_interface_t.transpose: A <- A^T in place (for a complex matrix
        the transpose, not the conjugate transpose)."""
        _check(self.iface.transpose(self.handle), "transpose")

    def norm(self):
        r = self.iface.norm(self.handle)
        if r != r:
            raise EngineError(f"norm failed: {load().hmx_last_error().decode()}")
        return r

    def scale(self, alpha):
        a = self._scalar(alpha)
        _check(self.iface.scale(C.cast(C.byref(a), _vp), self.handle), "scale")

    def add_identity(self, alpha):
        a = self._scalar(alpha)
        _check(self.iface.add_identity(self.handle, C.cast(C.byref(a), _vp)), "add_identity")

    def truncate(self):
        _check(self.iface.truncate(self.handle), "truncate")

    def axpy(self, alpha, x):
        """self <- self + alpha x (x of the same partition)."""
        a = self._scalar(alpha)
        _check(self.iface.axpy(C.cast(C.byref(a), _vp), x.handle, self.handle), "axpy")

    def gemm(self, ta, tb, alpha, a, b, beta):
        """self <- beta self + alpha op(a) op(b), H-matrices over the same trees."""
        al, be = self._scalar(alpha), self._scalar(beta)
        _check(self.iface.gemm(ta.encode() if isinstance(ta, str) else ta,
                               tb.encode() if isinstance(tb, str) else tb,
                               C.cast(C.byref(al), _vp), a.handle, b.handle,
                               C.cast(C.byref(be), _vp), self.handle), "gemm")

    def solve_mat(self, b, lower=False, upper=False):
        """b <- self^-1 b for a factorised self and an H-matrix b (in place);
        lower / upper: L^-1 b / L^-T b only (solve_generic with nr_rhs = 0)."""
        if not lower and not upper:
            _check(self.iface.solve_mat(self.handle, b.handle), "solve_mat")
            return
        ctx = hmat_solve_context_t()
        load().hmat_solve_context_init(C.byref(ctx))
        ctx.values, ctx.nr_rhs, ctx.lower, ctx.upper = b.handle, 0, int(lower), int(upper)
        _check(self.iface.solve_generic(self.handle, C.byref(ctx)), "solve_generic")

    def extract_diagonal(self):
        d = np.zeros(self.n, dtype=np.float64 if self.value_type else np.float32)
        _check(self.iface.extract_diagonal(self.handle, d.ctypes.data, self.n), "extract_diagonal")
        return d

    def solve_lower(self, b, transpose=False, permute=True):
        """L x = b (or L^T x = b) with the LLt factor, x overwriting a copy of b."""
        x = np.array(b, dtype=np.float64 if self.value_type else np.float32, order="F", copy=True)
        nrhs = 1 if x.ndim == 1 else x.shape[1]
        fn = self.iface.solve_lower_triangular if permute else self.iface.solve_lower_triangular_dense
        _check(fn(self.handle, 1 if transpose else 0, x.ctypes.data, nrhs), "solve_lower_triangular")
        return x

    # ---- serialization (hmat-oss stream format) ---------------------------------
    def _write(self, slot):
        chunks = []

        @IOSTREAM
        def sink(buf, n, ud):
            if n:
                chunks.append(C.string_at(buf, n))
        getattr(self.iface, slot)(self.handle, sink, None)
        err = load().hmx_last_error().decode()
        if err:
            raise EngineError(f"{slot} failed: {err}")
        return b"".join(chunks)

    def write_struct(self):
        """The matrix structure (type, factorisation, cluster trees, block tree)."""
        return self._write("write_struct")

    def write_data(self):
        """The leaves, in preorder."""
        return self._write("write_data")

    def read_data(self, data):
        pos = [0]

        source = _source(data, pos)
        self.iface.read_data(self.handle, source, None)
        err = load().hmx_last_error().decode()
        if err:
            raise EngineError(f"read_data failed: {err}")
        return pos[0]

    @staticmethod
    def read_struct(data, value_type=HMAT_DOUBLE_PRECISION):
        """A matrix (unassembled) from a serialized structure; its cluster trees
        are those of the stream, owned by the matrix."""
        iface = interface(value_type)
        pos = [0]

        source = _source(data, pos)
        h = iface.read_struct(source, None)
        if not h:
            raise EngineError(f"read_struct failed: {load().hmx_last_error().decode()}")
        r, c = _vp(), _vp()
        iface.get_cluster_trees(h, C.byref(r), C.byref(c))
        m = object.__new__(HMatrix)
        m.tree = _BorrowedTree(r.value)
        m.iface, m.value_type, m.handle, m._kernel = iface, value_type, h, None
        m.lower_symmetric = bool(r.value == c.value) and value_type not in (HMAT_SIMPLE_COMPLEX,
                                                                            HMAT_DOUBLE_COMPLEX)
        m._consumed = pos[0]
        return m

    def logdet(self):
        r = C.c_double()
        _check(self.iface.logdet(self.handle, C.byref(r)), "logdet")
        return r.value

    # device-pointer entries (hmat numbering), e.g. torch tensors' data_ptr()
    def gemv_device(self, x_ptr, y_ptr, nrhs, alpha=1.0, beta=0.0, precision=0):
        _check(load().hmx_gemv_device(self.handle, precision, alpha, x_ptr, beta, y_ptr, nrhs),
               "hmx_gemv_device")

    def gemv_device_noalloc(self, x_ptr, y_ptr, nrhs, alpha=1.0, beta=0.0, precision=0):
        """hmx_gemv_device with the temporaries in the library's fixed device
        scratch (allocates nothing; see include/hmat/hmat_mi355x.h)."""
        _check(load().hmx_gemv_device_noalloc(self.handle, precision, alpha, x_ptr, beta, y_ptr,
                                              nrhs), "hmx_gemv_device_noalloc")

    def solve_device(self, b_ptr, nrhs, precision=0):
        _check(load().hmx_solve_device(self.handle, precision, b_ptr, nrhs), "hmx_solve_device")

    def solve_refined_device(self, b_ptr, nrhs, tol=1e-10, max_iter=30):
        r = C.c_double()
        it = load().hmx_solve_refined_device(self.handle, b_ptr, nrhs, tol, max_iter, C.byref(r))
        if it < 0:
            raise EngineError(f"refined solve failed: {load().hmx_last_error().decode()}")
        return it, r.value

    def autotune_solve(self, nrhs=1):
        """hmx_autotune_solve: (rows of the largest inverse blocks kept, ms per solve)."""
        r, ms = C.c_int(0), C.c_double(0.0)
        _check(load().hmx_autotune_solve(self.handle, nrhs, C.byref(r), C.byref(ms)),
               "hmx_autotune_solve")
        return r.value, ms.value

    def set_solve_block_rows(self, rows):
        _check(load().hmx_set_solve_block_rows(self.handle, rows), "hmx_set_solve_block_rows")

    def set_large_block_rows(self, rows):
        """Complex blocks of at least this many rows and columns are
        compressed on the whole device (randomized SVD) at the next assembly."""
        _check(load().hmx_set_large_block_rows(self.handle, rows), "hmx_set_large_block_rows")

    def get_block(self, row_offset, row_size, col_offset, col_size, renumber_rows=False):
        """get_block: the dense block of rows [row_offset, +row_size) x cols
        [col_offset, +col_size) in hmat numbering (rows in user order with
        renumber_rows, for a full-height block)."""
        out = np.zeros((col_size, row_size), dtype=self.dtype)   # column-major storage
        ctx = hmat_get_values_context_t()
        ctx.matrix, ctx.values = self.handle, out.ctypes.data
        ctx.row_offset, ctx.row_size, ctx.col_offset, ctx.col_size = row_offset, row_size, col_offset, col_size
        ctx.renumber_rows = 1 if renumber_rows else 0
        _check(self.iface.get_block(C.byref(ctx)), "get_block")
        return out.T

    def get_values(self, rows, cols):
        """get_values: the entries at user indices rows x cols (0-based here;
        the C slot takes them 1-based)."""
        r = np.ascontiguousarray(np.asarray(rows, dtype=np.int32) + 1)
        c = np.ascontiguousarray(np.asarray(cols, dtype=np.int32) + 1)
        out = np.zeros((len(c), len(r)), dtype=self.dtype)
        ctx = hmat_get_values_context_t()
        ctx.matrix, ctx.values = self.handle, out.ctypes.data
        ctx.row_size, ctx.col_size = len(r), len(c)
        ctx.row_indices = r.ctypes.data_as(C.POINTER(C.c_int))
        ctx.col_indices = c.ctypes.data_as(C.POINTER(C.c_int))
        _check(self.iface.get_values(C.byref(ctx)), "get_values")
        return out.T

    def set_stream(self, stream_ptr):
        """Launch on a caller's HIP stream (int handle, e.g.
        torch.cuda.current_stream().cuda_stream); None restores the engine's."""
        _check(load().hmx_set_stream(self.handle, stream_ptr), "hmx_set_stream")

    def stats(self):
        out = (C.c_double * 8)()
        load().hmx_stats(self.handle, out)
        keys = ["t_assemble", "t_factor", "t_solve", "tasks", "waves", "launches", "segments",
                "temp_bytes"]
        return dict(zip(keys, list(out)))

    def close(self):
        if getattr(self, "handle", None):
            self.iface.destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def run_batch(problems, algo=FACT_LLT, mixed=False, tol=1e-10, threads=8):
    """Assemble, factorise and solve independent systems through hmx_run_batch.

    problems: list of dicts with 'tree' (ClusterTree), 'b' (N x nrhs, user
    numbering) and the assemble() keywords ('kind', 'length', 'nugget', 'eps',
    optionally 'lowrank_eps').  Returns (solutions, seconds per system,
    status per system); the H-matrices are destroyed afterwards."""
    lib = load()
    count = len(problems)
    mats, ctxs, comps, bs = [], [], [], []
    try:
        for p in problems:
            h = HMatrix(p["tree"], lowrank_eps=p.get("lowrank_eps", p.get("eps", 1e-4)))
            ctx, comp = h.assemble_context(kind=p.get("kind", KERNEL_EXPONENTIAL),
                                           length=p["length"], nugget=p["nugget"],
                                           eps=p.get("eps", 1e-4))
            b = np.array(p["b"], dtype=np.float64, order="F", copy=True)
            mats.append(h)
            ctxs.append(ctx)
            comps.append(comp)
            bs.append(b)
        handles = (_vp * count)(*[m.handle for m in mats])
        cptrs = (_vp * count)(*[C.cast(C.pointer(c), _vp).value for c in ctxs])
        bptrs = (_vp * count)(*[b.ctypes.data for b in bs])
        nrhs = (C.c_int * count)(*[1 if b.ndim == 1 else b.shape[1] for b in bs])
        secs = (C.c_double * count)()
        status = (C.c_int * count)()
        _check(lib.hmx_run_batch(handles, cptrs, count, algo, 1 if mixed else 0, float(tol), bptrs,
                                 nrhs, int(threads), secs, status), "hmx_run_batch")
        return bs, list(secs), list(status)
    finally:
        for c in comps:
            lib.hmat_delete_compression(c)
        for m in mats:
            m.close()
