"""ctypes mirror of the engine's task record (hmx::FTask, ftask.hpp) and
helpers to run one batched launch of task kernels through the test hook
hmx_run_tasks on device operands held in torch tensors (column-major: a
matrix of m rows and n columns is a contiguous (n, m) tensor, ld = m)."""
import ctypes as C

import numpy as np

T_GEMM, T_POTRF, T_TRSM, T_FADD, T_FADD_DENSE, T_RKRK, T_ACAFULL, T_ZERO, T_AXPY, T_GETRF, \
    T_SUMK, T_LDLT, T_DSCALE, T_HODLR = range(14)
TRSM_TRANS_A, TRSM_UNIT, TRSM_PIV = 4, 8, 16


class Dim(C.Structure):
    _fields_ = [("s", C.c_int), ("d", C.c_void_p)]


class RkRef(C.Structure):
    _fields_ = [("a", C.c_void_p), ("b", C.c_void_p), ("la", C.c_int), ("lb", C.c_int),
                ("m", C.c_int), ("n", C.c_int), ("cap", C.c_int), ("rank", C.c_void_p)]


class Part(C.Structure):
    _fields_ = [("rk", RkRef), ("roff", C.c_int), ("coff", C.c_int), ("alpha", C.c_double)]


class FTask(C.Structure):
    _fields_ = [("type", C.c_int), ("mode", C.c_int), ("ta", C.c_int), ("tb", C.c_int),
                ("m", Dim), ("n", Dim), ("p", Dim), ("alpha", C.c_double), ("eps", C.c_double),
                ("A", C.c_void_p), ("B", C.c_void_p), ("C", C.c_void_p),
                ("lda", C.c_int), ("ldb", C.c_int), ("ldc", C.c_int), ("nparts", C.c_int),
                ("pad", C.c_int), ("tgt", RkRef), ("parts", Part * 4), ("ws", C.c_int64)]


def lib():
    import hmat_amd
    L = hmat_amd.load()
    L.hmx_ftask_size.restype = C.c_int
    L.hmx_run_tasks.restype = C.c_int
    L.hmx_run_tasks.argtypes = [C.c_int, C.c_int, C.c_void_p, C.c_int, C.POINTER(C.c_int)]
    assert L.hmx_ftask_size() == C.sizeof(FTask), "task record layout differs from ftask.hpp"
    return L


def run(tasks, typ, fp32=False):
    """One launch of the given FTask records; returns the two flags."""
    L = lib()
    arr = (FTask * len(tasks))(*tasks)
    flags = (C.c_int * 2)()
    rc = L.hmx_run_tasks(1 if fp32 else 0, typ, C.cast(arr, C.c_void_p), len(tasks), flags)
    if rc != 0:
        import hmat_amd
        raise hmat_amd.EngineError(L.hmx_last_error().decode())
    return list(flags)


def dev(a, fp32=False):
    """A device copy of a host matrix (column-major, ld = rows)."""
    import torch
    a = np.asarray(a, dtype=np.float64)
    if a.ndim == 1:
        a = a[:, None]
    return torch.tensor(np.ascontiguousarray(a.T), dtype=torch.float32 if fp32 else torch.float64,
                        device="cuda")


def host(t, rows=None, cols=None):
    """The host matrix of a device copy (optionally its leading block)."""
    a = t.double().cpu().numpy().T
    return a[:rows, :cols] if rows is not None else a


def ints(v):
    import torch
    return torch.tensor(np.asarray(v, dtype=np.int32), device="cuda")


def rkref(a_t, b_t, m, n, cap, rank_t):
    r = RkRef()
    r.a, r.b = a_t.data_ptr(), b_t.data_ptr()
    r.la, r.lb, r.m, r.n, r.cap = m, n, m, n, cap
    r.rank = rank_t.data_ptr()
    return r


def find_k(sig, eps):
    """findK scalar_array.cpp:98-116: number of sigma_i > eps sigma_0."""
    if len(sig) == 0 or sig[0] == 0:
        return 0
    return int(np.sum(sig > eps * sig[0]))
