"""The persistent-sweep H-LLt solve (one launch per triangular sweep) against
the recorded-task solve of the same factor, on uniform and uneven-depth
partitions, Rk panels spanning many 256-row chunks, one and several
right-hand sides (more than one 8-column block), FP64 and FP32 factors.

Skipped: the sweep solve did not complete on the device (the first solve at
N=4000 ran past the test's time limit, twice, the second time with every wait
of the kernel bounded by the wall clock); its cause is not found, so the path
is switched off in the engine (asking for it raises an error)."""
import os
import sys

import numpy as np
import pytest

from oracle import hmat_oracle as ho

hmat_amd = pytest.importorskip("hmat_amd")

EPS = 1e-4


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


SKIP = pytest.mark.skip(reason="persistent-sweep solve switched off: it did not complete on the "
                               "device (N=4000, first solve), cause not found")


def _solve(h, b, sweep):
    if sweep:
        os.environ["HMX_SOLVE_SWEEP"] = "1"
    try:
        return h.solve(b, permute=False)
    finally:
        os.environ.pop("HMX_SOLVE_SWEEP", None)


CASES = [(4000, 1.6, 9, 128), (2050, 1.0, 5, 128), (1030, 0.6, 5, 64), (6000, 8.0, 3, 64)]


@SKIP
@pytest.mark.gpu
@pytest.mark.parametrize("n,height,seed,leaf", CASES)
def test_sweep_solve_equals_task_solve_fp64(n, height, seed, leaf):
    say(f"sweep vs task solve fp64 n={n} leaf={leaf}")
    pts = ho.cylinder_points(n, 1.0, height, seed)
    t = hmat_amd.ClusterTree(pts, leaf)
    h = hmat_amd.HMatrix(t)
    h.assemble(length=0.5, nugget=1e-2, eps=EPS)
    h.factorize(hmat_amd.FACT_LLT)
    L = np.tril(h.to_dense(mirror=False))
    rng = np.random.default_rng(seed)
    for nrhs in (1, 3, 9):
        b = rng.standard_normal((n, nrhs))
        xs = _solve(h, b, True)
        xt = _solve(h, b, False)
        # exact triangular solves with the stored factor
        xe = np.linalg.solve(L.T, np.linalg.solve(L, b))
        d = np.linalg.norm(xs - xt) / np.linalg.norm(xt)
        e = np.linalg.norm(xs - xe) / np.linalg.norm(xe)
        say(f"nrhs={nrhs}: sweep vs tasks {d:.3e}, sweep vs dense solve of the factor {e:.3e}")
        assert d <= 1e-12
        assert e <= 1e-12
        # a second sweep solve is bitwise the first (fixed summation order)
        assert np.array_equal(_solve(h, b, True), xs)
    h.close()


@SKIP
@pytest.mark.gpu
@pytest.mark.parametrize("n,height,seed,leaf", [CASES[0], CASES[3]])
def test_sweep_solve_equals_task_solve_fp32(n, height, seed, leaf):
    say(f"sweep vs task solve fp32 n={n} leaf={leaf}")
    pts = ho.cylinder_points(n, 1.0, height, seed)
    t = hmat_amd.ClusterTree(pts, leaf)
    h = hmat_amd.HMatrix(t, value_type=hmat_amd.HMAT_SIMPLE_PRECISION)
    h.assemble(length=0.5, nugget=1e-2, eps=EPS)
    h.factorize(hmat_amd.FACT_LLT)
    L = np.tril(h.to_dense(precision=1, mirror=False))
    b = np.random.default_rng(seed).standard_normal((n, 8)).astype(np.float32)
    xs = _solve(h, b, True).astype(np.float64)
    xt = _solve(h, b, False).astype(np.float64)
    xe = np.linalg.solve(L.T, np.linalg.solve(L, b.astype(np.float64)))
    d = np.linalg.norm(xs - xt) / np.linalg.norm(xt)
    e = np.linalg.norm(xs - xe) / np.linalg.norm(xe)
    say(f"sweep vs tasks {d:.3e}, sweep vs dense solve of the factor {e:.3e}")
    assert d <= 1e-5 and e <= 1e-5
    h.close()
