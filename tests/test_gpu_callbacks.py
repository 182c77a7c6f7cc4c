"""GPU test of assembly from the caller's host functions (assemble_generic with
simple_compute, or block_compute + prepare; assembly.cpp:75-292): every block
is evaluated on the host and compressed on the device.  The callbacks evaluate
the exponential covariance with numpy/math on the caller's side; the
assembled leaves are compared with the oracle's assembly of the same entries
(same partition, same ACA+ / SVD rules), and the H-LLt solve with the
oracle's."""
import ctypes as C
import math
import sys

import numpy as np
import pytest

from oracle import hmat_oracle as ho

hmat_amd = pytest.importorskip("hmat_amd")

EPS = 1e-4


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _compare_with_oracle(h, pts, leaf):
    root, idx = ho.build_cluster_tree(pts, leaf)
    bt = ho.build_block_tree(root, root, 2.0, True)
    k = ho.ExpKernel(pts, 0.5, 1e-2)
    ho.assemble(bt, idx, k)
    worst, mism, nrk = 0.0, 0, 0
    for ob, (m, n, r0, c0, kind, rank, a, b) in zip(ho.leaves(bt), h.leaves()):
        assert (ob.rows.offset, ob.m, ob.cols.offset, ob.n) == (r0, m, c0, n)
        if kind == 0:
            worst = max(worst, np.abs(a - ob.full).max() / max(np.abs(ob.full).max(), 1e-300))
        else:
            nrk += 1
            g = a @ b.T if rank > 0 else np.zeros((m, n))
            o = ob.rk.eval(m, n, np.float64)
            worst = max(worst, np.linalg.norm(g - o) / max(np.linalg.norm(o), 1e-300))
            mism += rank != ob.rk.rank
    return bt, idx, k, worst, mism, nrk


def _solve_vs_oracle(h, bt, idx, n, seed):
    h.factorize(hmat_amd.FACT_LLT)
    b = np.random.default_rng(seed).standard_normal((n, 1))
    x = h.solve(b)
    ho.llt(bt, EPS)
    xo = b[idx].copy()
    ho.solve_llt(bt, xo)
    xu = np.empty_like(xo)
    xu[idx] = xo
    return np.linalg.norm(x - xu) / np.linalg.norm(xu)


@pytest.mark.gpu
def test_simple_compute_callback():
    n, leaf = 1500, 64
    say(f"simple_compute callback n={n} leaf={leaf}")
    pts = ho.cylinder_points(n, 1.0, 0.6, 3)
    P = [tuple(p) for p in pts]

    @hmat_amd.INTERACTION_FUNC
    def interaction(ctx, i, j, out):
        xi, xj = P[i], P[j]
        dx, dy, dz = xi[0] - xj[0], xi[1] - xj[1], xi[2] - xj[2]
        v = math.exp(-math.sqrt(dx * dx + dy * dy + dz * dz) / 0.5) + (1e-2 if i == j else 0.0)
        C.cast(out, C.POINTER(C.c_double))[0] = v

    t = hmat_amd.ClusterTree(pts, leaf)
    h = hmat_amd.HMatrix(t)
    h.assemble_functions(simple=interaction, eps=EPS)
    bt, idx, k, worst, mism, nrk = _compare_with_oracle(h, pts, leaf)
    say(f"leaves vs oracle: worst {worst:.2e}, rank mismatches {mism} of {nrk}")
    assert mism == 0 and worst <= 1e-10
    rel = _solve_vs_oracle(h, bt, idx, n, 3)
    say(f"solution rel diff vs oracle {rel:.2e}")
    assert rel <= 1e-5
    h.close()


@pytest.mark.gpu
def test_block_compute_callback():
    n, leaf = 4000, 128
    say(f"block_compute + prepare callbacks n={n} leaf={leaf}")
    pts = ho.cylinder_points(n, 1.0, 1.6, 9)
    t = hmat_amd.ClusterTree(pts, leaf)
    user = t.indices
    blocks = {}
    counter = [0]

    @hmat_amd.PREPARE_FUNC
    def prepare(r0, rc, c0, cc, rows_h2c, rows_c2h, cols_h2c, cols_c2h, ctx, info):
        # remember the block's user indices; user_data is a key into `blocks`
        counter[0] += 1
        key = counter[0]
        blocks[key] = (np.array(rows_h2c[r0:r0 + rc]), np.array(cols_h2c[c0:c0 + cc]))
        info[0].user_data = key

    @hmat_amd.COMPUTE_FUNC
    def compute(data, r0, rc, c0, cc, out):
        ri, ci = blocks.pop(data)     # the engine asks for each whole block once
        ri, ci = ri[r0:r0 + rc], ci[c0:c0 + cc]
        d = pts[ri][:, None, :] - pts[ci][None, :, :]
        v = np.exp(-np.sqrt((d * d).sum(axis=2)) / 0.5) + 1e-2 * (ri[:, None] == ci[None, :])
        np.ctypeslib.as_array(C.cast(out, C.POINTER(C.c_double)), shape=(cc, rc))[:] = v.T

    h = hmat_amd.HMatrix(t)
    h.assemble_functions(block=compute, prepare=prepare, eps=EPS)
    assert not blocks   # every prepared block was computed once
    bt, idx, k, worst, mism, nrk = _compare_with_oracle(h, pts, leaf)
    assert np.array_equal(idx, user)
    say(f"leaves vs oracle: worst {worst:.2e}, rank mismatches {mism} of {nrk}")
    assert mism == 0 and worst <= 1e-10
    # the same matrix from the built-in device kernel
    g = hmat_amd.HMatrix(t)
    g.assemble(length=0.5, nugget=1e-2, eps=EPS)
    dd = np.linalg.norm(h.to_dense() - g.to_dense()) / np.linalg.norm(g.to_dense())
    say(f"host-function matrix vs device-kernel matrix {dd:.2e}")
    assert dd <= 1e-10
    g.close()
    rel = _solve_vs_oracle(h, bt, idx, n, 9)
    say(f"solution rel diff vs oracle {rel:.2e}")
    assert rel <= 1e-5
    h.close()
