"""GPU test of many independent systems run together (hmx_run_batch): Gaussian
covariance exp(-r^2/l^2) + 1e-2 I on seeded cylinders of radius U[0.5, 2] and
height 4 r, l ~ U[0.1, 1], leaves 16 to 256, ACA+ eps 1e-4, one N(0,1)
right-hand side each.  FP64 systems are compared with the oracle's H-LLt
solve on the same inputs and bitwise with the same system run alone; FP32 +
refinement systems reach 1e-10 against the dense H-matrix."""
import sys

import numpy as np
import pytest

from oracle import hmat_oracle as ho

hmat_amd = pytest.importorskip("hmat_amd")

EPS = 1e-4
SIZES = [(800, 16), (1200, 32), (2000, 64), (2500, 128), (1500, 256), (1000, 64)]


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _problems(seed=33):
    rng = np.random.default_rng(seed)
    out = []
    for n, leaf in SIZES:
        r, l = rng.uniform(0.5, 2.0), rng.uniform(0.1, 1.0)
        pts = ho.cylinder_points(n, r, 4 * r, 100 + n)
        b = np.random.default_rng(n).standard_normal((n, 1))
        out.append(dict(pts=pts, leaf=leaf, tree=hmat_amd.ClusterTree(pts, leaf), b=b,
                        kind=hmat_amd.KERNEL_GAUSSIAN, length=l, nugget=1e-2, eps=EPS))
    return out


@pytest.mark.gpu
def test_batch_fp64_matches_oracle():
    say("batch of 6 systems, FP64, 4 threads")
    probs = _problems()
    xs, secs, status = hmat_amd.run_batch(probs, mixed=False, threads=4)
    say("seconds", [round(s, 3) for s in secs], "status", status)
    assert status == [0] * len(probs)
    rows = []
    for p, x in zip(probs, xs):
        n = len(p["pts"])
        root, idx = ho.build_cluster_tree(p["pts"], p["leaf"])
        assert np.array_equal(idx, p["tree"].indices)
        bt = ho.build_block_tree(root, root, 2.0, True)
        k = ho.GaussKernel(p["pts"], p["length"], p["nugget"])
        ho.assemble(bt, idx, k)
        Ho = ho.to_dense(bt, n, n)
        ho.llt(bt, EPS)
        xo = p["b"][idx].copy()
        ho.solve_llt(bt, xo)
        xou = np.empty_like(xo)
        xou[idx] = xo
        # the oracle's own distance from the exact solution of its H-matrix
        xe = np.linalg.solve(Ho, p["b"][idx])
        dev_o = np.linalg.norm(xo - xe) / np.linalg.norm(xe)
        rel = np.linalg.norm(x - xou) / np.linalg.norm(xou)
        say(f"n={n} leaf={p['leaf']} l={p['length']:.3f}: solution rel diff vs oracle {rel:.2e}, "
            f"oracle vs exact H solve {dev_o:.2e}, cond(H) {np.linalg.cond(Ho):.2e}")
        rows.append((rel, dev_o))
    # within eps of the oracle's solution; where the factorisation itself moves
    # the solution by more than 10 eps (an ill-conditioned Gaussian system: the
    # oracle's own solution 2.9e-3 from the exact H solve at N=800, leaf 16),
    # within a tenth of the oracle's own distance (measured 1.1e-4 there)
    for rel, dev_o in rows:
        assert rel <= max(EPS, 0.1 * dev_o)
    # a system of the batch equals the same system run alone, bit for bit
    p = probs[2]
    h = hmat_amd.HMatrix(p["tree"])
    h.assemble(kind=p["kind"], length=p["length"], nugget=p["nugget"], eps=EPS)
    h.factorize(hmat_amd.FACT_LLT)
    x1 = h.solve(p["b"])
    h.close()
    assert np.array_equal(x1, xs[2])


@pytest.mark.gpu
def test_batch_mixed_precision_reaches_1e10():
    say("batch of 6 systems, FP32 + refinement, 3 threads")
    probs = _problems(seed=34)
    xs, secs, status = hmat_amd.run_batch(probs, mixed=True, tol=1e-10, threads=3)
    say("seconds", [round(s, 3) for s in secs], "status", status)
    assert status == [0] * len(probs)
    for p, x in zip(probs, xs):
        h = hmat_amd.HMatrix(p["tree"])
        h.assemble(kind=p["kind"], length=p["length"], nugget=p["nugget"], eps=EPS)
        H = h.to_dense()
        h.close()
        idx = p["tree"].indices
        xh, bh = x[idx], p["b"][idx]
        res = np.linalg.norm(H @ xh - bh) / np.linalg.norm(bh)
        xe = np.linalg.solve(H, bh)
        rel = np.linalg.norm(xh - xe) / np.linalg.norm(xe)
        say(f"n={len(idx)} leaf={p['leaf']}: residual vs dense H {res:.2e}, solution {rel:.2e}")
        assert res <= 1e-10
        assert rel <= 1e-6
