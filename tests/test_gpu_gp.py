"""GPU test of the Gaussian-process likelihood and its gradient (hmat_amd.gp):
N=8000 observations on the cylinder of radius 1 and height 4, exponential
covariance with l=0.5 and nugget 1e-2, y drawn from the GP (dense Cholesky,
seeded).  The forward and backward passes (FP32 H-LLt factor, FP64
refinement, log det from the factor's diagonal, derivative kernel assembled
by ACA+, 64 Rademacher probes) are compared with a dense FP64 reference: the
exact likelihood and gradient, the same trace estimate with the same probes,
and central finite differences (step 1e-5) of the exact likelihood."""
import math
import sys

import numpy as np
import pytest

from oracle import hmat_oracle as ho

hmat_amd = pytest.importorskip("hmat_amd")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _dense_kernel(P, length, nugget, deriv=False):
    import torch
    n = P.shape[0]
    K = torch.empty((n, n), dtype=torch.float64, device=P.device)
    for s in range(0, n, 1000):
        d = P[s:s + 1000, None, :] - P[None, :, :]
        r = torch.sqrt((d * d).sum(dim=2))
        e = torch.exp(-r / length)
        K[s:s + 1000] = r / (length * length) * e if deriv else e
    if nugget:
        K += nugget * torch.eye(n, dtype=torch.float64, device=P.device)
    return K


def _dense_nll(P, y, length, nugget):
    import torch
    K = _dense_kernel(P, length, nugget)
    L = torch.linalg.cholesky(K)
    a = torch.cholesky_solve(y[:, None], L)[:, 0]
    return 0.5 * float(y @ a) + float(torch.log(torch.diagonal(L)).sum()) + \
        0.5 * y.shape[0] * math.log(2 * math.pi)


def _to_user(h, tree, dev):
    """Dense matrix of an H-matrix in user numbering, on the device."""
    import torch
    H = torch.tensor(h.to_dense(), dtype=torch.float64, device=dev)
    idx = torch.as_tensor(tree.indices, device=dev)
    out = torch.empty_like(H)
    out[idx[:, None], idx[None, :]] = H
    return out


def _probe_std(A, probes):
    """Standard deviation of the mean of z^T A z over Rademacher probes:
    var(z^T A z) = 2 sum_{i != j} S_ij^2 with S the symmetric part of A."""
    import torch
    S = 0.5 * (A + A.T)
    off = float((S * S).sum()) - float((torch.diagonal(S) ** 2).sum())
    return math.sqrt(2.0 * off / probes)


@pytest.mark.gpu
def test_gp_likelihood_gradient_n8000():
    import torch
    from hmat_amd import gp
    n, length, nugget, probes = 8000, 0.5, 1e-2, 64
    say(f"GP likelihood and gradient N={n}")
    pts = ho.cylinder_points(n, 1.0, 4.0, 44)
    P = torch.tensor(pts, dtype=torch.float64, device="cuda")
    K = _dense_kernel(P, length, nugget)
    Lc = torch.linalg.cholesky(K)
    g = torch.Generator(device="cpu").manual_seed(7)
    y = (Lc @ torch.randn(n, dtype=torch.float64, generator=g).to("cuda")).contiguous()
    # engine: forward + backward
    tree = hmat_amd.ClusterTree(pts, 128)
    lt = torch.tensor(length, dtype=torch.float64, requires_grad=True)
    nt = torch.tensor(nugget, dtype=torch.float64, requires_grad=True)
    yt = y.clone().requires_grad_(True)
    nll = gp.gp_nll(lt, nt, yt, tree, probes=probes, seed=3)
    nll.backward()
    g_l, g_n, g_y = float(lt.grad), float(nt.grad), yt.grad.detach()
    Z = torch.tensor(gp.rademacher(n, probes, 3), dtype=torch.float64, device="cuda")

    def reference(Km, Dm):
        """Likelihood, exact gradient and the same-probe gradient, dense FP64."""
        L = torch.linalg.cholesky(Km)
        a = torch.cholesky_solve(y[:, None], L)[:, 0]
        f = 0.5 * float(y @ a) + float(torch.log(torch.diagonal(L)).sum()) + \
            0.5 * n * math.log(2 * math.pi)
        Ki = torch.cholesky_inverse(L)
        Da = Dm @ a
        gl = 0.5 * (float((Ki * Dm.T).sum()) - float(a @ Da))
        gn = 0.5 * (float(torch.trace(Ki)) - float(a @ a))
        S = Ki @ Z
        glp = 0.5 * (float((S * (Dm @ Z)).sum()) / probes - float(a @ Da))
        gnp = 0.5 * (float((S * Z).sum()) / probes - float(a @ a))
        return f, a, gl, gn, glp, gnp, Ki

    D = _dense_kernel(P, length, 0.0, deriv=True)
    f_k, a_k, gl_k, gn_k, glp_k, gnp_k, Ki = reference(K, D)
    # the same computation on the engine's own compressed matrices (assembled
    # with the same parameters: the assembly is deterministic)
    hK = hmat_amd.HMatrix(tree)
    hK.assemble(length=length, nugget=nugget, eps=1e-4)
    hD = hmat_amd.HMatrix(tree)
    hD.assemble(kind=hmat_amd.KERNEL_EXPONENTIAL_DLENGTH, length=length, nugget=0.0, eps=1e-4)
    KH, DH = _to_user(hK, tree, "cuda"), _to_user(hD, tree, "cuda")
    hK.close()
    hD.close()
    f_h, a_h, _, _, glp_h, gnp_h, _ = reference(KH, DH)
    # central finite differences of the exact likelihood (step 1e-5)
    h = 1e-5
    gl_fd = (_dense_nll(P, y, length + h, nugget) - _dense_nll(P, y, length - h, nugget)) / (2 * h)
    gn_fd = (_dense_nll(P, y, length, nugget + h) - _dense_nll(P, y, length, nugget - h)) / (2 * h)
    say(f"nll engine {float(nll):.8f} dense H {f_h:.8f} dense K {f_k:.8f}")
    say(f"dNLL/dl engine {g_l:.6f} same-probe dense H {glp_h:.6f} dense K {glp_k:.6f} "
        f"exact {gl_k:.6f} FD {gl_fd:.6f}")
    say(f"dNLL/dnugget engine {g_n:.6f} same-probe dense H {gnp_h:.6f} dense K {gnp_k:.6f} "
        f"exact {gn_k:.6f} FD {gn_fd:.6f}")
    ry_h = float(torch.linalg.norm(g_y - a_h) / torch.linalg.norm(a_h))
    ry_k = float(torch.linalg.norm(g_y - a_k) / torch.linalg.norm(a_k))
    say(f"dNLL/dy rel diff vs dense H {ry_h:.2e}, vs dense K {ry_k:.2e}")
    # 1. the dense reference is pinned by finite differences of the exact
    # likelihood; for the nugget the step's truncation error (h^2/6 times the
    # third derivative, about tr(K^-3) ~ N / nugget^3) is a few 1e-5 relative
    assert abs(gl_fd - gl_k) <= 1e-5 * abs(gl_k)
    assert abs(gn_fd - gn_k) <= 1e-4 * abs(gn_k)
    # 2. the engine equals the dense FP64 computation on its own H-matrices
    # with the same probes (FP32 factor + refinement to 1e-10; log det from the
    # FP32 factor)
    assert abs(float(nll) - f_h) <= 1e-6 * abs(f_h)
    assert ry_h <= 1e-8
    assert abs(g_l - glp_h) <= 1e-6 * abs(glp_h)
    assert abs(g_n - gnp_h) <= 1e-6 * abs(gnp_h)
    # 3. against the exact kernel matrix the compression (eps 1e-4) moves the
    # solution by about eps ||K^-1|| ||K|| (measured 1e-3) and the likelihood by 4e-5
    assert ry_k <= 1e-2
    assert abs(float(nll) - f_k) <= 1e-4 * abs(f_k)
    # 4. the 64-probe trace estimates lie within 4 standard deviations of the
    # exact traces (the deviation computed from the exact matrices)
    sd_n = 0.5 * _probe_std(Ki, probes)
    sd_l = 0.5 * _probe_std(Ki @ D, probes)
    say(f"probe std: dl {sd_l:.4f}, dnugget {sd_n:.4f}")
    assert abs(glp_k - gl_k) <= 4 * sd_l
    assert abs(gnp_k - gn_k) <= 4 * sd_n


@pytest.mark.gpu
def test_derivative_kernel_assembly():
    """The derivative kernel (r / l^2) exp(-r / l) assembled by ACA+ against
    its exact dense matrix."""
    say("derivative kernel assembly n=3000")
    n, length = 3000, 0.5
    pts = ho.cylinder_points(n, 1.0, 1.2, 9)
    t = hmat_amd.ClusterTree(pts, 128)
    h = hmat_amd.HMatrix(t)
    h.assemble(kind=hmat_amd.KERNEL_EXPONENTIAL_DLENGTH, length=length, nugget=0.0, eps=1e-4)
    idx = t.indices
    q = pts[idx]
    d = q[:, None, :] - q[None, :, :]
    r = np.sqrt((d * d).sum(axis=2))
    exact = r / length ** 2 * np.exp(-r / length)
    err = np.linalg.norm(h.to_dense() - exact) / np.linalg.norm(exact)
    say("derivative kernel rel err", err)
    assert err <= 1e-4
    h.close()
