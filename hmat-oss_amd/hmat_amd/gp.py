"""Gaussian-process negative log-likelihood with gradients through the C ABI.

    NLL(l, nugget, y) = 1/2 y^T K^{-1} y + 1/2 log det K + N/2 log(2 pi),
    K = k_l(x_i, x_j) + nugget * delta_ij

The forward pass assembles K as an H-matrix (ACA+), factorises an FP32 copy
(H-LLt on the device), solves alpha = K^{-1} y by refinement against the FP64
H-matrix, and takes log det K = 2 sum log diag(L) from the factor
(HMatrix::logdet h_matrix.cpp:2925-2936).  The backward pass reuses that
factor (it is kept, not differentiated through):

    dNLL/dy      = alpha
    dNLL/dnugget = 1/2 (tr(K^{-1}) - alpha^T alpha)
    dNLL/dl      = 1/2 (tr(K^{-1} dK/dl) - alpha^T (dK/dl) alpha)

with dK/dl assembled as a second H-matrix by ACA+ (the derivative kernel
evaluated on the device) and both traces estimated from P seeded Rademacher
probes z_j: S = K^{-1} Z by the same refined solve (K symmetric, so the
adjoint solve is the solve), tr(K^{-1}) ~ mean_j z_j . s_j and
tr(K^{-1} dK/dl) ~ mean_j s_j . (dK/dl z_j).
"""
from __future__ import annotations

import math

import numpy as np

from . import (FACT_LLT, HMatrix, KERNEL_EXPONENTIAL, KERNEL_EXPONENTIAL_DLENGTH, KERNEL_GAUSSIAN,
               KERNEL_GAUSSIAN_DLENGTH)

_DERIVATIVE = {KERNEL_EXPONENTIAL: KERNEL_EXPONENTIAL_DLENGTH,
               KERNEL_GAUSSIAN: KERNEL_GAUSSIAN_DLENGTH}


def rademacher(n, count, seed):
    """count seeded Rademacher vectors (n x count, entries +-1, float64)."""
    rng = np.random.default_rng(seed)
    return np.where(rng.integers(0, 2, size=(n, count)) == 0, -1.0, 1.0)


class _State:
    """The factorised H-matrix and the solution kept between passes."""

    def __init__(self, tree, kind, length, nugget, eps, tol):
        self.tree, self.kind, self.length, self.nugget = tree, kind, length, nugget
        self.eps, self.tol = eps, tol
        self.K = HMatrix(tree, lowrank_eps=eps)
        self.K.assemble(kind=kind, length=length, nugget=nugget, eps=eps)
        self.K.factorize(FACT_LLT, mixed=True)
        self.idx = tree.indices

    def solve(self, b_user, tol=None):
        """K^{-1} b (user numbering, N x nrhs float64 torch tensor on the device)."""
        import torch
        bh = b_user[torch.as_tensor(self.idx, device=b_user.device)]
        bd = bh.T.contiguous()
        it, rel = self.K.solve_refined_device(bd.data_ptr(), bd.shape[0], tol=tol or self.tol,
                                              max_iter=200)
        torch.cuda.synchronize()
        if rel > (tol or self.tol):
            raise RuntimeError(f"refined solve stopped at {rel:.2e} after {it} iterations")
        out = torch.empty_like(b_user)
        out[torch.as_tensor(self.idx, device=b_user.device)] = bd.T
        return out

    def close(self):
        self.K.close()


class GPNegLogLikelihood:
    """torch.autograd.Function wrapper (built lazily so that importing this
    module does not need torch)."""

    _fn = None

    @classmethod
    def function(cls):
        if cls._fn is not None:
            return cls._fn
        import torch

        class Fn(torch.autograd.Function):
            @staticmethod
            def forward(ctx, length, nugget, y, tree, kind, eps, tol, probes, seed):
                st = _State(tree, kind, float(length), float(nugget), eps, tol)
                yv = y.reshape(-1, 1).to(torch.float64)
                alpha = st.solve(yv)
                logdet = 2.0 * st.K.logdet()
                n = yv.shape[0]
                nll = 0.5 * float((yv * alpha).sum()) + 0.5 * logdet + 0.5 * n * math.log(2 * math.pi)
                ctx.st, ctx.alpha, ctx.probes, ctx.seed = st, alpha, probes, seed
                ctx.yshape = y.shape
                return torch.tensor(nll, dtype=torch.float64, device=y.device)

            @staticmethod
            def backward(ctx, g):
                st, alpha = ctx.st, ctx.alpha
                try:
                    dev = alpha.device
                    n = alpha.shape[0]
                    Z = torch.tensor(rademacher(n, ctx.probes, ctx.seed), dtype=torch.float64, device=dev)
                    S = st.solve(Z, tol=1e-8)
                    D = HMatrix(st.tree, lowrank_eps=st.eps)
                    D.assemble(kind=_DERIVATIVE[st.kind], length=st.length, nugget=0.0, eps=st.eps)
                    # dK/dl applied to [alpha, Z] in one H-gemv (hmat numbering)
                    X = torch.cat([alpha, Z], dim=1)
                    ih = torch.as_tensor(st.idx, device=dev)
                    xd = X[ih].T.contiguous()
                    yd = torch.zeros_like(xd)
                    D.gemv_device(xd.data_ptr(), yd.data_ptr(), xd.shape[0])
                    torch.cuda.synchronize()
                    DX = torch.empty_like(X)
                    DX[ih] = yd.T
                    D.close()
                    d_alpha = float((alpha[:, 0] * DX[:, 0]).sum())
                    tr_l = float((S * DX[:, 1:]).sum()) / ctx.probes
                    tr_n = float((Z * S).sum()) / ctx.probes
                    a2 = float((alpha * alpha).sum())
                    g_l = 0.5 * (tr_l - d_alpha)
                    g_n = 0.5 * (tr_n - a2)
                    gy = alpha.reshape(ctx.yshape)
                finally:
                    st.close()
                return (g * g_l, g * g_n, g * gy, None, None, None, None, None, None)

        cls._fn = Fn
        return Fn


def gp_nll(length, nugget, y, tree, kind=KERNEL_EXPONENTIAL, eps=1e-4, tol=1e-10, probes=64,
           seed=0):
    """NLL of y (float64 tensor on the device, user numbering) under the GP with
    the given kernel on tree's points; differentiable in length, nugget and y."""
    return GPNegLogLikelihood.function().apply(length, nugget, y, tree, kind, eps, tol, probes, seed)
