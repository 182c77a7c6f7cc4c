"""GPU test of the allocation-free H-gemv entry point (hmx_gemv_device_noalloc):
its result must equal hmx_gemv_device's bit for bit (same kernels, same
summation order, temporaries in the library's fixed device array instead of
allocated buffers), on two calls in a row, the second with a smaller input,
and a product whose temporaries exceed the array is refused with an error."""
import sys

import numpy as np
import pytest

from oracle import hmat_oracle as ho

hmat_amd = pytest.importorskip("hmat_amd")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _matrix(n, height, seed, leaf=128):
    pts = ho.cylinder_points(n, 1.0, height, seed)
    t = hmat_amd.ClusterTree(pts, leaf)
    h = hmat_amd.HMatrix(t)
    h.assemble(length=0.5, nugget=1e-2, eps=1e-4)
    return t, h


def _both(h, nrhs, seed):
    import torch
    rng = np.random.default_rng(seed)
    x = rng.standard_normal((h.n, nrhs))
    y0 = rng.standard_normal((h.n, nrhs))
    xd = torch.tensor(x.T.copy(), dtype=torch.float64, device="cuda").contiguous()
    ya = torch.tensor(y0.T.copy(), dtype=torch.float64, device="cuda").contiguous()
    yb = ya.clone()
    h.gemv_device(xd.data_ptr(), ya.data_ptr(), nrhs, alpha=0.7, beta=-0.3)
    h.gemv_device_noalloc(xd.data_ptr(), yb.data_ptr(), nrhs, alpha=0.7, beta=-0.3)
    torch.cuda.synchronize()
    return ya.cpu().numpy().T, yb.cpu().numpy().T, x, y0


@pytest.mark.gpu
def test_noalloc_gemv_equals_allocating_gemv():
    say("noalloc gemv: n=6000 nrhs=3, then n=2500 nrhs=1")
    t1, h1 = _matrix(6000, 2.4, 7)
    ya, yb, x, y0 = _both(h1, 3, 1)
    say("call 1 max |diff|", float(np.max(np.abs(ya - yb))))
    assert np.array_equal(ya, yb)
    # and it is the product of the stored leaves
    D = h1.to_dense()
    ref = 0.7 * (D @ x) - 0.3 * y0
    assert np.linalg.norm(yb - ref) / np.linalg.norm(ref) < 1e-13
    t2, h2 = _matrix(2500, 1.0, 8)
    ya, yb, _, _ = _both(h2, 1, 2)
    say("call 2 max |diff|", float(np.max(np.abs(ya - yb))))
    assert np.array_equal(ya, yb)
    # the first matrix again after the smaller call: the array's leftovers
    # from the second call must not leak into the result
    ya, yb, _, _ = _both(h1, 3, 1)
    assert np.array_equal(ya, yb)
    h1.close()
    h2.close()


@pytest.mark.gpu
def test_noalloc_gemv_rejects_oversized_product():
    import torch
    say("noalloc gemv: oversized nrhs refused")
    t, h = _matrix(2500, 1.0, 8)
    nrhs = 4096   # temporaries > 64 MiB for this matrix (N * nrhs doubles alone is 78 MiB)
    x = torch.zeros((nrhs, h.n), dtype=torch.float64, device="cuda")
    y = torch.zeros_like(x)
    with pytest.raises(hmat_amd.EngineError, match="scratch"):
        h.gemv_device_noalloc(x.data_ptr(), y.data_ptr(), nrhs)
    h.close()
