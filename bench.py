"""Benchmark of the MI355X H-matrix engine on the precision workload:
N=200,000 seeded points on a cylinder (radius 1, height 16), exponential
covariance exp(-r/0.5) + 1e-2 I, median clustering (leaf 128), eta 2, ACA+
eps 1e-4.

Timed (inputs resident in HBM):
  * H-LLt factor + solve of 8 N(0,1) right-hand sides: FP32 factor with FP64
    iterative refinement (residual <= 1e-10 against the FP64 H-matrix), and the
    plain FP64 factor + solve, when the factorisation is available;
  * H-gemv on the FP64 H-matrix (roofline line, HBM bound).

Prints ONE JSON line on rank 0.  Only one GPU is used: --gpus > 1 exits.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "hmat-oss_amd"))

import numpy as np  # noqa: E402

METRIC = "H-LLt factor+solve s @N=200k, resid<=1e-10: FP32+IR vs FP64; H-gemv GB/s"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def gemv_bytes(h, info, nrhs, itemsize=8):
    """Algorithmic bytes of one H-gemv: stored leaf data once plus, per leaf,
    x[cols] read once and y[rows] read and written."""
    vec = 0
    for m, n, r0, c0, kind in h.leaf_shapes():
        vec += n + 2 * m
    return itemsize * info["compressed_size"] + itemsize * nrhs * vec


def cpu_gemv_baseline(h, pts, leaf, nrhs, reps=2):
    """Oracle H-gemv (numpy restatement) on the engine's own leaves."""
    from oracle import hmat_oracle as ho
    from threadpoolctl import threadpool_limits
    t0 = time.time()
    root, idx = ho.build_cluster_tree(pts, leaf)
    bt = ho.build_block_tree(root, root, 2.0, True)
    lv = ho.leaves(bt)
    for ob, (m, n, r0, c0, kind, rank, a, b) in zip(lv, h.leaves()):
        if kind == 0:
            ob.full = a
        else:
            ob.rk = ho.Rk(a, b) if rank > 0 else ho.Rk(None, None)
    log(f"cpu baseline: oracle tree filled in {time.time() - t0:.1f}s")
    x = np.random.default_rng(0).standard_normal((h.n, nrhs))
    y = np.zeros_like(x)
    with threadpool_limits(1):
        ts = []
        for _ in range(reps):
            t = time.time()
            ho.gemv(bt, "N", 1.0, x, 0.0, y)
            ts.append(time.time() - t)
    return min(ts)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=200000)
    ap.add_argument("--height", type=float, default=16.0)
    ap.add_argument("--nrhs", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              "(the engine runs on a single MI355X)", file=sys.stderr)
        sys.exit(2)

    import torch
    import hmat_amd
    from oracle.hmat_oracle import cylinder_points  # input generator only

    assert torch.cuda.is_available(), "bench.py needs a HIP device"
    leaf = 128
    t0 = time.time()
    pts = cylinder_points(args.n, 1.0, args.height, 2)
    log(f"points {time.time() - t0:.1f}s")
    t0 = time.time()
    tree = hmat_amd.ClusterTree(pts, leaf)
    h = hmat_amd.HMatrix(tree, eta=2.0, lower_symmetric=True)
    log(f"partition {time.time() - t0:.1f}s")
    t0 = time.time()
    h.assemble(kind=hmat_amd.KERNEL_EXPONENTIAL, length=0.5, nugget=1e-2, eps=1e-4)
    t_asm = time.time() - t0
    info = h.info()
    log(f"assembly {t_asm:.2f}s info {info}")

    # ---- H-gemv (HBM roofline) -------------------------------------------
    stream = torch.cuda.current_stream()
    h.set_stream(stream.cuda_stream)
    nrhs = args.nrhs
    x = torch.randn(nrhs, h.n, dtype=torch.float64, device="cuda")
    y = torch.zeros(nrhs, h.n, dtype=torch.float64, device="cuda")
    for _ in range(args.warmup):
        h.gemv_device(x.data_ptr(), y.data_ptr(), nrhs)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t_wall = time.time()
    e0.record(stream)
    for _ in range(args.steps):
        h.gemv_device(x.data_ptr(), y.data_ptr(), nrhs)
    e1.record(stream)
    torch.cuda.synchronize()
    t_wall = time.time() - t_wall
    gemv_ms = e0.elapsed_time(e1) / args.steps
    nbytes = gemv_bytes(h, info, nrhs)
    gbs = nbytes / (gemv_ms * 1e-3) / 1e9
    log(f"gemv {gemv_ms:.3f} ms  {gbs:.0f} GB/s  ({nbytes / 1e9:.3f} GB algorithmic)")
    h.set_stream(None)

    cpu = None
    if not args.no_cpu:
        t_cpu = cpu_gemv_baseline(h, pts, leaf, nrhs)
        cpu = {"value": nbytes / t_cpu / 1e9, "unit": "GB/s", "cores": 1, "kind": "port",
               "sample": f"oracle H-gemv (numpy, 1 thread) on the same N={args.n} leaves, "
                         f"best of 2: {t_cpu:.2f} s per gemv"}

    out = {
        "metric": METRIC,
        "value": gbs,
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": gemv_ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded cylinder points, exponential covariance)",
        "config": {"workload": f"H-gemv, N={args.n} cylinder r=1 h={args.height}, exp kernel "
                               f"l=0.5 nugget 1e-2, leaf 128, eta 2, ACA+ 1e-4, nrhs {nrhs}",
                   "n": args.n, "nrhs": nrhs, "assembly_s": t_asm,
                   "compressed_size": info["compressed_size"],
                   "wall_s_timed": t_wall},
        "roofline": {"bound": "hbm", "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": gbs / HBM_PEAK_GBS, "traffic": None},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
