"""Builds libhmat_mi355x.so in-tree with hipcc for gfx950.

Host-only sources are compiled with g++ (-ffp-contract=off keeps the
bounding-box arithmetic of the partition free of fused multiply-adds); HIP
sources and the C ABI with hipcc.  Objects are compiled in parallel.
"""
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libhmat_mi355x.so")
BUILD = os.path.join(HERE, "_build")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

HOST_ONLY = ["tree.cpp"]
HIP_SRC = ["engine.cpp", "factor.cpp", "sweep.cpp", "ops.cpp", "serialize.cpp", "capi.cpp", "giant.cpp",
           "k_assembly.hip",
           "k_gemv.hip", "k_factor.hip", "k_sweep.hip", "k_ops.hip"]

COMMON = ["-O3", "-fPIC", "-std=c++17", f"-I{os.path.join(ROOT, 'include')}", f"-I{CSRC}"]
HIP_FLAGS = ["--offload-arch=gfx950", "-x", "hip", "-fvisibility=hidden",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wno-unused-command-line-argument"]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("build failed:\n" + " ".join(cmd) + "\n" + r.stdout)
    return r.stdout


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    jobs = []
    for src in HOST_ONLY:
        obj = os.path.join(BUILD, src + ".o")
        jobs.append(["g++", *COMMON, "-ffp-contract=off", "-fvisibility=hidden", "-c",
                     os.path.join(CSRC, src), "-o", obj])
    for src in HIP_SRC:
        if not os.path.exists(os.path.join(CSRC, src)):
            continue
        obj = os.path.join(BUILD, src + ".o")
        jobs.append([HIPCC, *COMMON, *HIP_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj])
    with cf.ThreadPoolExecutor(max_workers=min(8, len(jobs))) as ex:
        outs = list(ex.map(_run, jobs))
    if verbose:
        for o in outs:
            if o.strip():
                print(o)
    objs = [j[-1] for j in jobs]
    tmp = OUT + ".tmp"
    # rocBLAS / rocSOLVER: GEMM, QR and SVD of the whole-device compression of
    # very large blocks
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs,
          "-L/opt/rocm/lib", "-Wl,-rpath,/opt/rocm/lib", "-lrocsolver", "-lrocblas"])
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
