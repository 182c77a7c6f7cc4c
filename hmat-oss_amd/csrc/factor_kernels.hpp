// Launch wrappers and workspace sizes of the factorisation task kernels.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <stdexcept>

#include "dev_common.hpp"
#include "ftask.hpp"

namespace hmx {

// Workspace (doubles) of one task.
int64_t fadd_workspace(int m, int n, int kmax_concat);
int64_t fadd_dense_workspace(int m, int n);
int64_t rkrk_workspace(int k1, int k2);
int64_t hodlr_workspace(int cap);

// p[0..n) = v on stream s (a kernel, so it may be captured in a graph).
void launch_fill_int(int* p, size_t n, int v, hipStream_t s);

// One batched launch of 'count' tasks of 'type' (GEMM: 'ntiles' tiles).
// flags[0] |= 1 on an Rk capacity overflow, flags[1] |= 1 on a non-positive
// pivot in POTRF.
template <typename T>
void launch_task_kernel(int type, const FTask* tasks, int count, const GemmTile* tiles,
                        int ntiles, double* ws, int* flags, hipStream_t s);

}  // namespace hmx
