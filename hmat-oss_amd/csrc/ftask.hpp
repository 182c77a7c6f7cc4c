// Task records of the factorisation / solve executor.  The host recursion
// (factor.cpp) turns the H-LLt of hmat-oss (recursion.cpp:349-390 and the
// H-gemm / H-trsm it calls) into a list of leaf-level tasks; tasks of one
// dependency wave and one type run as one batched kernel launch.
//
// All operands are device pointers.  Sizes are static capacities; a dynamic
// size (the rank of an Rk panel, known only on the device) is given by a
// pointer to its rank slot and the effective size is min(static, *slot).
#pragma once
#include <cstdint>

namespace hmx {

enum TaskType : int {
  T_GEMM = 0,     // C = beta C + alpha op(A) op(B)
  T_POTRF,        // A <- chol(A) lower, upper triangle zeroed
  T_TRSM,         // mode: 0 X <- X L^{-T}, 1 X <- L^{-1} X, 2 X <- L^{-T} X
  T_FADD,         // Rk target (+)= sum_i alpha_i part_i, then truncate (mode 1: truncate
                  // the target alone, no parts)
  T_FADD_DENSE,   // Rk target += alpha * dense part, eval + SVD (mode 1: target := alpha * part)
  T_RKRK,         // Rk product core: (A1 B1^T)(A2 B2^T)
  T_ACAFULL,      // dense -> Rk by full-pivot ACA
  T_ZERO,         // C = 0 (mode 1: C = I)
  T_AXPY,         // C += alpha op(A) (mode 1: C = alpha op(A))
  T_GETRF,        // A <- P L U in place (L unit lower), row interchanges into B (int*)
  T_SUMK,         // C (+)= sum over p.s slices of A (ld lda, slice stride ldb): split-K GEMM
  T_LDLT,         // A <- L D L^T in place (L unit lower), D into B (T*)
  T_DSCALE,       // C = A scaled by the diagonal B (T*): mode kDscale* flags
  T_HODLR,        // r x r blocks of a symmetric HODLR node from the Grams A = a^T a,
                  // B = b^T b: X11 into C, KK into tgt.a, log det(M11) into tgt.b (double*)
  T_NTYPES
};

// T_TRSM mode: a base operation on a lower-triangular operand L
//   0: X <- X L^{-T}   1: X <- L^{-1} X   2: X <- L^{-T} X
// plus flags: kTrsmTransA takes L = A^T (A upper triangular, as U of an LU
// leaf), kTrsmUnit takes a unit diagonal, kTrsmPiv first applies the row
// interchanges B (int*, getrf order) to X (left modes).
constexpr int kTrsmBase = 3;
constexpr int kTrsmTransA = 4;
constexpr int kTrsmUnit = 8;
constexpr int kTrsmPiv = 16;

// T_DSCALE mode flags: scale columns (C_ij = A_ij d_j; rows otherwise,
// C_ij = A_ij d_i), by the inverse of the diagonal.
constexpr int kDscaleCols = 1;
constexpr int kDscaleInv = 2;

struct Dim {
  int s;             // static size / capacity
  const int* d;      // dynamic size slot or nullptr
};

struct RkRef {
  void* a;           // m x cap, ld la
  void* b;           // n x cap, ld lb
  int la, lb;
  int m, n;
  int cap;
  int* rank;         // rank slot
};

struct Part {
  RkRef rk;          // rows/cols are a subset of the target at (roff, coff)
  int roff, coff;
  double alpha;
};

struct FTask {
  int type;
  int mode;          // trsm mode, gemm beta (0/1)
  int ta, tb;        // gemm transposes
  Dim m, n, p;
  double alpha;
  double eps;
  void* A;
  void* B;           // GETRF / TRSM with kTrsmPiv: row interchanges (int*)
  void* C;
  int lda, ldb, ldc;
  int nparts;
  int pad;
  RkRef tgt;         // FADD / FADD_DENSE / RKRK / ACAFULL output
  Part parts[4];     // FADD inputs; RKRK uses parts[0] (A1 B1^T) and parts[1] (A2 B2^T)
  int64_t ws;        // workspace offset (doubles)
};

struct GemmTile {
  int task;
  int tm, tn;
};

}  // namespace hmx
