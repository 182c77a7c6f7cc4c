// Batched task kernels of the H-LLt factorisation and the H-triangular solves.
// Each launch processes the tasks of one type in one dependency wave; the
// tasks of a launch are independent.  Dense arithmetic runs in the storage
// type T (FP64 or FP32); QR/SVD of the recompressions run in FP64.
#include "dev_linalg.hpp"
#include "dev_panel.hpp"
#include "factor_kernels.hpp"

namespace hmx {

namespace {

__device__ inline int eff(const Dim& d) {
  if (d.d == nullptr) return d.s;
  const int v = *d.d;
  return v < d.s ? v : d.s;
}
__device__ inline int rank_of(const RkRef& r) {
  const int v = *r.rank;
  return v < r.cap ? v : r.cap;
}

// 16 x 16 x 4 matrix-core products in FP64 and FP32.  D[i][j] is returned to
// lane j + 16 (i mod 4) as element i / 4 in FP64, to lane j + 16 (i / 4) as
// element i mod 4 in FP32.
template <typename T>
struct MfmaAcc;
template <>
struct MfmaAcc<double> {
  typedef double type __attribute__((ext_vector_type(4)));
  __device__ static int row(int lane, int v) { return (lane >> 4) + 4 * v; }
};
template <>
struct MfmaAcc<float> {
  typedef float type __attribute__((ext_vector_type(4)));
  __device__ static int row(int lane, int v) { return 4 * (lane >> 4) + v; }
};
__device__ inline MfmaAcc<double>::type mfma16x16x4(double a, double b, MfmaAcc<double>::type c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}
__device__ inline MfmaAcc<float>::type mfma16x16x4(float a, float b, MfmaAcc<float>::type c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

// ---------------------------------------------------------------------------
// GEMM: C (m x n, ldc) = beta C + alpha op(A) op(B), 64 x 64 tiles.
// ---------------------------------------------------------------------------
constexpr int TM = 64, TN = 64, TK = 32;

template <typename T>
__global__ void __launch_bounds__(256) k_gemm(const FTask* tasks, const GemmTile* tiles,
                                              const int* fail) {
  __shared__ T As[TK][TM + 1];
  __shared__ T Bs[TK][TN + 1];
  if (*fail) return;
  const GemmTile tl = tiles[blockIdx.x];
  const FTask& t = tasks[tl.task];
  const int M = eff(t.m), N = eff(t.n), P = eff(t.p);
  const int i0 = tl.tm * TM, j0 = tl.tn * TN;
  if (i0 >= M || j0 >= N) return;
  const T* A = (const T*)t.A;
  const T* B = (const T*)t.B;
  T* C = (T*)t.C;
  // wavefront w computes the 32 x 32 quadrant (w & 1, w >> 1) of the tile as
  // 2 x 2 matrix-core tiles of 16 x 16: A[i][k] comes from lane i + 16 k,
  // B[k][j] from lane j + 16 k (inner dimension 4 per instruction)
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int li = lane & 15, lk = lane >> 4;
  const int qi = (w & 1) * 32, qj = (w >> 1) * 32;
  using Acc = typename MfmaAcc<T>::type;
  Acc acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) acc[a][b] = Acc{0, 0, 0, 0};
  // As[kk][ii] = op(A)(i0 + ii, k0 + kk), Bs[kk][jj] = op(B)(k0 + kk, j0 + jj):
  // each thread loads 8 + 8 entries of a k-step into registers; the loads of
  // step k0 + TK are in flight while step k0 is multiplied
  constexpr int kPer = TM * TK / 256;
  const long la = t.lda, lb = t.ldb;
  T ra[kPer], rb[kPer];
  auto load_step = [&](int k0) {
#pragma unroll
    for (int u = 0; u < kPer; ++u) {
      const int e = threadIdx.x + 256 * u, i = e % TM, j = e / TM;
      const bool ok = i < M - i0 && j < P - k0;
      ra[u] = ok ? A[t.ta ? (size_t)(k0 + j) + (size_t)(i0 + i) * la : (size_t)(i0 + i) + (size_t)(k0 + j) * la]
                 : (T)0;
    }
#pragma unroll
    for (int u = 0; u < kPer; ++u) {
      const int e = threadIdx.x + 256 * u, i = e % TN, j = e / TN;
      const bool ok = i < N - j0 && j < P - k0;
      rb[u] = ok ? B[t.tb ? (size_t)(j0 + i) + (size_t)(k0 + j) * lb : (size_t)(k0 + j) + (size_t)(j0 + i) * lb]
                 : (T)0;
    }
  };
  load_step(0);
  for (int k0 = 0; k0 < P; k0 += TK) {
#pragma unroll
    for (int u = 0; u < kPer; ++u) {
      const int e = threadIdx.x + 256 * u;
      As[e / TM][e % TM] = ra[u];
      Bs[e / TN][e % TN] = rb[u];
    }
    __syncthreads();
    if (k0 + TK < P) load_step(k0 + TK);
#pragma unroll
    for (int kk = 0; kk < TK; kk += 4) {
      T fa[2], fb[2];
#pragma unroll
      for (int a = 0; a < 2; ++a) fa[a] = As[kk + lk][qi + 16 * a + li];
#pragma unroll
      for (int b = 0; b < 2; ++b) fb[b] = Bs[kk + lk][qj + 16 * b + li];
#pragma unroll
      for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) acc[a][b] = mfma16x16x4(fa[a], fb[b], acc[a][b]);
    }
    __syncthreads();
  }
  const T alpha = (T)t.alpha;
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int i = i0 + qi + 16 * a + MfmaAcc<T>::row(lane, v), j = j0 + qj + 16 * b + li;
        if (i >= M || j >= N) continue;
        T* pc = C + i + (size_t)j * t.ldc;
        *pc = t.mode ? *pc + alpha * acc[a][b][v] : alpha * acc[a][b][v];
      }
}

// ---------------------------------------------------------------------------
// POTRF (lower, left-looking): column j is updated by one thread per row with
// the already factorised columns (row of L read from LDS, L[j,:] broadcast),
// then scaled.  Two barriers per column.  LDS holds the block when n <= 128.
// ---------------------------------------------------------------------------
constexpr int kLdsDense = 128;

constexpr int kNb = 16;  // panel width of the blocked POTRF / TRSM

// Right-looking blocked Cholesky: each 16-column panel is factorised
// left-looking (short dots over the panel only), then the trailing lower
// triangle gets the panel's rank-16 update with all threads.
template <typename T>
__device__ void potrf_body(T* L, int ld, int n, int* info) {
  for (int jb = 0; jb < n; jb += kNb) {
    const int je = jb + kNb < n ? jb + kNb : n;
    for (int j = jb; j < je; ++j) {
      for (int i = j + threadIdx.x; i < n; i += kBlock) {
        // the panel's earlier columns: a fixed 16 terms with a mask, so that
        // their loads issue together (column indices clamped into the matrix)
        T s = L[i + (size_t)j * ld];
#pragma unroll
        for (int q = 0; q < kNb; ++q) {
          const int l = jb + q < n ? jb + q : n - 1;
          const T a = L[i + (size_t)l * ld], b = L[j + (size_t)l * ld];
          s -= (jb + q < j) ? a * b : (T)0;
        }
        L[i + (size_t)j * ld] = s;
      }
      __syncthreads();
      const T d = L[j + (size_t)j * ld];
      const T piv = sqrt(d);
      for (int i = j + 1 + threadIdx.x; i < n; i += kBlock) L[i + (size_t)j * ld] /= piv;
      __syncthreads();
      if (threadIdx.x == 0) {
        if (!(d > (T)0)) atomicOr(info, 1);
        L[j + (size_t)j * ld] = piv;
      }
    }
    __syncthreads();
    const int r = n - je;
    for (int e = threadIdx.x; e < r * r; e += kBlock) {
      const int i = je + e % r, c = je + e / r;
      if (c > i) continue;
      T s = 0;
#pragma unroll
      for (int q = 0; q < kNb; ++q) {
        const int l = jb + q < n ? jb + q : n - 1;
        const T a = L[i + (size_t)l * ld], b = L[c + (size_t)l * ld];
        s += (jb + q < je) ? a * b : (T)0;
      }
      L[i + (size_t)c * ld] -= s;
    }
    __syncthreads();
  }
}

template <typename T>
__global__ void __launch_bounds__(kBlock) k_potrf(const FTask* tasks, int* info) {
  __shared__ T lds[kLdsDense * kLdsDense];
  const FTask& t = tasks[blockIdx.x];
  const int n = t.m.s;
  T* A = (T*)t.A;
  const int lda = t.lda;
  if (n <= kLdsDense) {
    stage_in(lds, n, A, 1L, (long)lda, n, n);
    __syncthreads();
    potrf_body(lds, n, n, info);
    for (int e = threadIdx.x; e < n * n; e += kBlock) {
      const int i = e % n, j = e / n;
      A[i + (size_t)j * lda] = (i >= j) ? lds[i + j * n] : (T)0;
    }
  } else {
    potrf_body(A, lda, n, info);
    for (int e = threadIdx.x; e < n * n; e += kBlock) {
      const int i = e % n, j = e / n;
      if (i < j) A[i + (size_t)j * lda] = (T)0;
    }
  }
}

// ---------------------------------------------------------------------------
// LDLt of a diagonal leaf (ScalarArray::ldltDecomposition scalar_array.cpp:
// 737-779): A = L D L^T in place, L unit lower with its diagonal set to 1 and
// the upper triangle zeroed, D into d.  Left-looking, one column per step:
// d_j = a_jj - sum_l l_jl^2 d_l, l_ij = (a_ij - sum_l l_il l_jl d_l) / d_j.
// A zero or non-finite pivot sets info (the reference throws).
// ---------------------------------------------------------------------------
template <typename T>
__device__ void ldlt_body(T* L, int ld, int n, T* d, int* info) {
  for (int j = 0; j < n; ++j) {
    for (int i = j + threadIdx.x; i < n; i += kBlock) {
      T s = L[i + (size_t)j * ld];
      for (int l = 0; l < j; ++l) s -= L[i + (size_t)l * ld] * (L[j + (size_t)l * ld] * d[l]);
      L[i + (size_t)j * ld] = s;
    }
    __syncthreads();
    const T dj = L[j + (size_t)j * ld];
    const T inv = (T)1 / dj;
    for (int i = j + 1 + threadIdx.x; i < n; i += kBlock) L[i + (size_t)j * ld] *= inv;
    __syncthreads();
    if (threadIdx.x == 0) {
      if (!(dj != (T)0) || !isfinite((double)dj)) atomicOr(info, 1);
      d[j] = dj;
      L[j + (size_t)j * ld] = (T)1;
    }
    __syncthreads();
  }
}

template <typename T>
__global__ void __launch_bounds__(kBlock) k_ldlt(const FTask* tasks, int* info) {
  __shared__ T lds[kLdsDense * kLdsDense];
  __shared__ T dl[kLdsDense];
  const FTask& t = tasks[blockIdx.x];
  const int n = t.m.s;
  T* A = (T*)t.A;
  T* D = (T*)t.B;
  const int lda = t.lda;
  if (n <= kLdsDense) {
    stage_in(lds, n, A, 1L, (long)lda, n, n);
    __syncthreads();
    ldlt_body(lds, n, n, dl, info);
    for (int e = threadIdx.x; e < n * n; e += kBlock) {
      const int i = e % n, j = e / n;
      A[i + (size_t)j * lda] = (i >= j) ? lds[i + j * n] : (T)0;
    }
    for (int i = threadIdx.x; i < n; i += kBlock) D[i] = dl[i];
  } else {
    ldlt_body(A, lda, n, D, info);
    for (int e = threadIdx.x; e < n * n; e += kBlock) {
      const int i = e % n, j = e / n;
      if (i < j) A[i + (size_t)j * lda] = (T)0;
    }
  }
}

// C = A times a diagonal (ScalarArray::multiplyWithDiagOrDiagInv
// scalar_array.cpp:1434-1470): the columns (kDscaleCols) or the rows of A
// scaled by d or, with kDscaleInv, by 1 / d.  A may be C.  kDscaleSplit
// workgroups share a task.
constexpr int kDscaleSplit = 8;
template <typename T>
__global__ void k_dscale(const FTask* tasks, const int* fail) {
  if (*fail) return;
  const FTask& t = tasks[blockIdx.x / kDscaleSplit];
  const int part = blockIdx.x % kDscaleSplit;
  const T* A = (const T*)t.A;
  const T* d = (const T*)t.B;
  T* C = (T*)t.C;
  const int m = eff(t.m), n = eff(t.n);
  const bool cols = t.mode & kDscaleCols, inv = t.mode & kDscaleInv;
  const int64_t total = (int64_t)m * n;
  for (int64_t e = part * (int64_t)blockDim.x + threadIdx.x; e < total;
       e += (int64_t)kDscaleSplit * blockDim.x) {
    const int i = (int)(e % m), j = (int)(e / m);
    const T f = d[cols ? j : i];
    C[i + (size_t)j * t.ldc] = A[i + (size_t)j * t.lda] * (inv ? (T)1 / f : f);
  }
}

// ---------------------------------------------------------------------------
// GETRF of a diagonal leaf (ScalarArray::luDecomposition, scalar_array.cpp:
// 694-705): A = P L U in place, L unit lower, U upper, partial pivoting as
// LAPACK getrf (pivot = first row of largest |a_ij| in the column), 0-based
// interchanges into piv[0..n).  A zero or non-finite pivot sets info (the
// reference throws on getrf's info > 0).  Right-looking, one column per step.
// ---------------------------------------------------------------------------
template <typename T>
__device__ void getrf_body(T* P, int ld, int n, int* piv, int* info) {
  __shared__ int pv;
  const int lane = threadIdx.x & (kWave - 1);
  for (int j = 0; j < n; ++j) {
    if (threadIdx.x < kWave) {
      double best = -1.0;
      int bi = n;
      for (int i = j + lane; i < n; i += kWave) {
        const double v = fabs((double)P[i + (size_t)j * ld]);
        if (v > best) {
          best = v;
          bi = i;
        }
      }
      for (int o = kWave / 2; o > 0; o >>= 1) {
        const double ob = __shfl_xor(best, o, kWave);
        const int oi = __shfl_xor(bi, o, kWave);
        if (ob > best || (ob == best && oi < bi)) {
          best = ob;
          bi = oi;
        }
      }
      if (bi >= n) bi = j;  // column of NaNs: keep the row, the pivot test below fails
      if (lane == 0) {
        pv = bi;
        piv[j] = bi;
      }
    }
    __syncthreads();
    const int p = pv;
    if (p != j)
      for (int c = threadIdx.x; c < n; c += kBlock) {
        const T a = P[j + (size_t)c * ld];
        P[j + (size_t)c * ld] = P[p + (size_t)c * ld];
        P[p + (size_t)c * ld] = a;
      }
    __syncthreads();
    const double d = (double)P[j + (size_t)j * ld];
    if (d == 0.0 || !isfinite(d)) {
      if (threadIdx.x == 0) atomicOr(info, 1);
      return;
    }
    for (int i = j + 1 + threadIdx.x; i < n; i += kBlock) P[i + (size_t)j * ld] = (T)((double)P[i + (size_t)j * ld] / d);
    __syncthreads();
    const int r = n - j - 1;
    for (int e = threadIdx.x; e < r * r; e += kBlock) {
      const int i = j + 1 + e % r, c = j + 1 + e / r;
      P[i + (size_t)c * ld] -= P[i + (size_t)j * ld] * P[j + (size_t)c * ld];
    }
    __syncthreads();
  }
}

template <typename T>
__global__ void __launch_bounds__(kBlock) k_getrf(const FTask* tasks, int* info) {
  __shared__ T lds[kLdsDense * kLdsDense];
  const FTask& t = tasks[blockIdx.x];
  const int n = t.m.s;
  T* A = (T*)t.A;
  const int lda = t.lda;
  int* piv = (int*)t.B;
  if (n <= kLdsDense) {
    stage_in(lds, n, A, 1L, (long)lda, n, n);
    __syncthreads();
    getrf_body(lds, n, n, piv, info);
    __syncthreads();
    for (int e = threadIdx.x; e < n * n; e += kBlock) A[e % n + (size_t)(e / n) * lda] = lds[e];
  } else {
    getrf_body(A, lda, n, piv, info);
  }
}

// ---------------------------------------------------------------------------
// The r x r blocks of a symmetric HODLR node (computeSymX11 + computeSymKK,
// hodlr.cpp:163-219, the Gram form the reference selects with
// HMAT_HODLR_NOQR; its default QR form is the same in exact arithmetic).
// With G = a^T a and P = b^T b of the node's transformed Rk block a b^T:
//   La La^T = G,  M11 M11^T = I - La^T P La,  log det M11 into tgt.b,
//   X11 = La^-T (M11 - I) La^-1 into C,  KK = (I + X11 G)^-1 X11 into tgt.a.
// FP64 in the task's workspace (5 r^2 + r values), one workgroup per node.
// ---------------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(kBlock) k_hodlr(const FTask* tasks, double* ws_base, int* info) {
  const FTask& t = tasks[blockIdx.x];
  const int n = eff(t.m), ld = t.lda;
  const T* G = (const T*)t.A;
  const T* P = (const T*)t.B;
  T* X11 = (T*)t.C;
  T* KK = (T*)t.tgt.a;
  double* logdet = (double*)t.tgt.b;
  if (n == 0) {
    if (threadIdx.x == 0) *logdet = 0.0;
    return;
  }
  const size_t nn = (size_t)n * n;
  double* la = ws_base + t.ws;
  double* x = la + nn;
  double* tmp = x + nn;
  double* g = tmp + nn;
  double* y = g + nn;
  int* piv = (int*)(y + nn);
  for (size_t e = threadIdx.x; e < nn; e += kBlock) {
    const int i = (int)(e % n), j = (int)(e / n);
    g[e] = (double)G[i + (size_t)j * ld];
    la[e] = g[e];
    tmp[e] = (double)P[i + (size_t)j * ld];
  }
  __syncthreads();
  potrf_body<double>(la, n, n, info);
  __syncthreads();
  for (size_t e = threadIdx.x; e < nn; e += kBlock)
    if ((int)(e % n) < (int)(e / n)) la[e] = 0.0;
  __syncthreads();
  // x = -P La (La lower)
  for (size_t e = threadIdx.x; e < nn; e += kBlock) {
    const int i = (int)(e % n), j = (int)(e / n);
    double s = 0.0;
    for (int k = j; k < n; ++k) s += tmp[i + (size_t)k * n] * la[k + (size_t)j * n];
    x[e] = -s;
  }
  __syncthreads();
  // tmp = I + La^T x
  for (size_t e = threadIdx.x; e < nn; e += kBlock) {
    const int i = (int)(e % n), j = (int)(e / n);
    double s = 0.0;
    for (int k = i; k < n; ++k) s += la[k + (size_t)i * n] * x[k + (size_t)j * n];
    tmp[e] = s + (i == j ? 1.0 : 0.0);
  }
  __syncthreads();
  potrf_body<double>(tmp, n, n, info);  // M11 in the lower triangle
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0.0;
    for (int i = 0; i < n; ++i) s += log(tmp[i + (size_t)i * n]);
    *logdet = s;
  }
  // x = M11 - I (lower)
  for (size_t e = threadIdx.x; e < nn; e += kBlock) {
    const int i = (int)(e % n), j = (int)(e / n);
    x[e] = i >= j ? tmp[e] - (i == j ? 1.0 : 0.0) : 0.0;
  }
  __syncthreads();
  // x <- x La^-1 (each row on its own)
  for (int i = threadIdx.x; i < n; i += kBlock)
    for (int j = n - 1; j >= 0; --j) {
      double s = x[i + (size_t)j * n];
      for (int k = j + 1; k < n; ++k) s -= x[i + (size_t)k * n] * la[k + (size_t)j * n];
      x[i + (size_t)j * n] = s / la[j + (size_t)j * n];
    }
  __syncthreads();
  // x <- La^-T x (each column on its own)
  for (int j = threadIdx.x; j < n; j += kBlock)
    for (int i = n - 1; i >= 0; --i) {
      double s = x[i + (size_t)j * n];
      for (int k = i + 1; k < n; ++k) s -= la[k + (size_t)i * n] * x[k + (size_t)j * n];
      x[i + (size_t)j * n] = s / la[i + (size_t)i * n];
    }
  __syncthreads();
  // X11 out; tmp = I + X11 G; y = X11
  for (size_t e = threadIdx.x; e < nn; e += kBlock) {
    const int i = (int)(e % n), j = (int)(e / n);
    X11[i + (size_t)j * ld] = (T)x[e];
    double s = 0.0;
    for (int k = 0; k < n; ++k) s += x[i + (size_t)k * n] * g[k + (size_t)j * n];
    tmp[e] = s + (i == j ? 1.0 : 0.0);
    y[e] = x[e];
  }
  __syncthreads();
  getrf_body<double>(tmp, n, n, piv, info);
  __syncthreads();
  // KK = tmp^-1 X11: interchanges, unit lower, upper
  for (int j = 0; j < n; ++j) {
    const int p = piv[j];
    if (p != j && p >= 0 && p < n)
      for (int c = threadIdx.x; c < n; c += kBlock) {
        const double a = y[j + (size_t)c * n];
        y[j + (size_t)c * n] = y[p + (size_t)c * n];
        y[p + (size_t)c * n] = a;
      }
    __syncthreads();
  }
  for (int c = threadIdx.x; c < n; c += kBlock) {
    for (int i = 0; i < n; ++i) {
      double s = y[i + (size_t)c * n];
      for (int k = 0; k < i; ++k) s -= tmp[i + (size_t)k * n] * y[k + (size_t)c * n];
      y[i + (size_t)c * n] = s;
    }
    for (int i = n - 1; i >= 0; --i) {
      double s = y[i + (size_t)c * n];
      for (int k = i + 1; k < n; ++k) s -= tmp[i + (size_t)k * n] * y[k + (size_t)c * n];
      y[i + (size_t)c * n] = s / tmp[i + (size_t)i * n];
    }
    for (int i = 0; i < n; ++i) KK[i + (size_t)c * ld] = (T)y[i + (size_t)c * n];
  }
}

// ---------------------------------------------------------------------------
// TRSM with a lower-triangular leaf L (n x n), L staged in LDS:
//   mode 0: X (m x n) <- X L^{-T}     (solveUpperTriangularRight, LLt)
//   mode 1: X (n x k) <- L^{-1} X     (solveLowerTriangularLeft)
//   mode 2: X (n x k) <- L^{-T} X     (solveUpperTriangularLeft, Uplo LOWER)
// Flags (ftask.hpp): kTrsmTransA reads L as the transpose of an upper
// triangle U (so 0 is X U^{-1}, 1 is U^{-T} X, 2 is U^{-1} X: the LU cases of
// scalar_array.cpp:852-874), kTrsmUnit assumes a unit diagonal, kTrsmPiv
// applies getrf's row interchanges to X first (laswp, scalar_array.cpp:855).
// Rows (mode 0) / columns (modes 1, 2) of X are independent: one thread runs
// the substitution of one of them, reading L from LDS (broadcast) and its own
// row / column from LDS when it fits, so no barrier is needed inside.
// ---------------------------------------------------------------------------
constexpr int kShDoubles = 19456;  // 152 KiB of LDS per workgroup

// Substitution with one thread per independent row / column (L or X too
// large for the blocked-inverse kernel below).
template <typename T>
__device__ void trsm_rowwise(const FTask& t, double* sh) {
  const int n = t.n.s;  // order of L
  const int other = eff(t.m);
  if (other <= 0) return;
  const T* Lg = (const T*)t.A;
  const int ldl = t.lda;
  T* X = (T*)t.C;
  const int ldx = t.ldc;
  const int mode = t.mode & kTrsmBase;
  const bool transa = (t.mode & kTrsmTransA) != 0, unit = (t.mode & kTrsmUnit) != 0;
  const bool lds_l = n * n <= kShDoubles;
  double* Ls = sh;
  if (lds_l)
    for (int e = threadIdx.x; e < n * n; e += kBlock) {
      const int i = e % n, j = e / n;
      Ls[e] = transa ? (double)Lg[j + (size_t)i * ldl] : (double)Lg[i + (size_t)j * ldl];
    }
  // X block staged in LDS behind L when it fits: rows = n for modes 1/2
  const int xr = mode == 0 ? other : n, xc = mode == 0 ? n : other;
  const int xl = xr | 1;  // odd leading dimension: fewer bank conflicts for column walks
  const bool lds_x = lds_l && n * n + xl * xc <= kShDoubles;
  double* Xs = sh + n * n;
  if (lds_x)
    for (int e = threadIdx.x; e < xr * xc; e += kBlock) {
      const int i = e % xr, j = e / xr;
      Xs[i + j * xl] = (double)X[i + (size_t)j * ldx];
    }
  __syncthreads();
  auto Lat = [&](int i, int j) -> double {
    if (lds_l) return Ls[i + j * n];
    return transa ? (double)Lg[j + (size_t)i * ldl] : (double)Lg[i + (size_t)j * ldl];
  };
  auto Ldiv = [&](double s, int j) -> double { return unit ? s : s / Lat(j, j); };
  auto Xg = [&](int i, int j) -> double { return lds_x ? Xs[i + j * xl] : (double)X[i + (size_t)j * ldx]; };
  auto Xp = [&](int i, int j, double v) {
    if (lds_x)
      Xs[i + j * xl] = v;
    else
      X[i + (size_t)j * ldx] = (T)v;
  };
  if ((t.mode & kTrsmPiv) && mode != 0) {
    // each thread replays the interchanges on its own columns
    const int* piv = (const int*)t.B;
    for (int c = threadIdx.x; c < other; c += kBlock)
      for (int i = 0; i < n; ++i) {
        int p = piv[i];
        p = p < 0 ? 0 : (p >= n ? n - 1 : p);
        if (p != i) {
          const double a = Xg(i, c), b = Xg(p, c);
          Xp(i, c, b);
          Xp(p, c, a);
        }
      }
    __syncthreads();
  }
  // Blocked substitution: a 16-wide diagonal block is solved with one thread per
  // independent row / column, then the rest gets a rank-16 update from all threads.
  if (mode == 0) {
    for (int jb = 0; jb < n; jb += kNb) {
      const int je = jb + kNb < n ? jb + kNb : n;
      for (int r = threadIdx.x; r < other; r += kBlock)
        for (int c = jb; c < je; ++c) {
          double s = Xg(r, c);
          for (int l = jb; l < c; ++l) s -= Lat(c, l) * Xg(r, l);
          Xp(r, c, Ldiv(s, c));
        }
      __syncthreads();
      const int nc = n - je;
      for (int e = threadIdx.x; e < other * nc; e += kBlock) {
        const int r = e % other, c = je + e / other;
        double s = Xg(r, c);
        for (int l = jb; l < je; ++l) s -= Xg(r, l) * Lat(c, l);
        Xp(r, c, s);
      }
      __syncthreads();
    }
  } else if (mode == 1) {
    for (int jb = 0; jb < n; jb += kNb) {
      const int je = jb + kNb < n ? jb + kNb : n;
      for (int c = threadIdx.x; c < other; c += kBlock)
        for (int j = jb; j < je; ++j) {
          double s = Xg(j, c);
          for (int l = jb; l < j; ++l) s -= Lat(j, l) * Xg(l, c);
          Xp(j, c, Ldiv(s, j));
        }
      __syncthreads();
      const int nr = n - je;
      for (int e = threadIdx.x; e < nr * other; e += kBlock) {
        const int i = je + e % nr, c = e / nr;
        double s = Xg(i, c);
        for (int l = jb; l < je; ++l) s -= Lat(i, l) * Xg(l, c);
        Xp(i, c, s);
      }
      __syncthreads();
    }
  } else {
    for (int je = n; je > 0; je -= kNb) {
      const int jb = je - kNb > 0 ? je - kNb : 0;
      for (int c = threadIdx.x; c < other; c += kBlock)
        for (int j = je - 1; j >= jb; --j) {
          double s = Xg(j, c);
          for (int l = j + 1; l < je; ++l) s -= Lat(l, j) * Xg(l, c);
          Xp(j, c, Ldiv(s, j));
        }
      __syncthreads();
      for (int e = threadIdx.x; e < jb * other; e += kBlock) {
        const int i = e % jb, c = e / jb;
        double s = Xg(i, c);
        for (int l = jb; l < je; ++l) s -= Lat(l, i) * Xg(l, c);
        Xp(i, c, s);
      }
      __syncthreads();
    }
  }
  if (lds_x) {
    __syncthreads();
    for (int e = threadIdx.x; e < xr * xc; e += kBlock) {
      const int i = e % xr, j = e / xr;
      X[i + (size_t)j * ldx] = (T)Xs[i + j * xl];
    }
  }
}

constexpr int kTrsmRegs = 8;  // block-step results per thread: 16 x 8 x 256 covers 128 columns

// Blocked TRSM with L (n x n) and X staged in LDS.  The 16 x 16 diagonal
// blocks of L are inverted first (one thread per column of each block), then
// every 16-row (16-column for mode 0) step is a small dense product with that
// inverse followed by a rank-16 update of the rest, each spread over the whole
// workgroup: the serial chain is 3 barriers per block instead of one thread
// walking the whole triangle.
// Two workgroups per task: the columns of X (its rows in mode 0) are
// independent, so workgroup h of a task solves the h-th half of them.
constexpr int kTrsmSplit = 2;

template <typename T>
__global__ void __launch_bounds__(kBlock) k_trsm(const FTask* tasks, const int* fail) {
  __shared__ double sh[kShDoubles];
  if (*fail) return;
  FTask t = tasks[blockIdx.x / kTrsmSplit];
  {
    const int all = eff(t.m), half = blockIdx.x % kTrsmSplit;
    const int per = (all + kTrsmSplit - 1) / kTrsmSplit;
    const int lo = half * per, cnt = all - lo < per ? all - lo : per;
    if (cnt <= 0) return;
    t.m.s = cnt;
    t.m.d = nullptr;
    t.C = (void*)((T*)t.C + ((t.mode & kTrsmBase) == 0 ? (size_t)lo : (size_t)lo * t.ldc));
  }
  const int n = t.n.s;
  const int other = eff(t.m);
  if (other <= 0) return;
  const int mode = t.mode & kTrsmBase;
  const bool transa = (t.mode & kTrsmTransA) != 0, unit = (t.mode & kTrsmUnit) != 0;
  const int nb = (n + kNb - 1) / kNb;
  const int xr = mode == 0 ? other : n, xc = mode == 0 ? n : other;
  const int xl = xr | 1;
  // every block step keeps at most kTrsmRegs results per thread in registers
  const int64_t need = (int64_t)n * n + (int64_t)nb * kNb * kNb + (int64_t)xl * xc;
  if (need > kShDoubles || other > kNb * kTrsmRegs) {
    trsm_rowwise<T>(t, sh);
    return;
  }
  double* Ls = sh;                          // n x n, ld n (L, or U^T with transa)
  double* Inv = Ls + (size_t)n * n;         // nb blocks of 16 x 16 (ld 16)
  double* Xs = Inv + (size_t)nb * kNb * kNb;  // xr x xc, ld xl
  const T* Lg = (const T*)t.A;
  const int ldl = t.lda;
  T* X = (T*)t.C;
  const int ldx = t.ldc;
  if (transa)
    stage_in(Ls, n, Lg, (long)ldl, 1L, n, n);
  else
    stage_in(Ls, n, Lg, 1L, (long)ldl, n, n);
  __syncthreads();
  stage_in(Xs, xl, X, 1L, (long)ldx, xr, xc);
  __syncthreads();
  if ((t.mode & kTrsmPiv) && mode != 0) {
    const int* piv = (const int*)t.B;
    for (int c = threadIdx.x; c < other; c += kBlock)
      for (int i = 0; i < n; ++i) {
        int p = piv[i];
        p = p < 0 ? 0 : (p >= n ? n - 1 : p);
        if (p != i) {
          const double a = Xs[i + c * xl];
          Xs[i + c * xl] = Xs[p + c * xl];
          Xs[p + c * xl] = a;
        }
      }
  }
  // inverses of the diagonal blocks: thread (b, c) forms column c of block b
  // by forward substitution, the column held in registers (rows beyond the
  // block read a clamped entry and are discarded)
  for (int w = threadIdx.x; w < nb * kNb; w += kBlock) {
    const int b = w / kNb, c = w % kNb, r0 = b * kNb;
    const int bs = n - r0 < kNb ? n - r0 : kNb;
    double* inv = Inv + (size_t)b * kNb * kNb;
    if (c >= bs) continue;
    double x[kNb];
#pragma unroll
    for (int r = 0; r < kNb; ++r) {
      const int rr = r < bs ? r : bs - 1;
      const double* lrow = Ls + (r0 + rr);
      double s = r == c ? 1.0 : 0.0;
#pragma unroll
      for (int l = 0; l < r; ++l) {
        // x[l] == 0 for l < c; columns past the matrix only feed discarded rows
        const int cl = r0 + l < n ? r0 + l : n - 1;
        s -= lrow[(size_t)cl * n] * x[l];
      }
      const double d = unit ? 1.0 : lrow[(size_t)(r0 + rr) * n];
      x[r] = r < c ? 0.0 : s / d;
    }
#pragma unroll
    for (int r = 0; r < kNb; ++r)
      if (r < bs) inv[r + c * kNb] = x[r];
  }
  __syncthreads();
  // Every block step is two small products on the matrix cores (16 x 16
  // tiles, inner dimension the block's bs <= 16 columns of L): the block of
  // X times the inverted diagonal block, held in registers across a barrier,
  // then the rank-bs update of the remaining rows (columns in mode 0).
  const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
  const int ot = (other + 15) / 16;  // 16-wide tiles along 'other' (<= 8)
  for (int step = 0; step < nb; ++step) {
    const int b = mode == 2 ? nb - 1 - step : step;
    const int r0 = b * kNb, bs = n - r0 < kNb ? n - r0 : kNb;
    const double* inv = Inv + (size_t)b * kNb * kNb;
    hmx_double4 hold[2];
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      const int tt = w + kWaves * q;
      if (tt >= ot) continue;
      const int o0 = 16 * tt;
      if (mode == 1)  // X_b <- inv X_b
        hold[q] = mfma_tile16(
            [&](int ii, int k) { const double v = inv[(ii < bs ? ii : bs - 1) + k * kNb]; return ii < bs ? v : 0.0; },
            [&](int k, int jj) { const int c = o0 + jj < other ? o0 + jj : other - 1; const double v = Xs[(r0 + k) + c * xl]; return o0 + jj < other ? v : 0.0; }, bs);
      else if (mode == 2)  // X_b <- inv^T X_b
        hold[q] = mfma_tile16(
            [&](int ii, int k) { const double v = inv[k + (ii < bs ? ii : bs - 1) * kNb]; return ii < bs ? v : 0.0; },
            [&](int k, int jj) { const int c = o0 + jj < other ? o0 + jj : other - 1; const double v = Xs[(r0 + k) + c * xl]; return o0 + jj < other ? v : 0.0; }, bs);
      else  // X_b <- X_b inv^T (rows of X along 'other')
        hold[q] = mfma_tile16(
            [&](int ii, int k) { const int r = o0 + ii < other ? o0 + ii : other - 1; const double v = Xs[r + (r0 + k) * xl]; return o0 + ii < other ? v : 0.0; },
            [&](int k, int jj) { const double v = inv[(jj < bs ? jj : bs - 1) + k * kNb]; return jj < bs ? v : 0.0; }, bs);
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      const int tt = w + kWaves * q;
      if (tt >= ot) continue;
      const int o0 = 16 * tt, jj = lane & 15;
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int ii = (lane >> 4) + 4 * v;
        if (mode == 0) {
          if (o0 + ii < other && jj < bs) Xs[(o0 + ii) + (r0 + jj) * xl] = hold[q][v];
        } else {
          if (ii < bs && o0 + jj < other) Xs[(r0 + ii) + (o0 + jj) * xl] = hold[q][v];
        }
      }
    }
    __syncthreads();
    // the rest: rows [r0 + bs, n) (mode 1), [0, r0) (mode 2), columns [r0 + bs, n) (mode 0)
    const int lo = mode == 2 ? 0 : r0 + bs, rest = mode == 2 ? r0 : n - r0 - bs;
    const int rt = (rest + 15) / 16;
    for (int tt = w; tt < rt * ot; tt += kWaves) {
      const int x0 = lo + 16 * (tt % rt), o0 = 16 * (tt / rt);
      hmx_double4 acc;
      if (mode == 1)  // X_rest -= L(rest, b) X_b
        acc = mfma_tile16(
            [&](int ii, int k) { const int r = x0 + ii < lo + rest ? x0 + ii : lo + rest - 1; const double v = Ls[r + (size_t)(r0 + k) * n]; return x0 + ii < lo + rest ? v : 0.0; },
            [&](int k, int jj) { const int c = o0 + jj < other ? o0 + jj : other - 1; const double v = Xs[(r0 + k) + c * xl]; return o0 + jj < other ? v : 0.0; }, bs);
      else if (mode == 2)  // X_rest -= L(b, rest)^T X_b
        acc = mfma_tile16(
            [&](int ii, int k) { const int r = x0 + ii < lo + rest ? x0 + ii : lo + rest - 1; const double v = Ls[(r0 + k) + (size_t)r * n]; return x0 + ii < lo + rest ? v : 0.0; },
            [&](int k, int jj) { const int c = o0 + jj < other ? o0 + jj : other - 1; const double v = Xs[(r0 + k) + c * xl]; return o0 + jj < other ? v : 0.0; }, bs);
      else  // X(:, rest) -= X_b L(rest, b)^T, rows of X along 'other'
        acc = mfma_tile16(
            [&](int ii, int k) { const int r = o0 + ii < other ? o0 + ii : other - 1; const double v = Xs[r + (r0 + k) * xl]; return o0 + ii < other ? v : 0.0; },
            [&](int k, int jj) { const int c = x0 + jj < lo + rest ? x0 + jj : lo + rest - 1; const double v = Ls[c + (size_t)(r0 + k) * n]; return x0 + jj < lo + rest ? v : 0.0; }, bs);
      const int jj = lane & 15;
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int ii = (lane >> 4) + 4 * v;
        if (mode == 0) {
          if (o0 + ii < other && x0 + jj < lo + rest) Xs[(o0 + ii) + (x0 + jj) * xl] -= acc[v];
        } else {
          if (x0 + ii < lo + rest && o0 + jj < other) Xs[(x0 + ii) + (o0 + jj) * xl] -= acc[v];
        }
      }
    }
    __syncthreads();
  }
  for (int e = threadIdx.x; e < xr * xc; e += kBlock) {
    const int i = e % xr, j = e / xr;
    X[i + (size_t)j * ldx] = (T)Xs[i + j * xl];
  }
}

// ---------------------------------------------------------------------------
// SUMK: C (m x n) = (mode ? C : 0) + sum_q W_q, the partial products of a
// GEMM split along its inner dimension, added in a fixed order.
// ---------------------------------------------------------------------------
template <typename T>
__global__ void k_sumk(const FTask* tasks, const int* fail) {
  if (*fail) return;
  const FTask& t = tasks[blockIdx.x];
  const int M = eff(t.m), N = eff(t.n), S = t.p.s;
  const T* W = (const T*)t.A;
  T* C = (T*)t.C;
  for (int e = threadIdx.x; e < M * N; e += blockDim.x) {
    const int i = e % M, j = e / M;
    T* pc = C + i + (size_t)j * t.ldc;
    double s = t.mode ? (double)*pc : 0.0;
    for (int q = 0; q < S; ++q) s += (double)W[i + (size_t)j * t.lda + (size_t)q * t.ldb];
    *pc = (T)s;
  }
}

// ---------------------------------------------------------------------------
// zero fill and axpy
// ---------------------------------------------------------------------------
template <typename T>
__global__ void k_zero(const FTask* tasks, const int* fail) {
  if (*fail) return;
  const FTask& t = tasks[blockIdx.x];
  T* C = (T*)t.C;
  const int m = t.m.s, n = t.n.s;
  // mode 1: the identity
  for (int e = threadIdx.x; e < m * n; e += blockDim.x)
    C[e % m + (size_t)(e / m) * t.ldc] = (t.mode == 1 && e % m == e / m) ? (T)1 : (T)0;
}

template <typename T>
__global__ void k_axpy(const FTask* tasks, const int* fail) {
  if (*fail) return;
  const FTask& t = tasks[blockIdx.x];
  T* C = (T*)t.C;
  const T* A = (const T*)t.A;
  const int m = eff(t.m), n = eff(t.n);
  const T al = (T)t.alpha;
  for (int e = threadIdx.x; e < m * n; e += blockDim.x) {
    const int i = e % m, j = e / m;
    const T v = t.ta ? A[j + (size_t)i * t.lda] : A[i + (size_t)j * t.lda];
    if (t.mode)
      C[i + (size_t)j * t.ldc] = al * v;
    else
      C[i + (size_t)j * t.ldc] += al * v;
  }
}

// ---------------------------------------------------------------------------
// Recompression helpers (FP64; operands in LDS when they fit)
// ---------------------------------------------------------------------------
// G (gr x gc, ld gr) holds D (tr == false) or D^T (tr == true) of the dense
// block to compress; V, sig, order scratch (gc^2, gc, gc).  Truncated SVD at
// eps into out (A = U sqrt(S) on the row side of D).
template <typename T>
__device__ int dense_to_rk(double* G, int gr, int gc, bool tr, double* V, double* sig, int* order,
                           double eps, RkRef out, int* overflow, int* flag) {
  jacobi_svd(G, gr, gr, gc, V, gc, flag);
  sorted_norms(G, gr, gr, gc, sig, order);
  int k = find_k(sig, order, gc, eps);
  if (k > out.cap) {
    if (threadIdx.x == 0) atomicOr(overflow, 1);
    k = out.cap;
  }
  T* A = (T*)out.a;
  T* B = (T*)out.b;
  for (int e = threadIdx.x; e < k * gr; e += kBlock) {
    const int i = e % gr, c = e / gr, j = order[c];
    const double v = G[i + (size_t)j * gr] / sqrt(sig[j]);
    if (tr)
      B[i + (size_t)c * out.lb] = (T)v;
    else
      A[i + (size_t)c * out.la] = (T)v;
  }
  for (int e = threadIdx.x; e < k * gc; e += kBlock) {
    const int i = e % gc, c = e / gc, j = order[c];
    const double v = V[i + (size_t)j * gc] * sqrt(sig[j]);
    if (tr)
      A[i + (size_t)c * out.la] = (T)v;
    else
      B[i + (size_t)c * out.lb] = (T)v;
  }
  __syncthreads();
  if (threadIdx.x == 0) *out.rank = k;
  return k;
}

// Widest core SVD done by one wavefront; wider cores use the workgroup (a
// barrier per round, but a quarter of the rows per lane).
constexpr int kWaveJacobiMax = 16;

// Scratch of the QR recompression (all K x K blocks ld K).
struct QrScratch {
  double *taua, *taub;  // K each
  double *g, *v;        // core SVD: G = Ra Rb^T -> U Sigma, and its V
  double *ta, *tb;      // block-reflector factors of the two panels
  double *x, *w;        // K x nk work blocks of the back-transformation
  double *sig, *scale;  // K each
  int* order;           // K
};

// doubles of the scratch above
__device__ inline int64_t qr_core(int K) { return 6 * (int64_t)K * K + 5 * (int64_t)K; }

__device__ inline QrScratch carve_qr(double* core, int K) {
  const size_t K2 = (size_t)K * K;
  QrScratch w;
  w.taua = core;
  w.taub = core + K;
  w.sig = core + 2 * K;
  w.scale = core + 3 * K;
  w.order = (int*)(core + 4 * K);
  w.g = core + 5 * K;
  w.v = w.g + K2;
  w.ta = w.v + K2;
  w.tb = w.ta + K2;
  w.x = w.tb + K2;
  w.w = w.x + K2;
  return w;
}

// Rows of one side of the formatted sum, concatenated: the target's panel
// followed by every live part's panel placed at its row offset (zero
// elsewhere); alpha_i scales the A side.  dst is rows x K, ld ld.
template <typename T>
__device__ void concat_side(double* dst, int ld, int rows, bool bside, const FTask& t, int kt) {
  const RkRef& tg = t.tgt;
  const T* src = (const T*)(bside ? tg.b : tg.a);
  const int lds = bside ? tg.lb : tg.la;
  stage_in(dst, ld, src, 1L, (long)lds, rows, kt);
  int off = kt;
  for (int q = 0; q < t.nparts; ++q) {
    const Part& p = t.parts[q];
    const int kq = rank_of(p.rk);
    if (!(kq > 0 && p.alpha != 0.0 && p.rk.m > 0 && p.rk.n > 0)) continue;
    const T* P = (const T*)(bside ? p.rk.b : p.rk.a);
    const int ldp = bside ? p.rk.lb : p.rk.la;
    const int o = bside ? p.coff : p.roff, pr = bside ? p.rk.n : p.rk.m;
    const double f = bside ? 1.0 : p.alpha;
    // the part covers rows [o, o + pr) of the target; zero elsewhere
    const int r_lo = o > 0 ? o : 0, r_hi = o + pr < rows ? o + pr : rows;
    double* col0 = dst + (size_t)off * ld;
    if (r_hi > r_lo)
      stage_in(col0 + r_lo, ld, P + (r_lo - o), 1L, (long)ldp, r_hi - r_lo, kq, r_hi - r_lo, kq, f);
    const int nz = rows - (r_hi > r_lo ? r_hi - r_lo : 0);
    for (int e = threadIdx.x; e < nz * kq; e += kBlock) {
      int i = e % nz;
      const int c = e / nz;
      if (r_hi > r_lo && i >= r_lo) i += r_hi - r_lo;
      col0[i + (size_t)c * ld] = 0.0;
    }
    off += kq;
  }
}

// G (K x K) = Ra Rb^T with Ra, Rb the upper triangles of two panels.
__device__ inline void core_product(const double* Ra, int lra, const double* Rb, int lrb, int K,
                                    double* G) {
  for (int e = threadIdx.x; e < K * K; e += kBlock) {
    const int i = e % K, j = e / K;
    double s = 0.0;
    const int lo = i > j ? i : j;
    for (int l = lo; l < K; ++l) s += Ra[i + (size_t)l * lra] * Rb[j + (size_t)l * lrb];
    G[i + (size_t)j * K] = s;
  }
}

// out (rows x nk, ld ldo) = Q [U; 0] with Q = I - V Tq V^T (rows x K panel of
// reflectors V, ld ldv) and U[i, c] = src[i, order[c]] * scale[c] (K x nk,
// ld K): X = V1^T U, W = Tq X (K x nk, from the unit lower K x K top V1 of V
// staged in w.w), then out = [U; 0] - V W with each thread owning rows and
// streaming a row of V once per eight output columns.
template <typename T>
__device__ void wy_apply_out(const double* V, int ldv, int rows, int K, const double* Tq,
                             const double* src, const QrScratch& w, int nk, T* out, int ldo) {
  for (int e = threadIdx.x; e < K * K; e += kBlock) {
    const int i = e % K, l = e / K;
    w.w[e] = i > l ? V[i + (size_t)l * ldv] : (i == l ? 1.0 : 0.0);
  }
  __syncthreads();
  for (int e = threadIdx.x; e < K * nk; e += kBlock) {
    const int l = e % K, c = e / K;
    const double* u = src + (size_t)w.order[c] * K;
    const double* vl = w.w + (size_t)l * K;
    double s = 0.0;
    for (int i = l; i < K; ++i) s += vl[i] * u[i];
    w.x[l + (size_t)c * K] = s * w.scale[c];
  }
  __syncthreads();
  for (int e = threadIdx.x; e < K * nk; e += kBlock) {
    const int l = e % K, c = e / K;
    double s = 0.0;
    for (int r = l; r < K; ++r) s += Tq[l + (size_t)r * K] * w.x[r + (size_t)c * K];
    w.w[l + (size_t)c * K] = s;  // the staged V1 is no longer read
  }
  __syncthreads();
  constexpr int kC = 8;
  for (int c0 = 0; c0 < nk; c0 += kC) {
    const double* wc[kC];
#pragma unroll
    for (int b = 0; b < kC; ++b) wc[b] = w.w + (size_t)(c0 + b < nk ? c0 + b : nk - 1) * K;
    // two rows per thread at a time (the second one kBlock further), so that
    // 16 loads of V are in flight per round; rows past the panel read a
    // clamped row and are not stored
    for (int i0 = threadIdx.x; i0 < rows; i0 += 2 * kBlock) {
      const int ia = i0, ib = i0 + kBlock < rows ? i0 + kBlock : rows - 1;
      const bool has_b = i0 + kBlock < rows;
      double acc[2][kC];
#pragma unroll
      for (int b = 0; b < kC; ++b) acc[0][b] = acc[1][b] = 0.0;
      // entries strictly below the unit diagonal: l < min(i, K)
      const int la = ia < K ? ia : K, lb = ib < K ? ib : K, lmax = la > lb ? la : lb;
      const double* va = V + ia;
      const double* vb = V + ib;
      int l = 0;
      for (; l + 8 <= lmax; l += 8) {
        double xa[8], xb[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) {
          xa[q] = va[(size_t)(l + q) * ldv];
          xb[q] = vb[(size_t)(l + q) * ldv];
        }
#pragma unroll
        for (int q = 0; q < 8; ++q) {
          const double ya = l + q < la ? xa[q] : 0.0, yb = l + q < lb ? xb[q] : 0.0;
#pragma unroll
          for (int b = 0; b < kC; ++b) {
            acc[0][b] += ya * wc[b][l + q];
            acc[1][b] += yb * wc[b][l + q];
          }
        }
      }
      for (; l < lmax; ++l) {
        const double ya = l < la ? va[(size_t)l * ldv] : 0.0, yb = l < lb ? vb[(size_t)l * ldv] : 0.0;
#pragma unroll
        for (int b = 0; b < kC; ++b) {
          acc[0][b] += ya * wc[b][l];
          acc[1][b] += yb * wc[b][l];
        }
      }
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        if (h == 1 && !has_b) break;
        const int i = h == 0 ? ia : ib;
#pragma unroll
        for (int b = 0; b < kC; ++b) {
          if (c0 + b >= nk) break;
          double y = -acc[h][b];
          if (i < K) y += src[i + (size_t)w.order[c0 + b] * K] * w.scale[c0 + b] - wc[b][i];
          out[i + (size_t)(c0 + b) * ldo] = (T)y;
        }
      }
    }
  }
  __syncthreads();
}

// After QR of both panels (reflectors Va, Vb, factors ta, tb) and G = Ra Rb^T
// in w.g: SVD of G, cut at eps, write Qa U sqrt(S) and Qb V sqrt(S) to out.
// stage: free LDS of kShDoubles doubles when the core lives in global memory
// (nullptr otherwise): a core of at most kShDoubles entries is rotated there.
template <typename T>
__device__ void qr_finish(const double* Va, int lda, int m, const double* Vb, int ldb, int n, int K,
                          const QrScratch& w, bool wave_core, double eps, RkRef out, int* overflow,
                          int* flag, int* sweeps_out, double* stage = nullptr) {
  __shared__ int sweeps;
  // one-sided Jacobi on G without V: G_in (kept in w.v) gives the right side
  // afterwards as G_in^T U, so the rounds only rotate G
  for (int e = threadIdx.x; e < K * K; e += kBlock) w.v[e] = w.g[e];
  __syncthreads();
  if (stage != nullptr && K * K <= kShDoubles) {
    // the rounds' loads and stores hit LDS instead of the L2
    for (int e = threadIdx.x; e < K * K; e += kBlock) stage[e] = w.g[e];
    __syncthreads();
    jacobi_svd(stage, K, K, K, nullptr, K, flag);
    for (int e = threadIdx.x; e < K * K; e += kBlock) w.g[e] = stage[e];
    if (threadIdx.x == 0) sweeps = 0;
    __syncthreads();
  } else if (wave_core) {
    if (threadIdx.x < kWave) {
      const int sw = wave_jacobi_svd(w.g, K, K, K, nullptr, K);
      if (threadIdx.x == 0) sweeps = sw;
    }
    __syncthreads();
  } else {
    jacobi_svd(w.g, K, K, K, nullptr, K, flag);
    if (threadIdx.x == 0) sweeps = 0;
    __syncthreads();
  }
  if (sweeps_out) *sweeps_out = sweeps;
  sorted_norms(w.g, K, K, K, w.sig, w.order);
  int nk = find_k(w.sig, w.order, K, eps);
  if (nk > out.cap) {
    if (threadIdx.x == 0) atomicOr(overflow, 1);
    nk = out.cap;
  }
  // A = Qa U sqrt(S): column j of the rotated G is sigma_j u_j
  for (int c = threadIdx.x; c < nk; c += kBlock) w.scale[c] = 1.0 / sqrt(w.sig[w.order[c]]);
  __syncthreads();
  wy_apply_out<T>(Va, lda, m, K, w.ta, w.g, w, nk, (T*)out.a, out.la);
  // B = Qb V sqrt(S) with v_j sigma_j = G_in^T u_j, u_j = g_j / sigma_j: column j
  // of w.ta (the A side's block factor is done) gets G_in^T u_j, and the same
  // scale 1 / sqrt(sigma_j) as the A side applies
  for (int e = threadIdx.x; e < K * nk; e += kBlock) {
    const int i = e % K, j = w.order[e / K];
    const double* gi = w.v + (size_t)i * K;
    const double* gj = w.g + (size_t)j * K;
    double s = 0.0;
    for (int l = 0; l < K; ++l) s += gi[l] * gj[l];
    w.ta[i + (size_t)j * K] = s / w.sig[j];
  }
  __syncthreads();
  wy_apply_out<T>(Vb, ldb, n, K, w.tb, w.ta, w, nk, (T*)out.b, out.lb);
  if (threadIdx.x == 0) *out.rank = nk;
}

// ---------------------------------------------------------------------------
// Gram recompression of single-precision panels.  The formatted sum is
// [A] [B]^T with A (m x K), B (n x K) the concatenated panels (FP64 copies).
// Instead of Householder QR of each panel (one column step at a time), the
// Grams A^T A and B^T B are formed on the FP64 matrix cores in one pass over
// the panels and factored by pivoted Cholesky, A P = Q [R11 R12] with Q =
// A P1 R11^-1, stopping where the remaining diagonal falls below 1e-13 of the
// largest (columns below 3e-7 of the largest norm, far under the FP32 data's
// own rounding relative to eps).  The core R~a R~b^T (R with its columns back
// in panel order) gets the same Jacobi SVD and eps cut as the QR form of
// RkMatrix::truncate (rk_matrix.cpp:490-585), and the new panels are
// A' = A P1 R11^-1 U sqrt(S), B' = B P1 R11^-1 V sqrt(S).  Q is orthonormal to
// about 1e-16 / 1e-13 = 1e-3, so the truncated product stays within eps.
// ---------------------------------------------------------------------------
constexpr int kGramMaxK = 32;
constexpr double kGramTol = 1e-13;
constexpr double kGramJacobiTol2 = 1e-16;

// Pivoted Cholesky of G (K x K, ld K, symmetric, K <= 64) by one wavefront:
// returns the rank r; rows 0..r-1 of the upper triangle hold R in the pivoted
// column order, piv[0..K) the order.
__device__ inline int wave_pchol(double* G, int K, int* piv) {
  const int lane = threadIdx.x & (kWave - 1);
  if (lane < K) piv[lane] = lane;
  const double dmax = wave_argmax(ArgMax{lane < K ? G[lane + (size_t)lane * K] : -1.0, lane < K ? lane : -1}).v;
  const double tol = kGramTol * dmax;
  __builtin_amdgcn_wave_barrier();
  int r = 0;
  for (int j = 0; j < K; ++j) {
    const bool in = lane >= j && lane < K;
    const ArgMax a = wave_argmax(ArgMax{in ? G[lane + (size_t)lane * K] : -1.0, in ? lane : -1});
    if (!(a.v > tol)) break;  // also on NaN
    const int p = a.i;
    if (p != j) {
      if (lane < K) {
        const double x = G[j + (size_t)lane * K];
        G[j + (size_t)lane * K] = G[p + (size_t)lane * K];
        G[p + (size_t)lane * K] = x;
      }
      __builtin_amdgcn_wave_barrier();
      if (lane < K) {
        const double x = G[lane + (size_t)j * K];
        G[lane + (size_t)j * K] = G[lane + (size_t)p * K];
        G[lane + (size_t)p * K] = x;
      }
      if (lane == 0) {
        const int q = piv[j];
        piv[j] = piv[p];
        piv[p] = q;
      }
      __builtin_amdgcn_wave_barrier();
    }
    const double d = sqrt(G[j + (size_t)j * K]);
    __builtin_amdgcn_wave_barrier();
    if (lane > j && lane < K) G[j + (size_t)lane * K] /= d;
    if (lane == j) G[j + (size_t)j * K] = d;
    __builtin_amdgcn_wave_barrier();
    const int tr = K - j - 1;
    for (int e = lane; e < tr * tr; e += kWave) {
      const int i = j + 1 + e % tr, l = j + 1 + e / tr;
      G[i + (size_t)l * K] -= G[j + (size_t)i * K] * G[j + (size_t)l * K];
    }
    __builtin_amdgcn_wave_barrier();
    r = j + 1;
  }
  return r;
}

// LDS doubles of gram_recompress for K columns
__device__ inline int gram_lds(int K) { return 6 * K * K + 4 * K + xty_part(K, K); }

template <typename T>
__device__ void gram_recompress(const FTask& t, int kt, int K, double* gws, double* sh, double eps,
                                int* overflow, int* flag) {
  __shared__ int ranks[2];
  const RkRef tg = t.tgt;
  const int m = tg.m, n = tg.n;
  double* Va = gws;
  double* Vb = gws + (size_t)m * K;
  concat_side<T>(Va, m, m, false, t, kt);
  concat_side<T>(Vb, n, n, true, t, kt);
  __syncthreads();
  const int K2 = K * K;
  double* Ga = sh;
  double* Gb = Ga + K2;
  double* Cm = Gb + K2;
  double* Cin = Cm + K2;
  double* Ma = Cin + K2;
  double* Mb = Ma + K2;
  double* sig = Mb + K2;
  int* order = (int*)(sig + K);
  int* pa = order + K;
  int* pb = pa + K;
  double* part = sig + 4 * K;
  wg_xty_any<double>(Va, m, K, Va, m, K, m, Ga, K, false, part);
  wg_xty_any<double>(Vb, n, K, Vb, n, K, n, Gb, K, false, part);
  const int wid = threadIdx.x / kWave, lane = threadIdx.x & (kWave - 1);
  if (wid == 0) {
    const int r = wave_pchol(Ga, K, pa);
    if (lane == 0) ranks[0] = r;
  } else if (wid == 1) {
    const int r = wave_pchol(Gb, K, pb);
    if (lane == 0) ranks[1] = r;
  }
  __syncthreads();
  const int ra = ranks[0], rb = ranks[1];
  if (ra == 0 || rb == 0) {
    if (threadIdx.x == 0) *tg.rank = 0;
    return;
  }
  // R~a (ra x K) and R~b (rb x K), ld K, columns in panel order
  for (int e = threadIdx.x; e < K2; e += kBlock) Ma[e] = Mb[e] = 0.0;
  __syncthreads();
  for (int e = threadIdx.x; e < ra * K; e += kBlock) {
    const int x = e % ra, i = e / ra;
    if (i >= x) Ma[x + (size_t)pa[i] * K] = Ga[x + (size_t)i * K];
  }
  for (int e = threadIdx.x; e < rb * K; e += kBlock) {
    const int y = e % rb, i = e / rb;
    if (i >= y) Mb[y + (size_t)pb[i] * K] = Gb[y + (size_t)i * K];
  }
  __syncthreads();
  // core (ra x rb, ld ra) = R~a R~b^T, kept in Cin for the right side
  for (int e = threadIdx.x; e < ra * rb; e += kBlock) {
    const int x = e % ra, y = e / ra;
    double s = 0.0;
    for (int c = 0; c < K; ++c) s += Ma[x + (size_t)c * K] * Mb[y + (size_t)c * K];
    Cm[e] = s;
    Cin[e] = s;
  }
  __syncthreads();
  // columns orthogonal to 1e-8 of their norms: far below the FP32 data's
  // rounding once scaled by the eps cut
  if (rb <= kWaveJacobiMax) {
    if (wid == 0) wave_jacobi_svd(Cm, ra, ra, rb, nullptr, 0, 40, kGramJacobiTol2);
    __syncthreads();
  } else {
    jacobi_svd(Cm, ra, ra, rb, nullptr, 0, flag, 40, kGramJacobiTol2);
  }
  sorted_norms(Cm, ra, ra, rb, sig, order);
  int nk = find_k(sig, order, rb, eps);
  if (nk > tg.cap) {
    if (threadIdx.x == 0) atomicOr(overflow, 1);
    nk = tg.cap;
  }
  // left U sqrt(S) (ra x nk) and right V sqrt(S) = Cin^T U / sqrt(S) (rb x nk), ld K
  for (int e = threadIdx.x; e < ra * nk; e += kBlock) {
    const int x = e % ra, oc = order[e / ra];
    Ma[x + (size_t)(e / ra) * K] = Cm[x + (size_t)oc * ra] / sqrt(sig[oc]);
  }
  for (int e = threadIdx.x; e < rb * nk; e += kBlock) {
    const int y = e % rb, oc = order[e / rb];
    double s = 0.0;
    for (int x = 0; x < ra; ++x) s += Cin[x + (size_t)y * ra] * Cm[x + (size_t)oc * ra];
    Mb[y + (size_t)(e / rb) * K] = s / (sig[oc] * sqrt(sig[oc]));
  }
  __syncthreads();
  // R11^-1 of each side, one column per thread (A: threads 0.., B: 128..)
  const int half = kBlock / 2;
  {
    const bool bside = threadIdx.x >= half;
    const int c = bside ? threadIdx.x - half : threadIdx.x;
    const int r = bside ? rb : ra;
    const double* R = bside ? Gb : Ga;
    double* X = bside ? Mb : Ma;
    if (c < nk)
      for (int x = r - 1; x >= 0; --x) {
        double v = X[x + (size_t)c * K];
        for (int i = x + 1; i < r; ++i) v -= R[x + (size_t)i * K] * X[i + (size_t)c * K];
        X[x + (size_t)c * K] = v / R[x + (size_t)x * K];
      }
  }
  __syncthreads();
  // scatter to panel order (rows piv[i]) into Cm / Cin, K x nk, ld K
  for (int e = threadIdx.x; e < K * nk; e += kBlock) Cm[e] = Cin[e] = 0.0;
  __syncthreads();
  for (int e = threadIdx.x; e < ra * nk; e += kBlock)
    Cm[pa[e % ra] + (size_t)(e / ra) * K] = Ma[e % ra + (size_t)(e / ra) * K];
  for (int e = threadIdx.x; e < rb * nk; e += kBlock)
    Cin[pb[e % rb] + (size_t)(e / rb) * K] = Mb[e % rb + (size_t)(e / rb) * K];
  __syncthreads();
  rows_times_small<double, T>(Va, m, m, K, Cm, K, nk, (T*)tg.a, tg.la);
  rows_times_small<double, T>(Vb, n, n, K, Cin, K, nk, (T*)tg.b, tg.lb);
  if (threadIdx.x == 0) *tg.rank = nk;
}

// ---------------------------------------------------------------------------
// FADD: RkMatrix::formattedAddParts (rk_matrix.cpp:846-981)
// ---------------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(kBlock) k_fadd(const FTask* tasks, double* wsbase, int* overflow) {
  __shared__ double sh[kShDoubles];
  __shared__ int flag;
  if (overflow[1]) return;
  const FTask& t = tasks[blockIdx.x];
  const RkRef tg = t.tgt;
  const int m = tg.m, n = tg.n;
  const int kt = rank_of(tg);
  int K = kt, np = 0, last = -1;
  for (int q = 0; q < t.nparts; ++q) {
    const int kq = rank_of(t.parts[q].rk);
    if (kq > 0 && t.parts[q].alpha != 0.0 && t.parts[q].rk.m > 0 && t.parts[q].rk.n > 0) {
      K += kq;
      ++np;
      last = q;
    }
  }
  // mode 1 recompresses the target alone (RkMatrix::truncate)
  if (np == 0 && !(t.mode == 1 && kt > 0)) return;
  T* TA = (T*)tg.a;
  T* TB = (T*)tg.b;
  if (kt == 0 && np == 1) {
    // single part: copy (scaled), no recompression
    const Part& p = t.parts[last];
    int k = rank_of(p.rk);
    if (k > tg.cap) {
      if (threadIdx.x == 0) atomicOr(overflow, 1);
      k = tg.cap;
    }
    const T* PA = (const T*)p.rk.a;
    const T* PB = (const T*)p.rk.b;
    for (int e = threadIdx.x; e < m * k; e += kBlock) {
      const int i = e % m, c = e / m;
      const int ii = i - p.roff;
      TA[i + (size_t)c * tg.la] =
          (ii >= 0 && ii < p.rk.m) ? (T)(p.alpha * (double)PA[ii + (size_t)c * p.rk.la]) : (T)0;
    }
    for (int e = threadIdx.x; e < n * k; e += kBlock) {
      const int i = e % n, c = e / n;
      const int ii = i - p.coff;
      TB[i + (size_t)c * tg.lb] = (ii >= 0 && ii < p.rk.n) ? PB[ii + (size_t)c * p.rk.lb] : (T)0;
    }
    __syncthreads();
    if (threadIdx.x == 0) *tg.rank = k;
    return;
  }
  double* gws = wsbase + t.ws;
  const int mn = m < n ? m : n;
  if (K >= mn) {
    // dense path: eval everything (oriented so that the Jacobi works on min(m,n) columns)
    const bool tr = m < n;
    const int gr = tr ? n : m, gc = mn;
    const int64_t need = (int64_t)m * n + (int64_t)gc * gc + 2 * gc;
    double* G = need <= kShDoubles ? sh : gws;
    for (int e = threadIdx.x; e < m * n; e += kBlock) {
      const int i = e % m, j = e / m;
      double s = 0.0;
      for (int l = 0; l < kt; ++l) s += (double)TA[i + (size_t)l * tg.la] * (double)TB[j + (size_t)l * tg.lb];
      for (int q = 0; q < t.nparts; ++q) {
        const Part& p = t.parts[q];
        const int ii = i - p.roff, jj = j - p.coff;
        if (ii < 0 || ii >= p.rk.m || jj < 0 || jj >= p.rk.n || p.alpha == 0.0) continue;
        const int kq = rank_of(p.rk);
        const T* PA = (const T*)p.rk.a;
        const T* PB = (const T*)p.rk.b;
        double sp = 0.0;
        for (int l = 0; l < kq; ++l) sp += (double)PA[ii + (size_t)l * p.rk.la] * (double)PB[jj + (size_t)l * p.rk.lb];
        s += p.alpha * sp;
      }
      if (tr)
        G[j + (size_t)i * gr] = s;
      else
        G[i + (size_t)j * gr] = s;
    }
    __syncthreads();
    double* V = G + (size_t)m * n;
    dense_to_rk<T>(G, gr, gc, tr, V, V + (size_t)gc * gc, (int*)(V + (size_t)gc * gc + gc),
                   fabs(t.eps), tg, overflow, &flag);
    return;
  }
  if constexpr (sizeof(T) == 4) {
    if (K <= kGramMaxK) {
      gram_recompress<T>(t, kt, K, gws, sh, fabs(t.eps), overflow, &flag);
      return;
    }
  }
  // QR path (truncate, rk_matrix.cpp:490-585) on the concatenated panels
  // [target, alpha_i parts].  Three placements:
  //  fast   both panels and the scratch in LDS; wavefront 0 reduces A while
  //         wavefront 1 reduces B;
  //  staged one panel at a time in LDS (workgroup QR), reflectors parked in
  //         the workspace;
  //  global panels in the workspace, scratch in LDS when it fits.
  // Odd leading dimensions in LDS keep the strided lane accesses spread over
  // the banks.
  const int64_t core = qr_core(K);
  const int ldAf = m | 1, ldBf = n | 1, ldM = (m > n ? m : n) | 1;
  int mode;
  double *Va, *Vb, *cr;
  int lda, ldb;
  if ((int64_t)(ldAf + ldBf) * K + core <= kShDoubles) {
    mode = 5;
    Va = sh;
    lda = ldAf;
    Vb = sh + (size_t)ldAf * K;
    ldb = ldBf;
    cr = Vb + (size_t)ldBf * K;
  } else if ((int64_t)ldM * K + core <= kShDoubles) {
    mode = 6;
    Va = gws;
    lda = m;
    Vb = gws + (size_t)m * K;
    ldb = n;
    cr = sh + (size_t)ldM * K;
  } else {
    Va = gws;
    lda = m;
    Vb = gws + (size_t)m * K;
    ldb = n;
    cr = core <= kShDoubles ? sh : gws + (size_t)(m + n) * K;
    mode = cr == sh ? 3 : 4;
  }
  const QrScratch w = carve_qr(cr, K);
  if (mode == 5) {
    concat_side<T>(Va, lda, m, false, t, kt);
    concat_side<T>(Vb, ldb, n, true, t, kt);
    __syncthreads();
    const int wid = threadIdx.x / kWave;
    if (wid == 0) {
      wave_house_qr(Va, lda, m, K, w.taua);
      team_larft<kWave>(Va, lda, m, K, w.taua, w.x, w.ta);
    } else if (wid == 1) {
      wave_house_qr(Vb, ldb, n, K, w.taub);
      team_larft<kWave>(Vb, ldb, n, K, w.taub, w.w, w.tb);
    }
    __syncthreads();
    core_product(Va, lda, Vb, ldb, K, w.g);
  } else if (mode == 6) {
    double* P = sh;
    concat_side<T>(P, ldM, m, false, t, kt);
    __syncthreads();
    house_qr(P, ldM, m, K, w.taua);
    team_larft<kBlock>(P, ldM, m, K, w.taua, w.x, w.ta);
    for (int e = threadIdx.x; e < m * K; e += kBlock) {
      const int i = e % m, c = e / m;
      Va[i + (size_t)c * lda] = P[i + (size_t)c * ldM];
      if (i < K) w.w[i + (size_t)c * K] = P[i + (size_t)c * ldM];  // Ra
    }
    __syncthreads();
    concat_side<T>(P, ldM, n, true, t, kt);
    __syncthreads();
    house_qr(P, ldM, n, K, w.taub);
    team_larft<kBlock>(P, ldM, n, K, w.taub, w.x, w.tb);
    for (int e = threadIdx.x; e < n * K; e += kBlock) {
      const int i = e % n, c = e / n;
      Vb[i + (size_t)c * ldb] = P[i + (size_t)c * ldM];
    }
    core_product(w.w, K, P, ldM, K, w.g);
  } else {
    concat_side<T>(Va, lda, m, false, t, kt);
    concat_side<T>(Vb, ldb, n, true, t, kt);
    __syncthreads();
    // V^T V of the global panels on the matrix cores when its LDS staging fits
    double* part = cr == sh ? sh + core : sh;
    const bool mfma_gram = K <= 64 && (part - sh) + gram_part(K) <= kShDoubles;
    // both panels at once, one per half of the workgroup, when each has at
    // most 16 rows per thread
    const bool pair = m <= 16 * (kBlock / 2) && n <= 16 * (kBlock / 2);
    if (pair)
      house_qr_pair(Va, lda, m, w.taua, Vb, ldb, n, w.taub, K);
    else
      house_qr(Va, lda, m, K, w.taua);
    if (mfma_gram) {
      wg_reflector_gram_any(Va, lda, m, K, w.x, part);
      larft_recurrence<kBlock>(K, w.taua, w.x, w.ta);
    } else {
      team_larft<kBlock>(Va, lda, m, K, w.taua, w.x, w.ta);
    }
    if (!pair) house_qr(Vb, ldb, n, K, w.taub);
    if (mfma_gram) {
      wg_reflector_gram_any(Vb, ldb, n, K, w.x, part);
      larft_recurrence<kBlock>(K, w.taub, w.x, w.tb);
    } else {
      team_larft<kBlock>(Vb, ldb, n, K, w.taub, w.x, w.tb);
    }
    core_product(Va, lda, Vb, ldb, K, w.g);
  }
  __syncthreads();
  int sweeps = 0;
  // the core SVD runs on one wavefront when its scratch is in LDS
  const bool core_lds = mode == 5 || mode == 6 || cr == sh;
  qr_finish<T>(Va, lda, m, Vb, ldb, n, K, w, core_lds && K <= kWaveJacobiMax, fabs(t.eps), tg, overflow, &flag,
               &sweeps, core_lds ? nullptr : sh);
}

// FADD_DENSE: RkMatrix::formattedAddParts with a full part (rk_matrix.cpp:984-1010)
template <typename T>
__global__ void __launch_bounds__(kBlock) k_fadd_dense(const FTask* tasks, double* wsbase, int* overflow) {
  __shared__ double sh[kShDoubles];
  __shared__ int flag;
  if (overflow[1]) return;
  const FTask& t = tasks[blockIdx.x];
  const RkRef tg = t.tgt;
  // mode 1 replaces the target (its current panels are ignored)
  const int m = tg.m, n = tg.n, kt = t.mode ? 0 : rank_of(tg);
  const T* TA = (const T*)tg.a;
  const T* TB = (const T*)tg.b;
  const T* P = (const T*)t.A;
  const int roff = t.parts[0].roff, coff = t.parts[0].coff;
  const int pm = t.m.s, pn = t.n.s;
  const bool tr = m < n;
  const int mn = tr ? m : n, gr = tr ? n : m, gc = mn;
  const int64_t need = (int64_t)m * n + (int64_t)gc * gc + 2 * gc;
  double* G = need <= kShDoubles ? sh : wsbase + t.ws;
  for (int e = threadIdx.x; e < m * n; e += kBlock) {
    const int i = e % m, j = e / m;
    double s = 0.0;
    for (int l = 0; l < kt; ++l) s += (double)TA[i + (size_t)l * tg.la] * (double)TB[j + (size_t)l * tg.lb];
    const int ii = i - roff, jj = j - coff;
    if (ii >= 0 && ii < pm && jj >= 0 && jj < pn) s += t.alpha * (double)P[ii + (size_t)jj * t.lda];
    if (tr)
      G[j + (size_t)i * gr] = s;
    else
      G[i + (size_t)j * gr] = s;
  }
  __syncthreads();
  double* V = G + (size_t)m * n;
  dense_to_rk<T>(G, gr, gc, tr, V, V + (size_t)gc * gc, (int*)(V + (size_t)gc * gc + gc), t.eps, tg,
                 overflow, &flag);
}

// ---------------------------------------------------------------------------
// RKRK: RkMatrix::multiplyRkRk (rk_matrix.cpp:1245-1333)
//   parts[0] = (A1, B1) op-applied left factor, parts[1] = (A2, B2) right factor
//   tmp = B1^T A2, truncated SVD, newA = A1 U sqrt(S), newB = B2 V sqrt(S)
// ---------------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(kBlock) k_rkrk(const FTask* tasks, double* wsbase, int* overflow) {
  __shared__ double sh[kShDoubles];
  __shared__ int flag;
  if (overflow[1]) return;
  const FTask& t = tasks[blockIdx.x];
  const RkRef r1 = t.parts[0].rk, r2 = t.parts[1].rk;
  RkRef out = t.tgt;
  const int k1 = rank_of(r1), k2 = rank_of(r2);
  if (k1 == 0 || k2 == 0) {
    if (threadIdx.x == 0) *out.rank = 0;
    return;
  }
  const int p = r1.n;  // common dimension (rows of B1 == rows of A2)
  const bool tr = k1 < k2;
  const int gr = tr ? k2 : k1, gc = tr ? k1 : k2;
  const int64_t need = (int64_t)gr * gc + (int64_t)gc * gc + 2 * gc;
  // SVD scratch, then the output factors (at most (gr + gc) x gc)
  const int64_t need_full = need + 1 + (int64_t)(gr + gc) * gc;
  double* ws = need_full <= kShDoubles ? sh : wsbase + t.ws;
  double* G = ws; ws += (size_t)gr * gc;
  double* V = ws; ws += (size_t)gc * gc;
  double* sig = ws; ws += gc;
  int* order = (int*)ws;
  const T* A1 = (const T*)r1.a;
  const T* B1 = (const T*)r1.b;
  const T* A2 = (const T*)r2.a;
  const T* B2 = (const T*)r2.b;
  // core B1^T A2 (k1 x k2; stored transposed when tr): on the matrix cores
  // when both ranks are <= 32 and the LDS staging fits, else a wavefront per
  // entry
  double* part = G == sh ? sh + need + 1 : sh;  // reused by the output factors afterwards
  const bool mfma = k1 <= 32 && k2 <= 32 && (part - sh) + xty_part(k1, k2) <= kShDoubles &&
                    wg_xty_any<T>(B1, r1.lb, k1, A2, r2.la, k2, p, G, gr, tr, part);
  if (!mfma) {
    const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
    for (int e = w; e < k1 * k2; e += kWaves) {
      const int i = e % k1, j = e / k1;
      double s = 0.0;
      for (int r = lane; r < p; r += kWave) s += (double)B1[r + (size_t)i * r1.lb] * (double)A2[r + (size_t)j * r2.la];
      s = wave_sum(s);
      if (lane == 0) {
        if (tr)
          G[j + (size_t)i * gr] = s;
        else
          G[i + (size_t)j * gr] = s;
      }
    }
    __syncthreads();
  }
  if (G == sh && gc <= kWaveJacobiMax) {
    if (threadIdx.x < kWave) wave_jacobi_svd(G, gr, gr, gc, V, gc);
    __syncthreads();
  } else {
    jacobi_svd(G, gr, gr, gc, V, gc, &flag);
  }
  sorted_norms(G, gr, gr, gc, sig, order);
  int nk = find_k(sig, order, gc, t.eps);
  if (nk > out.cap) {
    if (threadIdx.x == 0) atomicOr(overflow, 1);
    nk = out.cap;
  }
  // tmp = U S V^T with U = G(:,j)/sig (k1 side), V (k2 side); swapped when tr.
  // The small factors Uc (k1 x nk) = U sqrt(S) and Vc (k2 x nk) = V sqrt(S)
  // go after the SVD scratch, then the outputs stream A1 and B2 once per
  // eight columns.
  double* Uc = (double*)(order + gc + (gc & 1));
  double* Vc = Uc + (size_t)k1 * nk;
  for (int e = threadIdx.x; e < k1 * nk; e += kBlock) {
    const int l = e % k1, c = e / k1, j = order[c];
    const double sq = sqrt(sig[j]);
    Uc[e] = tr ? V[l + (size_t)j * gc] * sq : G[l + (size_t)j * gr] / sq;
  }
  for (int e = threadIdx.x; e < k2 * nk; e += kBlock) {
    const int l = e % k2, c = e / k2, j = order[c];
    const double sq = sqrt(sig[j]);
    Vc[e] = tr ? G[l + (size_t)j * gr] / sq : V[l + (size_t)j * gc] * sq;
  }
  __syncthreads();
  const int m = r1.m, n = r2.n;
  rows_times_small<T>(A1, r1.la, m, k1, Uc, k1, nk, (T*)out.a, out.la);
  rows_times_small<T>(B2, r2.lb, n, k2, Vc, k2, nk, (T*)out.b, out.lb);
  if (threadIdx.x == 0) *out.rank = nk;
}

// ---------------------------------------------------------------------------
// ACAFULL: acaFull (compression.cpp:270-355) on a dense temp (modified)
constexpr int kAcaLds = 16384;  // entries of the LDS copy (a 128 x 128 leaf product)
// ---------------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(kBlock) k_acafull(const FTask* tasks, const int* fail) {
  __shared__ double rv[kWaves];
  __shared__ int ri[kWaves];
  __shared__ double red[kWaves];
  if (*fail) return;
  const FTask& t = tasks[blockIdx.x];
  RkRef out = t.tgt;
  const int m = t.m.s, n = t.n.s;
  // the dense block is a temporary: work on an LDS copy (in the storage
  // precision) when it fits, on the block itself otherwise
  __shared__ T dl[kAcaLds];
  T* D = (T*)t.A;
  int ld = t.lda;
  if ((int64_t)m * n <= kAcaLds) {
    stage_in(dl, m, (const T*)t.A, 1L, (long)t.lda, m, n);
    __syncthreads();
    D = dl;
    ld = m;
  }
  int maxk = m < n ? m : n;
  if (maxk > out.cap) maxk = out.cap;
  T* OA = (T*)out.a;
  T* OB = (T*)out.b;
  double est = 0.0;
  int nu = 0;
  for (; nu < maxk; ++nu) {
    ArgMax a{-1.0, -1};
    for (int e = threadIdx.x; e < m * n; e += kBlock)
      a = better_max(a, ArgMax{fabs((double)D[e % m + (size_t)(e / m) * ld]), e});
    a = block_argmax(a, rv, ri);
    if (a.i < 0) break;  // all-NaN block
    const int i = a.i % m, j = a.i / m;
    const double delta = (double)D[i + (size_t)j * ld];
    if (delta == 0.0) break;
    T* va = OA + (size_t)nu * out.la;
    T* vb = OB + (size_t)nu * out.lb;
    for (int r = threadIdx.x; r < m; r += kBlock) va[r] = D[r + (size_t)j * ld];
    for (int c = threadIdx.x; c < n; c += kBlock) vb[c] = (T)((double)D[i + (size_t)c * ld] / delta);
    __syncthreads();
    for (int e = threadIdx.x; e < m * n; e += kBlock) {
      const int r = e % m, c = e / m;
      D[r + (size_t)c * ld] -= va[r] * vb[c];
    }
    // estimate (the reference sums l < nu - 1): each wavefront takes the terms
    // l = w, w + 4, ... with wavefront sums, then one workgroup sum
    double cross = 0.0;
    {
      const int lane = threadIdx.x & (kWave - 1), wv = threadIdx.x / kWave;
      double part = 0.0;
      for (int l = wv; l + 1 < nu; l += kWaves) {
        double da = 0.0, db = 0.0;
        for (int r = lane; r < m; r += kWave) da += (double)va[r] * (double)OA[r + (size_t)l * out.la];
        for (int c = lane; c < n; c += kWave) db += (double)vb[c] * (double)OB[c + (size_t)l * out.lb];
        part += wave_sum(da) * wave_sum(db);
      }
      cross = block_sum(lane == 0 ? part : 0.0, red);
    }
    double an = 0.0, bn = 0.0;
    for (int r = threadIdx.x; r < m; r += kBlock) an += (double)va[r] * (double)va[r];
    for (int c = threadIdx.x; c < n; c += kBlock) bn += (double)vb[c] * (double)vb[c];
    an = block_sum(an, red);
    bn = block_sum(bn, red);
    est += 2.0 * cross + an * bn;
    if (an * bn < t.eps * t.eps * est) break;
  }
  __syncthreads();
  if (threadIdx.x == 0) *out.rank = nu;
}

}  // namespace

int64_t fadd_workspace(int m, int n, int K) {
  // dense path only when K >= min(m, n); QR path otherwise (K actual < min)
  const int mn = m < n ? m : n;
  int64_t dense = K >= mn ? (int64_t)m * n + (int64_t)mn * mn + 2 * (int64_t)mn : 0;
  const int kq = K < mn ? K : mn - 1;
  int64_t qr = (int64_t)(m + n) * kq + 6 * (int64_t)kq * kq + 5 * (int64_t)kq;
  return (std::max(dense, qr) + 127) / 64 * 64;
}
int64_t fadd_dense_workspace(int m, int n) {
  const int mn = m < n ? m : n;
  return ((int64_t)m * n + (int64_t)mn * mn + 2 * (int64_t)mn + 127) / 64 * 64;
}
int64_t hodlr_workspace(int cap) { return ((int64_t)5 * cap * cap + cap + 127) / 64 * 64; }
int64_t rkrk_workspace(int k1, int k2) {
  const int64_t a = k1 > k2 ? k1 : k2, b = k1 < k2 ? k1 : k2;
  // G, V, sig, order (+ pad), then the two output factors (a + b) x b
  return (a * b + b * b + 2 * b + 1 + (a + b) * b + 127) / 64 * 64;
}

namespace {
__global__ void k_fill_int(int* p, size_t n, int v) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    p[i] = v;
}
}  // namespace

void launch_fill_int(int* p, size_t n, int v, hipStream_t s) {
  if (n == 0) return;
  const size_t blocks = std::min<size_t>((n + 255) / 256, 4096);
  hipLaunchKernelGGL(k_fill_int, dim3((unsigned)blocks), dim3(256), 0, s, p, n, v);
  HMX_CHECK(hipGetLastError());
}

template <typename T>
void launch_task_kernel(int type, const FTask* tasks, int count, const GemmTile* tiles,
                        int ntiles, double* ws, int* flags, hipStream_t s) {
  if (count == 0) return;
  switch (type) {
    case T_GEMM:
      hipLaunchKernelGGL(k_gemm<T>, dim3(ntiles), dim3(256), 0, s, tasks, tiles, flags + 1);
      break;
    case T_POTRF:
      hipLaunchKernelGGL(k_potrf<T>, dim3(count), dim3(kBlock), 0, s, tasks, flags + 1);
      break;
    case T_TRSM:
      hipLaunchKernelGGL(k_trsm<T>, dim3(count * kTrsmSplit), dim3(kBlock), 0, s, tasks, flags + 1);
      break;
    case T_FADD:
      hipLaunchKernelGGL(k_fadd<T>, dim3(count), dim3(kBlock), 0, s, tasks, ws, flags);
      break;
    case T_FADD_DENSE:
      hipLaunchKernelGGL(k_fadd_dense<T>, dim3(count), dim3(kBlock), 0, s, tasks, ws, flags);
      break;
    case T_RKRK:
      hipLaunchKernelGGL(k_rkrk<T>, dim3(count), dim3(kBlock), 0, s, tasks, ws, flags);
      break;
    case T_ACAFULL:
      hipLaunchKernelGGL(k_acafull<T>, dim3(count), dim3(kBlock), 0, s, tasks, flags + 1);
      break;
    case T_ZERO:
      hipLaunchKernelGGL(k_zero<T>, dim3(count), dim3(256), 0, s, tasks, flags + 1);
      break;
    case T_AXPY:
      hipLaunchKernelGGL(k_axpy<T>, dim3(count), dim3(256), 0, s, tasks, flags + 1);
      break;
    case T_GETRF:
      hipLaunchKernelGGL(k_getrf<T>, dim3(count), dim3(kBlock), 0, s, tasks, flags + 1);
      break;
    case T_SUMK:
      hipLaunchKernelGGL(k_sumk<T>, dim3(count), dim3(256), 0, s, tasks, flags + 1);
      break;
    case T_LDLT:
      hipLaunchKernelGGL(k_ldlt<T>, dim3(count), dim3(kBlock), 0, s, tasks, flags + 1);
      break;
    case T_DSCALE:
      hipLaunchKernelGGL(k_dscale<T>, dim3(count * kDscaleSplit), dim3(256), 0, s, tasks, flags + 1);
      break;
    case T_HODLR:
      hipLaunchKernelGGL(k_hodlr<T>, dim3(count), dim3(kBlock), 0, s, tasks, ws, flags + 1);
      break;
    default:
      throw std::runtime_error("unknown task type");
  }
  HMX_CHECK(hipGetLastError());
}

template void launch_task_kernel<double>(int, const FTask*, int, const GemmTile*, int, double*, int*, hipStream_t);
template void launch_task_kernel<float>(int, const FTask*, int, const GemmTile*, int, double*, int*, hipStream_t);

}  // namespace hmx
