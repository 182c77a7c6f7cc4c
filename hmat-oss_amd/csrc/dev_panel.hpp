// Matrix-core helpers of the recompression and Rk-product kernels for one
// workgroup (kBlock threads): the Gram matrix of Householder reflectors, the
// cross product X^T Y of two tall panels, single 16 x 16 tiles, and a row-
// streamed product of a tall panel with a small matrix.
#pragma once
#include "dev_common.hpp"

namespace hmx {

// ---------------------------------------------------------------------------
// Gram matrix of Householder reflectors on the matrix cores
// ---------------------------------------------------------------------------
typedef double hmx_double4 __attribute__((ext_vector_type(4)));

// Entry (r, c) of the unit lower trapezoidal reflector matrix stored in V
// (m x K, ld ldv: reflector entries below the diagonal), zero outside.
__device__ inline double refl_entry(const double* V, int ldv, int m, int K, int r, int c) {
  const int rc = r < m ? r : m - 1, cc = c < K ? c : K - 1;
  const double x = V[rc + (size_t)cc * ldv];
  return (r >= m || c >= K) ? 0.0 : (r > c ? x : (r == c ? 1.0 : 0.0));
}

// S = V^T V (strict upper triangle, K x K, ld K) for the reflectors in V,
// with v_mfma_f64_16x16x4f64 over 4-row slabs: lane i + 16k holds entry
// (k, i) of a slab's 16-column block, which is both the A (= V^T) and the B
// operand; tile (a, b) of 16 x 16 comes back with D[i][j] in lane
// j + 16 (i mod 4), element i / 4.  The four wavefronts take interleaved
// slabs; their tiles go through 'part' (4 * kNb (kNb + 1) / 2 * 256
// doubles of LDS) and are summed in a fixed order.
template <int kNb>
__device__ void wg_reflector_gram(const double* V, int ldv, int m, int K, double* S, double* part) {
  constexpr int kTiles = kNb * (kNb + 1) / 2;
  const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
  const int kk = lane >> 4, ii = lane & 15;
  hmx_double4 acc[kTiles];
#pragma unroll
  for (int t = 0; t < kTiles; ++t) acc[t] = hmx_double4{0.0, 0.0, 0.0, 0.0};
  constexpr int kU = 4;  // slabs per step, loads issued together
  for (int r0 = 4 * w; r0 < m; r0 += 16 * kU) {
    double x[kU][kNb];
#pragma unroll
    for (int u = 0; u < kU; ++u)
#pragma unroll
      for (int b = 0; b < kNb; ++b) x[u][b] = refl_entry(V, ldv, m, K, r0 + 16 * u + kk, 16 * b + ii);
#pragma unroll
    for (int u = 0; u < kU; ++u) {
      int t = 0;
#pragma unroll
      for (int a = 0; a < kNb; ++a)
#pragma unroll
        for (int b = a; b < kNb; ++b, ++t)
          acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(x[u][a], x[u][b], acc[t], 0, 0, 0);
    }
  }
  __syncthreads();
#pragma unroll
  for (int t = 0; t < kTiles; ++t)
#pragma unroll
    for (int v = 0; v < 4; ++v) part[((size_t)(w * kTiles + t) * 4 + v) * kWave + lane] = acc[t][v];
  __syncthreads();
  for (int e = threadIdx.x; e < K * K; e += kBlock) {
    const int i = e % K, j = e / K;
    if (i >= j) continue;
    const int a = i >> 4, b = j >> 4, si = i & 15, sj = j & 15;
    const int t = a * kNb - a * (a - 1) / 2 + (b - a);
    const int ln = sj + 16 * (si & 3), v = si >> 2;
    double s = 0.0;
#pragma unroll
    for (int q = 0; q < kWaves; ++q) s += part[((size_t)(q * kTiles + t) * 4 + v) * kWave + ln];
    S[i + (size_t)j * K] = s;
  }
  __syncthreads();
}

// doubles of 'part' for K columns
__host__ __device__ constexpr int gram_part(int K) {
  return 4 * (((K + 15) / 16) * ((K + 15) / 16 + 1) / 2) * 256;
}

__device__ inline void wg_reflector_gram_any(const double* V, int ldv, int m, int K, double* S, double* part) {
  switch ((K + 15) / 16) {
    case 1: wg_reflector_gram<1>(V, ldv, m, K, S, part); break;
    case 2: wg_reflector_gram<2>(V, ldv, m, K, S, part); break;
    case 3: wg_reflector_gram<3>(V, ldv, m, K, S, part); break;
    default: wg_reflector_gram<4>(V, ldv, m, K, S, part); break;
  }
}

}  // namespace hmx

namespace hmx {

// G = X^T Y (kx x ky; G[i + j ldg], or G[j + i ldg] when 'trans') for X
// (p x kx, ld ldx) and Y (p x ky, ld ldy) of element type T, on the matrix
// cores as wg_reflector_gram: 16-column blocks of X are the A operand, of Y
// the B operand, tile (a, b) accumulates over 4-row slabs, the wavefronts'
// tiles meet in 'part' (4 kNa kNb 256 doubles of LDS).
template <typename T, int kNa, int kNb>
__device__ void wg_xty(const T* X, int ldx, int kx, const T* Y, int ldy, int ky, int p, double* G, int ldg,
                       bool trans, double* part) {
  constexpr int kTiles = kNa * kNb;
  const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
  const int kk = lane >> 4, ii = lane & 15;
  hmx_double4 acc[kTiles];
#pragma unroll
  for (int t = 0; t < kTiles; ++t) acc[t] = hmx_double4{0.0, 0.0, 0.0, 0.0};
  constexpr int kU = 4;
  for (int r0 = 4 * w; r0 < p; r0 += 16 * kU) {
    double x[kU][kNa], y[kU][kNb];
#pragma unroll
    for (int u = 0; u < kU; ++u) {
      const int r = r0 + 16 * u + kk, rc = r < p ? r : p - 1;
#pragma unroll
      for (int a = 0; a < kNa; ++a) {
        const int c = 16 * a + ii, cc = c < kx ? c : kx - 1;
        const double v = (double)X[rc + (size_t)cc * ldx];
        x[u][a] = (r < p && c < kx) ? v : 0.0;
      }
#pragma unroll
      for (int b = 0; b < kNb; ++b) {
        const int c = 16 * b + ii, cc = c < ky ? c : ky - 1;
        const double v = (double)Y[rc + (size_t)cc * ldy];
        y[u][b] = (r < p && c < ky) ? v : 0.0;
      }
    }
#pragma unroll
    for (int u = 0; u < kU; ++u)
#pragma unroll
      for (int a = 0; a < kNa; ++a)
#pragma unroll
        for (int b = 0; b < kNb; ++b)
          acc[a * kNb + b] = __builtin_amdgcn_mfma_f64_16x16x4f64(x[u][a], y[u][b], acc[a * kNb + b], 0, 0, 0);
  }
  __syncthreads();
#pragma unroll
  for (int t = 0; t < kTiles; ++t)
#pragma unroll
    for (int v = 0; v < 4; ++v) part[((size_t)(w * kTiles + t) * 4 + v) * kWave + lane] = acc[t][v];
  __syncthreads();
  for (int e = threadIdx.x; e < kx * ky; e += kBlock) {
    const int i = e % kx, j = e / kx;
    const int t = (i >> 4) * kNb + (j >> 4), si = i & 15, sj = j & 15;
    const int ln = sj + 16 * (si & 3), v = si >> 2;
    double s = 0.0;
#pragma unroll
    for (int q = 0; q < kWaves; ++q) s += part[((size_t)(q * kTiles + t) * 4 + v) * kWave + ln];
    G[trans ? j + (size_t)i * ldg : i + (size_t)j * ldg] = s;
  }
  __syncthreads();
}

// doubles of 'part' for wg_xty_any
__host__ __device__ constexpr int xty_part(int kx, int ky) { return 4 * ((kx + 15) / 16) * ((ky + 15) / 16) * 256; }

// kx, ky <= 32; returns false (nothing done) otherwise.
template <typename T>
__device__ inline bool wg_xty_any(const T* X, int ldx, int kx, const T* Y, int ldy, int ky, int p, double* G,
                                  int ldg, bool trans, double* part) {
  const int na = (kx + 15) / 16, nb = (ky + 15) / 16;
  if (na == 1 && nb == 1) wg_xty<T, 1, 1>(X, ldx, kx, Y, ldy, ky, p, G, ldg, trans, part);
  else if (na == 1 && nb == 2) wg_xty<T, 1, 2>(X, ldx, kx, Y, ldy, ky, p, G, ldg, trans, part);
  else if (na == 2 && nb == 1) wg_xty<T, 2, 1>(X, ldx, kx, Y, ldy, ky, p, G, ldg, trans, part);
  else if (na == 2 && nb == 2) wg_xty<T, 2, 2>(X, ldx, kx, Y, ldy, ky, p, G, ldg, trans, part);
  else return false;
  return true;
}

// out (rows x nk, ld ldo) = M (rows x k, ld ldm) C (k x nk, ld ldc; LDS or
// global): each thread owns rows and streams a row of M once per eight
// output columns.
template <typename T, typename TO = T>
__device__ void rows_times_small(const T* M, int ldm, int rows, int k, const double* C, int ldc, int nk, TO* out,
                                 int ldo) {
  constexpr int kC = 8;
  for (int c0 = 0; c0 < nk; c0 += kC) {
    const double* cc[kC];
#pragma unroll
    for (int b = 0; b < kC; ++b) cc[b] = C + (size_t)(c0 + b < nk ? c0 + b : nk - 1) * ldc;
    for (int i = threadIdx.x; i < rows; i += kBlock) {
      double acc[kC];
#pragma unroll
      for (int b = 0; b < kC; ++b) acc[b] = 0.0;
      const T* mi = M + i;
      int l = 0;
      for (; l + 8 <= k; l += 8) {
        double v[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) v[q] = (double)mi[(size_t)(l + q) * ldm];
#pragma unroll
        for (int q = 0; q < 8; ++q)
#pragma unroll
          for (int b = 0; b < kC; ++b) acc[b] += v[q] * cc[b][l + q];
      }
      for (; l < k; ++l) {
        const double v = (double)mi[(size_t)l * ldm];
#pragma unroll
        for (int b = 0; b < kC; ++b) acc[b] += v * cc[b][l];
      }
#pragma unroll
      for (int b = 0; b < kC; ++b)
        if (c0 + b < nk) out[i + (size_t)(c0 + b) * ldo] = (TO)acc[b];
    }
  }
  __syncthreads();
}

}  // namespace hmx

namespace hmx {

// One 16 x 16 tile D = A B on the matrix cores, inner dimension kmax <= 16
// (entries k >= kmax count as zero; 1 <= kmax): a(ii, k) gives A[ii][k],
// b(k, jj) gives B[k][jj] for k < kmax (both must return 0 outside their
// matrices, best by a clamped index and a select so that the loads stay
// unconditional).  The wavefront's
// lane jj + 16 (ii mod 4) receives D[ii][jj] in element ii / 4.
template <typename FA, typename FB>
__device__ inline hmx_double4 mfma_tile16(FA a, FB b, int kmax) {
  const int lane = threadIdx.x & (kWave - 1), ii = lane & 15, kk = lane >> 4;
  // all eight operand loads first (clamped k, then a select), then the MFMAs
  double av[4], bv[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int k = 4 * q + kk, kc = k < kmax ? k : kmax - 1;
    av[q] = a(ii, kc);
    bv[q] = b(kc, ii);
  }
  hmx_double4 acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const bool in = 4 * q + kk < kmax;
    acc = __builtin_amdgcn_mfma_f64_16x16x4f64(in ? av[q] : 0.0, in ? bv[q] : 0.0, acc, 0, 0, 0);
  }
  return acc;
}

}  // namespace hmx
