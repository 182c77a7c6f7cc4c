// One-workgroup dense linear algebra used inside the batched task kernels:
// Householder QR of a tall panel, application of its Q, one-sided (Hestenes)
// Jacobi SVD, and the singular-value ordering / truncation rule.
//
// All routines are executed cooperatively by a whole workgroup of kBlock
// threads, work on column-major matrices that may live in LDS or in global
// memory (generic pointers), and end with a __syncthreads().  Arithmetic is
// FP64; callers convert at the boundary.
#pragma once
#include "dev_common.hpp"

namespace hmx {

// Householder QR, LAPACK dgeqrf conventions (dlarfg: beta = -sign(alpha)*norm,
// tau = (beta - alpha) / beta, v = x / (alpha - beta), v[0] = 1 implicit).
// On exit R is in the upper triangle, v_j below the diagonal, tau[j] set.
__device__ inline void house_qr(double* A, int lda, int m, int k, double* tau, double* red) {
  const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
  const int steps = m < k ? m : k;
  for (int j = 0; j < steps; ++j) {
    double* col = A + (size_t)j * lda;
    double s = 0.0;
    for (int i = j + 1 + threadIdx.x; i < m; i += kBlock) s += col[i] * col[i];
    s = block_sum(s, red);
    const double alpha = col[j];
    double t = 0.0, scal = 0.0, beta = alpha;
    if (s != 0.0) {
      beta = -copysign(sqrt(alpha * alpha + s), alpha);
      t = (beta - alpha) / beta;
      scal = 1.0 / (alpha - beta);
    }
    __syncthreads();
    if (s != 0.0)
      for (int i = j + 1 + threadIdx.x; i < m; i += kBlock) col[i] *= scal;
    if (threadIdx.x == 0) {
      col[j] = beta;
      tau[j] = t;
    }
    __syncthreads();
    if (t != 0.0) {
      for (int c = j + 1 + w; c < k; c += kWaves) {
        double* ac = A + (size_t)c * lda;
        double d = 0.0;
        for (int i = j + 1 + lane; i < m; i += kWave) d += col[i] * ac[i];
        d = wave_sum(d) + ac[j];
        const double f = t * d;
        if (lane == 0) ac[j] -= f;
        for (int i = j + 1 + lane; i < m; i += kWave) ac[i] -= f * col[i];
      }
    }
    __syncthreads();
  }
  if (threadIdx.x == 0)
    for (int j = steps; j < k; ++j) tau[j] = 0.0;
  __syncthreads();
}

// Y (m x r, ldy) <- Q Y with Q = H_0 ... H_{nref-1} from house_qr.
__device__ inline void apply_q(const double* V, int ldv, int m, int nref, const double* tau,
                               double* Y, int ldy, int r) {
  const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
  for (int j = nref - 1; j >= 0; --j) {
    const double t = tau[j];
    if (t != 0.0) {
      const double* v = V + (size_t)j * ldv;
      for (int c = w; c < r; c += kWaves) {
        double* y = Y + (size_t)c * ldy;
        double d = 0.0;
        for (int i = j + 1 + lane; i < m; i += kWave) d += v[i] * y[i];
        d = wave_sum(d) + y[j];
        const double f = t * d;
        if (lane == 0) y[j] -= f;
        for (int i = j + 1 + lane; i < m; i += kWave) y[i] -= f * v[i];
      }
    }
    __syncthreads();
  }
}

// Sum over an aligned group of g lanes (g a power of two <= 64).
template <typename T>
__device__ inline T group_sum(T v, int g) {
  for (int o = g >> 1; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
  return v;
}

// One-sided Jacobi SVD of G (rows x cols, ldg): on exit G = U * Sigma
// (columns mutually orthogonal, unsorted) and V (cols x cols, ldv) holds the
// accumulated rotations so that G_in = (U Sigma) V^T.  'flag' is a shared int.
// Column pairs of one round (round-robin ordering) are independent; each is
// handled by a group of g lanes sized to the column length, so that a
// wavefront works on 64 / g pairs at once.
__device__ inline void jacobi_svd(double* G, int ldg, int rows, int cols, double* V, int ldv,
                                  int* flag, int max_sweeps = 40) {
  for (int e = threadIdx.x; e < cols * cols; e += kBlock) {
    const int i = e % cols, j = e / cols;
    V[i + (size_t)j * ldv] = (i == j) ? 1.0 : 0.0;
  }
  __syncthreads();
  if (cols < 2) return;
  const int longest = rows > cols ? rows : cols;
  int g = 8;
  while (g < 64 && g * 4 < longest) g <<= 1;
  const int gl = threadIdx.x % g, grp = threadIdx.x / g, ngrp = kBlock / g;
  const int np = (cols + 1) & ~1;  // players, padded to even
  const int rounds = np - 1, half = np / 2;
  const double tol = 1e-15;
  for (int sweep = 0; sweep < max_sweeps; ++sweep) {
    if (threadIdx.x == 0) *flag = 0;
    __syncthreads();
    for (int rd = 0; rd < rounds; ++rd) {
      for (int pk0 = 0; pk0 < half; pk0 += ngrp) {
        const int pk = pk0 + grp;
        int p = -1, q = -1;
        if (pk < half) {
          if (pk == 0) {
            p = np - 1;
            q = rd;
          } else {
            p = (rd + pk) % (np - 1);
            q = (rd - pk + (np - 1)) % (np - 1);
          }
          if (p > q) {
            const int tmp = p;
            p = q;
            q = tmp;
          }
        }
        const bool active = pk < half && q < cols;
        double a = 0.0, b = 0.0, c2 = 0.0;
        double* gp = G + (size_t)(active ? p : 0) * ldg;
        double* gq = G + (size_t)(active ? q : 0) * ldg;
        if (active)
          for (int i = gl; i < rows; i += g) {
            const double x = gp[i], y = gq[i];
            a += x * x;
            b += y * y;
            c2 += x * y;
          }
        // every lane of the wavefront takes part in the shuffles
        a = group_sum(a, g);
        b = group_sum(b, g);
        c2 = group_sum(c2, g);
        if (!active || c2 == 0.0 || !(fabs(c2) > tol * sqrt(a * b))) continue;
        const double zeta = (b - a) / (2.0 * c2);
        const double t = copysign(1.0, zeta) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
        const double c = 1.0 / sqrt(1.0 + t * t), s = c * t;
        for (int i = gl; i < rows; i += g) {
          const double x = gp[i], y = gq[i];
          gp[i] = c * x - s * y;
          gq[i] = s * x + c * y;
        }
        double* vp = V + (size_t)p * ldv;
        double* vq = V + (size_t)q * ldv;
        for (int i = gl; i < cols; i += g) {
          const double x = vp[i], y = vq[i];
          vp[i] = c * x - s * y;
          vq[i] = s * x + c * y;
        }
        if (gl == 0) *flag = 1;
      }
      __syncthreads();
    }
    if (*flag == 0) break;
    __syncthreads();
  }
  __syncthreads();
}

// Column norms of G into sig[0..cols), then order[] = permutation sorting
// sig descending (ties by index).  Returns nothing; ends with a barrier.
__device__ inline void sorted_norms(const double* G, int ldg, int rows, int cols, double* sig,
                                    int* order) {
  const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
  for (int j = w; j < cols; j += kWaves) {
    const double* g = G + (size_t)j * ldg;
    double s = 0.0;
    for (int i = lane; i < rows; i += kWave) s += g[i] * g[i];
    s = wave_sum(s);
    // a non-finite norm (NaN input) is ordered as zero so that 'order' is
    // always a permutation of [0, cols)
    if (lane == 0) sig[j] = (s >= 0.0 && s <= 1.79e308) ? sqrt(s) : 0.0;
  }
  __syncthreads();
  for (int j = threadIdx.x; j < cols; j += kBlock) {
    const double sj = sig[j];
    int pos = 0;
    for (int i = 0; i < cols; ++i) {
      const double si = sig[i];
      pos += (si > sj) || (si == sj && i < j);
    }
    order[pos] = j;
  }
  __syncthreads();
}

// findK scalar_array.cpp:98-116: number of sigma (descending) with
// sigma[i] > eps * sigma[0].
__device__ inline int find_k(const double* sig, const int* order, int n, double eps) {
  if (n == 0) return 0;
  const double thr = sig[order[0]] * eps;
  int k = 0;
  while (k < n && !(sig[order[k]] <= thr)) ++k;
  return k;
}

}  // namespace hmx
