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

constexpr int kQrChunk = 16;  // trailing columns reduced together

// w[c] = sum_{i >= j} v_i A[i, c] for c in [c0, c1), v_j = 1, v_i = V[i] below:
// every thread accumulates its rows for a chunk of columns, one wave_sum per
// column and a cross-wave sum through LDS 'red' (kWaves * kQrChunk doubles).
__device__ inline void house_dots(const double* v, int j, int m, const double* A, int lda, int c0,
                                  int c1, double* w, double* red) {
  const int lane = threadIdx.x & (kWave - 1), wv = threadIdx.x / kWave;
  for (int cb = c0; cb < c1; cb += kQrChunk) {
    const int nc = c1 - cb < kQrChunk ? c1 - cb : kQrChunk;
    double part[kQrChunk];
#pragma unroll
    for (int q = 0; q < kQrChunk; ++q) part[q] = 0.0;
    for (int i = j + threadIdx.x; i < m; i += kBlock) {
      const double vi = (i == j) ? 1.0 : v[i];
#pragma unroll
      for (int q = 0; q < kQrChunk; ++q)
        if (q < nc) part[q] += vi * A[i + (size_t)(cb + q) * lda];
    }
#pragma unroll
    for (int q = 0; q < kQrChunk; ++q) {
      const double sum = wave_sum(part[q]);
      if (lane == 0 && q < nc) red[wv * kQrChunk + q] = sum;
    }
    __syncthreads();
    for (int q = threadIdx.x; q < nc; q += kBlock) {
      double sum = 0.0;
#pragma unroll
      for (int ww = 0; ww < kWaves; ++ww) sum += red[ww * kQrChunk + q];
      w[cb + q] = sum;
    }
    __syncthreads();
  }
}

// Householder QR, LAPACK dgeqrf conventions (dlarfg: beta = -sign(alpha)*norm,
// tau = (beta - alpha) / beta, v = x / (alpha - beta), v[0] = 1 implicit).
// On exit R is in the upper triangle, v_j below the diagonal, tau[j] set.
// 'w' is scratch of k doubles, 'red' of kWaves * kQrChunk doubles.
__device__ inline void house_qr(double* A, int lda, int m, int k, double* tau, double* red,
                                double* w) {
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
    if (t != 0.0 && j + 1 < k) {
      house_dots(col, j, m, A, lda, j + 1, k, w, red);
      const int nc = k - j - 1, len = m - j;
      for (int e = threadIdx.x; e < len * nc; e += kBlock) {
        const int i = j + e % len, c = j + 1 + e / len;
        const double vi = (i == j) ? 1.0 : col[i];
        A[i + (size_t)c * lda] -= t * vi * w[c];
      }
      __syncthreads();
    }
  }
  if (threadIdx.x == 0)
    for (int j = steps; j < k; ++j) tau[j] = 0.0;
  __syncthreads();
}

// Y (m x r, ldy) <- Q Y with Q = H_0 ... H_{nref-1} from house_qr.
// 'w' scratch of r doubles, 'red' of kWaves * kQrChunk doubles.
__device__ inline void apply_q(const double* V, int ldv, int m, int nref, const double* tau,
                               double* Y, int ldy, int r, double* red, double* w) {
  for (int j = nref - 1; j >= 0; --j) {
    const double t = tau[j];
    if (t == 0.0) continue;
    const double* v = V + (size_t)j * ldv;
    house_dots(v, j, m, Y, ldy, 0, r, w, red);
    const int len = m - j;
    for (int e = threadIdx.x; e < len * r; e += kBlock) {
      const int i = j + e % len, c = e / len;
      const double vi = (i == j) ? 1.0 : v[i];
      Y[i + (size_t)c * ldy] -= t * vi * w[c];
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
  // ||G||_F^2: columns below 1e-15 ||G|| carry no information for any
  // truncation threshold used here; pairs of them are not rotated (a purely
  // relative test never settles on such noise-level columns)
  __shared__ double fro_s;
  {
    double f = 0.0;
    for (int e = threadIdx.x; e < rows * cols; e += kBlock) {
      const double x = G[e % rows + (size_t)(e / rows) * ldg];
      f += x * x;
    }
    f = wave_sum(f);
    if (threadIdx.x == 0) fro_s = 0.0;
    __syncthreads();
    if ((threadIdx.x & (kWave - 1)) == 0) atomicAdd(&fro_s, f);
    __syncthreads();
  }
  const double floor2 = 1e-30 * fro_s;
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
        if (!active || c2 == 0.0 || !(fabs(c2) > tol * sqrt(a * b)) || !(fabs(c2) > floor2)) continue;
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
