// One-workgroup dense linear algebra used inside the batched task kernels:
// Householder QR of a tall panel, application of its Q, one-sided (Hestenes)
// Jacobi SVD, and the singular-value ordering / truncation rule.
//
// The routines work on column-major matrices that may live in LDS or in
// global memory (generic pointers) and end with a synchronisation of the
// threads that ran them.  Arithmetic is FP64; callers convert at the boundary.
#pragma once
#include "dev_common.hpp"

namespace hmx {

constexpr int kQrChunk = 16;  // size factor of the reduction scratch kept by callers

// The routines below are written for a "team" of kLanes threads: one
// wavefront (kLanes == kWave, synchronised by wave_barrier: LDS accesses of
// one wavefront complete in program order) or the whole workgroup
// (kLanes == kBlock, __syncthreads).  Work items (columns, column pairs) are
// spread over aligned groups of g lanes, g as large as the item count allows;
// a group reduces its partial sums with group_sum, so a reduction costs
// log2(g) register exchanges instead of a round trip through LDS.

template <int kLanes>
__device__ inline void team_sync() {
  if constexpr (kLanes == kWave)
    __builtin_amdgcn_wave_barrier();
  else
    __syncthreads();
}

// Thread index within its team; teams are aligned groups of kLanes threads
// (a power of two: a wavefront, half or all of the workgroup).  Teams
// smaller than the workgroup but larger than a wavefront synchronise with
// workgroup barriers, so all teams must pass the same number of them.
template <int kLanes>
__device__ inline int team_tid() {
  return (int)(threadIdx.x & (kLanes - 1));
}

// Largest power of two g <= kWave with g * items <= lanes (at least 1).
__device__ inline int group_lanes(int items, int lanes) {
  int g = kWave;
  while (g > 1 && g * items > lanes) g >>= 1;
  return g;
}

// sum_{i = i0, i0+g, ... < m} x[i] y[i]; four loads in flight per operand.
__device__ inline double strided_dot(const double* x, const double* y, int i0, int m, int g) {
  double s0 = 0.0, s1 = 0.0;
  int i = i0;
  for (; i + 3 * g < m; i += 4 * g) {
    const double x0 = x[i], x1 = x[i + g], x2 = x[i + 2 * g], x3 = x[i + 3 * g];
    const double y0 = y[i], y1 = y[i + g], y2 = y[i + 2 * g], y3 = y[i + 3 * g];
    s0 += x0 * y0;
    s1 += x1 * y1;
    s0 += x2 * y2;
    s1 += x3 * y3;
  }
  for (; i < m; i += g) s0 += x[i] * y[i];
  return s0 + s1;
}

// y[i] -= f x[i] for i = i0, i0+g, ... < m (x and y distinct columns).
__device__ inline void strided_axpy(double* y, double f, const double* x, int i0, int m, int g) {
  int i = i0;
  for (; i + 3 * g < m; i += 4 * g) {
    const double x0 = x[i], x1 = x[i + g], x2 = x[i + 2 * g], x3 = x[i + 3 * g];
    const double y0 = y[i], y1 = y[i + g], y2 = y[i + 2 * g], y3 = y[i + 3 * g];
    y[i] = y0 - f * x0;
    y[i + g] = y1 - f * x1;
    y[i + 2 * g] = y2 - f * x2;
    y[i + 3 * g] = y3 - f * x3;
  }
  for (; i < m; i += g) y[i] -= f * x[i];
}

// Sum over the whole team; every thread gets the same value.  'red' holds
// kWaves doubles (workgroup team only).
template <int kLanes>
__device__ inline double team_sum(double v, double* red) {
  v = wave_sum(v);
  if constexpr (kLanes == kWave) {
    return v;
  } else {
    const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
    __syncthreads();
    if (lane == 0) red[w] = v;
    __syncthreads();
    double s = 0.0;
#pragma unroll
    for (int q = 0; q < kWaves; ++q) s += red[q];
    return s;
  }
}

// Householder QR, LAPACK dgeqrf conventions (dlarfg: beta = -sign(alpha)*norm,
// tau = (beta - alpha) / beta, v = x / (alpha - beta), v[0] = 1 implicit).
// On exit R is in the upper triangle, v_j below the diagonal, tau[j] set.
// The whole team forms each reflector (norm, scaling); trailing column c
// belongs to lane group c mod (kLanes / g), which applies the reflector to it.
template <int kLanes>
__device__ inline void team_house_qr(double* A, int lda, int m, int k, double* tau) {
  __shared__ double red[kWaves];
  const int tid = team_tid<kLanes>();
  const int steps = m < k ? m : k;
  const int g = group_lanes(k, kLanes), ngrp = kLanes / g, grp = tid / g, sub = tid % g;
  for (int j = 0; j < steps; ++j) {
    double* v = A + (size_t)j * lda;
    double s = strided_dot(v, v, j + 1 + tid, m, kLanes);
    s = team_sum<kLanes>(s, red);
    const double alpha = v[j];
    double t = 0.0, beta = alpha;
    if (s != 0.0) {
      beta = -copysign(sqrt(alpha * alpha + s), alpha);
      t = (beta - alpha) / beta;
      const double scal = 1.0 / (alpha - beta);
      for (int i = j + 1 + tid; i < m; i += kLanes) v[i] *= scal;
    }
    team_sync<kLanes>();
    if (tid == 0) {
      tau[j] = t;
      v[j] = beta;
    }
    team_sync<kLanes>();
    if (t != 0.0) {
      int c = j + 1 + ((grp - (j + 1) % ngrp) + ngrp) % ngrp;
      for (; c < k; c += ngrp) {
        double* a = A + (size_t)c * lda;
        double d = strided_dot(v, a, j + 1 + sub, m, g);
        d = group_sum(d, g) + a[j];
        const double f = t * d;
        strided_axpy(a, f, v, j + 1 + sub, m, g);
        if (sub == 0) a[j] -= f;
      }
    }
    team_sync<kLanes>();
  }
  if (tid == 0)
    for (int j = steps; j < k; ++j) tau[j] = 0.0;
  team_sync<kLanes>();
}

// Householder QR as team_house_qr with row ownership: thread t of the team
// keeps rows t + kLanes * r (r < kQrRows) of the current column in registers,
// and trailing columns are updated four at a time, their loads issued
// together and their four dot products reduced in one exchange.  A step thus
// costs a few memory round trips whether the panel lives in LDS or in global
// memory.  Requires m <= R * kLanes; 'red' holds 4 * kWaves doubles of
// shared scratch (workgroup team only).
constexpr int kQrRows = 8;

template <int kLanes, int kC>
__device__ inline void team_sum_n(double (&v)[kC], double* red) {
#pragma unroll
  for (int c = 0; c < kC; ++c) v[c] = wave_sum(v[c]);
  if constexpr (kLanes != kWave) {
    constexpr int kTeamWaves = kLanes / kWave;
    const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
    const int w0 = w - w % kTeamWaves;  // first wavefront of the team
    __syncthreads();
    if (lane == 0)
#pragma unroll
      for (int c = 0; c < kC; ++c) red[w * kC + c] = v[c];
    __syncthreads();
#pragma unroll
    for (int c = 0; c < kC; ++c) {
      double s = 0.0;
#pragma unroll
      for (int q = 0; q < kTeamWaves; ++q) s += red[(w0 + q) * kC + c];
      v[c] = s;
    }
  }
}

// Element type of a panel known to be in global memory: loads and stores then
// use a uniform base and 32-bit offsets instead of 64-bit flat addresses.
typedef __attribute__((address_space(1))) double gdouble;

template <int kLanes, int R, typename E = double>
__device__ inline void team_house_qr_rows(E* A, int lda, int m, int k, double* tau) {
  __shared__ double red[4 * kWaves];
  const int tid = team_tid<kLanes>();
  const int steps = m < k ? m : k;
  // rows of this thread; loads use a clamped row and a select instead of a
  // branch, so all of them are in flight together
  int row[R], rc[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    row[r] = tid + kLanes * r;
    rc[r] = row[r] < m ? row[r] : m - 1;
  }
  for (int j = 0; j < steps; ++j) {
    E* vcol = A + (size_t)j * lda;
    double v[R];
    double s1[1] = {0.0};
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const double x = vcol[rc[r]];
      v[r] = (row[r] > j && row[r] < m) ? x : 0.0;
      s1[0] += v[r] * v[r];
    }
    team_sum_n<kLanes, 1>(s1, red);
    const double s = s1[0];
    const double alpha = vcol[j];
    double t = 0.0, beta = alpha;
    if (s != 0.0) {
      beta = -copysign(sqrt(alpha * alpha + s), alpha);
      t = (beta - alpha) / beta;
      const double scal = 1.0 / (alpha - beta);
#pragma unroll
      for (int r = 0; r < R; ++r) {
        v[r] *= scal;
        if (row[r] > j && row[r] < m) vcol[row[r]] = v[r];
      }
    }
    team_sync<kLanes>();
    if (tid == 0) {
      tau[j] = t;
      vcol[j] = beta;
    }
    {
      // also when t == 0 (a no-op then), so that every team passes the same barriers
      for (int c0 = j + 1; c0 < k; c0 += 4) {
        double a[4][R];
        double d[4], aj[4];
#pragma unroll
        for (int c = 0; c < 4; ++c) {
          const int cc = c0 + c < k ? c0 + c : k - 1;
          const E* ac = A + (size_t)cc * lda;
          aj[c] = ac[j];  // read by all before thread 0 rewrites it
#pragma unroll
          for (int r = 0; r < R; ++r) a[c][r] = ac[rc[r]];
        }
#pragma unroll
        for (int c = 0; c < 4; ++c) {
          d[c] = 0.0;
#pragma unroll
          for (int r = 0; r < R; ++r) d[c] += v[r] * a[c][r];  // v is 0 outside rows (j, m)
        }
        team_sum_n<kLanes, 4>(d, red);
        team_sync<kLanes>();
#pragma unroll
        for (int c = 0; c < 4; ++c) {
          if (c0 + c >= k) continue;
          E* ac = A + (size_t)(c0 + c) * lda;
          const double f = t * (d[c] + aj[c]);
#pragma unroll
          for (int r = 0; r < R; ++r)
            if (row[r] > j && row[r] < m) ac[row[r]] = a[c][r] - f * v[r];
          if (tid == 0) ac[j] = aj[c] - f;
        }
      }
    }
    team_sync<kLanes>();
  }
  if (tid == 0)
    for (int j = steps; j < k; ++j) tau[j] = 0.0;
  team_sync<kLanes>();
}

// Picks the rows-per-thread count for m (m <= kQrRows * kLanes for a
// wavefront, 2 * kQrRows * kLanes for larger teams).
template <int kLanes>
__device__ inline void team_house_qr_rows(double* A, int lda, int m, int k, double* tau) {
  if (m <= 2 * kLanes)
    team_house_qr_rows<kLanes, 2>(A, lda, m, k, tau);
  else if (m <= 4 * kLanes)
    team_house_qr_rows<kLanes, 4>(A, lda, m, k, tau);
  else if (kLanes == kWave || m <= 8 * kLanes)
    team_house_qr_rows<kLanes, kQrRows>(A, lda, m, k, tau);
  else
    team_house_qr_rows<kLanes, 2 * kQrRows>(A, lda, m, k, tau);
}

// The rows QR of a panel in global memory (m <= 16 kLanes).
template <int kLanes>
__device__ inline void team_house_qr_rows_global(double* A, int lda, int m, int k, double* tau) {
  gdouble* G = (gdouble*)A;
  if (m <= 2 * kLanes)
    team_house_qr_rows<kLanes, 2, gdouble>(G, lda, m, k, tau);
  else if (m <= 4 * kLanes)
    team_house_qr_rows<kLanes, 4, gdouble>(G, lda, m, k, tau);
  else if (m <= 8 * kLanes)
    team_house_qr_rows<kLanes, kQrRows, gdouble>(G, lda, m, k, tau);
  else
    team_house_qr_rows<kLanes, 2 * kQrRows, gdouble>(G, lda, m, k, tau);
}

// Householder QR of two global panels at once, one per half of the
// workgroup (rows <= 16 * 128 each, k columns each, k <= min(m, n)).  Both
// halves take the same k steps, so their workgroup barriers pair up.
__device__ inline void house_qr_pair(double* A, int lda, int m, double* tau_a, double* B, int ldb, int n,
                                     double* tau_b, int k) {
  if (threadIdx.x < kBlock / 2)
    team_house_qr_rows_global<kBlock / 2>(A, lda, m, k, tau_a);
  else
    team_house_qr_rows_global<kBlock / 2>(B, ldb, n, k, tau_b);
  __syncthreads();
}

// Y (m x r, ldy) <- Q Y with Q = H_0 ... H_{nref-1} from team_house_qr;
// column c of Y belongs to lane group c mod (kLanes / g).
template <int kLanes>
__device__ inline void team_apply_q(const double* V, int ldv, int m, int nref, const double* tau,
                                    double* Y, int ldy, int r) {
  const int tid = team_tid<kLanes>();
  const int g = group_lanes(r, kLanes), ngrp = kLanes / g, grp = tid / g, sub = tid % g;
  for (int c = grp; c < r; c += ngrp) {
    double* y = Y + (size_t)c * ldy;
    for (int j = nref - 1; j >= 0; --j) {
      const double t = tau[j];
      if (t == 0.0) continue;
      const double* v = V + (size_t)j * ldv;
      double d = strided_dot(v, y, j + 1 + sub, m, g);
      d = group_sum(d, g) + y[j];
      const double f = t * d;
      strided_axpy(y, f, v, j + 1 + sub, m, g);
      if (sub == 0) y[j] -= f;
      if (g > 1) __builtin_amdgcn_wave_barrier();
    }
  }
  team_sync<kLanes>();
}

// T of team_larft from the strict upper triangle of S = V^T V: T[j, j] = tau_j,
// T[l, j] = -tau_j sum_{r=l}^{j-1} T[l, r] S[r, j].  The recurrence runs on
// the team's first wavefront; row l of T is written by lane l mod 64 only.
template <int kLanes>
__device__ inline void larft_recurrence(int k, const double* tau, const double* S, double* T) {
  const int tid = team_tid<kLanes>();
  if (tid < kWave) {
    for (int j = 0; j < k; ++j) {
      const double tj = tau[j];
      for (int l = tid; l < j; l += kWave) {
        double a = 0.0;
        for (int r = l; r < j; ++r) a += T[l + (size_t)r * k] * S[r + (size_t)j * k];
        T[l + (size_t)j * k] = -tj * a;
      }
      if (tid == (j & (kWave - 1))) T[j + (size_t)j * k] = tj;
    }
  }
  team_sync<kLanes>();
}

// Triangular factor T (k x k, ld k, upper) of the block reflector
// H_0 H_1 ... H_{k-1} = I - V T V^T of team_house_qr (LAPACK dlarft, forward,
// columnwise): S = V^T V (strict upper part, into S, k x k), then
// T[l, j] = -tau_j sum_{r=l}^{j-1} T[l, r] S[r, j], T[j, j] = tau_j.
// The recurrence runs on the team's first wavefront; row l of T is written
// by lane l mod 64 only, so its steps need no cross-lane exchange.
template <int kLanes>
__device__ inline void team_larft(const double* V, int ldv, int m, int k, const double* tau,
                                  double* S, double* T) {
  const int tid = team_tid<kLanes>();
  const int npairs = k * (k - 1) / 2;
  if (npairs > 0) {
    const int g = group_lanes(npairs, kLanes), ngrp = kLanes / g, grp = tid / g, sub = tid % g;
    for (int p = grp; p < npairs; p += ngrp) {
      int j = (int)((1.0 + sqrt(1.0 + 8.0 * p)) * 0.5);
      while (j * (j - 1) / 2 > p) --j;
      while ((j + 1) * j / 2 <= p) ++j;
      const int l = p - j * (j - 1) / 2;
      const double* vl = V + (size_t)l * ldv;
      const double* vj = V + (size_t)j * ldv;
      double d = strided_dot(vl, vj, j + 1 + sub, m, g);
      d = group_sum(d, g) + vl[j];
      if (sub == 0) S[l + (size_t)j * k] = d;
    }
  }
  team_sync<kLanes>();
  larft_recurrence<kLanes>(k, tau, S, T);
}

// Workgroup-wide wrappers (the workgroup's own reduction scratch is unused).
__device__ inline void house_qr(double* A, int lda, int m, int k, double* tau, double* = nullptr,
                                double* = nullptr) {
  if (m <= 2 * kQrRows * kBlock)
    team_house_qr_rows<kBlock>(A, lda, m, k, tau);
  else
    team_house_qr<kBlock>(A, lda, m, k, tau);
}
__device__ inline void apply_q(const double* V, int ldv, int m, int nref, const double* tau,
                               double* Y, int ldy, int r, double* = nullptr, double* = nullptr) {
  team_apply_q<kBlock>(V, ldv, m, nref, tau, Y, ldy, r);
}
// Single-wavefront versions (operands in LDS).
__device__ inline void wave_house_qr(double* A, int lda, int m, int k, double* tau) {
  if (m <= kQrRows * kWave)
    team_house_qr_rows<kWave>(A, lda, m, k, tau);
  else
    team_house_qr<kWave>(A, lda, m, k, tau);
}
__device__ inline void wave_apply_q(const double* V, int ldv, int m, int nref, const double* tau,
                                    double* Y, int ldy, int r) {
  team_apply_q<kWave>(V, ldv, m, nref, tau, Y, ldy, r);
}

constexpr int kJacobiMaxCols = 256;  // columns whose norms the Jacobi keeps in scratch

// One-sided (Hestenes) Jacobi SVD of G (rows x cols, ldg): on exit G = U Sigma
// (columns mutually orthogonal, unsorted) and V (cols x cols, ldv) holds the
// accumulated rotations so that G_in = (U Sigma) V^T.  Round-robin ordering:
// the cols/2 pairs of a round are independent and each belongs to one lane
// group.  Squared column norms are kept in 'nrm' (cols doubles of shared
// scratch), recomputed at the start of every sweep and updated exactly by
// each rotation (a' = a - t g, b' = b + t g), so a pair costs one reduction.
// A pair is rotated while g^2 > tol2 a b (tol2 1e-26 by default) and
// |g| > 1e-30 ||G||_F^2 (columns
// at rounding level of ||G|| carry nothing for the truncation thresholds used
// here).  Returns the number of sweeps.  'sh2' is 2 + kWaves shared doubles of
// scratch (workgroup team only); 'nrm' holds cols doubles or is null.
// Rotation of one-sided Jacobi zeroing the inner product c2 of two columns
// with squared norms a and b: t = tan = sign(d) 2 c2 / (|d| + sqrt(d^2 + 4 c2^2))
// with d = b - a, c = 1 / sqrt(1 + t^2), s = c t.  d and 2 c2 are first
// scaled by a power of two (exact); the square roots and the quotient start
// from the hardware reciprocal and reciprocal square root and take Newton
// steps (two for c, so that c^2 + s^2 = 1 to a few ulp).
__device__ inline void jacobi_rot(double a, double b, double c2, double& t, double& c, double& s) {
  const double d = b - a, e = 2.0 * c2;
  const int ex = __builtin_amdgcn_frexp_exp(fmax(fabs(d), fabs(e)));
  const double ds = __builtin_amdgcn_ldexp(d, -ex), es = __builtin_amdgcn_ldexp(e, -ex);
  const double q = ds * ds + es * es;  // in [0.25, 2)
  double r = __builtin_amdgcn_rsq(q);
  r = r * (1.5 - 0.5 * q * r * r);
  const double den = fabs(ds) + q * r;  // |d| + sqrt(d^2 + e^2), scaled
  double inv = __builtin_amdgcn_rcp(den);
  inv = inv * (2.0 - den * inv);
  t = (d < 0.0 ? -es : es) * inv;
  const double w = 1.0 + t * t;  // in [1, 2]
  double y = __builtin_amdgcn_rsq(w);
  y = y * (1.5 - 0.5 * w * y * y);
  y = y * (1.5 - 0.5 * w * y * y);
  c = y;
  s = y * t;
}

template <int kLanes>
__device__ inline int team_jacobi(double* G, int ldg, int rows, int cols, double* V, int ldv,
                                  double* sh2, double* nrm, int max_sweeps, double tol2 = 1e-26) {
  const int tid = team_tid<kLanes>();
  if (V != nullptr)
    for (int e = tid; e < cols * cols; e += kLanes) {
      const int i = e % cols, j = e / cols;
      V[i + (size_t)j * ldv] = (i == j) ? 1.0 : 0.0;
    }
  if (cols < 2) {
    team_sync<kLanes>();
    return 0;
  }
  const int np = (cols + 1) & ~1, rounds = np - 1, half = np / 2;
  // long columns: at least 16 lanes per pair, so that a pair's loads are
  // coalesced and split 16 ways (fewer pairs side by side, each much faster
  // than a pair walked by one or two lanes)
  int g0 = group_lanes(half, kLanes);
  if (kLanes >= kWave && rows >= 128 && g0 < 16) g0 = 16;
  const int g = g0, ngrp = kLanes / g, grp = tid / g, sub = tid % g;
  // without scratch for every column the norms are recomputed pair by pair
  const bool track = nrm != nullptr && cols <= kJacobiMaxCols;
  int sweep = 0;
  for (; sweep < max_sweeps; ++sweep) {
    // column norms (and the Frobenius floor) afresh each sweep
    team_sync<kLanes>();
    double f = 0.0;
    if (track) {
      for (int c = grp; c < cols; c += ngrp) {
        double a = 0.0;
        for (int i = sub; i < rows; i += g) a += G[i + (size_t)c * ldg] * G[i + (size_t)c * ldg];
        a = group_sum(a, g);
        if (sub == 0) nrm[c] = a;
      }
      team_sync<kLanes>();
      for (int c = 0; c < cols; ++c) f += nrm[c];
    } else {
      for (int e = tid; e < rows * cols; e += kLanes) {
        const double x = G[e % rows + (size_t)(e / rows) * ldg];
        f += x * x;
      }
      f = team_sum<kLanes>(f, sh2 ? sh2 + 2 : nullptr);
    }
    const double floor2 = 1e-30 * f;
    int rotated = 0;
    for (int rd = 0; rd < rounds; ++rd) {
      for (int pk = grp; pk < half; pk += ngrp) {
        int p, q;
        if (pk == 0) {
          p = np - 1;
          q = rd;
        } else {
          p = rd + pk;
          if (p >= np - 1) p -= np - 1;
          q = rd - pk + (np - 1);
          if (q >= np - 1) q -= np - 1;
        }
        if (p > q) {
          const int tmp = p;
          p = q;
          q = tmp;
        }
        if (q >= cols) continue;
        double* gp = G + (size_t)p * ldg;
        double* gq = G + (size_t)q * ldg;
        double c2 = 0.0, a, b;
        if (track) {
          c2 = group_sum(strided_dot(gp, gq, sub, rows, g), g);
          a = nrm[p];
          b = nrm[q];
        } else {
          a = 0.0;
          b = 0.0;
          for (int i = sub; i < rows; i += g) {
            a += gp[i] * gp[i];
            b += gq[i] * gq[i];
            c2 += gp[i] * gq[i];
          }
          a = group_sum(a, g);
          b = group_sum(b, g);
          c2 = group_sum(c2, g);
        }
        if (!(c2 * c2 > tol2 * a * b) || !(fabs(c2) > floor2)) continue;
        // t = tan(theta) of the rotation zeroing the pair's inner product
        double t, c, s;
        jacobi_rot(a, b, c2, t, c, s);
        for (int i = sub; i < rows; i += g) {
          const double x = gp[i], y = gq[i];
          gp[i] = c * x - s * y;
          gq[i] = s * x + c * y;
        }
        if (V != nullptr) {
          double* vp = V + (size_t)p * ldv;
          double* vq = V + (size_t)q * ldv;
          for (int i = sub; i < cols; i += g) {
            const double x = vp[i], y = vq[i];
            vp[i] = c * x - s * y;
            vq[i] = s * x + c * y;
          }
        }
        if (track && sub == 0) {
          nrm[p] = a - t * c2;
          nrm[q] = b + t * c2;
        }
        rotated = 1;
      }
      team_sync<kLanes>();
    }
    int any;
    if constexpr (kLanes == kWave) {
      any = __ballot(rotated) != 0;
    } else {
      if (threadIdx.x == 0) sh2[1] = 0.0;
      __syncthreads();
      if (rotated) sh2[1] = 1.0;
      __syncthreads();
      any = sh2[1] != 0.0;
      __syncthreads();
    }
    if (!any) {
      ++sweep;
      break;
    }
  }
  team_sync<kLanes>();
  return sweep;
}

// Sum over aligned groups of kG lanes (compile-time kG <= 16, see group_sum).
template <int kG>
__device__ inline double group_sum_c(double v) {
  if constexpr (kG >= 2) v += dpp_swap<0xB1>(v);
  if constexpr (kG >= 4) v += dpp_swap<0x4E>(v);
  if constexpr (kG >= 8) v += dpp_swap<0x141>(v);
  if constexpr (kG >= 16) v += dpp_swap<0x140>(v);
  return v;
}

// team_jacobi with the pair's rows held in registers and a compile-time
// group of kG lanes per pair: lane 'sub' of a group owns rows sub + kG r
// (r < R) of G (and of V when kV); the loads use clamped rows and a select,
// the columns are read once per round and written back after the rotation.
// Requires rows (and cols when kV) <= R kG, kG == group_lanes(pairs, kLanes)
// and tracked norms.  Without V (kV false) the caller recovers the right
// singular vectors from the input, as G_in^T U.
template <int kLanes, int R, int kG, bool kV>
__device__ inline int team_jacobi_rows(double* G, int ldg, int rows, int cols, double* V, int ldv,
                                       double* sh2, double* nrm, int max_sweeps, double tol2 = 1e-26) {
  const int tid = team_tid<kLanes>();
  if constexpr (kV) {
    for (int e = tid; e < cols * cols; e += kLanes) {
      const int i = e % cols, j = e / cols;
      V[i + (size_t)j * ldv] = (i == j) ? 1.0 : 0.0;
    }
  }
  if (cols < 2) {
    team_sync<kLanes>();
    return 0;
  }
  const int np = (cols + 1) & ~1, rounds = np - 1, half = np / 2;
  constexpr int ngrp = kLanes / kG;
  const int grp = tid / kG, sub = tid % kG;
  int rg[R], rgc[R], rvc[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    rg[r] = sub + kG * r;
    rgc[r] = rg[r] < rows ? rg[r] : rows - 1;
    rvc[r] = rg[r] < cols ? rg[r] : cols - 1;
  }
  int sweep = 0;
  for (; sweep < max_sweeps; ++sweep) {
    team_sync<kLanes>();
    for (int c = grp; c < cols; c += ngrp) {
      double a = 0.0;
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const double x = G[rgc[r] + (size_t)c * ldg];
        a += rg[r] < rows ? x * x : 0.0;
      }
      a = group_sum_c<kG>(a);
      if (sub == 0) nrm[c] = a;
    }
    team_sync<kLanes>();
    double f = 0.0;
    for (int c = 0; c < cols; ++c) f += nrm[c];
    const double floor2 = 1e-30 * f;
    int rotated = 0;
    for (int rd = 0; rd < rounds; ++rd) {
      for (int pk = grp; pk < half; pk += ngrp) {
        int p, q;
        if (pk == 0) {
          p = np - 1;
          q = rd;
        } else {
          p = rd + pk;
          if (p >= np - 1) p -= np - 1;
          q = rd - pk + (np - 1);
          if (q >= np - 1) q -= np - 1;
        }
        if (p > q) {
          const int tmp = p;
          p = q;
          q = tmp;
        }
        if (q >= cols) continue;
        double* gp = G + (size_t)p * ldg;
        double* gq = G + (size_t)q * ldg;
        // all loads of the pair issued before the first wait
        double xp[R], xq[R], up[R], uq[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
          xp[r] = gp[rgc[r]];
          xq[r] = gq[rgc[r]];
        }
        if constexpr (kV) {
          const double* vp = V + (size_t)p * ldv;
          const double* vq = V + (size_t)q * ldv;
#pragma unroll
          for (int r = 0; r < R; ++r) {
            up[r] = vp[rvc[r]];
            uq[r] = vq[rvc[r]];
          }
        }
        const double a = nrm[p], b = nrm[q];
        double c2 = 0.0;
#pragma unroll
        for (int r = 0; r < R; ++r) c2 += rg[r] < rows ? xp[r] * xq[r] : 0.0;
        c2 = group_sum_c<kG>(c2);
        if (!(c2 * c2 > tol2 * a * b) || !(fabs(c2) > floor2)) continue;
        double t, c, s;
        jacobi_rot(a, b, c2, t, c, s);
#pragma unroll
        for (int r = 0; r < R; ++r) {
          if (rg[r] < rows) {
            gp[rg[r]] = c * xp[r] - s * xq[r];
            gq[rg[r]] = s * xp[r] + c * xq[r];
          }
        }
        if constexpr (kV) {
          double* vp = V + (size_t)p * ldv;
          double* vq = V + (size_t)q * ldv;
#pragma unroll
          for (int r = 0; r < R; ++r) {
            if (rg[r] < cols) {
              vp[rg[r]] = c * up[r] - s * uq[r];
              vq[rg[r]] = s * up[r] + c * uq[r];
            }
          }
        }
        if (sub == 0) {
          nrm[p] = a - t * c2;
          nrm[q] = b + t * c2;
        }
        rotated = 1;
      }
      team_sync<kLanes>();
    }
    int any;
    if constexpr (kLanes == kWave) {
      any = __ballot(rotated) != 0;
    } else {
      if (threadIdx.x == 0) sh2[1] = 0.0;
      __syncthreads();
      if (rotated) sh2[1] = 1.0;
      __syncthreads();
      any = sh2[1] != 0.0;
      __syncthreads();
    }
    if (!any) {
      ++sweep;
      break;
    }
  }
  team_sync<kLanes>();
  return sweep;
}

// Dispatch on the group size and rows per lane; V == nullptr skips the
// accumulation of V.  Shapes outside the instantiated ones use team_jacobi.
template <int kLanes, bool kV>
__device__ inline int team_jacobi_sel(double* G, int ldg, int rows, int cols, double* V, int ldv,
                                      double* sh2, double* nrm, int max_sweeps, double tol2 = 1e-26) {
  const int half = ((cols + 1) & ~1) / 2;
  const int g = group_lanes(half, kLanes);
  const int span = kV ? (rows > cols ? rows : cols) : rows;
  const int need = (span + g - 1) / g;
#define HMX_JR(G_, R_)                                                                             \
  if (g == G_ && need <= R_)                                                                       \
    return team_jacobi_rows<kLanes, R_, G_, kV>(G, ldg, rows, cols, V, ldv, sh2, nrm, max_sweeps, tol2);
  if constexpr (kLanes == kWave) {
    HMX_JR(16, 1)
    HMX_JR(8, 2)
    HMX_JR(4, 6)
    HMX_JR(4, 8)
  } else {
    HMX_JR(16, 2)
    HMX_JR(8, 8)
  }
#undef HMX_JR
  return team_jacobi<kLanes>(G, ldg, rows, cols, V, ldv, sh2, nrm, max_sweeps, tol2);
}

template <int kLanes>
__device__ inline int team_jacobi_any(double* G, int ldg, int rows, int cols, double* V, int ldv,
                                      double* sh2, double* nrm, int max_sweeps, double tol2 = 1e-26) {
  if (nrm != nullptr && cols <= kJacobiMaxCols && cols >= 2 && rows >= 1) {
    if (V != nullptr)
      return team_jacobi_sel<kLanes, true>(G, ldg, rows, cols, V, ldv, sh2, nrm, max_sweeps, tol2);
    return team_jacobi_sel<kLanes, false>(G, ldg, rows, cols, V, ldv, sh2, nrm, max_sweeps, tol2);
  }
  return team_jacobi<kLanes>(G, ldg, rows, cols, V, ldv, sh2, nrm, max_sweeps, tol2);
}

// Workgroup version used by the existing call sites (cols <= kJacobiMaxCols);
// V == nullptr leaves V out.
__device__ inline void jacobi_svd(double* G, int ldg, int rows, int cols, double* V, int ldv,
                                  int* /*flag*/, int max_sweeps = 40, double tol2 = 1e-26) {
  __shared__ double sh2[2 + kWaves];
  __shared__ double nrm[kJacobiMaxCols];
  team_jacobi_any<kBlock>(G, ldg, rows, cols, V, ldv, sh2, nrm, max_sweeps, tol2);
}

// Single-wavefront version (G, V in LDS; cols <= kJacobiMaxCols); V == nullptr
// leaves V out.
__device__ inline int wave_jacobi_svd(double* G, int ldg, int rows, int cols, double* V, int ldv,
                                      int max_sweeps = 40, double tol2 = 1e-26) {
  __shared__ double nrm[kJacobiMaxCols];
  return team_jacobi_any<kWave>(G, ldg, rows, cols, V, ldv, nullptr, nrm, max_sweeps, tol2);
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
