// Blocked Householder QR of tall panels for one workgroup (kBlock threads),
// with the panel's rows owned by threads: thread t holds rows t + kBlock r
// (r < R) in registers.  Columns are factored four at a time in registers
// and the trailing columns are updated four at a time with the four-column
// block reflector, so a step costs one batched reduction instead of one
// memory round trip and one reduction per trailing column.
//
// Output as LAPACK dgeqrf (R in the upper triangle, reflectors below the
// unit diagonal, tau) plus the 4 x 4 triangular factor of every block
// reflector (dlarft, forward columnwise), which wg_apply_q uses to form
// Q [U; 0] without a K x K factor.
#pragma once
#include "dev_common.hpp"

namespace hmx {

// ---------------------------------------------------------------------------
// Cross-lane exchange and batched workgroup sums
// ---------------------------------------------------------------------------
// Value held by lane ^ (1 << kBit) of the wavefront (all lanes active).
template <int kBit>
__device__ inline double lane_xor(double v) {
  const int lo = __double2loint(v), hi = __double2hiint(v);
  if constexpr (kBit == 0) {
    return dpp_swap<0xB1>(v);  // quad_perm [1,0,3,2]
  } else if constexpr (kBit == 1) {
    return dpp_swap<0x4E>(v);  // quad_perm [2,3,0,1]
  } else if constexpr (kBit == 2 || kBit == 3) {
    // row_shr:n writes the lanes of banks with the bit set (source l - n),
    // then row_shl:n the others (source l + n); banks are 4-lane groups
    constexpr int kShl = kBit == 2 ? 0x104 : 0x108, kShr = kBit == 2 ? 0x114 : 0x118;
    constexpr int kUp = kBit == 2 ? 0xA : 0xC, kDown = kBit == 2 ? 0x5 : 0x3;
    int rlo = __builtin_amdgcn_update_dpp(0, lo, kShr, 0xF, kUp, false);
    int rhi = __builtin_amdgcn_update_dpp(0, hi, kShr, 0xF, kUp, false);
    rlo = __builtin_amdgcn_update_dpp(rlo, lo, kShl, 0xF, kDown, false);
    rhi = __builtin_amdgcn_update_dpp(rhi, hi, kShl, 0xF, kDown, false);
    return __hiloint2double(rhi, rlo);
  } else {
    // permlane{16,32}_swap(x, x): the first result holds the partner's value
    // in the upper lane of each pair, the second in the lower one
    const bool up = (threadIdx.x >> kBit) & 1;
    unsigned a0, a1, b0, b1;
    if constexpr (kBit == 4) {
      auto l = __builtin_amdgcn_permlane16_swap((unsigned)lo, (unsigned)lo, false, false);
      auto h = __builtin_amdgcn_permlane16_swap((unsigned)hi, (unsigned)hi, false, false);
      a0 = l[0]; a1 = h[0]; b0 = l[1]; b1 = h[1];
    } else {
      auto l = __builtin_amdgcn_permlane32_swap((unsigned)lo, (unsigned)lo, false, false);
      auto h = __builtin_amdgcn_permlane32_swap((unsigned)hi, (unsigned)hi, false, false);
      a0 = l[0]; a1 = h[0]; b0 = l[1]; b1 = h[1];
    }
    return up ? __hiloint2double((int)a1, (int)a0) : __hiloint2double((int)b1, (int)b0);
  }
}

// One reduce-scatter stage over lane bit kBit: the lane with the bit clear
// keeps 'lo' plus its partner's 'lo', the other 'hi' plus its partner's 'hi'.
template <int kBit>
__device__ inline double rs_stage(double lo, double hi) {
  if constexpr (kBit >= 4) {
    // permlane{16,32}_swap(lo, hi) leaves, in every lane, the two addends
    // of its kept half (own and partner's) in the two results
    const int l0 = __double2loint(lo), l1 = __double2hiint(lo);
    const int h0 = __double2loint(hi), h1 = __double2hiint(hi);
    unsigned a0, a1, b0, b1;
    if constexpr (kBit == 4) {
      auto x = __builtin_amdgcn_permlane16_swap((unsigned)l0, (unsigned)h0, false, false);
      auto y = __builtin_amdgcn_permlane16_swap((unsigned)l1, (unsigned)h1, false, false);
      a0 = x[0]; b0 = x[1]; a1 = y[0]; b1 = y[1];
    } else {
      auto x = __builtin_amdgcn_permlane32_swap((unsigned)l0, (unsigned)h0, false, false);
      auto y = __builtin_amdgcn_permlane32_swap((unsigned)l1, (unsigned)h1, false, false);
      a0 = x[0]; b0 = x[1]; a1 = y[0]; b1 = y[1];
    }
    return __hiloint2double((int)a1, (int)a0) + __hiloint2double((int)b1, (int)b0);
  } else {
    const bool up = (threadIdx.x >> kBit) & 1;
    const double send = up ? lo : hi, keep = up ? hi : lo;
    return keep + lane_xor<kBit>(send);
  }
}

// Workgroup barrier that orders LDS only: outstanding global stores are not
// waited for (callers exchange data through LDS; global rows are only read
// back by the thread that wrote them).
__device__ inline void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

constexpr int pow2_ceil(int n) { return n <= 1 ? 1 : 2 * pow2_ceil((n + 1) / 2); }
constexpr int log2_exact(int n) { return n <= 1 ? 0 : 1 + log2_exact(n / 2); }

template <int kBit, int kNp>
__device__ inline void rs_stages(double (&v)[kNp]) {
  if constexpr (kBit < 6) {
    constexpr int n = log2_exact(kNp);
    if constexpr (kBit < n) {
      constexpr int c = kNp >> kBit, h = c / 2;
#pragma unroll
      for (int i = 0; i < h; ++i) v[i] = rs_stage<kBit>(v[i], v[i + h]);
    } else {
      v[0] += lane_xor<kBit>(v[0]);
    }
    rs_stages<kBit + 1, kNp>(v);
  }
}

// Sums v[0..N) over the workgroup (every thread active); every thread gets
// the N totals, bitwise identical across threads.  Within a wavefront a
// reduce-scatter leaves each item's sum in one lane group, then the four
// wavefront sums go through LDS.
template <int N>
__device__ inline void wg_sum(double (&v)[N]) {
  constexpr int kNp = pow2_ceil(N), kN2 = log2_exact(kNp);
  __shared__ double red[kWaves * kNp + kNp];
  double u[kNp];
#pragma unroll
  for (int i = 0; i < kNp; ++i) u[i] = i < N ? v[i] : 0.0;
  rs_stages<0, kNp>(u);
  const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
  // item held after the stages: bit k of the lane selects the upper half at stage k
  int item = 0;
#pragma unroll
  for (int k = 0; k < kN2; ++k) item += ((lane >> k) & 1) * (kNp >> (k + 1));
  lds_barrier();
  if (lane < kNp) red[w * kNp + item] = u[0];
  lds_barrier();
  if constexpr (N <= 8) {
#pragma unroll
    for (int i = 0; i < N; ++i)
      v[i] = (red[i] + red[kNp + i]) + (red[2 * kNp + i] + red[3 * kNp + i]);
  } else {
    double* tot = red + kWaves * kNp;
    if (threadIdx.x < N)
      tot[threadIdx.x] = (red[threadIdx.x] + red[kNp + threadIdx.x]) +
                         (red[2 * kNp + threadIdx.x] + red[3 * kNp + threadIdx.x]);
    lds_barrier();
#pragma unroll
    for (int i = 0; i < N; ++i) v[i] = tot[i];
  }
}

// ---------------------------------------------------------------------------
// Block reflector of four columns applied to four columns held in registers
// ---------------------------------------------------------------------------
// Turns the four stored columns of a block (R entries on and above the
// diagonal, reflector entries below) into the reflectors: a unit entry at
// row j0 + a, zeros above.
template <int R>
__device__ inline void to_reflectors(double (&x)[4][R], const int (&row)[R], int j0) {
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int d = j0 + a;
      x[a][r] = row[r] > d ? x[a][r] : (row[r] == d ? 1.0 : 0.0);
    }
}

// y <- (I - V T V^T)^T y (kTrans) or (I - V T V^T) y, for the four columns
// of y; V the reflectors (to_reflectors), T upper triangular 4 x 4,
// column-major.
template <int R, bool kTrans>
__device__ inline void apply_block(const double (&v)[4][R], const double (&T)[16], double (&y)[4][R]) {
  double z[16];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      double s = 0.0;
#pragma unroll
      for (int r = 0; r < R; ++r) s += v[a][r] * y[b][r];
      z[a + 4 * b] = s;
    }
  wg_sum<16>(z);
  // W = T^T Z (in place, rows from the bottom) or T Z (from the top)
#pragma unroll
  for (int b = 0; b < 4; ++b) {
    if constexpr (kTrans) {
#pragma unroll
      for (int a = 3; a >= 0; --a) {
        double s = 0.0;
#pragma unroll
        for (int k = 0; k <= a; ++k) s += T[k + 4 * a] * z[k + 4 * b];
        z[a + 4 * b] = s;
      }
    } else {
#pragma unroll
      for (int a = 0; a < 4; ++a) {
        double s = 0.0;
#pragma unroll
        for (int k = a; k < 4; ++k) s += T[a + 4 * k] * z[k + 4 * b];
        z[a + 4 * b] = s;
      }
    }
  }
#pragma unroll
  for (int b = 0; b < 4; ++b)
#pragma unroll
    for (int r = 0; r < R; ++r) {
      double s = 0.0;
#pragma unroll
      for (int a = 0; a < 4; ++a) s += v[a][r] * z[a + 4 * b];
      y[b][r] -= s;
    }
}

// Four columns c0.. of P (rows of this thread; zero outside m x K).
template <int R>
__device__ inline void load_cols(const double* P, int ldp, int m, int K, int c0, const int (&row)[R],
                                 const int (&rc)[R], double (&y)[4][R]) {
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    const int cc = c0 + c < K ? c0 + c : K - 1;
    const double* col = P + (size_t)cc * ldp;
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const double t = col[rc[r]];
      y[c][r] = (row[r] < m && c0 + c < K) ? t : 0.0;
    }
  }
}

// Stores rows >= r0 of the valid columns.
template <int R>
__device__ inline void store_cols(double* P, int ldp, int m, int K, int c0, const int (&row)[R], int r0,
                                  const double (&y)[4][R]) {
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    if (c0 + c >= K) continue;
    double* col = P + (size_t)(c0 + c) * ldp;
#pragma unroll
    for (int r = 0; r < R; ++r)
      if (row[r] >= r0 && row[r] < m) col[row[r]] = y[c][r];
  }
}

// ---------------------------------------------------------------------------
// QR (dgeqrf): P (m x K, ld ldp, LDS or global), m <= R * kBlock.
// tau[0..K) and Tp[16 * (K + 3) / 4] (block factors, column-major 4 x 4).
// Every thread of the workgroup calls it; ends with a barrier.
// ---------------------------------------------------------------------------
template <int R>
__device__ void wg_qr(double* P, int ldp, int m, int K, double* tau, double* Tp) {
  const int tid = threadIdx.x;
  int row[R], rc[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    row[r] = tid + kBlock * r;
    rc[r] = row[r] < m ? row[r] : m - 1;
  }
  for (int j0 = 0; j0 < K; j0 += 4) {
    const int nb = K - j0 < 4 ? K - j0 : 4;
    double x[4][R];
    load_cols(P, ldp, m, K, j0, row, rc, x);
    double tl[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int jj = 0; jj < 4; ++jj) {
      if (jj >= nb) break;
      const int j = j0 + jj;
      // below-diagonal dots of column jj with columns jj.. and row j of them
      double s8[8];
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        double d = 0.0, e = 0.0;
        if (c >= jj) {
#pragma unroll
          for (int r = 0; r < R; ++r) {
            d += row[r] > j ? x[jj][r] * x[c][r] : 0.0;
            e += row[r] == j ? x[c][r] : 0.0;
          }
        }
        s8[c] = d;
        s8[4 + c] = e;
      }
      wg_sum<8>(s8);
      const double s = s8[jj], alpha = s8[4 + jj];
      double t = 0.0, beta = alpha, scal = 0.0;
      if (s != 0.0) {
        beta = -copysign(sqrt(alpha * alpha + s), alpha);
        t = (beta - alpha) / beta;
        scal = 1.0 / (alpha - beta);
      }
      tl[jj] = t;
#pragma unroll
      for (int r = 0; r < R; ++r)
        if (row[r] > j) x[jj][r] *= scal;
#pragma unroll
      for (int c = jj + 1; c < 4; ++c) {
        // f = tau v^T a_c with v = [1; x_jj scal]
        const double f = t * (s8[4 + c] + scal * s8[c]);
#pragma unroll
        for (int r = 0; r < R; ++r) {
          if (row[r] > j)
            x[c][r] -= f * x[jj][r];
          else if (row[r] == j)
            x[c][r] -= f;
        }
      }
#pragma unroll
      for (int r = 0; r < R; ++r)
        if (row[r] == j) x[jj][r] = beta;
    }
    store_cols(P, ldp, m, K, j0, row, j0, x);
    to_reflectors(x, row, j0);
    // T of the block: T[a][a] = tau_a, T[a][b] = -tau_b sum_{r=a}^{b-1} T[a][r] (v_r^T v_b)
    double sv[8];
    {
      int q = 0;
#pragma unroll
      for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = a + 1; b < 4; ++b) {
          double d = 0.0;
#pragma unroll
          for (int r = 0; r < R; ++r) d += x[a][r] * x[b][r];
          sv[q++] = d;
        }
      sv[6] = sv[7] = 0.0;
    }
    wg_sum<8>(sv);
    double T[16];
#pragma unroll
    for (int e = 0; e < 16; ++e) T[e] = 0.0;
    {
      // sv index of pair (a, b), a < b: (0,1)0 (0,2)1 (0,3)2 (1,2)3 (1,3)4 (2,3)5
      const int pidx[4][4] = {{-1, 0, 1, 2}, {-1, -1, 3, 4}, {-1, -1, -1, 5}, {-1, -1, -1, -1}};
#pragma unroll
      for (int b = 0; b < 4; ++b) {
        T[b + 4 * b] = tl[b];
#pragma unroll
        for (int a = 0; a < b; ++a) {
          double acc = 0.0;
#pragma unroll
          for (int r = a; r < b; ++r) acc += T[a + 4 * r] * sv[pidx[r][b]];
          T[a + 4 * b] = -tl[b] * acc;
        }
      }
    }
    if (tid == 0) {
#pragma unroll
      for (int e = 0; e < 16; ++e) Tp[16 * (j0 / 4) + e] = T[e];
      for (int c = 0; c < nb; ++c) tau[j0 + c] = tl[c];
    }
    for (int c0 = j0 + 4; c0 < K; c0 += 4) {
      double y[4][R];
      load_cols(P, ldp, m, K, c0, row, rc, y);
      apply_block<R, true>(x, T, y);
      store_cols(P, ldp, m, K, c0, row, j0, y);
    }
  }
  __syncthreads();
}

// out (m x nk, ld ldo) = Q [U; 0] with Q from wg_qr (reflectors in V, ld ldv,
// block factors Tp) and U[i, c] = src[i + order[c] K] scale[c] (K rows).
// Columns go four at a time through the blocks in reverse order.
template <int R, typename Tout>
__device__ void wg_apply_q(const double* V, int ldv, int m, int K, const double* Tp, const double* src,
                           const int* order, const double* scale, int nk, Tout* out, int ldo) {
  const int tid = threadIdx.x;
  int row[R], rc[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    row[r] = tid + kBlock * r;
    rc[r] = row[r] < m ? row[r] : m - 1;
  }
  const int nblk = (K + 3) / 4;
  for (int c0 = 0; c0 < nk; c0 += 4) {
    double y[4][R];
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      const int c = c0 + b < nk ? c0 + b : nk - 1;
      const double* u = src + (size_t)order[c] * K;
      const double sc = scale[c];
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const double t = u[row[r] < K ? row[r] : K - 1];
        y[b][r] = (row[r] < K && c0 + b < nk) ? t * sc : 0.0;
      }
    }
    for (int p = nblk - 1; p >= 0; --p) {
      double x[4][R];
      load_cols(V, ldv, m, K, 4 * p, row, rc, x);
      to_reflectors(x, row, 4 * p);
      double T[16];
#pragma unroll
      for (int e = 0; e < 16; ++e) T[e] = Tp[16 * p + e];
      apply_block<R, false>(x, T, y);
    }
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      if (c0 + b >= nk) continue;
#pragma unroll
      for (int r = 0; r < R; ++r)
        if (row[r] < m) out[row[r] + (size_t)(c0 + b) * ldo] = (Tout)y[b][r];
    }
  }
  __syncthreads();
}

// Rows per thread for m rows: 1, 2, 4 or 8 (m <= 8 kBlock); 0 when too tall.
__host__ __device__ constexpr int wg_rows(int m) {
  return m <= kBlock ? 1 : m <= 2 * kBlock ? 2 : m <= 4 * kBlock ? 4 : m <= 8 * kBlock ? 8 : 0;
}

__device__ inline void wg_qr_any(double* P, int ldp, int m, int K, double* tau, double* Tp) {
  switch (wg_rows(m)) {
    case 1: wg_qr<1>(P, ldp, m, K, tau, Tp); break;
    case 2: wg_qr<2>(P, ldp, m, K, tau, Tp); break;
    case 4: wg_qr<4>(P, ldp, m, K, tau, Tp); break;
    default: wg_qr<8>(P, ldp, m, K, tau, Tp); break;
  }
}

template <typename Tout>
__device__ inline void wg_apply_q_any(const double* V, int ldv, int m, int K, const double* Tp,
                                      const double* src, const int* order, const double* scale, int nk,
                                      Tout* out, int ldo) {
  switch (wg_rows(m)) {
    case 1: wg_apply_q<1, Tout>(V, ldv, m, K, Tp, src, order, scale, nk, out, ldo); break;
    case 2: wg_apply_q<2, Tout>(V, ldv, m, K, Tp, src, order, scale, nk, out, ldo); break;
    case 4: wg_apply_q<4, Tout>(V, ldv, m, K, Tp, src, order, scale, nk, out, ldo); break;
    default: wg_apply_q<8, Tout>(V, ldv, m, K, Tp, src, order, scale, nk, out, ldo); break;
  }
}

}  // namespace hmx

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
template <typename T>
__device__ void rows_times_small(const T* M, int ldm, int rows, int k, const double* C, int ldc, int nk, T* out,
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
        if (c0 + b < nk) out[i + (size_t)(c0 + b) * ldo] = (T)acc[b];
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
