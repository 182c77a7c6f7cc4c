// Batched task kernels of the H-LLt factorisation and the H-triangular solves.
// Each launch processes the tasks of one type in one dependency wave; the
// tasks of a launch are independent.  Dense arithmetic runs in the storage
// type T (FP64 or FP32); QR/SVD of the recompressions run in FP64.
#include "dev_linalg.hpp"
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

// ---------------------------------------------------------------------------
// GEMM: C (m x n, ldc) = beta C + alpha op(A) op(B), 64 x 64 tiles.
// ---------------------------------------------------------------------------
constexpr int TM = 64, TN = 64, TK = 16;

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
  const int tr = threadIdx.x % 16, tc = threadIdx.x / 16;
  T acc[4][4];
#pragma unroll
  for (int r = 0; r < 4; ++r)
#pragma unroll
    for (int c = 0; c < 4; ++c) acc[r][c] = 0;
  for (int k0 = 0; k0 < P; k0 += TK) {
    for (int e = threadIdx.x; e < TK * TM; e += 256) {
      const int ii = e % TM, kk = e / TM;
      const int i = i0 + ii, k = k0 + kk;
      T v = 0;
      if (i < M && k < P) v = t.ta ? A[k + (size_t)i * t.lda] : A[i + (size_t)k * t.lda];
      As[kk][ii] = v;
    }
    for (int e = threadIdx.x; e < TK * TN; e += 256) {
      const int kk = e % TK, jj = e / TK;
      const int j = j0 + jj, k = k0 + kk;
      T v = 0;
      if (j < N && k < P) v = t.tb ? B[j + (size_t)k * t.ldb] : B[k + (size_t)j * t.ldb];
      Bs[kk][jj] = v;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < TK; ++kk) {
      T a[4], b[4];
#pragma unroll
      for (int r = 0; r < 4; ++r) a[r] = As[kk][tr + 16 * r];
#pragma unroll
      for (int c = 0; c < 4; ++c) b[c] = Bs[kk][tc + 16 * c];
#pragma unroll
      for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int c = 0; c < 4; ++c) acc[r][c] += a[r] * b[c];
    }
    __syncthreads();
  }
  const T alpha = (T)t.alpha;
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    const int j = j0 + tc + 16 * c;
    if (j >= N) continue;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int i = i0 + tr + 16 * r;
      if (i >= M) continue;
      T* pc = C + i + (size_t)j * t.ldc;
      *pc = t.mode ? *pc + alpha * acc[r][c] : alpha * acc[r][c];
    }
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
        T s = L[i + (size_t)j * ld];
        for (int l = jb; l < j; ++l) s -= L[i + (size_t)l * ld] * L[j + (size_t)l * ld];
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
      for (int l = jb; l < je; ++l) s += L[i + (size_t)l * ld] * L[c + (size_t)l * ld];
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
    for (int e = threadIdx.x; e < n * n; e += kBlock) {
      const int i = e % n, j = e / n;
      lds[i + j * n] = A[i + (size_t)j * lda];
    }
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
// TRSM with a lower-triangular leaf L (n x n), L staged in LDS:
//   mode 0: X (m x n) <- X L^{-T}     (solveUpperTriangularRight, LLt)
//   mode 1: X (n x k) <- L^{-1} X     (solveLowerTriangularLeft)
//   mode 2: X (n x k) <- L^{-T} X     (solveUpperTriangularLeft, Uplo LOWER)
// Rows (mode 0) / columns (modes 1, 2) of X are independent: one thread runs
// the substitution of one of them, reading L from LDS (broadcast) and its own
// row / column from LDS when it fits, so no barrier is needed inside.
// ---------------------------------------------------------------------------
constexpr int kShDoubles = 19456;  // 152 KiB of LDS per workgroup

template <typename T>
__global__ void __launch_bounds__(kBlock) k_trsm(const FTask* tasks, const int* fail) {
  __shared__ double sh[kShDoubles];
  if (*fail) return;
  const FTask& t = tasks[blockIdx.x];
  const int n = t.n.s;  // order of L
  const int other = eff(t.m);
  if (other <= 0) return;
  const T* Lg = (const T*)t.A;
  const int ldl = t.lda;
  T* X = (T*)t.C;
  const int ldx = t.ldc;
  const bool lds_l = n * n <= kShDoubles;
  double* Ls = sh;
  if (lds_l)
    for (int e = threadIdx.x; e < n * n; e += kBlock) Ls[e] = (double)Lg[e % n + (size_t)(e / n) * ldl];
  // X block staged in LDS behind L when it fits: rows = n for modes 1/2
  const int xr = t.mode == 0 ? other : n, xc = t.mode == 0 ? n : other;
  const int xl = xr | 1;  // odd leading dimension: fewer bank conflicts for column walks
  const bool lds_x = lds_l && n * n + xl * xc <= kShDoubles;
  double* Xs = sh + n * n;
  if (lds_x)
    for (int e = threadIdx.x; e < xr * xc; e += kBlock) {
      const int i = e % xr, j = e / xr;
      Xs[i + j * xl] = (double)X[i + (size_t)j * ldx];
    }
  __syncthreads();
  auto Lat = [&](int i, int j) -> double { return lds_l ? Ls[i + j * n] : (double)Lg[i + (size_t)j * ldl]; };
  auto Xg = [&](int i, int j) -> double { return lds_x ? Xs[i + j * xl] : (double)X[i + (size_t)j * ldx]; };
  auto Xp = [&](int i, int j, double v) {
    if (lds_x)
      Xs[i + j * xl] = v;
    else
      X[i + (size_t)j * ldx] = (T)v;
  };
  // Blocked substitution: a 16-wide diagonal block is solved with one thread per
  // independent row / column, then the rest gets a rank-16 update from all threads.
  if (t.mode == 0) {
    for (int jb = 0; jb < n; jb += kNb) {
      const int je = jb + kNb < n ? jb + kNb : n;
      for (int r = threadIdx.x; r < other; r += kBlock)
        for (int c = jb; c < je; ++c) {
          double s = Xg(r, c);
          for (int l = jb; l < c; ++l) s -= Lat(c, l) * Xg(r, l);
          Xp(r, c, s / Lat(c, c));
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
  } else if (t.mode == 1) {
    for (int jb = 0; jb < n; jb += kNb) {
      const int je = jb + kNb < n ? jb + kNb : n;
      for (int c = threadIdx.x; c < other; c += kBlock)
        for (int j = jb; j < je; ++j) {
          double s = Xg(j, c);
          for (int l = jb; l < j; ++l) s -= Lat(j, l) * Xg(l, c);
          Xp(j, c, s / Lat(j, j));
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
          Xp(j, c, s / Lat(j, j));
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

// ---------------------------------------------------------------------------
// zero fill and axpy
// ---------------------------------------------------------------------------
template <typename T>
__global__ void k_zero(const FTask* tasks, const int* fail) {
  if (*fail) return;
  const FTask& t = tasks[blockIdx.x];
  T* C = (T*)t.C;
  const int m = t.m.s, n = t.n.s;
  for (int e = threadIdx.x; e < m * n; e += blockDim.x) C[e % m + (size_t)(e / m) * t.ldc] = (T)0;
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

struct QrScratch {
  double *taua, *taub, *g, *v, *sig;
  int* order;
  double* wq;  // K doubles: Householder dot products
  double* y;   // max(m, n) x K, reused for A then B
};

// Panels Ac (m x K, ld m), Bc (n x K, ld n), K < min(m, n): QR both, SVD of
// Ra Rb^T, cut at eps, write Qa U sqrt(S), Qb V sqrt(S) to out.
template <typename T>
__device__ void qr_truncate(double* Ac, int m, double* Bc, int n, int K, const QrScratch& w,
                            double eps, RkRef out, int* overflow, double* red, int* flag,
                            double* pr = nullptr) {
  const long q0 = wall_clock64();
  house_qr(Ac, m, m, K, w.taua, red, w.wq);
  house_qr(Bc, n, n, K, w.taub, red, w.wq);
  const long q1 = wall_clock64();
  for (int e = threadIdx.x; e < K * K; e += kBlock) {
    const int i = e % K, j = e / K;
    double s = 0.0;
    const int lo = i > j ? i : j;
    for (int l = lo; l < K; ++l) s += Ac[i + (size_t)l * m] * Bc[j + (size_t)l * n];
    w.g[i + (size_t)j * K] = s;
  }
  __syncthreads();
  jacobi_svd(w.g, K, K, K, w.v, K, flag);
  const long q2 = wall_clock64();
  if (pr && threadIdx.x == 0) {
    pr[6] = (double)(q1 - q0);
    pr[7] = (double)(q2 - q1);
  }
  sorted_norms(w.g, K, K, K, w.sig, w.order);
  int nk = find_k(w.sig, w.order, K, eps);
  if (nk > out.cap) {
    if (threadIdx.x == 0) atomicOr(overflow, 1);
    nk = out.cap;
  }
  T* A = (T*)out.a;
  T* B = (T*)out.b;
  for (int e = threadIdx.x; e < nk * m; e += kBlock) {
    const int i = e % m, c = e / m, j = w.order[c];
    w.y[e] = (i < K) ? w.g[i + (size_t)j * K] / sqrt(w.sig[j]) : 0.0;
  }
  __syncthreads();
  apply_q(Ac, m, m, K, w.taua, w.y, m, nk, red, w.wq);
  for (int e = threadIdx.x; e < nk * m; e += kBlock) A[e % m + (size_t)(e / m) * out.la] = (T)w.y[e];
  __syncthreads();
  for (int e = threadIdx.x; e < nk * n; e += kBlock) {
    const int i = e % n, c = e / n, j = w.order[c];
    w.y[e] = (i < K) ? w.v[i + (size_t)j * K] * sqrt(w.sig[j]) : 0.0;
  }
  __syncthreads();
  apply_q(Bc, n, n, K, w.taub, w.y, n, nk, red, w.wq);
  for (int e = threadIdx.x; e < nk * n; e += kBlock) B[e % n + (size_t)(e / n) * out.lb] = (T)w.y[e];
  __syncthreads();
  if (threadIdx.x == 0) *out.rank = nk;
}

// doubles of the QR path: panels, reusable y, core
__device__ inline int64_t qr_panels(int m, int n, int K) { return (int64_t)(m + n) * K; }
__device__ inline int64_t qr_y(int m, int n, int K) { return (int64_t)(m > n ? m : n) * K; }
__device__ inline int64_t qr_core(int K) { return 2 * (int64_t)K * K + 5 * (int64_t)K; }

__device__ inline QrScratch carve_qr(double* core, double* y, int K) {
  QrScratch w;
  w.taua = core;
  w.taub = core + K;
  w.g = core + 2 * K;
  w.v = w.g + (size_t)K * K;
  w.sig = w.v + (size_t)K * K;
  w.order = (int*)(w.sig + K);
  w.wq = w.sig + 2 * K;
  w.y = y;
  return w;
}

// ---------------------------------------------------------------------------
// FADD: RkMatrix::formattedAddParts (rk_matrix.cpp:846-981)
// ---------------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(kBlock) k_fadd(const FTask* tasks, double* wsbase, int* overflow,
                                                 double* prof) {
  __shared__ double sh[kShDoubles];
  __shared__ double red[kWaves * kQrChunk];
  __shared__ int flag;
  if (overflow[1]) return;
  const FTask& t = tasks[blockIdx.x];
  const long t0 = wall_clock64();
  double* pr = prof ? prof + (size_t)8 * t.pad : nullptr;
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
  if (np == 0) return;
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
    if (pr && threadIdx.x == 0) {
      pr[0] = 0; pr[1] = m; pr[2] = n; pr[3] = K; pr[4] = (double)(wall_clock64() - t0);
    }
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
    const long t1 = wall_clock64();
    dense_to_rk<T>(G, gr, gc, tr, V, V + (size_t)gc * gc, (int*)(V + (size_t)gc * gc + gc),
                   fabs(t.eps), tg, overflow, &flag);
    if (pr && threadIdx.x == 0) {
      pr[0] = 1; pr[1] = m; pr[2] = n; pr[3] = K; pr[4] = (double)(wall_clock64() - t0);
      pr[5] = (double)(t1 - t0); pr[6] = (G == sh);
    }
    return;
  }
  // QR path: concatenate [target, alpha_i * parts] (zero padded)
  const int64_t pan = qr_panels(m, n, K), yy = qr_y(m, n, K), core = qr_core(K);
  double *Ac, *y, *cr;
  if (pan + yy + core <= kShDoubles) {
    Ac = sh;
    y = sh + pan;
    cr = y + yy;
  } else if (core <= kShDoubles) {
    Ac = gws;
    y = gws + pan;
    cr = sh;
  } else {
    Ac = gws;
    y = gws + pan;
    cr = y + yy;
  }
  double* Bc = Ac + (size_t)m * K;
  for (int e = threadIdx.x; e < m * kt; e += kBlock)
    Ac[e] = (double)TA[e % m + (size_t)(e / m) * tg.la];
  for (int e = threadIdx.x; e < n * kt; e += kBlock)
    Bc[e] = (double)TB[e % n + (size_t)(e / n) * tg.lb];
  int off = kt;
  for (int q = 0; q < t.nparts; ++q) {
    const Part& p = t.parts[q];
    const int kq = rank_of(p.rk);
    if (!(kq > 0 && p.alpha != 0.0 && p.rk.m > 0 && p.rk.n > 0)) continue;
    const T* PA = (const T*)p.rk.a;
    const T* PB = (const T*)p.rk.b;
    for (int e = threadIdx.x; e < m * kq; e += kBlock) {
      const int i = e % m, c = e / m, ii = i - p.roff;
      Ac[i + (size_t)(off + c) * m] =
          (ii >= 0 && ii < p.rk.m) ? p.alpha * (double)PA[ii + (size_t)c * p.rk.la] : 0.0;
    }
    for (int e = threadIdx.x; e < n * kq; e += kBlock) {
      const int i = e % n, c = e / n, ii = i - p.coff;
      Bc[i + (size_t)(off + c) * n] = (ii >= 0 && ii < p.rk.n) ? (double)PB[ii + (size_t)c * p.rk.lb] : 0.0;
    }
    off += kq;
  }
  __syncthreads();
  const long t1 = wall_clock64();
  qr_truncate<T>(Ac, m, Bc, n, K, carve_qr(cr, y, K), t.eps, tg, overflow, red, &flag, pr);
  if (pr && threadIdx.x == 0) {
    pr[0] = 2 + (Ac == sh ? 0 : (cr == sh ? 1 : 2)); pr[1] = m; pr[2] = n; pr[3] = K;
    pr[4] = (double)(wall_clock64() - t0); pr[5] = (double)(t1 - t0);
  }
}

// FADD_DENSE: RkMatrix::formattedAddParts with a full part (rk_matrix.cpp:984-1010)
template <typename T>
__global__ void __launch_bounds__(kBlock) k_fadd_dense(const FTask* tasks, double* wsbase, int* overflow) {
  __shared__ double sh[kShDoubles];
  __shared__ int flag;
  if (overflow[1]) return;
  const FTask& t = tasks[blockIdx.x];
  const RkRef tg = t.tgt;
  const int m = tg.m, n = tg.n, kt = rank_of(tg);
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
  double* ws = need <= kShDoubles ? sh : wsbase + t.ws;
  double* G = ws; ws += (size_t)gr * gc;
  double* V = ws; ws += (size_t)gc * gc;
  double* sig = ws; ws += gc;
  int* order = (int*)ws;
  const T* A1 = (const T*)r1.a;
  const T* B1 = (const T*)r1.b;
  const T* A2 = (const T*)r2.a;
  const T* B2 = (const T*)r2.b;
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
  jacobi_svd(G, gr, gr, gc, V, gc, &flag);
  sorted_norms(G, gr, gr, gc, sig, order);
  int nk = find_k(sig, order, gc, t.eps);
  if (nk > out.cap) {
    if (threadIdx.x == 0) atomicOr(overflow, 1);
    nk = out.cap;
  }
  // tmp = U S V^T with U = G(:,j)/sig (k1 side), V (k2 side); swapped when tr
  T* OA = (T*)out.a;
  T* OB = (T*)out.b;
  const int m = r1.m, n = r2.n;
  for (int e = threadIdx.x; e < m * nk; e += kBlock) {
    const int i = e % m, c = e / m, j = order[c];
    const double sq = sqrt(sig[j]);
    double s = 0.0;
    for (int l = 0; l < k1; ++l) {
      const double u = tr ? V[l + (size_t)j * gc] * sq : G[l + (size_t)j * gr] / sq;
      s += (double)A1[i + (size_t)l * r1.la] * u;
    }
    OA[i + (size_t)c * out.la] = (T)s;
  }
  for (int e = threadIdx.x; e < n * nk; e += kBlock) {
    const int i = e % n, c = e / n, j = order[c];
    const double sq = sqrt(sig[j]);
    double s = 0.0;
    for (int l = 0; l < k2; ++l) {
      const double vv = tr ? G[l + (size_t)j * gr] / sq : V[l + (size_t)j * gc] * sq;
      s += (double)B2[i + (size_t)l * r2.lb] * vv;
    }
    OB[i + (size_t)c * out.lb] = (T)s;
  }
  __syncthreads();
  if (threadIdx.x == 0) *out.rank = nk;
}

// ---------------------------------------------------------------------------
// ACAFULL: acaFull (compression.cpp:270-355) on a dense temp (modified)
// ---------------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(kBlock) k_acafull(const FTask* tasks, const int* fail) {
  __shared__ double rv[kWaves];
  __shared__ int ri[kWaves];
  __shared__ double red[kWaves];
  if (*fail) return;
  const FTask& t = tasks[blockIdx.x];
  RkRef out = t.tgt;
  T* D = (T*)t.A;
  const int m = t.m.s, n = t.n.s, ld = t.lda;
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
    // estimate (the reference sums l < nu - 1)
    double cross = 0.0;
    for (int l = 0; l + 1 < nu; ++l) {
      double da = 0.0, db = 0.0;
      for (int r = threadIdx.x; r < m; r += kBlock) da += (double)va[r] * (double)OA[r + (size_t)l * out.la];
      for (int c = threadIdx.x; c < n; c += kBlock) db += (double)vb[c] * (double)OB[c + (size_t)l * out.lb];
      da = block_sum(da, red);
      db = block_sum(db, red);
      cross += da * db;
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
  int64_t qr = (int64_t)(m + n) * kq + (int64_t)(m > n ? m : n) * kq + 2 * (int64_t)kq * kq + 5 * (int64_t)kq;
  return (std::max(dense, qr) + 127) / 64 * 64;
}
int64_t fadd_dense_workspace(int m, int n) {
  const int mn = m < n ? m : n;
  return ((int64_t)m * n + (int64_t)mn * mn + 2 * (int64_t)mn + 127) / 64 * 64;
}
int64_t rkrk_workspace(int k1, int k2) {
  const int64_t a = k1 > k2 ? k1 : k2, b = k1 < k2 ? k1 : k2;
  return (a * b + b * b + 2 * b + 127) / 64 * 64;
}

double* g_fadd_prof = nullptr;

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
      hipLaunchKernelGGL(k_trsm<T>, dim3(count), dim3(kBlock), 0, s, tasks, flags + 1);
      break;
    case T_FADD:
      hipLaunchKernelGGL(k_fadd<T>, dim3(count), dim3(kBlock), 0, s, tasks, ws, flags, g_fadd_prof);
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
    default:
      throw std::runtime_error("unknown task type");
  }
  HMX_CHECK(hipGetLastError());
}

template void launch_task_kernel<double>(int, const FTask*, int, const GemmTile*, int, double*, int*, hipStream_t);
template void launch_task_kernel<float>(int, const FTask*, int, const GemmTile*, int, double*, int*, hipStream_t);

}  // namespace hmx
