// Assembly kernels: full-leaf evaluation, SVD compression of small admissible
// leaves, ACA+ / ACA-partial compression followed by Rk truncation.
// One workgroup per leaf (batched over a list of leaves).
#include "dev_linalg.hpp"
#include "kernels.hpp"

namespace hmx {

namespace {

__device__ inline double entry(const PointSet& rp, const PointSet& cp, const KernelParams& kp,
                               int i, int j) {
  const double dx = rp.xyz[i] - cp.xyz[j];
  const double dy = rp.xyz[rp.n + i] - cp.xyz[cp.n + j];
  const double dz = rp.xyz[2 * rp.n + i] - cp.xyz[2 * cp.n + j];
  if (kp.kind == kHelmholtz) return helmholtz_embedded(kp, dx, dy, dz, rp.user[i], cp.user[j]);
  return kernel_value(kp, dx, dy, dz, rp.user[i] == cp.user[j]);
}

template <typename T>
__global__ void __launch_bounds__(kBlock) k_assemble_full(const LeafDesc* desc, const int* list,
                                                          T* arena, PointSet rp, PointSet cp,
                                                          KernelParams kp) {
  const LeafDesc d = desc[list[blockIdx.x]];
  T* out = arena + d.a;
  const int total = d.m * d.n;
  for (int e = threadIdx.x; e < total; e += kBlock) {
    const int i = e % d.m, j = e / d.m;
    out[e] = (T)entry(rp, cp, kp, d.r0 + i, d.c0 + j);
  }
}

// Full leaves whose entries were evaluated on the host (user callbacks): the
// staged block (column-major, ld m, doubles) converted into the leaf storage.
template <typename T>
__global__ void __launch_bounds__(kBlock) k_table_full(const LeafDesc* desc, const int* list,
                                                       const int64_t* tab_off, const double* table,
                                                       T* arena) {
  const LeafDesc d = desc[list[blockIdx.x]];
  const double* src = table + tab_off[blockIdx.x];
  T* out = arena + d.a;
  const int total = d.m * d.n;
  for (int e = threadIdx.x; e < total; e += kBlock) out[e] = (T)src[e];
}

constexpr int kSvdMax = 99;  // compressionMinLeafSize - 1

template <typename T>
__global__ void __launch_bounds__(kBlock) k_svd_compress(const LeafDesc* desc, const int* list,
                                                         T* arena, int* ranks, int* overflow,
                                                         PointSet rp, PointSet cp,
                                                         KernelParams kp, double eps,
                                                         const int64_t* tab_off,
                                                         const double* table) {
  __shared__ double G[kSvdMax * kSvdMax];
  __shared__ double V[kSvdMax * kSvdMax];
  __shared__ double sig[kSvdMax];
  __shared__ int order[kSvdMax];
  __shared__ int flag;
  const int leaf = list[blockIdx.x];
  const LeafDesc d = desc[leaf];
  // G = A (m >= n) or A^T (m < n), so that G has min(m, n) columns.
  const bool tr = d.m < d.n;
  const int gr = tr ? d.n : d.m, gc = tr ? d.m : d.n;
  // entries from the kernel, or from a block staged by the host (tab_off)
  const double* tab = tab_off ? table + tab_off[blockIdx.x] : nullptr;
  for (int e = threadIdx.x; e < d.m * d.n; e += kBlock) {
    const int i = e % d.m, j = e / d.m;
    const double v = tab ? tab[e] : entry(rp, cp, kp, d.r0 + i, d.c0 + j);
    if (tr)
      G[j + i * gr] = v;
    else
      G[i + j * gr] = v;
  }
  __syncthreads();
  jacobi_svd(G, gr, gr, gc, V, gc, &flag);
  sorted_norms(G, gr, gr, gc, sig, order);
  int k = find_k(sig, order, gc, eps);
  if (k > d.cap) {
    if (threadIdx.x == 0) atomicOr(overflow, 1);
    k = d.cap;
  }
  // G = U S V^T.  A = U sqrt(S), B = V sqrt(S) (or swapped when transposed).
  T* A = arena + d.a;
  T* B = arena + d.b;
  for (int e = threadIdx.x; e < k * gr; e += kBlock) {
    const int i = e % gr, c = e / gr;
    const int j = order[c];
    const double v = G[i + j * gr] / sqrt(sig[j]);
    if (tr)
      B[i + (size_t)c * d.n] = (T)v;
    else
      A[i + (size_t)c * d.m] = (T)v;
  }
  for (int e = threadIdx.x; e < k * gc; e += kBlock) {
    const int i = e % gc, c = e / gc;
    const int j = order[c];
    const double v = V[i + j * gc] * sqrt(sig[j]);
    if (tr)
      A[i + (size_t)c * d.m] = (T)v;
    else
      B[i + (size_t)c * d.n] = (T)v;
  }
  if (threadIdx.x == 0) ranks[leaf] = k;
}

// ---------------------------------------------------------------------------
// ACA
// ---------------------------------------------------------------------------
struct AcaWs {
  double *acols, *bcols, *aref, *bref, *avec, *bvec;
  int *rowfree, *colfree;
  // truncation
  double *ya, *yb, *g, *v, *taua, *taub, *sig;
  int* order;
};

__device__ inline AcaWs carve(double* base, int m, int n, int kmax) {
  AcaWs w;
  double* p = base;
  w.acols = p; p += (size_t)m * kmax;
  w.bcols = p; p += (size_t)n * kmax;
  w.aref = p; p += m;
  w.bref = p; p += n;
  w.avec = p; p += m;
  w.bvec = p; p += n;
  w.ya = p; p += (size_t)m * kmax;
  w.yb = p; p += (size_t)n * kmax;
  w.g = p; p += (size_t)kmax * kmax;
  w.v = p; p += (size_t)kmax * kmax;
  w.taua = p; p += kmax;
  w.taub = p; p += kmax;
  w.sig = p; p += kmax;
  w.order = (int*)p; p += kmax;
  w.rowfree = (int*)p; p += m;
  w.colfree = (int*)p; p += n;
  return w;
}

struct AcaCtx {
  const PointSet* rp;
  const PointSet* cp;
  const KernelParams* kp;
  int r0, c0, m, n;
  const double* tab;  // the block staged by the host (ld m), or null: evaluate the kernel
};

__device__ inline void get_row(const AcaCtx& c, int i, double* out) {
  if (c.tab) {
    for (int j = threadIdx.x; j < c.n; j += kBlock) out[j] = c.tab[i + (size_t)j * c.m];
    return;
  }
  for (int j = threadIdx.x; j < c.n; j += kBlock) out[j] = entry(*c.rp, *c.cp, *c.kp, c.r0 + i, c.c0 + j);
}
__device__ inline void get_col(const AcaCtx& c, int j, double* out) {
  if (c.tab) {
    for (int i = threadIdx.x; i < c.m; i += kBlock) out[i] = c.tab[i + (size_t)j * c.m];
    return;
  }
  for (int i = threadIdx.x; i < c.m; i += kBlock) out[i] = entry(*c.rp, *c.cp, *c.kp, c.r0 + i, c.c0 + j);
}
// vec[t] -= sum_{l<k} other_l[idx] * same_l[t]   (updateRow compression.cpp:82-89)
__device__ inline void update(double* vec, int len, int idx, const double* same, int ld_same,
                              const double* other, int ld_other, int k) {
  __syncthreads();
  for (int t = threadIdx.x; t < len; t += kBlock) {
    double v = vec[t];
    for (int l = 0; l < k; ++l) v -= other[(size_t)l * ld_other + idx] * same[(size_t)l * ld_same + t];
    vec[t] = v;
  }
  __syncthreads();
}
__device__ inline bool any_nonzero(const double* v, int len, double* red) {
  double s = 0.0;
  for (int t = threadIdx.x; t < len; t += kBlock) s += (v[t] != 0.0) ? 1.0 : 0.0;
  return block_sum(s, red) > 0.0;
}
__device__ inline int amax(const double* v, int len, double* rv, int* ri) {
  ArgMax a{-1.0, -1};
  for (int t = threadIdx.x; t < len; t += kBlock) a = better_max(a, ArgMax{fabs(v[t]), t});
  const int i = block_argmax(a, rv, ri).i;
  return i < 0 ? 0 : i;  // all-NaN vector: any in-range index
}
__device__ inline double dot(const double* a, const double* b, int len, double* red) {
  double s = 0.0;
  for (int t = threadIdx.x; t < len; t += kBlock) s += a[t] * b[t];
  return block_sum(s, red);
}
__device__ inline void copy_vec(double* dst, const double* src, int len) {
  for (int t = threadIdx.x; t < len; t += kBlock) dst[t] = src[t];
}
__device__ inline void scale_vec(double* v, int len, double f) {
  for (int t = threadIdx.x; t < len; t += kBlock) v[t] *= f;
}
// est increment: sum_{l<k} dot(a, A_l) dot(b, B_l); waves split l.
__device__ inline double cross_sum(const double* a, const double* A, int m, const double* b,
                                   const double* B, int n, int k, double* red, double* tmp) {
  const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
  for (int l = w; l < k; l += kWaves) {
    double da = 0.0, db = 0.0;
    for (int t = lane; t < m; t += kWave) da += a[t] * A[(size_t)l * m + t];
    for (int t = lane; t < n; t += kWave) db += b[t] * B[(size_t)l * n + t];
    da = wave_sum(da);
    db = wave_sum(db);
    if (lane == 0) tmp[l] = da * db;
  }
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0)
    for (int l = 0; l < k; ++l) s += tmp[l];
  return block_sum(s, red);
}

// findCol compression.cpp:129-147: first free column that is not zero.
__device__ int find_col(const AcaCtx& c, int* colfree, double* col, double* red) {
  for (int j = 0; j < c.n; ++j) {
    if (colfree[j]) {
      __syncthreads();
      get_col(c, j, col);
      __syncthreads();
      if (threadIdx.x == 0) colfree[j] = 0;
      if (any_nonzero(col, c.m, red)) return j;
    }
  }
  return -1;
}

// findMinRow compression.cpp:150-184
__device__ int find_min_row(const AcaCtx& c, int* rowfree, const AcaWs& w, int k,
                            const double* aref, double* row, double* red, double* rv, int* ri) {
  while (true) {
    __syncthreads();
    ArgMax a{0.0, -1};
    for (int i = threadIdx.x; i < c.m; i += kBlock)
      if (rowfree[i]) a = better_min(a, ArgMax{aref[i] * aref[i], i});
    const int iref = block_argmin(a, rv, ri).i;
    if (iref < 0) return -1;
    get_row(c, iref, row);
    update(row, c.n, iref, w.bcols, c.n, w.acols, c.m, k);
    if (threadIdx.x == 0) rowfree[iref] = 0;
    if (any_nonzero(row, c.n, red)) return iref;
  }
}

// findMinCol compression.cpp:187-221
__device__ int find_min_col(const AcaCtx& c, int* colfree, const AcaWs& w, int k,
                            const double* bref, double* col, double* red, double* rv, int* ri) {
  while (true) {
    __syncthreads();
    ArgMax a{0.0, -1};
    for (int j = threadIdx.x; j < c.n; j += kBlock)
      if (colfree[j]) a = better_min(a, ArgMax{bref[j] * bref[j], j});
    const int jref = block_argmin(a, rv, ri).i;
    if (jref < 0) return -1;
    get_col(c, jref, col);
    update(col, c.m, jref, w.acols, c.m, w.bcols, c.n, k);
    if (threadIdx.x == 0) colfree[jref] = 0;
    if (any_nonzero(col, c.m, red)) return jref;
  }
}

// doCompressionAcaPlus compression.cpp:555-724 (without null-row hints).
__device__ int aca_plus(const AcaCtx& c, const AcaWs& w, int kmax, double eps, double* red,
                        double* rv, int* ri, double* tmp) {
  const int m = c.m, n = c.n;
  int maxk = m < n ? m : n;
  if (maxk > kmax) maxk = kmax;
  for (int i = threadIdx.x; i < m; i += kBlock) { w.rowfree[i] = 1; w.aref[i] = 0.0; }
  for (int j = threadIdx.x; j < n; j += kBlock) { w.colfree[j] = 1; w.bref[j] = 0.0; }
  __syncthreads();
  int jref = find_col(c, w.colfree, w.aref, red);
  if (jref < 0) return 0;
  int iref = find_min_row(c, w.rowfree, w, 0, w.aref, w.bref, red, rv, ri);
  if (iref < 0) {
    for (int j = threadIdx.x; j < n; j += kBlock) w.bref[j] = 0.0;
    __syncthreads();
  }
  double est = 0.0;
  int k = 0;
  do {
    double* a = w.acols + (size_t)k * m;
    double* b = w.bcols + (size_t)k * n;
    __syncthreads();
    int istar = amax(w.aref, m, rv, ri);
    int jstar = amax(w.bref, n, rv, ri);
    const double ai = w.aref[istar], bj = w.bref[jstar];
    if (ai * ai > bj * bj) {
      get_row(c, istar, b);
      update(b, n, istar, w.bcols, n, w.acols, m, k);
      jstar = amax(b, n, rv, ri);
      const double piv = b[jstar];
      __syncthreads();
      get_col(c, jstar, a);
      update(a, m, jstar, w.acols, m, w.bcols, n, k);
      if (piv != 0.0) scale_vec(a, m, 1.0 / piv);
    } else {
      get_col(c, jstar, a);
      update(a, m, jstar, w.acols, m, w.bcols, n, k);
      istar = amax(a, m, rv, ri);
      const double piv = a[istar];
      __syncthreads();
      get_row(c, istar, b);
      update(b, n, istar, w.bcols, n, w.acols, m, k);
      if (piv != 0.0) scale_vec(b, n, 1.0 / piv);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      w.rowfree[istar] = 0;
      w.colfree[jstar] = 0;
    }
    const double cross = cross_sum(a, w.acols, m, b, w.bcols, n, k, red, tmp);
    est += 2.0 * cross;
    const double an = dot(a, a, m, red), bn = dot(b, b, n, red);
    const double ab = an * bn;
    est += ab;
    ++k;
    if (ab < eps * eps * est) break;
    // update of the reference row and column
    const double fa = b[jref], fb = iref >= 0 ? a[iref] : 0.0;
    __syncthreads();
    for (int i = threadIdx.x; i < m; i += kBlock) w.aref[i] -= fa * a[i];
    for (int j = threadIdx.x; j < n; j += kBlock) w.bref[j] -= fb * b[j];
    __syncthreads();
    const bool needA = (jstar == jref) || !any_nonzero(w.aref, m, red);
    const bool needB = (istar == iref) || !any_nonzero(w.bref, n, red);
    if (needA && needB) {
      bool found = false;
      while (!found) {
        __syncthreads();
        jref = find_col(c, w.colfree, w.aref, red);
        if (jref < 0) break;
        update(w.aref, m, jref, w.acols, m, w.bcols, n, k);
        found = any_nonzero(w.aref, m, red);
      }
      if (!found) break;
      iref = find_min_row(c, w.rowfree, w, k, w.aref, w.bref, red, rv, ri);
      if (iref < 0) break;
    } else if (needB) {
      iref = find_min_row(c, w.rowfree, w, k, w.aref, w.bref, red, rv, ri);
      if (iref < 0) break;
    } else if (needA) {
      jref = find_min_col(c, w.colfree, w, k, w.bref, w.aref, red, rv, ri);
      if (jref < 0) break;
    }
  } while (k < maxk);
  __syncthreads();
  return k;
}

// doCompressionAcaPartial compression.cpp:384-533 (no random pivots).
__device__ int aca_partial(const AcaCtx& c, const AcaWs& w, int kmax, double eps, double* red,
                           double* rv, int* ri, double* tmp) {
  const int m = c.m, n = c.n;
  int maxk = m < n ? m : n;
  for (int i = threadIdx.x; i < m; i += kBlock) w.rowfree[i] = 1;
  for (int j = threadIdx.x; j < n; j += kBlock) w.colfree[j] = 1;
  __syncthreads();
  int row = 0, J = 0, k = 0, count = 0;
  double est = 0.0;
  do {
    if (k >= kmax) break;
    double* a = w.acols + (size_t)k * m;
    double* b = w.bcols + (size_t)k * n;
    __syncthreads();
    get_row(c, row, b);
    update(b, n, row, w.bcols, n, w.acols, m, k);
    if (threadIdx.x == 0) w.rowfree[row] = 0;
    __syncthreads();
    {
      ArgMax am{-1.0, -1};
      for (int j = threadIdx.x; j < n; j += kBlock) {
        const double v = b[j] * b[j];
        if (w.colfree[j] && v > 0.0) am = better_max(am, ArgMax{v, j});
      }
      const int jj = block_argmax(am, rv, ri).i;
      if (jj >= 0) J = jj;
    }
    if (b[J] == 0.0) {
      __syncthreads();
      row = 0;
      while (row < m && !w.rowfree[row]) ++row;
    } else {
      const double piv = 1.0 / b[J];
      __syncthreads();
      scale_vec(b, n, piv);
      __syncthreads();
      get_col(c, J, a);
      update(a, m, J, w.acols, m, w.bcols, n, k);
      if (threadIdx.x == 0) w.colfree[J] = 0;
      __syncthreads();
      {
        ArgMax am{-1.0, -1};
        for (int i = threadIdx.x; i < m; i += kBlock) {
          const double v = a[i] * a[i];
          if (w.rowfree[i] && v > 0.0) am = better_max(am, ArgMax{v, i});
        }
        const int ii = block_argmax(am, rv, ri).i;
        if (ii >= 0) row = ii;
      }
      const double cross = cross_sum(a, w.acols, m, b, w.bcols, n, k, red, tmp);
      est += 2.0 * cross;
      const double ab = dot(a, a, m, red) * dot(b, b, n, red);
      est += ab;
      ++k;
      if (ab < eps * eps * est) break;
    }
    ++count;
  } while (count < maxk && row < m);
  __syncthreads();
  return k;
}

// RkMatrix::truncate rk_matrix.cpp:490-585 on panels (double, ld m / n) with
// rank k <= min(m, n): QR of both panels, SVD of Ra Rb^T, cut, apply Q.
template <typename T>
__device__ int truncate_panels(double* A, int m, double* B, int n, int k, const AcaWs& w,
                               double eps, T* outA, T* outB, int cap, int* overflow, double* red,
                               int* flag) {
  house_qr(A, m, m, k, w.taua, red, w.sig);
  house_qr(B, n, n, k, w.taub, red, w.sig);
  // g = Ra Rb^T (k x k), Ra, Rb upper triangular
  for (int e = threadIdx.x; e < k * k; e += kBlock) {
    const int i = e % k, j = e / k;
    double s = 0.0;
    const int lo = i > j ? i : j;
    for (int l = lo; l < k; ++l) s += A[i + (size_t)l * m] * B[j + (size_t)l * n];
    w.g[i + (size_t)j * k] = s;
  }
  __syncthreads();
  jacobi_svd(w.g, k, k, k, w.v, k, flag);
  sorted_norms(w.g, k, k, k, w.sig, w.order);
  int nk = find_k(w.sig, w.order, k, eps);
  if (nk > cap) {
    if (threadIdx.x == 0) atomicOr(overflow, 1);
    nk = cap;
  }
  // ya = [U sqrt(S); 0], yb = [V sqrt(S); 0]
  for (int e = threadIdx.x; e < nk * m; e += kBlock) {
    const int i = e % m, c = e / m, j = w.order[c];
    w.ya[e] = (i < k) ? w.g[i + (size_t)j * k] / sqrt(w.sig[j]) : 0.0;
  }
  for (int e = threadIdx.x; e < nk * n; e += kBlock) {
    const int i = e % n, c = e / n, j = w.order[c];
    w.yb[e] = (i < k) ? w.v[i + (size_t)j * k] * sqrt(w.sig[j]) : 0.0;
  }
  __syncthreads();
  apply_q(A, m, m, k, w.taua, w.ya, m, nk, red, w.v);
  apply_q(B, n, n, k, w.taub, w.yb, n, nk, red, w.v);
  for (int e = threadIdx.x; e < nk * m; e += kBlock) outA[e] = (T)w.ya[e];
  for (int e = threadIdx.x; e < nk * n; e += kBlock) outB[e] = (T)w.yb[e];
  __syncthreads();
  return nk;
}

template <typename T>
__global__ void __launch_bounds__(kBlock) k_aca(const LeafDesc* desc, const AcaTask* tasks,
                                                T* arena, int* ranks, int* overflow, double* ws,
                                                PointSet rp, PointSet cp, KernelParams kp,
                                                double aca_eps, double trunc_eps, int partial,
                                                const double* table) {
  __shared__ double red[kWaves * kQrChunk];
  __shared__ double rv[kWaves];
  __shared__ int ri[kWaves];
  __shared__ int flag;
  __shared__ double tmp[1024];
  const AcaTask t = tasks[blockIdx.x];
  const LeafDesc d = desc[t.leaf];
  AcaWs w = carve(ws + t.ws, d.m, d.n, t.kmax);
  AcaCtx c{&rp, &cp, &kp, d.r0, d.c0, d.m, d.n, t.tab >= 0 ? table + t.tab : nullptr};
  int k;
  if (partial || (long)d.n * 100 < (long)d.m)
    k = aca_partial(c, w, t.kmax, aca_eps, red, rv, ri, tmp);
  else
    k = aca_plus(c, w, t.kmax, aca_eps, red, rv, ri, tmp);
  // stopped by the panel room rather than by the criterion: the host grows
  // the capacities and assembles again
  const int room = d.m < d.n ? d.m : d.n;
  if (k >= t.kmax && t.kmax < (room < 1024 ? room : 1024) && threadIdx.x == 0) atomicOr(overflow, 1);
  int nk = 0;
  if (k > 0)
    nk = truncate_panels<T>(w.acols, d.m, w.bcols, d.n, k, w, trunc_eps, arena + d.a,
                            arena + d.b, d.cap, overflow, red, &flag);
  if (threadIdx.x == 0) ranks[t.leaf] = nk;
}

}  // namespace

int64_t aca_workspace(int m, int n, int kmax) {
  int64_t s = 0;
  s += (int64_t)(m + n) * kmax * 2;  // panels + y
  s += 2 * (int64_t)(m + n);         // refs + vecs
  s += 2 * (int64_t)kmax * kmax;     // g, v
  s += 4 * (int64_t)kmax;            // taus, sig, order
  s += (int64_t)(m + n);             // free flags (ints, one double slot each)
  return (s + 63) / 64 * 64;
}

template <typename T>
void launch_assemble_full(const LeafDesc* desc, const int* list, int count, T* arena,
                          PointSet rp, PointSet cp, KernelParams kp, hipStream_t s) {
  if (count == 0) return;
  hipLaunchKernelGGL(k_assemble_full<T>, dim3(count), dim3(kBlock), 0, s, desc, list, arena, rp,
                     cp, kp);
  HMX_CHECK(hipGetLastError());
}

template <typename T>
void launch_svd_compress(const LeafDesc* desc, const int* list, int count, T* arena, int* ranks,
                         int* overflow, PointSet rp, PointSet cp, KernelParams kp, double eps,
                         hipStream_t s, const int64_t* tab_off, const double* table) {
  if (count == 0) return;
  hipLaunchKernelGGL(k_svd_compress<T>, dim3(count), dim3(kBlock), 0, s, desc, list, arena, ranks,
                     overflow, rp, cp, kp, eps, tab_off, table);
  HMX_CHECK(hipGetLastError());
}

template <typename T>
void launch_aca(const LeafDesc* desc, const AcaTask* tasks, int count, T* arena, int* ranks,
                int* overflow, double* ws, PointSet rp, PointSet cp, KernelParams kp,
                double aca_eps, double trunc_eps, int partial, hipStream_t s,
                const double* table) {
  if (count == 0) return;
  hipLaunchKernelGGL(k_aca<T>, dim3(count), dim3(kBlock), 0, s, desc, tasks, arena, ranks,
                     overflow, ws, rp, cp, kp, aca_eps, trunc_eps, partial, table);
  HMX_CHECK(hipGetLastError());
}

template <typename T>
void launch_table_full(const LeafDesc* desc, const int* list, const int64_t* tab_off, int count,
                       const double* table, T* arena, hipStream_t s) {
  if (count == 0) return;
  hipLaunchKernelGGL(k_table_full<T>, dim3(count), dim3(kBlock), 0, s, desc, list, tab_off, table,
                     arena);
  HMX_CHECK(hipGetLastError());
}

#define HMX_INST(T)                                                                          \
  template void launch_assemble_full<T>(const LeafDesc*, const int*, int, T*, PointSet,       \
                                        PointSet, KernelParams, hipStream_t);                 \
  template void launch_svd_compress<T>(const LeafDesc*, const int*, int, T*, int*, int*,      \
                                       PointSet, PointSet, KernelParams, double, hipStream_t,  \
                                       const int64_t*, const double*);                         \
  template void launch_aca<T>(const LeafDesc*, const AcaTask*, int, T*, int*, int*, double*,  \
                              PointSet, PointSet, KernelParams, double, double, int,          \
                              hipStream_t, const double*);                                    \
  template void launch_table_full<T>(const LeafDesc*, const int*, const int64_t*, int,        \
                                     const double*, T*, hipStream_t);
HMX_INST(double)
HMX_INST(float)

}  // namespace hmx
