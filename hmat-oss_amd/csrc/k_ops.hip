// Whole-matrix leaf operations behind the remaining hmat_interface_t slots:
// squared Frobenius norms of leaves (HMatrix::normSqr h_matrix.cpp:624-646,
// RkMatrix::normSqr, FullMatrix::normSqr), scaling (HMatrix::scale), adding a
// multiple of the identity (HMatrix::addIdentity) and reading the diagonal
// (HMatrix::extractDiagonal).  One workgroup per leaf of a list.  Also the
// element-wise helpers of the whole-device compression of very large blocks
// (seeded normal numbers, scaled eigenvector columns).
#include "kernels.hpp"

namespace hmx {

namespace {

// out[i] = ||leaf||_F^2 for the i-th listed leaf: full, sum of squares;
// Rk, sum over p, q of (A_p . A_q)(B_p . B_q) (the trace of (A^T A)(B^T B)).
template <typename T>
__global__ void __launch_bounds__(kBlock) k_leaf_norm2(const LeafDesc* desc, const int* ranks,
                                                       const T* arena, const int* list,
                                                       double* out) {
  __shared__ double red[kWaves];
  const int l = list[blockIdx.x];
  const LeafDesc L = desc[l];
  if (L.kind == 0) {
    const T* a = arena + L.a;
    const long cnt = (long)L.m * L.n;
    double s = 0.0;
    for (long e = threadIdx.x; e < cnt; e += kBlock) {
      const double v = (double)a[e];
      s += v * v;
    }
    s = block_sum(s, red);
    if (threadIdx.x == 0) out[blockIdx.x] = s;
    return;
  }
  const int k = ranks[l];
  const T* A = arena + L.a;
  const T* B = arena + L.b;
  double tot = 0.0;
  for (int p = 0; p < k; ++p)
    for (int q = p; q < k; ++q) {
      double sa = 0.0, sb = 0.0;
      for (int i = threadIdx.x; i < L.m; i += kBlock)
        sa += (double)A[i + (long)p * L.m] * (double)A[i + (long)q * L.m];
      for (int i = threadIdx.x; i < L.n; i += kBlock)
        sb += (double)B[i + (long)p * L.n] * (double)B[i + (long)q * L.n];
      sa = block_sum(sa, red);
      sb = block_sum(sb, red);
      tot += (p == q ? 1.0 : 2.0) * sa * sb;
    }
  if (threadIdx.x == 0) out[blockIdx.x] = tot;
}

// Full leaf: every entry times alpha; Rk leaf: its A panel times alpha.
template <typename T>
__global__ void k_leaf_scale(const LeafDesc* desc, const int* list, T* arena, double alpha) {
  const LeafDesc L = desc[list[blockIdx.x]];
  T* a = arena + L.a;
  const long cnt = (long)L.m * (L.kind == 0 ? L.n : L.cap);
  for (long e = threadIdx.x; e < cnt; e += blockDim.x) a[e] = (T)(alpha * (double)a[e]);
}

// Diagonal full leaves: a_ii += alpha.
template <typename T>
__global__ void k_leaf_add_diag(const LeafDesc* desc, const int* list, T* arena, double alpha) {
  const LeafDesc L = desc[list[blockIdx.x]];
  T* a = arena + L.a;
  for (int i = threadIdx.x; i < L.m; i += blockDim.x) a[i + (long)i * L.m] += (T)alpha;
}

// Diagonal full leaves: out[r0 + i] = a_ii (hmat numbering).
template <typename T>
__global__ void k_leaf_get_diag(const LeafDesc* desc, const int* list, const T* arena,
                                double* out) {
  const LeafDesc L = desc[list[blockIdx.x]];
  const T* a = arena + L.a;
  for (int i = threadIdx.x; i < L.m; i += blockDim.x) out[L.r0 + i] = (double)a[i + (long)i * L.m];
}

}  // namespace

template <typename T>
void launch_leaf_norm2(const LeafDesc* desc, const int* ranks, const T* arena, const int* list,
                       int count, double* out, hipStream_t s) {
  if (count == 0) return;
  hipLaunchKernelGGL(k_leaf_norm2<T>, dim3(count), dim3(kBlock), 0, s, desc, ranks, arena, list, out);
  HMX_CHECK(hipGetLastError());
}

template <typename T>
void launch_leaf_scale(const LeafDesc* desc, const int* list, int count, T* arena, double alpha,
                       hipStream_t s) {
  if (count == 0) return;
  hipLaunchKernelGGL(k_leaf_scale<T>, dim3(count), dim3(256), 0, s, desc, list, arena, alpha);
  HMX_CHECK(hipGetLastError());
}

template <typename T>
void launch_leaf_add_diag(const LeafDesc* desc, const int* list, int count, T* arena,
                          double alpha, hipStream_t s) {
  if (count == 0) return;
  hipLaunchKernelGGL(k_leaf_add_diag<T>, dim3(count), dim3(256), 0, s, desc, list, arena, alpha);
  HMX_CHECK(hipGetLastError());
}

template <typename T>
void launch_leaf_get_diag(const LeafDesc* desc, const int* list, int count, const T* arena,
                          double* out, hipStream_t s) {
  if (count == 0) return;
  hipLaunchKernelGGL(k_leaf_get_diag<T>, dim3(count), dim3(256), 0, s, desc, list, arena, out);
  HMX_CHECK(hipGetLastError());
}

namespace {
// splitmix64 finaliser (Steele, Lea, Flood: "Fast splittable pseudorandom
// number generators", OOPSLA 2014)
__device__ inline uint64_t mix64(uint64_t x) {
  x ^= x >> 30;
  x *= 0xBF58476D1CE4E5B9ull;
  x ^= x >> 27;
  x *= 0x94D049BB133111EBull;
  x ^= x >> 31;
  return x;
}

// standard normal numbers, entry i a function of (seed, i) only: two
// uniforms from the counter by the finaliser, then Box-Muller
__global__ void k_gauss(double* out, size_t n, uint64_t seed) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    const uint64_t z = seed + 0x9E3779B97F4A7C15ull * (i + 1);
    const uint64_t a = mix64(z), b = mix64(z ^ 0xD1B54A32D192ED03ull);
    const double u1 = ((double)(a >> 11) + 0.5) * 0x1.0p-53;
    const double u2 = (double)(b >> 11) * 0x1.0p-53;
    out[i] = sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
  }
}

// out (rows x count, ld rows) column c = V column first + count - 1 - c
// (ld ldv) times lam[first + count - 1 - c]^e: the last `count` columns of an
// eigenvector matrix (eigenvalues ascending), largest first, scaled
__global__ void k_pick_scale(const double* V, int ldv, int rows, int first, int count,
                             const double* lam, double e, double* out) {
  const size_t total = (size_t)rows * count;
  for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < total;
       q += (size_t)gridDim.x * blockDim.x) {
    const int i = (int)(q % rows), c = (int)(q / rows);
    const int j = first + count - 1 - c;
    out[q] = V[i + (size_t)j * ldv] * pow(lam[j], e);
  }
}

// HMatrix::transpose: one workgroup per item.  kind 0: dst (m x n, ld m) =
// src^T (src n x m, ld n), entries whose row and column parities differ
// negated when conj (the real representation of a complex transpose); kind
// 1: rows 2i + 1 of the m x n panel at dst negated in place.
template <typename T>
__global__ void k_leaf_transpose(const TransposeItem* items, T* arena, const T* copy) {
  const TransposeItem it = items[blockIdx.x];
  T* dst = arena + it.dst;
  const size_t total = (size_t)it.m * it.n;
  if (it.kind == 0) {
    const T* src = copy + it.src;
    for (size_t e = threadIdx.x; e < total; e += blockDim.x) {
      const int i = (int)(e % it.m), j = (int)(e / it.m);
      const T v = src[j + (size_t)i * it.n];
      dst[e] = (it.conj && ((i ^ j) & 1)) ? -v : v;
    }
  } else {
    for (size_t e = threadIdx.x; e < total; e += blockDim.x)
      if ((e % it.m) & 1) dst[e] = -dst[e];
  }
}

unsigned grid_for(size_t n) {
  const size_t b = (n + 255) / 256;
  return (unsigned)(b < 65536 ? (b ? b : 1) : 65536);
}
}  // namespace

void launch_gauss(double* out, size_t n, uint64_t seed, hipStream_t s) {
  hipLaunchKernelGGL(k_gauss, dim3(grid_for(n)), dim3(256), 0, s, out, n, seed);
  HMX_CHECK(hipGetLastError());
}

template <typename T>
void launch_leaf_transpose(const TransposeItem* items, int count, T* arena, const T* copy,
                           hipStream_t s) {
  if (count == 0) return;
  hipLaunchKernelGGL(k_leaf_transpose<T>, dim3(count), dim3(256), 0, s, items, arena, copy);
  HMX_CHECK(hipGetLastError());
}
template void launch_leaf_transpose<double>(const TransposeItem*, int, double*, const double*, hipStream_t);
template void launch_leaf_transpose<float>(const TransposeItem*, int, float*, const float*, hipStream_t);

void launch_pick_scale(const double* V, int ldv, int rows, int first, int count,
                       const double* lam, double e, double* out, hipStream_t s) {
  hipLaunchKernelGGL(k_pick_scale, dim3(grid_for((size_t)rows * count)), dim3(256), 0, s, V, ldv,
                     rows, first, count, lam, e, out);
  HMX_CHECK(hipGetLastError());
}

#define HMX_OPS_INST(T)                                                                          \
  template void launch_leaf_norm2<T>(const LeafDesc*, const int*, const T*, const int*, int,     \
                                     double*, hipStream_t);                                      \
  template void launch_leaf_scale<T>(const LeafDesc*, const int*, int, T*, double, hipStream_t); \
  template void launch_leaf_add_diag<T>(const LeafDesc*, const int*, int, T*, double,            \
                                        hipStream_t);                                            \
  template void launch_leaf_get_diag<T>(const LeafDesc*, const int*, int, const T*, double*,     \
                                        hipStream_t);
HMX_OPS_INST(double)
HMX_OPS_INST(float)

}  // namespace hmx
