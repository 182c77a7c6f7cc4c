// Whole-device compression of very large admissible blocks.
//
// One workgroup per block (the ACA kernels) suits the thousands of small and
// middle-sized blocks of a partition; a block of tens of thousands of rows
// with a rank of several hundred keeps one compute unit busy for minutes while
// the rest of the device idles (the partial ACA re-reads every earlier vector
// at each of its k iterations, and the truncation's QR streams k^2 m / 2
// entries).  Such blocks are evaluated densely (64-column slices, one
// workgroup each) and compressed by a randomized SVD on the whole device
// (Halko, Martinsson, Tropp, "Finding structure with randomness", SIAM Review
// 53(2), 2011: Gaussian test matrix, one power iteration, algorithms 4.4 and
// 5.1) whose every step is a GEMM (rocBLAS) or a small symmetric eigenproblem
// (rocSOLVER syevd):
//
//   Q = orth(A orth(A^T orth(A W)))     m x p,  W n x p Gaussian
//   Z = A^T Q,  Z^T Z = U L U^T          (Q^T A = Z^T has singular values L^1/2)
//   A ~ Q Q^T A = (Q U L^1/4) (Z U L^-1/4)^T
//
// orth(Y) is an eigen-decomposition QR taken twice: Y^T Y = V M V^T, Y <-
// Y V M^-1/2 over the eigenvalues above 1e-13 of the largest (directions
// below that are far under any cut and are dropped; the second pass restores
// orthogonality to rounding).  The cut is where the singular values fall to
// eps * s_0 (findK scalar_array.cpp:98-116, the rule of the reference's SVD
// compression, compression.cpp:224-239).  The sketch width p starts at 256
// and grows (fourfold when no singular value fell below the cut, else
// twofold) until the cut leaves at least p / 8 (and 32) columns of
// oversampling, the basis lost directions, or p reaches min(m, n).
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <cmath>
#include <stdexcept>
#include <string>
#include <vector>

#include "engine.hpp"
#include "kernels.hpp"

namespace hmx {

namespace {
void rb_check(rocblas_status s, const char* what) {
  if (s != rocblas_status_success)
    throw std::runtime_error(std::string("large-block compression: ") + what +
                             " failed (rocblas status " + std::to_string((int)s) + ")");
}

struct BlasHandle {
  rocblas_handle h = nullptr;
  explicit BlasHandle(hipStream_t s) {
    rb_check(rocblas_create_handle(&h), "rocblas_create_handle");
    rb_check(rocblas_set_stream(h, s), "rocblas_set_stream");
    rb_check(rocblas_set_pointer_mode(h, rocblas_pointer_mode_host), "rocblas_set_pointer_mode");
  }
  ~BlasHandle() {
    if (h) rocblas_destroy_handle(h);
  }
};

// C (rows x cols, ld rows) = op(X) op(Y)
void gemm(rocblas_handle h, bool tx, bool ty, int rows, int cols, int inner, const double* X,
          int ldx, const double* Y, int ldy, double* C) {
  const double one = 1.0, zero = 0.0;
  rb_check(rocblas_dgemm(h, tx ? rocblas_operation_transpose : rocblas_operation_none,
                         ty ? rocblas_operation_transpose : rocblas_operation_none, rows, cols,
                         inner, &one, X, ldx, Y, ldy, &zero, C, rows),
           "dgemm");
}

// eigenvalues (ascending) and eigenvectors of the p x p symmetric G (in place)
void syevd(rocblas_handle h, double* G, int p, double* lam, double* work, int* info,
           hipStream_t st) {
  rb_check(rocsolver_dsyevd(h, rocblas_evect_original, rocblas_fill_upper, p, G, p, lam, work,
                            info),
           "syevd");
  int fail = 0;
  HMX_CHECK(hipMemcpyAsync(&fail, info, sizeof(int), hipMemcpyDeviceToHost, st));
  HMX_CHECK(hipStreamSynchronize(st));
  if (fail != 0) throw std::runtime_error("large-block compression: the eigensolver did not converge");
}

struct Work {
  DevBuf G, V, lam, e, info, T;
};

// Y (rows x p, ld rows) -> an orthonormal basis of its range in T (rows x r,
// ld rows), r <= p (see the header); returns r
int orth(rocblas_handle h, const double* Y, int rows, int p, Work& w, hipStream_t st) {
  w.G.ensure(sizeof(double) * (size_t)p * p);
  w.lam.ensure(sizeof(double) * (size_t)p);
  w.e.ensure(sizeof(double) * (size_t)p);
  w.info.ensure(sizeof(int));
  w.V.ensure(sizeof(double) * (size_t)p * p);
  w.T.ensure(sizeof(double) * (size_t)rows * p);
  const double* src = Y;
  int r = p;
  for (int pass = 0; pass < 2; ++pass) {
    gemm(h, true, false, r, r, rows, src, rows, src, rows, w.G.as<double>());
    syevd(h, w.G.as<double>(), r, w.lam.as<double>(), w.e.as<double>(), w.info.as<int>(), st);
    std::vector<double> lam(r);
    HMX_CHECK(hipMemcpyAsync(lam.data(), w.lam.p, sizeof(double) * r, hipMemcpyDeviceToHost, st));
    HMX_CHECK(hipStreamSynchronize(st));
    int first = 0;
    while (first < r && !(lam[first] > 1e-13 * lam[r - 1])) ++first;
    const int keep = r - first;
    if (keep == 0) return 0;
    launch_pick_scale(w.G.as<double>(), r, r, first, keep, w.lam.as<double>(), -0.5,
                      w.V.as<double>(), st);
    // the product goes to the buffer the pass does not read
    double* dst = pass == 0 ? w.T.as<double>() : const_cast<double*>(Y);
    gemm(h, false, false, rows, keep, r, src, rows, w.V.as<double>(), r, dst);
    src = dst;
    r = keep;
  }
  // the result sits in Y's buffer after the second pass: copy it to T
  HMX_CHECK(hipMemcpyAsync(w.T.p, Y, sizeof(double) * (size_t)rows * r, hipMemcpyDeviceToDevice, st));
  return r;
}
}  // namespace

void Matrix::compress_giants(const std::vector<int>& list, PointSet rp, PointSet cp,
                             KernelParams kp, double eps) {
  if (list.empty()) return;
  BlasHandle H(stream);
  DevBuf A, Y, Z, sd, sl;
  Work w;
  int done = 0;
  for (int l : list) {
    report_assembly(done++, (int)bt->leaves.size());
    if (giant_.count(l)) continue;
    const LeafDesc L = s64.hdesc[l];
    const int m = L.m, n = L.n, mn = std::min(m, n);
    A.ensure(sizeof(double) * (size_t)m * n);
    {
      std::vector<LeafDesc> d;
      std::vector<int> lst;
      for (int j0 = 0; j0 < n; j0 += 64) {
        LeafDesc s = L;
        s.kind = 0;
        s.cap = 0;
        s.n = std::min(64, n - j0);
        s.c0 = L.c0 + j0;
        s.a = (int64_t)j0 * m;
        s.b = -1;
        lst.push_back((int)d.size());
        d.push_back(s);
      }
      upload(sd, d.data(), d.size() * sizeof(LeafDesc), stream);
      upload(sl, lst.data(), lst.size() * sizeof(int), stream);
      launch_assemble_full<double>(sd.as<LeafDesc>(), sl.as<int>(), (int)d.size(), A.as<double>(),
                                   rp, cp, kp, stream);
    }
    const double* a = A.as<double>();
    int p = std::min(mn, 256), k = 0, r = 0;
    for (;;) {
      Y.ensure(sizeof(double) * (size_t)std::max(m, n) * p);
      Z.ensure(sizeof(double) * (size_t)std::max(m, n) * p);
      double *y = Y.as<double>(), *z = Z.as<double>();
      launch_gauss(z, (size_t)n * p, 0x5EED0000ull + (uint64_t)l * 4096 + (uint64_t)p, stream);
      gemm(H.h, false, false, m, p, n, a, m, z, n, y);    // A W
      // a basis that loses every direction means a numerically zero block
      k = 0;
      if ((r = orth(H.h, y, m, p, w, stream)) == 0) break;
      gemm(H.h, true, false, n, r, m, a, m, w.T.as<double>(), m, z);  // A^T Q
      if ((r = orth(H.h, z, n, r, w, stream)) == 0) break;
      gemm(H.h, false, false, m, r, n, a, m, w.T.as<double>(), n, y);  // A Q'
      if ((r = orth(H.h, y, m, r, w, stream)) == 0) break;
      HMX_CHECK(hipMemcpyAsync(y, w.T.p, sizeof(double) * (size_t)m * r, hipMemcpyDeviceToDevice,
                               stream));                  // Q
      gemm(H.h, true, false, n, r, m, a, m, y, m, z);      // Z = A^T Q
      w.G.ensure(sizeof(double) * (size_t)r * r);
      gemm(H.h, true, false, r, r, n, z, n, z, n, w.G.as<double>());
      syevd(H.h, w.G.as<double>(), r, w.lam.as<double>(), w.e.as<double>(), w.info.as<int>(),
            stream);
      std::vector<double> lam(r);
      HMX_CHECK(hipMemcpyAsync(lam.data(), w.lam.p, sizeof(double) * r, hipMemcpyDeviceToHost,
                               stream));
      HMX_CHECK(hipStreamSynchronize(stream));
      // singular values of Q^T A, largest first: sqrt(lam[r-1]), sqrt(lam[r-2]), ...
      const double s0 = r > 0 && lam[r - 1] > 0.0 ? std::sqrt(lam[r - 1]) : 0.0;
      k = 0;
      if (s0 > 0.0)
        while (k < r && lam[r - 1 - k] > 0.0 && std::sqrt(lam[r - 1 - k]) > eps * s0) ++k;
      if (p >= mn || r < p || k + std::max(p / 8, 32) <= p) break;
      p = std::min(k == r ? 4 * p : 2 * p, mn);
    }
    auto g = std::make_unique<GiantPanels>();
    g->rank = k;
    if (k > 0) {
      // a = Q U_k L^1/4, b = Z U_k L^-1/4 (U_k: the k largest, in w.G)
      w.V.ensure(sizeof(double) * (size_t)r * k);
      g->a.alloc(sizeof(double) * (size_t)m * k);
      g->b.alloc(sizeof(double) * (size_t)n * k);
      launch_pick_scale(w.G.as<double>(), r, r, r - k, k, w.lam.as<double>(), 0.25,
                        w.V.as<double>(), stream);
      gemm(H.h, false, false, m, k, r, Y.as<double>(), m, w.V.as<double>(), r, g->a.as<double>());
      launch_pick_scale(w.G.as<double>(), r, r, r - k, k, w.lam.as<double>(), -0.25,
                        w.V.as<double>(), stream);
      gemm(H.h, false, false, n, k, r, Z.as<double>(), n, w.V.as<double>(), r, g->b.as<double>());
    }
    HMX_CHECK(hipStreamSynchronize(stream));
    giant_[l] = std::move(g);
  }
}

}  // namespace hmx
