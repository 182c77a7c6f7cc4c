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
// 5.1), with library GEMMs (rocBLAS) and QR / SVD (rocSOLVER):
//
//   Q = orth(A (orth(A^T orth(A W))))       m x p,  W n x p Gaussian
//   A^T Q = Q2 R                            Householder QR, R p x p
//   R = U S V^T                             SVD
//   A ~ Q Q^T A = Q R^T Q2^T = (Q V) S (Q2 U)^T
//
// cut where the singular values fall to eps * s_0 (findK scalar_array.cpp:
// 98-116, the rule of the reference's SVD compression, compression.cpp:
// 224-239), the panels Q V sqrt(S) and Q2 U sqrt(S).  The sketch width p
// starts at 256 and doubles until the cut leaves at least p / 8 (and 32)
// columns of oversampling, or p reaches min(m, n).
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <stdexcept>
#include <string>

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

// Y (m x p) replaced by the orthonormal factor of its Householder QR
void orth(rocblas_handle h, double* Y, int m, int p, double* tau) {
  rb_check(rocsolver_dgeqrf(h, m, p, Y, m, tau), "geqrf");
  rb_check(rocsolver_dorgqr(h, m, p, p, Y, m, tau), "orgqr");
}
}  // namespace

void Matrix::compress_giants(const std::vector<int>& list, PointSet rp, PointSet cp,
                             KernelParams kp, double eps) {
  if (list.empty()) return;
  BlasHandle H(stream);
  DevBuf A, W, Y, Z, tau, R, S, U, Vt, E, info, sd, sl;
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
    int p = std::min(mn, 256), k = 0;
    std::vector<double> s;
    for (;;) {
      W.ensure(sizeof(double) * (size_t)n * p);
      Y.ensure(sizeof(double) * (size_t)m * p);
      Z.ensure(sizeof(double) * (size_t)n * p);
      tau.ensure(sizeof(double) * (size_t)p);
      R.ensure(sizeof(double) * (size_t)p * p);
      U.ensure(sizeof(double) * (size_t)p * p);
      Vt.ensure(sizeof(double) * (size_t)p * p);
      S.ensure(sizeof(double) * (size_t)p);
      E.ensure(sizeof(double) * (size_t)std::max(p, 2));
      info.ensure(sizeof(int));
      const double* a = A.as<double>();
      double *w = W.as<double>(), *y = Y.as<double>(), *z = Z.as<double>(), *t = tau.as<double>();
      launch_gauss(w, (size_t)n * p, 0x5EED0000ull + (uint64_t)l * 4096 + (uint64_t)p, stream);
      gemm(H.h, false, false, m, p, n, a, m, w, n, y);  // A W
      orth(H.h, y, m, p, t);
      gemm(H.h, true, false, n, p, m, a, m, y, m, z);   // A^T Q
      orth(H.h, z, n, p, t);
      gemm(H.h, false, false, m, p, n, a, m, z, n, y);  // A Q'
      orth(H.h, y, m, p, t);                            // Q
      gemm(H.h, true, false, n, p, m, a, m, y, m, z);   // A^T Q
      rb_check(rocsolver_dgeqrf(H.h, n, p, z, n, t), "geqrf");
      launch_upper(z, n, R.as<double>(), p, stream);
      rb_check(rocsolver_dorgqr(H.h, n, p, p, z, n, t), "orgqr");  // Q2
      rb_check(rocsolver_dgesvd(H.h, rocblas_svect_singular, rocblas_svect_singular, p, p,
                                R.as<double>(), p, S.as<double>(), U.as<double>(), p,
                                Vt.as<double>(), p, E.as<double>(), rocblas_outofplace,
                                info.as<int>()),
               "gesvd");
      s.assign(p, 0.0);
      int fail = 0;
      HMX_CHECK(hipMemcpyAsync(s.data(), S.p, sizeof(double) * p, hipMemcpyDeviceToHost, stream));
      HMX_CHECK(hipMemcpyAsync(&fail, info.p, sizeof(int), hipMemcpyDeviceToHost, stream));
      HMX_CHECK(hipStreamSynchronize(stream));
      if (fail != 0) throw std::runtime_error("large-block compression: the SVD did not converge");
      k = 0;
      if (s[0] > 0.0)
        while (k < p && s[k] > eps * s[0]) ++k;
      if (p >= mn || k + std::max(p / 8, 32) <= p) break;
      p = std::min(2 * p, mn);
    }
    auto g = std::make_unique<GiantPanels>();
    g->rank = k;
    if (k > 0) {
      // Q V sqrt(S): rows j < k of V^T scaled; Q2 U sqrt(S): columns j < k of U
      launch_scale_sqrt(Vt.as<double>(), p, p, p, S.as<double>(), k, true, stream);
      launch_scale_sqrt(U.as<double>(), p, p, p, S.as<double>(), k, false, stream);
      g->a.alloc(sizeof(double) * (size_t)m * k);
      g->b.alloc(sizeof(double) * (size_t)n * k);
      gemm(H.h, false, true, m, k, p, Y.as<double>(), m, Vt.as<double>(), p, g->a.as<double>());
      gemm(H.h, false, false, n, k, p, Z.as<double>(), n, U.as<double>(), p, g->b.as<double>());
    }
    HMX_CHECK(hipStreamSynchronize(stream));
    giant_[l] = std::move(g);
  }
}

}  // namespace hmx
