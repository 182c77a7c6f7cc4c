// Whole-matrix operations behind the remaining hmat_interface_t slots: copy
// and copy_struct (HMatInterface::copy, c_wrapping.hpp:98-108), norm
// (HMatrix::normSqr h_matrix.cpp:624-646: off-diagonal blocks of symmetric
// storage count twice), scale (c_wrapping.hpp:431), add_identity
// (c_wrapping.hpp:393) and extract_diagonal (c_wrapping.hpp:760-773).
#include <cmath>
#include <cstring>
#include <stdexcept>

#include "engine.hpp"

namespace hmx {

namespace {

template <typename T>
void copy_store(Store<T>& dst, const Store<T>& src, hipStream_t s) {
  HMX_CHECK(hipMemcpyAsync(dst.arena.p, src.arena.p, src.elems * sizeof(T),
                           hipMemcpyDeviceToDevice, s));
  HMX_CHECK(hipMemcpyAsync(dst.ranks.p, src.ranks.p, src.hdesc.size() * sizeof(int),
                           hipMemcpyDeviceToDevice, s));
  dst.valid = true;
}

std::vector<int> leaves_where(const BlockTree& bt, bool diagonal_only) {
  std::vector<int> out;
  for (const Block* b : bt.leaves)
    if (!diagonal_only || (b->kind == Kind::Full && b->rows == b->cols)) out.push_back(b->leaf);
  return out;
}

}  // namespace

Matrix* Matrix::clone(bool structure_only) {
  if (!assembled && !structure_only) throw std::runtime_error("copy: matrix not assembled");
  auto* c = embedded ? create_complex(zrows, zcols, adm, rk_cap_max / 2, erows, ecols)
                     : new Matrix(rows, cols, lower_sym, adm, value_type, rk_cap_max);
  try {
    c->lowrank_eps = lowrank_eps;
    c->device_init();
    if (structure_only) {
      // a null matrix of the same structure: zero full leaves, rank-0 Rk leaves
      if (value_type == HMAT_SIMPLE_PRECISION) {
        c->init_store(c->s32);
        HMX_CHECK(hipMemsetAsync(c->s32.arena.p, 0, c->s32.elems * sizeof(float), c->stream));
        c->s32.valid = true;
      } else {
        c->init_store(c->s64);
        HMX_CHECK(hipMemsetAsync(c->s64.arena.p, 0, c->s64.elems * sizeof(double), c->stream));
        c->s64.valid = true;
      }
    } else {
      HMX_CHECK(hipStreamSynchronize(stream));
      if (s64.valid) {
        c->init_store(c->s64);
        copy_store(c->s64, s64, c->stream);
      }
      if (s32.valid) {
        c->init_store(c->s32);
        copy_store(c->s32, s32, c->stream);
      }
      if (factorization == hmat_factorization_hodlrsym || factorization == hmat_factorization_hodlr)
        throw std::runtime_error("copy: HODLR factors cannot be copied");
      c->factorization = factorization;
      c->factor32 = factor32;
      if (pivots.p) {
        c->pivots.alloc(pivots.bytes);
        HMX_CHECK(hipMemcpyAsync(c->pivots.p, pivots.p, pivots.bytes, hipMemcpyDeviceToDevice,
                                 c->stream));
      }
      if (ldlt_diag.p) {
        c->ldlt_diag.alloc(ldlt_diag.bytes);
        HMX_CHECK(hipMemcpyAsync(c->ldlt_diag.p, ldlt_diag.p, ldlt_diag.bytes,
                                 hipMemcpyDeviceToDevice, c->stream));
      }
    }
    HMX_CHECK(hipStreamSynchronize(c->stream));
    c->assembled = true;
  } catch (...) {
    delete c;
    throw;
  }
  return c;
}

double Matrix::norm() {
  if (!assembled) throw std::runtime_error("norm: matrix not assembled");
  std::vector<int> list = leaves_where(*bt, false);
  DevBuf dl, dout;
  upload(dl, list.data(), list.size() * sizeof(int), stream);
  dout.alloc(std::max<size_t>(list.size(), 1) * sizeof(double));
  if (s64.valid)
    launch_leaf_norm2<double>(s64.desc.as<LeafDesc>(), s64.ranks.as<int>(), s64.data(),
                              dl.as<int>(), (int)list.size(), dout.as<double>(), stream);
  else if (s32.valid)
    launch_leaf_norm2<float>(s32.desc.as<LeafDesc>(), s32.ranks.as<int>(), s32.data(),
                             dl.as<int>(), (int)list.size(), dout.as<double>(), stream);
  else
    throw std::runtime_error("norm: no matrix data");
  std::vector<double> h(list.size());
  HMX_CHECK(hipMemcpyAsync(h.data(), dout.p, h.size() * sizeof(double), hipMemcpyDeviceToHost,
                           stream));
  HMX_CHECK(hipStreamSynchronize(stream));
  double r = 0.0;
  for (size_t i = 0; i < list.size(); ++i) {
    const Block* b = bt->leaves[list[i]];
    r += (lower_sym && b->mirrored() ? 2.0 : 1.0) * h[i];
  }
  // the 2 x 2 real representation holds every complex entry twice in modulus
  return std::sqrt(embedded ? 0.5 * r : r);
}

void Matrix::scale(double alpha) {
  if (!assembled) throw std::runtime_error("scale: matrix not assembled");
  if (factorization != -1) throw std::runtime_error("scale: matrix has been factorised");
  std::vector<int> list = leaves_where(*bt, false);
  DevBuf dl;
  upload(dl, list.data(), list.size() * sizeof(int), stream);
  if (s64.valid)
    launch_leaf_scale<double>(s64.desc.as<LeafDesc>(), dl.as<int>(), (int)list.size(), s64.data(),
                              alpha, stream);
  if (s32.valid)
    launch_leaf_scale<float>(s32.desc.as<LeafDesc>(), dl.as<int>(), (int)list.size(), s32.data(),
                             alpha, stream);
  HMX_CHECK(hipStreamSynchronize(stream));
}

void Matrix::add_identity(double alpha) {
  if (!assembled) throw std::runtime_error("add_identity: matrix not assembled");
  if (factorization != -1) throw std::runtime_error("add_identity: matrix has been factorised");
  if (rows != cols) throw std::runtime_error("add_identity: the matrix is not square");
  std::vector<int> list = leaves_where(*bt, true);
  DevBuf dl;
  upload(dl, list.data(), list.size() * sizeof(int), stream);
  if (s64.valid)
    launch_leaf_add_diag<double>(s64.desc.as<LeafDesc>(), dl.as<int>(), (int)list.size(),
                                 s64.data(), alpha, stream);
  if (s32.valid)
    launch_leaf_add_diag<float>(s32.desc.as<LeafDesc>(), dl.as<int>(), (int)list.size(),
                                s32.data(), alpha, stream);
  HMX_CHECK(hipStreamSynchronize(stream));
}

void Matrix::extract_diagonal(double* diag) {
  if (!assembled) throw std::runtime_error("extract_diagonal: matrix not assembled");
  if (rows != cols) throw std::runtime_error("extract_diagonal: the matrix is not square");
  if (factorization == hmat_factorization_ldlt) {
    // an LDLt factor gives D (HMatrix::extractDiagonal h_matrix.cpp:2904-2906)
    const int n = n_rows();
    if (factor32) {
      std::vector<float> d(n);
      HMX_CHECK(hipMemcpy(d.data(), ldlt_diag.p, n * sizeof(float), hipMemcpyDeviceToHost));
      for (int i = 0; i < n; ++i) diag[i] = d[i];
    } else {
      HMX_CHECK(hipMemcpy(diag, ldlt_diag.p, n * sizeof(double), hipMemcpyDeviceToHost));
    }
    return;
  }
  std::vector<int> list = leaves_where(*bt, true);
  DevBuf dl, dout;
  upload(dl, list.data(), list.size() * sizeof(int), stream);
  dout.alloc((size_t)n_rows() * sizeof(double));
  // the factor's precision when factorised, else the matrix's
  const bool use32 = factorization != -1 ? factor32 : !s64.valid;
  if (use32)
    launch_leaf_get_diag<float>(s32.desc.as<LeafDesc>(), dl.as<int>(), (int)list.size(),
                                s32.data(), dout.as<double>(), stream);
  else
    launch_leaf_get_diag<double>(s64.desc.as<LeafDesc>(), dl.as<int>(), (int)list.size(),
                                 s64.data(), dout.as<double>(), stream);
  HMX_CHECK(hipMemcpyAsync(diag, dout.p, (size_t)n_rows() * sizeof(double),
                           hipMemcpyDeviceToHost, stream));
  HMX_CHECK(hipStreamSynchronize(stream));
}

}  // namespace hmx
