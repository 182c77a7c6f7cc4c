// Whole-matrix operations behind the remaining hmat_interface_t slots: copy
// and copy_struct (HMatInterface::copy, c_wrapping.hpp:98-108), norm
// (HMatrix::normSqr h_matrix.cpp:624-646: off-diagonal blocks of symmetric
// storage count twice), scale (c_wrapping.hpp:431), add_identity
// (c_wrapping.hpp:393), extract_diagonal (c_wrapping.hpp:760-773), transpose
// (c_wrapping.hpp:574-583) and the entries of get_block / get_values
// (c_wrapping.hpp:832-890).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <stdexcept>
#include <type_traits>
#include <utility>

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
  auto* c = embedded ? create_complex(zrows, zcols, adm, rk_cap_max / 2, erows, ecols, value_type)
                     : new Matrix(rows, cols, lower_sym, adm, value_type, rk_cap_max);
  try {
    c->lowrank_eps = lowrank_eps;
    // the same leaf layout (capacities raised by an assembly included)
    c->cap_hint_ = cap_hint_;
    c->s64.hdesc = s64.hdesc;
    c->s32.hdesc = s32.hdesc;
    c->s64.elems = s64.elems;
    c->s32.elems = s32.elems;
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

// HMatrix::transpose h_matrix.cpp:2036-2039 (transposeData + transposeMeta)
// for a square matrix over one cluster tree: the admissibility is symmetric,
// so leaf (I, J) has the partner (J, I) of the same kind.  An Rk leaf takes
// its partner's panels exchanged, (A B^T)^T = B A^T (descriptor offsets,
// capacities and ranks swap; no data moves), a full leaf its partner's block
// transposed.  Lower-symmetric storage is its own transpose.  A complex
// matrix gets its transpose (not the conjugate transpose): in the real
// representation the real transpose followed by the conjugation P R P,
// P = diag(1, -1, 1, -1, ...), i.e. odd rows of both Rk panels negated.
void Matrix::transpose() {
  if (!assembled) throw std::runtime_error("transpose: matrix not assembled");
  if (factorization != -1) throw std::runtime_error("transpose: factorised matrices are not provided");
  if (lower_sym) return;
  if (rows != cols) throw std::runtime_error("transpose: only square matrices over one cluster tree");
  std::map<std::pair<const Cluster*, const Cluster*>, int> at;
  for (const Block* b : bt->leaves) at[{b->rows, b->cols}] = b->leaf;
  std::vector<int> partner(bt->leaves.size());
  for (const Block* b : bt->leaves) {
    auto it = at.find({b->cols, b->rows});
    if (it == at.end() || bt->leaves[it->second]->kind != b->kind)
      throw std::runtime_error("transpose: the partition is not symmetric");
    partner[b->leaf] = it->second;
  }
  HMX_CHECK(hipStreamSynchronize(stream));
  auto run = [&](auto* tag, auto& st) {
    using T = std::remove_pointer_t<decltype(tag)>;
    if (!st.valid) return;
    std::vector<int> rk(st.hdesc.size());
    HMX_CHECK(hipMemcpy(rk.data(), st.ranks.p, rk.size() * sizeof(int), hipMemcpyDeviceToHost));
    std::vector<LeafDesc> nd = st.hdesc;
    std::vector<int> nrk = rk;
    std::vector<TransposeItem> items;
    for (size_t l = 0; l < nd.size(); ++l) {
      const LeafDesc& P = st.hdesc[partner[l]];
      LeafDesc& L = nd[l];
      if (L.kind == 0) {
        items.push_back(TransposeItem{L.a, P.a, L.m, L.n, 0, embedded ? 1 : 0});
      } else {
        L.a = P.b;
        L.b = P.a;
        L.cap = P.cap;
        nrk[l] = rk[partner[l]];
      }
    }
    if (embedded)
      for (size_t l = 0; l < nd.size(); ++l)
        if (nd[l].kind == 1 && nrk[l] > 0) {
          items.push_back(TransposeItem{nd[l].a, 0, nd[l].m, nrk[l], 1, 1});
          items.push_back(TransposeItem{nd[l].b, 0, nd[l].n, nrk[l], 1, 1});
        }
    DevBuf copy, ditems;
    copy.alloc(st.elems * sizeof(T));
    HMX_CHECK(hipMemcpyAsync(copy.p, st.arena.p, st.elems * sizeof(T), hipMemcpyDeviceToDevice, stream));
    upload(ditems, items.data(), items.size() * sizeof(TransposeItem), stream);
    launch_leaf_transpose<T>(ditems.as<TransposeItem>(), (int)items.size(), st.data(),
                             copy.as<T>(), stream);
    st.hdesc = nd;
    upload(st.desc, nd.data(), nd.size() * sizeof(LeafDesc), stream);
    HMX_CHECK(hipMemcpyAsync(st.ranks.p, nrk.data(), nrk.size() * sizeof(int), hipMemcpyHostToDevice,
                             stream));
    HMX_CHECK(hipStreamSynchronize(stream));
  };
  run((double*)nullptr, s64);
  run((float*)nullptr, s32);
  // the other store's layout follows (both share one arena layout)
  if (s64.valid && !s32.valid) s32.hdesc = s64.hdesc;
  if (s32.valid && !s64.valid) s64.hdesc = s32.hdesc;
  if (!cap_hint_.empty()) {
    std::vector<int> h(cap_hint_.size());
    for (size_t l = 0; l < h.size(); ++l) h[l] = cap_hint_[partner[l]];
    cap_hint_ = h;
  }
  plan_.built = false;
  plan_.nrhs = 0;
  build_gemv_plan();
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

namespace {
// (position, query index) pairs sorted by position
std::vector<std::pair<int, int>> sorted_query(const std::vector<int>& pos) {
  std::vector<std::pair<int, int>> q(pos.size());
  for (size_t i = 0; i < pos.size(); ++i) q[i] = {pos[i], (int)i};
  std::sort(q.begin(), q.end());
  return q;
}
// the query entries with positions in [lo, lo + len)
std::pair<size_t, size_t> query_range(const std::vector<std::pair<int, int>>& q, int lo, int len) {
  const auto b = std::lower_bound(q.begin(), q.end(), std::make_pair(lo, -1));
  const auto e = std::lower_bound(q.begin(), q.end(), std::make_pair(lo + len, -1));
  return {(size_t)(b - q.begin()), (size_t)(e - q.begin())};
}
}  // namespace

template <typename T>
void Matrix::get_entries_t(const Store<T>& st, const std::vector<int>& rpos,
                           const std::vector<int>& cpos, double* out, int ld) {
  const auto qr = sorted_query(rpos), qc = sorted_query(cpos);
  std::vector<int> ranks(st.hdesc.size());
  HMX_CHECK(hipMemcpy(ranks.data(), st.ranks.p, ranks.size() * sizeof(int), hipMemcpyDeviceToHost));
  for (size_t i = 0; i < rpos.size(); ++i)
    for (size_t j = 0; j < cpos.size(); ++j) out[i + (size_t)j * ld] = 0.0;
  std::vector<T> a, b;
  for (const Block* blk : bt->leaves) {
    const LeafDesc& L = st.hdesc[blk->leaf];
    // a factor is lower triangular (isLower off after the factorisation,
    // h_matrix.cpp:2562-2563): its blocks are not mirrored
    const bool mir = lower_sym && factorization == -1 && blk->mirrored();
    const auto rr = query_range(qr, L.r0, L.m), cc = query_range(qc, L.c0, L.n);
    const auto rt = mir ? query_range(qr, L.c0, L.n) : std::make_pair<size_t, size_t>(0, 0);
    const auto ct = mir ? query_range(qc, L.r0, L.m) : std::make_pair<size_t, size_t>(0, 0);
    const bool direct = rr.first < rr.second && cc.first < cc.second;
    const bool mirror = rt.first < rt.second && ct.first < ct.second;
    if (!direct && !mirror) continue;
    // the leaf's data on the host (full: m x n; Rk: A m x k, B n x k)
    const int k = L.kind == 0 ? 0 : std::max(ranks[blk->leaf], 0);
    if (L.kind == 0) {
      a.resize((size_t)L.m * L.n);
      HMX_CHECK(hipMemcpy(a.data(), st.data() + L.a, a.size() * sizeof(T), hipMemcpyDeviceToHost));
    } else {
      a.resize((size_t)L.m * k);
      b.resize((size_t)L.n * k);
      if (k) {
        HMX_CHECK(hipMemcpy(a.data(), st.data() + L.a, a.size() * sizeof(T), hipMemcpyDeviceToHost));
        HMX_CHECK(hipMemcpy(b.data(), st.data() + L.b, b.size() * sizeof(T), hipMemcpyDeviceToHost));
      }
    }
    auto value = [&](int i, int j) {  // leaf-local row i, column j
      if (L.kind == 0) return (double)a[i + (size_t)j * L.m];
      double v = 0.0;
      for (int l = 0; l < k; ++l) v += (double)a[i + (size_t)l * L.m] * (double)b[j + (size_t)l * L.n];
      return v;
    };
    if (direct)
      for (size_t x = rr.first; x < rr.second; ++x)
        for (size_t y = cc.first; y < cc.second; ++y)
          out[qr[x].second + (size_t)qc[y].second * ld] = value(qr[x].first - L.r0, qc[y].first - L.c0);
    if (mirror)
      for (size_t x = rt.first; x < rt.second; ++x)
        for (size_t y = ct.first; y < ct.second; ++y)
          out[qr[x].second + (size_t)qc[y].second * ld] = value(qc[y].first - L.r0, qr[x].first - L.c0);
  }
}

void Matrix::get_entries(const std::vector<int>& rpos, const std::vector<int>& cpos, double* out,
                         int ld) {
  if (!assembled) throw std::runtime_error("get_values: matrix not assembled");
  if (factorization == hmat_factorization_hodlrsym || factorization == hmat_factorization_hodlr)
    throw std::runtime_error("get_values: Unsupported operation (HODLR factor)");
  for (int p : rpos)
    if (p < 0 || p >= n_rows()) throw std::runtime_error("get_values: row index out of range");
  for (int p : cpos)
    if (p < 0 || p >= n_cols()) throw std::runtime_error("get_values: column index out of range");
  device_init();
  HMX_CHECK(hipStreamSynchronize(stream));
  // the factor's precision when factorised, else the matrix's
  const bool use32 = factorization != -1 ? factor32 : !s64.valid;
  if (use32)
    get_entries_t<float>(s32, rpos, cpos, out, ld);
  else
    get_entries_t<double>(s64, rpos, cpos, out, ld);
}

}  // namespace hmx
