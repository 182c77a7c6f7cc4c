// Serialization in the stream format of hmat-oss (serialization.cpp:34-394):
// write_struct / read_struct carry the scalar type, the factorisation, both
// cluster trees (coordinates, permutation, nodes) and the block tree;
// write_data / read_data carry the leaves in preorder.  Everything goes
// through the caller's hmat_iostream, so a structure or a factor written here
// is read by the reference and the other way round.
#include <cstring>
#include <functional>
#include <stdexcept>
#include <vector>

#include "engine.hpp"

namespace hmx {

namespace {

// Block kinds in the stream (DefaultRank, h_matrix.hpp:61)
constexpr int kUninitialized = -3, kNonLeaf = -2, kFullBlock = -1;
constexpr int kApproxRank = 25;  // StandardAdmissibilityCondition::getApproximateRank

struct Out {
  hmat_iostream f;
  void* ud;
  template <typename V>
  void put(V v) {
    f(&v, sizeof(v), ud);
  }
  void bytes(const void* p, size_t n) { f(const_cast<void*>(p), n, ud); }
};

struct In {
  hmat_iostream f;
  void* ud;
  template <typename V>
  V get() {
    V v{};
    f(&v, sizeof(v), ud);
    return v;
  }
  void bytes(void* p, size_t n) { f(p, n, ud); }
  void end_marker() {  // the zero-length read the reference issues after a tree
    int dummy = 0;
    f(&dummy, 0, ud);
  }
};

// ScalarTypes data_types.hpp:37-46
int type_code(int value_type) {
  if (value_type == HMAT_SIMPLE_PRECISION) return 0;
  if (value_type == HMAT_DOUBLE_PRECISION) return 1;
  if (value_type == HMAT_SIMPLE_COMPLEX) return 2;
  if (value_type == HMAT_DOUBLE_COMPLEX) return 3;
  throw std::runtime_error("serialization: unknown scalar type");
}

// MatrixStructMarshaller::write(ClusterTree) serialization.cpp:34-53
void write_cluster_tree(Out& o, const ClusterTree& t) {
  o.put<int>(t.n);
  o.put<int>(t.dim);
  for (int j = 0; j < t.n; ++j)
    for (int i = 0; i < t.dim; ++i) o.put<double>(t.coords[(size_t)j * t.dim + i]);
  o.put<int>(0);  // no group index
  o.bytes(t.indices.data(), sizeof(int) * (size_t)t.n);
  std::function<void(const Cluster*)> rec = [&](const Cluster* c) {
    o.put<int>(c->offset);
    o.put<int>(c->size);
    o.put<int>((int)c->children.size());
    for (const Cluster* k : c->children) rec(k);
  };
  rec(t.root);
}

std::unique_ptr<ClusterTree> read_cluster_tree(In& in) {
  const int n = in.get<int>();
  const int dim = in.get<int>();
  if (n < 1 || dim < 1 || dim > 3) throw std::runtime_error("read_struct: bad cluster tree header");
  std::vector<double> coords((size_t)n * dim);
  in.bytes(coords.data(), coords.size() * sizeof(double));
  if (in.get<int>()) throw std::runtime_error("read_struct: group indices are not supported");
  std::vector<int> idx(n);
  in.bytes(idx.data(), sizeof(int) * (size_t)n);
  std::vector<ClusterNodeDesc> nodes;
  std::function<void()> rec = [&]() {
    ClusterNodeDesc d;
    d.offset = in.get<int>();
    if (d.offset == -1) throw std::runtime_error("read_struct: null cluster node");
    d.size = in.get<int>();
    d.nchild = in.get<int>();
    if (d.nchild < 0 || d.nchild > 2) throw std::runtime_error("read_struct: bad cluster node");
    nodes.push_back(d);
    for (int k = 0; k < d.nchild; ++k) rec();
  };
  rec();
  return cluster_tree_from(coords.data(), dim, n, idx, nodes);
}

bool same_tree(const ClusterTree& a, const ClusterTree& b) {
  if (a.n != b.n || a.dim != b.dim || a.coords != b.coords || a.indices != b.indices ||
      a.pool.size() != b.pool.size())
    return false;
  for (size_t i = 0; i < a.pool.size(); ++i)
    if (a.pool[i].offset != b.pool[i].offset || a.pool[i].size != b.pool[i].size ||
        a.pool[i].children.size() != b.pool[i].children.size())
      return false;
  return true;
}

}  // namespace

struct Serializer {
  // the store whose leaves are written: the factor, or the matrix
  static bool use32(const Matrix& M) {
    if (M.factorization != -1) return M.factor32 && M.value_type == HMAT_SIMPLE_PRECISION;
    return M.value_type == HMAT_SIMPLE_PRECISION;
  }
  static int written_factorization(const Matrix& M) {
    // the HODLR factor's node blocks have no place in the stream (the
    // reference keeps them outside the matrix and cannot read them back)
    if (M.factorization == hmat_factorization_hodlrsym || M.factorization == hmat_factorization_hodlr)
      throw std::runtime_error("serialization of HODLR factors is not supported");
    // an FP32 factor of an FP64 matrix (mixed precision) is not a factor of the
    // stored matrix's type: the FP64 matrix is written unfactorised
    if (M.factorization != -1 && M.factor32 && M.value_type == HMAT_DOUBLE_PRECISION) return -1;
    return M.factorization;
  }

  static std::vector<int> ranks(const Matrix& M) {
    const DevBuf& rb = use32(M) ? M.s32.ranks : M.s64.ranks;
    std::vector<int> r(M.bt->leaves.size());
    HMX_CHECK(hipMemcpy(r.data(), rb.p, r.size() * sizeof(int), hipMemcpyDeviceToHost));
    return r;
  }

  // A complex matrix is held as its real representation (DoF 2u + c of the
  // embedded numbering is component c of DoF u, every entry a 2 x 2 block
  // [[Re, -Im], [Im, Re]]).  The stream holds the complex matrix: the caller's
  // cluster trees, the same block tree, complex leaves.  An embedded Rk leaf
  // A B^T of rank k is written as the complex Rk (A_even + i A_odd) B_even^T of
  // rank k (the first column of each 2 x 2 block; the two columns agree to the
  // compression), a full leaf as the average of the two places holding Re and
  // Im.  Factors are not written: the row interchanges of the real
  // representation's LU are not interchanges of complex rows.
  static void check_complex(const Matrix& M) {
    if (M.embedded && M.factorization != -1)
      throw std::runtime_error("serialization of a factorised complex matrix is not provided");
  }

  // MatrixStructMarshaller::write(HMatrix) serialization.cpp:63-130
  static void write_struct(Matrix& M, hmat_iostream f, void* ud) {
    if (!M.assembled) throw std::runtime_error("write_struct: matrix not assembled");
    check_complex(M);
    Out o{f, ud};
    const int fact = written_factorization(M);
    o.put<int>(type_code(!M.embedded ? M.value_type
                                     : M.value_type == HMAT_SIMPLE_PRECISION ? HMAT_SIMPLE_COMPLEX
                                                                             : HMAT_DOUBLE_COMPLEX));
    o.put<int>(fact);
    write_cluster_tree(o, M.embedded ? *M.zrows : *M.rows);
    write_cluster_tree(o, M.embedded ? *M.zcols : *M.cols);
    const std::vector<int> rk = ranks(M);
    // lltDecomposition / ldltDecomposition flags, h_matrix.cpp:2562, 2766
    const bool tri = fact == hmat_factorization_llt || fact == hmat_factorization_ldlt;
    std::function<void(const Block*)> rec = [&](const Block* b) {
      if (!b) {
        o.put<unsigned char>(1 << 7);
        return;
      }
      const bool diag = b->rows == b->cols && M.lower_sym;
      unsigned char bits = 0;
      if (!b->is_leaf() && b->sym && !tri) bits |= 1 << 1;  // isLower
      if (diag && tri) bits |= 1 << 3;                     // isTriLower
      const bool same_r = b->is_leaf() || b->nrow == 1, same_c = b->is_leaf() || b->ncol == 1;
      if (same_r) bits |= 1 << 4;
      if (same_c) bits |= 1 << 5;
      o.put<unsigned char>(bits);
      o.put<int>(b->is_leaf() ? kApproxRank : kUninitialized);
      o.put<int>(b->kind == Kind::Rk ? std::max(rk[b->leaf], 0)
                                     : (b->kind == Kind::Full ? kFullBlock : kNonLeaf));
      o.put<double>(M.lowrank_eps);
      int nchild = 0;
      if (!b->is_leaf())
        for (int k = 0; k < b->nrow * b->ncol; ++k)
          if (b->child[k]) nchild = k + 1;
      o.put<int>(nchild);
      for (int k = 0; k < nchild; ++k) rec(b->child[k]);
    };
    rec(M.bt->root);
  }

  // MatrixDataMarshaller::write serialization.cpp:241-303 (leaves in preorder)
  template <typename T>
  static void write_leaves(Matrix& M, Out& o, const Store<T>& st, const std::vector<int>& rk) {
    std::vector<int> piv;
    if (written_factorization(M) == hmat_factorization_lu) {
      piv.resize(M.n_rows());
      HMX_CHECK(hipMemcpy(piv.data(), M.pivots.p, piv.size() * sizeof(int), hipMemcpyDeviceToHost));
    }
    std::vector<T> dg;
    if (written_factorization(M) == hmat_factorization_ldlt) {
      dg.resize(M.n_rows());
      HMX_CHECK(hipMemcpy(dg.data(), M.ldlt_diag.p, dg.size() * sizeof(T), hipMemcpyDeviceToHost));
    }
    std::vector<T> buf;
    for (const Block* b : M.bt->leaves) {
      const LeafDesc& L = st.hdesc[b->leaf];
      if (L.kind == 1) {
        const int k = std::max(rk[b->leaf], 0);
        o.put<int>(k);
        if (k > 0) {
          buf.resize((size_t)L.m * k);
          HMX_CHECK(hipMemcpy(buf.data(), st.data() + L.a, buf.size() * sizeof(T), hipMemcpyDeviceToHost));
          o.bytes(buf.data(), buf.size() * sizeof(T));
          buf.resize((size_t)L.n * k);
          HMX_CHECK(hipMemcpy(buf.data(), st.data() + L.b, buf.size() * sizeof(T), hipMemcpyDeviceToHost));
          o.bytes(buf.data(), buf.size() * sizeof(T));
          o.put<int>(0);  // orthogonality flags of A and B: not tracked
          o.put<int>(0);
        }
        continue;
      }
      const bool has_piv = !piv.empty() && b->rows == b->cols;
      const bool has_diag = !dg.empty() && b->rows == b->cols;
      o.put<int>((has_piv ? 2 : 0) | (has_diag ? 4 : 0));
      buf.resize((size_t)L.m * L.n);
      HMX_CHECK(hipMemcpy(buf.data(), st.data() + L.a, buf.size() * sizeof(T), hipMemcpyDeviceToHost));
      o.bytes(buf.data(), buf.size() * sizeof(T));
      if (has_piv) {
        std::vector<int> p(L.m);
        for (int i = 0; i < L.m; ++i) p[i] = piv[L.r0 + i] + 1;  // LAPACK getrf numbering
        o.bytes(p.data(), p.size() * sizeof(int));
      }
      if (has_diag) o.bytes(dg.data() + L.r0, sizeof(T) * (size_t)L.m);
    }
  }

  // the complex leaves of an embedded matrix (see check_complex), components
  // of type T (the store's)
  template <typename T>
  static void write_complex_leaves(Matrix& M, Out& o, const Store<T>& st, const std::vector<int>& rk) {
    std::vector<T> buf, z;
    for (const Block* b : M.bt->leaves) {
      const LeafDesc& L = st.hdesc[b->leaf];
      const int m = L.m / 2, n = L.n / 2;
      if (L.kind == 1) {
        const int k = std::max(rk[b->leaf], 0);
        o.put<int>(k);
        if (k > 0) {
          buf.resize((size_t)L.m * k);
          HMX_CHECK(hipMemcpy(buf.data(), st.data() + L.a, buf.size() * sizeof(T), hipMemcpyDeviceToHost));
          z.assign(2 * (size_t)m * k, T(0));
          for (int l = 0; l < k; ++l)
            for (int u = 0; u < m; ++u) {
              z[2 * (u + (size_t)l * m)] = buf[2 * u + (size_t)l * L.m];
              z[2 * (u + (size_t)l * m) + 1] = buf[2 * u + 1 + (size_t)l * L.m];
            }
          o.bytes(z.data(), z.size() * sizeof(T));
          buf.resize((size_t)L.n * k);
          HMX_CHECK(hipMemcpy(buf.data(), st.data() + L.b, buf.size() * sizeof(T), hipMemcpyDeviceToHost));
          z.assign(2 * (size_t)n * k, T(0));
          for (int l = 0; l < k; ++l)
            for (int v = 0; v < n; ++v) z[2 * (v + (size_t)l * n)] = buf[2 * v + (size_t)l * L.n];
          o.bytes(z.data(), z.size() * sizeof(T));
          o.put<int>(0);
          o.put<int>(0);
        }
        continue;
      }
      o.put<int>(0);
      buf.resize((size_t)L.m * L.n);
      HMX_CHECK(hipMemcpy(buf.data(), st.data() + L.a, buf.size() * sizeof(T), hipMemcpyDeviceToHost));
      z.resize(2 * (size_t)m * n);
      auto at = [&](int i, int j) { return (double)buf[i + (size_t)j * L.m]; };
      for (int v = 0; v < n; ++v)
        for (int u = 0; u < m; ++u) {
          z[2 * (u + (size_t)v * m)] = (T)(0.5 * (at(2 * u, 2 * v) + at(2 * u + 1, 2 * v + 1)));
          z[2 * (u + (size_t)v * m) + 1] = (T)(0.5 * (at(2 * u + 1, 2 * v) - at(2 * u, 2 * v + 1)));
        }
      o.bytes(z.data(), z.size() * sizeof(T));
    }
  }

  static void write_data(Matrix& M, hmat_iostream f, void* ud) {
    if (!M.assembled) throw std::runtime_error("write_data: matrix not assembled");
    check_complex(M);
    Out o{f, ud};
    const std::vector<int> rk = ranks(M);
    if (M.embedded && use32(M))
      write_complex_leaves<float>(M, o, M.s32, rk);
    else if (M.embedded)
      write_complex_leaves<double>(M, o, M.s64, rk);
    else if (use32(M))
      write_leaves<float>(M, o, M.s32, rk);
    else
      write_leaves<double>(M, o, M.s64, rk);
  }

  // MatrixStructUnmarshaller::read serialization.cpp:206-232
  static Matrix* read_struct(hmat_iostream f, void* ud, int value_type, int rank_cap) {
    In in{f, ud};
    const int type = in.get<int>();
    if (type != type_code(value_type))
      throw std::runtime_error("read_struct: type mismatch (stream " + std::to_string(type) +
                               ", interface " + std::to_string(type_code(value_type)) + ")");
    const int fact = in.get<int>();
    if (fact != -1 && fact != hmat_factorization_lu && fact != hmat_factorization_llt &&
        fact != hmat_factorization_ldlt)
      throw std::runtime_error("read_struct: only unfactorised, LU, LDLt and LLt matrices are supported");
    const bool complex = value_type == HMAT_DOUBLE_COMPLEX || value_type == HMAT_SIMPLE_COMPLEX;
    const int real_type = value_type == HMAT_SIMPLE_COMPLEX ? HMAT_SIMPLE_PRECISION : HMAT_DOUBLE_PRECISION;
    if (complex && fact != -1)
      throw std::runtime_error("read_struct: a factorised complex matrix is not provided");
    std::unique_ptr<ClusterTree> rt = read_cluster_tree(in);
    in.end_marker();
    std::unique_ptr<ClusterTree> ct = read_cluster_tree(in);
    in.end_marker();
    std::vector<BlockNodeDesc> nodes;
    double eps = 1e-4;
    bool first = true;
    std::function<void()> rec = [&]() {
      BlockNodeDesc d{};
      d.bits = in.get<unsigned char>();
      if (d.bits & (1 << 7)) {
        d.null = true;
        nodes.push_back(d);
        return;
      }
      in.get<int>();  // approximate rank
      d.rank = in.get<int>();
      const double e = in.get<double>();
      if (first) eps = e;
      first = false;
      d.nchild = in.get<int>();
      if (d.nchild < 0 || d.nchild > 4) throw std::runtime_error("read_struct: bad block node");
      nodes.push_back(d);
      for (int k = 0; k < d.nchild; ++k) rec();
    };
    rec();
    in.end_marker();
    const bool sym_tree = same_tree(*rt, *ct);
    const ClusterTree* rows = rt.get();
    const ClusterTree* cols = sym_tree ? rt.get() : ct.get();
    std::unique_ptr<BlockTree> bt = block_tree_from(*rows, *cols, nodes);
    const unsigned char root_bits = nodes[0].bits;
    const bool lower = sym_tree && (root_bits & ((1 << 1) | (1 << 3))) &&
                       (bt->root->is_leaf() || bt->root->sym);
    std::vector<int> hint(bt->leaves.size(), 0);
    {
      // ranks of the leaves, in preorder as the leaves are numbered
      size_t q = 0;
      for (const BlockNodeDesc& d : nodes)
        if (!d.null && d.nchild == 0) {
          if (q < hint.size()) hint[q] = std::max(d.rank, 0);
          ++q;
        }
    }
    if (complex) {
      // the real representation of the stream's complex matrix: embedded
      // trees and block tree, a complex rank k taking real rank 2k
      std::shared_ptr<ClusterTree> er(embed_cluster_tree(*rows));
      std::shared_ptr<ClusterTree> ec = sym_tree ? er : std::shared_ptr<ClusterTree>(embed_cluster_tree(*cols));
      std::unique_ptr<BlockTree> et = embed_block_tree(*bt, *rows, *er, *cols, *ec);
      for (int& h : hint) h *= 2;
      auto* M = new Matrix(er.get(), ec.get(), false, std::move(et), real_type, 2 * rank_cap, hint);
      M->embedded = true;
      M->zrows = rows;
      M->zcols = cols;
      M->erows = er;
      M->ecols = ec;
      M->own_rows = true;
      M->own_cols = true;
      rt.release();
      if (!sym_tree) ct.release();
      M->lowrank_eps = eps;
      return M;
    }
    auto* M = new Matrix(rows, cols, lower, std::move(bt), value_type, rank_cap, hint);
    M->own_rows = true;
    M->own_cols = true;
    rt.release();
    if (!sym_tree) ct.release();
    M->lowrank_eps = eps;
    M->factorization = fact;
    M->factor32 = fact != -1 && value_type == HMAT_SIMPLE_PRECISION;
    return M;
  }

  // MatrixDataUnmarshaller::read serialization.cpp:305-381
  template <typename T>
  static void read_leaves(Matrix& M, In& in, Store<T>& st) {
    std::vector<T> image(st.elems, T(0));
    std::vector<int> rk(st.hdesc.size());
    std::vector<int> piv;
    if (M.factorization == hmat_factorization_lu) piv.assign(M.n_rows(), 0);
    std::vector<T> dg;
    if (M.factorization == hmat_factorization_ldlt) dg.assign(M.n_rows(), T(0));
    for (const Block* b : M.bt->leaves) {
      const LeafDesc& L = st.hdesc[b->leaf];
      const int header = in.get<int>();
      if (L.kind == 1) {
        const int k = header;
        if (k < 0 || k > L.cap)
          throw std::runtime_error("read_data: Rk rank " + std::to_string(k) + " above the leaf capacity");
        rk[b->leaf] = k;
        if (k > 0) {
          in.bytes(image.data() + L.a, sizeof(T) * (size_t)L.m * k);
          in.bytes(image.data() + L.b, sizeof(T) * (size_t)L.n * k);
          in.get<int>();
          in.get<int>();
        }
        continue;
      }
      rk[b->leaf] = -1;
      if (header & 1) continue;  // null full block
      if ((header & 4) && (dg.empty() || b->rows != b->cols))
        throw std::runtime_error("read_data: an LDLt diagonal outside an LDLt diagonal leaf");
      in.bytes(image.data() + L.a, sizeof(T) * (size_t)L.m * L.n);
      if (header & 2) {
        std::vector<int> p(L.m);
        in.bytes(p.data(), p.size() * sizeof(int));
        for (int i = 0; i < L.m; ++i)
          if (p[i] < 1 || p[i] > L.m) throw std::runtime_error("read_data: pivot out of range");
        if (!piv.empty())
          for (int i = 0; i < L.m; ++i) piv[L.r0 + i] = p[i] - 1;
      }
      if (header & 4) in.bytes(dg.data() + L.r0, sizeof(T) * (size_t)L.m);
    }
    in.end_marker();
    M.init_store(st);
    HMX_CHECK(hipMemcpy(st.arena.p, image.data(), image.size() * sizeof(T), hipMemcpyHostToDevice));
    HMX_CHECK(hipMemcpy(st.ranks.p, rk.data(), rk.size() * sizeof(int), hipMemcpyHostToDevice));
    st.valid = true;
    if (!dg.empty()) {
      M.ldlt_diag.alloc(sizeof(T) * dg.size());
      HMX_CHECK(hipMemcpy(M.ldlt_diag.p, dg.data(), dg.size() * sizeof(T), hipMemcpyHostToDevice));
    }
    if (!piv.empty()) {
      M.pivots.alloc(sizeof(int) * piv.size());
      HMX_CHECK(hipMemcpy(M.pivots.p, piv.data(), piv.size() * sizeof(int), hipMemcpyHostToDevice));
    }
  }

  // complex leaves into the real representation: an Rk (A, B) of complex
  // rank k becomes the real Rk (R(A), R(conj B)) of rank 2k, R(X) the 2m x 2k
  // real matrix of the 2 x 2 blocks [[Re x, -Im x], [Im x, Re x]] (R(A B^T) =
  // R(A) R(conj B)^T); a full leaf Z becomes R(Z)
  template <typename T>
  static void read_complex_leaves(Matrix& M, In& in, Store<T>& st) {
    std::vector<T> image(st.elems, T(0));
    std::vector<int> rk(st.hdesc.size());
    std::vector<T> z;
    auto embed = [](const T* z, int rows, int cols, bool conj, T* out, int ld) {
      for (int l = 0; l < cols; ++l)
        for (int u = 0; u < rows; ++u) {
          const T re = z[2 * (u + (size_t)l * rows)];
          const T im = conj ? -z[2 * (u + (size_t)l * rows) + 1] : z[2 * (u + (size_t)l * rows) + 1];
          out[2 * u + (size_t)(2 * l) * ld] = re;
          out[2 * u + 1 + (size_t)(2 * l) * ld] = im;
          out[2 * u + (size_t)(2 * l + 1) * ld] = -im;
          out[2 * u + 1 + (size_t)(2 * l + 1) * ld] = re;
        }
    };
    for (const Block* b : M.bt->leaves) {
      const LeafDesc& L = st.hdesc[b->leaf];
      const int m = L.m / 2, n = L.n / 2;
      const int header = in.get<int>();
      if (L.kind == 1) {
        const int k = header;
        if (k < 0 || 2 * k > L.cap)
          throw std::runtime_error("read_data: Rk rank " + std::to_string(k) + " above the leaf capacity");
        rk[b->leaf] = 2 * k;
        if (k > 0) {
          z.resize(2 * (size_t)m * k);
          in.bytes(z.data(), z.size() * sizeof(T));
          embed(z.data(), m, k, false, image.data() + L.a, L.m);
          z.resize(2 * (size_t)n * k);
          in.bytes(z.data(), z.size() * sizeof(T));
          embed(z.data(), n, k, true, image.data() + L.b, L.n);
          in.get<int>();
          in.get<int>();
        }
        continue;
      }
      rk[b->leaf] = -1;
      if (header & 1) continue;  // null full block
      if (header & 6) throw std::runtime_error("read_data: factor data in an unfactorised complex matrix");
      z.resize(2 * (size_t)m * n);
      in.bytes(z.data(), z.size() * sizeof(T));
      embed(z.data(), m, n, false, image.data() + L.a, L.m);
    }
    in.end_marker();
    M.init_store(st);
    HMX_CHECK(hipMemcpy(st.arena.p, image.data(), image.size() * sizeof(T), hipMemcpyHostToDevice));
    HMX_CHECK(hipMemcpy(st.ranks.p, rk.data(), rk.size() * sizeof(int), hipMemcpyHostToDevice));
    st.valid = true;
  }

  static void read_data(Matrix& M, hmat_iostream f, void* ud) {
    In in{f, ud};
    M.device_init();
    if (M.embedded && M.value_type == HMAT_SIMPLE_PRECISION)
      read_complex_leaves<float>(M, in, M.s32);
    else if (M.embedded)
      read_complex_leaves<double>(M, in, M.s64);
    else if (M.value_type == HMAT_SIMPLE_PRECISION)
      read_leaves<float>(M, in, M.s32);
    else
      read_leaves<double>(M, in, M.s64);
    M.assembled = true;
    M.plan_.built = false;
    M.plan_.nrhs = 0;
    if (M.factorization == -1) M.build_gemv_plan();
  }
};

void serialize_write_struct(Matrix& M, hmat_iostream f, void* ud) { Serializer::write_struct(M, f, ud); }
void serialize_write_data(Matrix& M, hmat_iostream f, void* ud) { Serializer::write_data(M, f, ud); }
Matrix* serialize_read_struct(hmat_iostream f, void* ud, int value_type, int rank_cap) {
  return Serializer::read_struct(f, ud, value_type, rank_cap);
}
void serialize_read_data(Matrix& M, hmat_iostream f, void* ud) { Serializer::read_data(M, f, ud); }

}  // namespace hmx
