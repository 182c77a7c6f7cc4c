// C ABI: the hmat.h free functions and the hmat_interface_t function table
// backed by the MI355X engine (replaces c_default_interface.cpp and
// c_wrapping.hpp of hmat-oss).  Exceptions never cross the boundary: table
// entries print the message on stderr and return 1, as c_wrapping.hpp does.
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstring>
#include <algorithm>
#include <atomic>
#include <chrono>
#include <functional>
#include <string>
#include <thread>
#include <vector>

#include "engine.hpp"
#include "factor_kernels.hpp"
#include "hmat/hmat.h"
#include "hmat/hmat_mi355x.h"

using hmx::Matrix;

namespace {

thread_local std::string g_last_error;
int g_rank_cap = 160;
hmat_settings_t g_settings = {100, 1e-4, 200, 0, 0, 0, 0, 0, 0, 0, 0.0};

struct ClusteringAlgo {
  int kind = 0;        // median
  int max_leaf = -1;   // -1: settings.maxLeafSize
  int divider = 2;
};

struct Compression {
  int method;
  double eps;
};

struct Admissibility {
  hmx::AdmissibilityParams p;
};

int fail(const std::exception& e) {
  g_last_error = e.what();
  fprintf(stderr, "%s\n", e.what());
  return 1;
}

Matrix* M(hmat_matrix_t* h) { return reinterpret_cast<Matrix*>(h); }

template <typename F>
int guard(F&& f) {
  try {
    g_last_error.clear();
    f();
    return 0;
  } catch (const std::exception& e) {
    return fail(e);
  }
}

// ---- table entries -----------------------------------------------------------
template <int VT>
hmat_matrix_t* create_empty(const hmat_cluster_tree_t* r, const hmat_cluster_tree_t* c,
                            int lower_sym, hmat_admissibility_t* cond) {
  try {
    const auto* rt = reinterpret_cast<const hmx::ClusterTree*>(r);
    const auto* ct = reinterpret_cast<const hmx::ClusterTree*>(c);
    hmx::AdmissibilityParams p;
    if (cond) p = reinterpret_cast<Admissibility*>(cond)->p;
    auto* m = new Matrix(rt, ct, lower_sym != 0, p, VT, g_rank_cap);
    return reinterpret_cast<hmat_matrix_t*>(m);
  } catch (const std::exception& e) {
    fail(e);
    return nullptr;
  }
}

void own_cluster_trees(hmat_matrix_t* h, int owns_row, int owns_col) {
  M(h)->own_rows = owns_row != 0;
  M(h)->own_cols = owns_col != 0;
}

void set_low_rank_epsilon(hmat_matrix_t* h, double eps) { M(h)->lowrank_eps = eps; }

int assemble_generic(hmat_matrix_t* h, hmat_assemble_context_t* ctx) {
  return guard([&] {
    Matrix* m = M(h);
    if (!ctx->compression) throw std::runtime_error("No compression algorithm defined in hmat_assemble_context_t");
    if (ctx->lower_symmetric && !m->lower_sym)
      throw std::runtime_error("lower_symmetric assembly needs a matrix created lower-symmetric");
    if (ctx->assembly)
      throw std::runtime_error("assembly: C++ assembly objects are not supported by this engine");
    const auto* c = reinterpret_cast<const Compression*>(ctx->compression);
    hmx::CompressionSpec cs{c->method, c->eps};
    // the context's progress reporter for this assembly (and the
    // factorisation it may ask for); the matrix's own one afterwards
    hmat_progress_t* saved = m->progress;
    if (ctx->progress) m->progress = ctx->progress;
    struct Restore {
      Matrix* m;
      hmat_progress_t* p;
      ~Restore() { m->progress = p; }
    } restore{m, saved};
    if (ctx->simple_compute != hmat_mi355x_kernel_compute) {
      // the caller's host functions: blocks evaluated on the host, compressed
      // on the device (only one of advanced / block / simple is used, advanced
      // first, as assembly.cpp and c_wrapping.hpp:66-95 do)
      hmx::HostFunctions f;
      f.user_context = ctx->user_context;
      f.prepare = ctx->prepare;
      if (ctx->advanced_compute)
        f.advanced = ctx->advanced_compute;
      else if (ctx->block_compute)
        f.block = ctx->block_compute;
      else
        f.simple = ctx->simple_compute;
      m->assemble_host(f, cs);
      if (ctx->factorization != hmat_factorization_none) m->factorize(ctx->factorization, false);
      return;
    }
    const auto* k = static_cast<const hmat_mi355x_kernel_t*>(ctx->user_context);
    hmx::KernelSpec spec;
    spec.kind = k->kind;
    spec.length = k->length;
    spec.nugget = k->nugget;
    spec.shift = k->kind == HMAT_MI355X_KERNEL_SHIFTED_EXPONENTIAL ? k->shift : 0.0;
    if (k->kind < 0 || k->kind > HMAT_MI355X_KERNEL_HELMHOLTZ)
      throw std::runtime_error("hmat_mi355x_kernel_t: unknown kernel kind");
    spec.dim = k->dimension;
    if (m->embedded) {
      // point u is DoF 2u + c (c = 0, 1) of the embedded numbering
      if (k->kind == HMAT_MI355X_KERNEL_HELMHOLTZ && !(k->regularization > 0.0))
        throw std::runtime_error("hmat_mi355x_kernel_t: the Helmholtz kernel needs regularization > 0");
      spec.wavenumber = k->wavenumber;
      spec.reg = k->regularization;
      spec.n = 2 * k->n;
      spec.points.resize((size_t)spec.n * k->dimension);
      for (int u = 0; u < k->n; ++u)
        for (int c = 0; c < 2; ++c)
          std::copy(k->points + (size_t)u * k->dimension, k->points + (size_t)(u + 1) * k->dimension,
                    spec.points.begin() + (size_t)(2 * u + c) * k->dimension);
    } else {
      spec.n = k->n;
      spec.points.assign(k->points, k->points + (size_t)k->n * k->dimension);
    }
    m->assemble(spec, cs);
    if (ctx->factorization != hmat_factorization_none) m->factorize(ctx->factorization, false);
  });
}

int destroy(hmat_matrix_t* h) {
  Matrix* m = M(h);
  const hmx::ClusterTree* r = m->own_rows ? m->rows : nullptr;
  const hmx::ClusterTree* c = m->own_cols ? m->cols : nullptr;
  delete m;
  if (r) delete r;
  if (c && c != r) delete c;
  return 0;
}

// factorize_generic c_wrapping.hpp:139-145: the context's progress reporter
// (HMatInterface::factorize sets its max to N, hmat_cpp_interface.cpp:74-83)
int factorize_generic(hmat_matrix_t* h, hmat_factorization_context_t* ctx) {
  return guard([&] {
    Matrix* m = M(h);
    hmat_progress_t* saved = m->progress;
    if (ctx->progress) m->progress = ctx->progress;
    struct Restore {
      Matrix* m;
      hmat_progress_t* p;
      ~Restore() { m->progress = p; }
    } restore{m, saved};
    m->factorize(ctx->factorization, false);
  });
}

int logdet(hmat_matrix_t* h, double* res) {
  return guard([&] { *res = M(h)->logdet(); });
}

int solve_systems(hmat_matrix_t* h, void* b, int nrhs) {
  return guard([&] { M(h)->solve_host(b, nrhs, true); });
}

int solve_dense(hmat_matrix_t* h, void* b, int nrhs) {
  return guard([&] { M(h)->solve_host(b, nrhs, false); });
}

// solve_generic c_wrapping.hpp:500-540: lower / upper select one triangular
// sweep of an LLt factor (solveLower(b, transpose)); nr_rhs == 0 means that
// values is an H-matrix
int solve_generic(hmat_matrix_t* h, const struct hmat_solve_context_t* ctx) {
  return guard([&] {
    if (ctx->lower && ctx->upper)
      throw std::runtime_error("lower and upper cannot both be true in hmat_solve_context_t");
    if (ctx->nr_rhs == 0) {
      M(h)->solve_mat(*M(reinterpret_cast<hmat_matrix_t*>(ctx->values)),
                      ctx->lower ? 1 : (ctx->upper ? 2 : 0));
      return;
    }
    M(h)->solve_host(ctx->values, ctx->nr_rhs, ctx->no_permutation == 0,
                     ctx->lower ? 1 : (ctx->upper ? 2 : 0));
  });
}

// solve_lower_triangular(_dense) c_wrapping.hpp:793-820
// transpose c_wrapping.hpp:574-583 (HMatrix::transpose)
int transpose(hmat_matrix_t* h) {
  return guard([&] { M(h)->transpose(); });
}

int solve_lower_triangular(hmat_matrix_t* h, int transpose, void* b, int nrhs) {
  return guard([&] { M(h)->solve_host(b, nrhs, true, transpose ? 2 : 1); });
}
int solve_lower_triangular_dense(hmat_matrix_t* h, int transpose, void* b, int nrhs) {
  return guard([&] { M(h)->solve_host(b, nrhs, false, transpose ? 2 : 1); });
}

// copy / copy_struct c_wrapping.hpp:98-108 (the copy shares the cluster trees)
hmat_matrix_t* copy_matrix(hmat_matrix_t* h) {
  hmat_matrix_t* r = nullptr;
  guard([&] { r = reinterpret_cast<hmat_matrix_t*>(M(h)->clone(false)); });
  return r;
}
hmat_matrix_t* copy_struct(hmat_matrix_t* h) {
  hmat_matrix_t* r = nullptr;
  guard([&] { r = reinterpret_cast<hmat_matrix_t*>(M(h)->clone(true)); });
  return r;
}

// norm c_wrapping.hpp:412 (HMatrix::norm, Frobenius)
double norm_matrix(hmat_matrix_t* h) {
  double r = NAN;
  guard([&] { r = M(h)->norm(); });
  return r;
}

template <typename T>
int scale_matrix(void* alpha, hmat_matrix_t* h) {
  return guard([&] { M(h)->scale((double)*(const T*)alpha); });
}
template <typename T>
int add_identity(hmat_matrix_t* h, void* alpha) {
  return guard([&] { M(h)->add_identity((double)*(const T*)alpha); });
}
int truncate_matrix(hmat_matrix_t* h) {
  return guard([&] { M(h)->truncate_all(); });
}
// axpy c_wrapping.hpp:218-230: y <- y + a x (same partition)
template <typename T>
int axpy_matrix(void* a, hmat_matrix_t* x, hmat_matrix_t* y) {
  return guard([&] { M(y)->axpy_from((double)*(const T*)a, *M(x)); });
}

// gemm c_wrapping.hpp:202-215 (HMatrix::gemm h_matrix.cpp:1570-1620):
// C <- beta C + alpha op(A) op(B) on the device
template <typename T>
int gemm_h(char ta, char tb, const void* alpha, hmat_matrix_t* a, hmat_matrix_t* b,
           const void* beta, hmat_matrix_t* c) {
  return guard([&] {
    M(c)->gemm_from(ta, tb, (double)*(const T*)alpha, *M(a), *M(b), (double)*(const T*)beta);
  });
}

// extract_diagonal c_wrapping.hpp:760-773: the diagonal, in user numbering
template <typename T>
int extract_diagonal(hmat_matrix_t* h, void* diag, int size) {
  return guard([&] {
    Matrix* m = M(h);
    const int n = m->n_rows();
    if (size < n) throw std::runtime_error("extract_diagonal: the output holds fewer than N values");
    std::vector<double> hd(n);
    m->extract_diagonal(hd.data());
    T* out = static_cast<T*>(diag);
    for (int i = 0; i < n; ++i) out[m->cols->indices[i]] = (T)hd[i];
  });
}

template <typename T>
void reorder_axis(T* v, int nrows, int ncols, const int* idx, int axis, bool restore) {
  const int n = axis == 0 ? nrows : ncols;
  std::vector<T> tmp(n);
  if (axis == 0) {
    for (int c = 0; c < ncols; ++c) {
      T* col = v + (size_t)c * nrows;
      for (int i = 0; i < n; ++i) {
        if (restore)
          tmp[idx[i]] = col[i];
        else
          tmp[i] = col[idx[i]];
      }
      std::memcpy(col, tmp.data(), n * sizeof(T));
    }
  } else {
    for (int r = 0; r < nrows; ++r) {
      for (int i = 0; i < n; ++i) {
        if (restore)
          tmp[idx[i]] = v[r + (size_t)i * nrows];
        else
          tmp[i] = v[r + (size_t)idx[i] * nrows];
      }
      for (int i = 0; i < n; ++i) v[r + (size_t)i * nrows] = tmp[i];
    }
  }
}

// reorderVector / restoreVectorOrder h_matrix.cpp:75-141 via c_wrapping.hpp:455-500
template <typename T, bool Restore>
int vector_order(void* vec, const hmat_cluster_tree_t* rows_ct, int rows,
                 const hmat_cluster_tree_t* cols_ct, int cols) {
  return guard([&] {
    if (!rows_ct && rows == 0) throw std::runtime_error("either row cluster tree or rows must be non null");
    if (!cols_ct && cols == 0) throw std::runtime_error("either col cluster tree or cols must be non null");
    const auto* rt = reinterpret_cast<const hmx::ClusterTree*>(rows_ct);
    const auto* ct = reinterpret_cast<const hmx::ClusterTree*>(cols_ct);
    const int nr = rt ? rt->n : rows, nc = ct ? ct->n : cols;
    if (rt) reorder_axis<T>((T*)vec, nr, nc, rt->indices.data(), 0, Restore);
    if (ct) reorder_axis<T>((T*)vec, nr, nc, ct->indices.data(), 1, Restore);
  });
}

int gemv(char trans, void* alpha, hmat_matrix_t* h, void* x, void* beta, void* y, int nrhs) {
  return guard([&] { M(h)->gemv_host(trans, alpha, x, beta, y, nrhs, true); });
}

int gemm_scalar(char trans, void* alpha, hmat_matrix_t* h, void* x, void* beta, void* y, int nrhs) {
  return guard([&] { M(h)->gemv_host(trans, alpha, x, beta, y, nrhs, false); });
}

int gemm_dense(char trans_b, char trans_x, char side, const void* alpha, hmat_matrix_t* h,
               void* x, const void* beta, void* y, int nrhs) {
  return guard([&] {
    if (side != 'L' || trans_x != 'N')
      throw std::runtime_error("gemm_dense: only side 'L' with trans_x 'N' is supported");
    M(h)->gemv_host(trans_b, alpha, x, beta, y, nrhs, false);
  });
}

int init_engine() {
  return guard([] {
    int n = 0;
    HMX_CHECK(hipGetDeviceCount(&n));
    if (n < 1) throw std::runtime_error("no HIP device");
  });
}
int finalize_engine() { return 0; }

int get_info(hmat_matrix_t* h, hmat_info_t* info) {
  return guard([&] { M(h)->info(*info); });
}

int get_cluster_trees(hmat_matrix_t* h, const hmat_cluster_tree_t** r, const hmat_cluster_tree_t** c) {
  if (r) *r = reinterpret_cast<const hmat_cluster_tree_t*>(M(h)->rows);
  if (c) *c = reinterpret_cast<const hmat_cluster_tree_t*>(M(h)->cols);
  return 0;
}

// ---- unsupported slots -------------------------------------------------------
int unsupported(const char* what) {
  g_last_error = std::string(what) + ": not provided by the MI355X engine";
  fprintf(stderr, "%s\n", g_last_error.c_str());
  return 1;
}
int u_inverse(hmat_matrix_t*) { return unsupported("inverse"); }
hmat_matrix_t* u_get_child(hmat_matrix_t*, int, int) { unsupported("get_child"); return nullptr; }
int u_h(hmat_matrix_t*) { return unsupported("operation"); }
// solve_mat c_wrapping.hpp:547-557 (HMatrix::solve(HMatrix*) h_matrix.cpp:2937-2966)
int solve_mat_h(hmat_matrix_t* h, hmat_matrix_t* b) {
  return guard([&] { M(h)->solve_mat(*M(b)); });
}
int u_trsm(char, char, char, char, int, int, void*, hmat_matrix_t*, int, void*) { return unsupported("trsm"); }
int u_full_gemm(char, char, int, int, void*, void*, void*, hmat_matrix_t*, void*) { return unsupported("full_gemm"); }
int u_profile(hmat_matrix_t*, hmat_profile_t*) { return unsupported("get_profile"); }
int u_ratio(hmat_matrix_t*, hmat_FPCompressionRatio_t*) { return unsupported("get_ratio"); }
hmat_fp_settings_t u_get_fp(hmat_matrix_t*) { unsupported("GetFPCompressionSettings"); return hmat_fp_settings_t{}; }
int u_set_fp(hmat_matrix_t*, hmat_fp_settings_t) { return unsupported("SetFPCompressionSettings"); }
int u_set_fp2(hmat_matrix_t*, float, int, hmat_FPcompress_t, bool, bool) { return unsupported("SetFPCompressionSettingsParams"); }
int u_dump(hmat_matrix_t*, char*) { return unsupported("dump_info"); }
int u_set_ct(hmat_matrix_t*, const hmat_cluster_tree_t*, const hmat_cluster_tree_t*) { return unsupported("set_cluster_trees"); }

// get_block c_wrapping.hpp:832-868 (UncompressedBlock): the rows [row_offset,
// +row_size) x cols [col_offset, +col_size) block in hmat numbering,
// column-major, ld row_size; renumber_rows puts the rows of a full-height
// block in user order; row_indices / col_indices point at the trees'
// numbering of the block
template <typename T>
int get_block(struct hmat_get_values_context_t* ctx) {
  return guard([&] {
    Matrix* m = M(ctx->matrix);
    std::vector<int> rp(ctx->row_size), cp(ctx->col_size);
    for (int i = 0; i < ctx->row_size; ++i) rp[i] = ctx->row_offset + i;
    for (int j = 0; j < ctx->col_size; ++j) cp[j] = ctx->col_offset + j;
    std::vector<double> v((size_t)ctx->row_size * ctx->col_size);
    m->get_entries(rp, cp, v.data(), ctx->row_size);
    T* out = static_cast<T*>(ctx->values);
    if (ctx->renumber_rows) {
      if (ctx->row_offset != 0 || ctx->row_size != m->n_rows())
        throw std::runtime_error("Cannot renumber");
      const int* idx = m->rows->indices.data();
      for (int j = 0; j < ctx->col_size; ++j)
        for (int i = 0; i < ctx->row_size; ++i)
          out[idx[i] + (size_t)j * ctx->row_size] = (T)v[i + (size_t)j * ctx->row_size];
    } else {
      for (size_t e = 0; e < v.size(); ++e) out[e] = (T)v[e];
    }
    ctx->row_indices = const_cast<int*>(m->rows->indices.data()) + ctx->row_offset;
    ctx->col_indices = const_cast<int*>(m->cols->indices.data()) + ctx->col_offset;
  });
}

// get_values c_wrapping.hpp:870-890 (UncompressedValues): the entries of the
// 1-based user indices row_indices x col_indices, column-major, ld row_size
template <typename T>
int get_values(struct hmat_get_values_context_t* ctx) {
  return guard([&] {
    Matrix* m = M(ctx->matrix);
    std::vector<int> rp(ctx->row_size), cp(ctx->col_size);
    for (int i = 0; i < ctx->row_size; ++i) {
      const int u = ctx->row_indices[i] - 1;
      if (u < 0 || u >= m->n_rows()) throw std::runtime_error("get_values: row index out of range");
      rp[i] = m->rows->indices_rev[u];
    }
    for (int j = 0; j < ctx->col_size; ++j) {
      const int u = ctx->col_indices[j] - 1;
      if (u < 0 || u >= m->n_cols()) throw std::runtime_error("get_values: column index out of range");
      cp[j] = m->cols->indices_rev[u];
    }
    std::vector<double> v((size_t)ctx->row_size * ctx->col_size);
    m->get_entries(rp, cp, v.data(), ctx->row_size);
    T* out = static_cast<T*>(ctx->values);
    for (size_t e = 0; e < v.size(); ++e) out[e] = (T)v[e];
  });
}
int u_walk(hmat_matrix_t*, hmat_procedure_t*) { return unsupported("walk"); }
int u_leaf(hmat_matrix_t*, const hmat_leaf_procedure_t*) { return unsupported("apply_on_leaf"); }
// read_struct / write_struct / read_data / write_data c_wrapping.hpp:921-946
template <int VT>
hmat_matrix_t* read_struct(hmat_iostream f, void* ud) {
  hmat_matrix_t* r = nullptr;
  guard([&] { r = reinterpret_cast<hmat_matrix_t*>(hmx::serialize_read_struct(f, ud, VT, g_rank_cap)); });
  return r;
}
void write_struct(hmat_matrix_t* h, hmat_iostream f, void* ud) {
  guard([&] { hmx::serialize_write_struct(*M(h), f, ud); });
}
void read_data(hmat_matrix_t* h, hmat_iostream f, void* ud) {
  guard([&] { hmx::serialize_read_data(*M(h), f, ud); });
}
void write_data(hmat_matrix_t* h, hmat_iostream f, void* ud) {
  guard([&] { hmx::serialize_write_data(*M(h), f, ud); });
}
// set_progressbar c_wrapping.hpp:949-951: the reporter used by later factorisations
void set_progressbar(hmat_matrix_t* h, hmat_progress_t* p) { M(h)->progress = p; }
int u_diag_block(hmat_matrix_t*, int, void*) { return unsupported("extract_diagonal_block"); }

template <int VT, typename T>
void fill(hmat_interface_t* i) {
  std::memset(i, 0, sizeof(*i));
  i->create_empty_hmatrix_admissibility = create_empty<VT>;
  i->own_cluster_trees = own_cluster_trees;
  i->set_low_rank_epsilon = set_low_rank_epsilon;
  i->assemble_generic = assemble_generic;
  i->copy = copy_matrix;
  i->copy_struct = copy_struct;
  i->norm = norm_matrix;
  i->logdet = logdet;
  i->inverse = u_inverse;
  i->factorize_generic = factorize_generic;
  i->destroy = destroy;
  i->get_child = u_get_child;
  i->destroy_child = u_h;
  i->solve_generic = solve_generic;
  i->solve_mat = solve_mat_h;
  i->solve_systems = solve_systems;
  i->solve_dense = solve_dense;
  i->transpose = transpose;
  i->scale = scale_matrix<T>;
  i->truncate = truncate_matrix;
  i->vector_reorder = vector_order<T, false>;
  i->vector_restore = vector_order<T, true>;
  i->gemm = gemm_h<T>;
  i->axpy = axpy_matrix<T>;
  i->trsm = u_trsm;
  i->gemv = gemv;
  i->gemm_scalar = gemm_scalar;
  i->full_gemm = u_full_gemm;
  i->gemm_dense = gemm_dense;
  i->add_identity = add_identity<T>;
  i->init = init_engine;
  i->finalize = finalize_engine;
  i->get_info = get_info;
  i->get_profile = u_profile;
  i->get_ratio = u_ratio;
  i->FPcompress = u_h;
  i->FPdecompress = u_h;
  i->GetFPCompressionSettings = u_get_fp;
  i->SetFPCompressionSettings = u_set_fp;
  i->SetFPCompressionSettingsParams = u_set_fp2;
  i->dump_info = u_dump;
  i->get_cluster_trees = get_cluster_trees;
  i->set_cluster_trees = u_set_ct;
  i->extract_diagonal = extract_diagonal<T>;
  i->solve_lower_triangular = solve_lower_triangular;
  i->solve_lower_triangular_dense = solve_lower_triangular_dense;
  i->get_block = get_block<T>;
  i->get_values = get_values<T>;
  i->walk = u_walk;
  i->apply_on_leaf = u_leaf;
  i->value_type = (hmat_value_t)VT;
  i->internal = nullptr;
  i->read_struct = read_struct<VT>;
  i->write_struct = write_struct;
  i->read_data = read_data;
  i->write_data = write_data;
  i->set_progressbar = set_progressbar;
  i->extract_diagonal_block = u_diag_block;
}

// ---- complex (Z) entries --------------------------------------------------------
// The engine's matrix is the 2 x 2 real representation of the complex one
// (Matrix::create_complex).  A complex array of n entries (interleaved re, im)
// is the real array of the embedded numbering, so vectors pass through as they
// are; complex scalars with a nonzero imaginary part are applied on the host
// (gemv) or refused where the real engine would have to take them.
using zd = std::complex<double>;
using zc = std::complex<float>;   // single-precision complex (C)

template <typename Z>
double real_scalar(const void* a, const char* what) {
  const Z v = *static_cast<const Z*>(a);
  if (v.imag() != 0.0)
    throw std::runtime_error(std::string(what) +
                             ": a complex scalar with a nonzero imaginary part is not provided");
  return v.real();
}

Matrix* MZ(hmat_matrix_t* h) {
  Matrix* m = M(h);
  if (!m->embedded) throw std::runtime_error("not a complex matrix");
  return m;
}

template <int VT>
hmat_matrix_t* create_empty_z(const hmat_cluster_tree_t* r, const hmat_cluster_tree_t* c,
                              int lower_sym, hmat_admissibility_t* cond) {
  try {
    if (lower_sym)
      throw std::runtime_error("complex matrices are stored whole (H-LU); lower-symmetric storage "
                               "of complex matrices is not provided");
    hmx::AdmissibilityParams p;
    if (cond) p = reinterpret_cast<Admissibility*>(cond)->p;
    auto* m = Matrix::create_complex(reinterpret_cast<const hmx::ClusterTree*>(r),
                                     reinterpret_cast<const hmx::ClusterTree*>(c), p, g_rank_cap,
                                     nullptr, nullptr, VT);
    return reinterpret_cast<hmat_matrix_t*>(m);
  } catch (const std::exception& e) {
    fail(e);
    return nullptr;
  }
}

int factorize_z(hmat_matrix_t* h, hmat_factorization_context_t* ctx) {
  return guard([&] {
    if (ctx->factorization != hmat_factorization_lu)
      throw std::runtime_error("complex matrices: only hmat_factorization_lu is provided");
    Matrix* m = MZ(h);
    hmat_progress_t* saved = m->progress;
    if (ctx->progress) m->progress = ctx->progress;
    struct Restore {
      Matrix* m;
      hmat_progress_t* p;
      ~Restore() { m->progress = p; }
    } restore{m, saved};
    m->factorize(ctx->factorization, false);
  });
}

// y = alpha A x + beta y with complex alpha, beta: real ones go to the device
// as they are, others through t = A x on the device and the combination on
// the host
template <typename Z>
void gemv_z_run(Matrix* m, char trans, const void* alpha, void* x, const void* beta, void* y,
                int nrhs, bool permute) {
  using R = typename Z::value_type;
  if (trans != 'N') throw std::runtime_error("gemv: only 'N' for complex matrices");
  const Z a = *static_cast<const Z*>(alpha), b = *static_cast<const Z*>(beta);
  if (a.imag() == R(0) && b.imag() == R(0)) {
    const R ar = a.real(), br = b.real();
    m->gemv_host('N', &ar, x, &br, y, nrhs, permute);
    return;
  }
  std::vector<Z> t((size_t)m->zrows->n * nrhs);
  const R one = 1, zero = 0;
  m->gemv_host('N', &one, x, &zero, t.data(), nrhs, permute);
  Z* yz = static_cast<Z*>(y);
  for (size_t i = 0; i < t.size(); ++i) yz[i] = a * t[i] + b * yz[i];
}
template <typename Z>
int gemv_z(char trans, void* alpha, hmat_matrix_t* h, void* x, void* beta, void* y, int nrhs) {
  return guard([&] { gemv_z_run<Z>(MZ(h), trans, alpha, x, beta, y, nrhs, true); });
}
template <typename Z>
int gemm_scalar_z(char trans, void* alpha, hmat_matrix_t* h, void* x, void* beta, void* y, int nrhs) {
  return guard([&] { gemv_z_run<Z>(MZ(h), trans, alpha, x, beta, y, nrhs, false); });
}
template <typename Z>
int gemm_dense_z(char trans_b, char trans_x, char side, const void* alpha, hmat_matrix_t* h,
                 void* x, const void* beta, void* y, int nrhs) {
  return guard([&] {
    if (side != 'L' || trans_x != 'N')
      throw std::runtime_error("gemm_dense: only side 'L' with trans_x 'N' is supported");
    gemv_z_run<Z>(MZ(h), trans_b, alpha, x, beta, y, nrhs, false);
  });
}
template <typename Z>
int scale_z(void* alpha, hmat_matrix_t* h) {
  return guard([&] { MZ(h)->scale(real_scalar<Z>(alpha, "scale")); });
}
template <typename Z>
int add_identity_z(hmat_matrix_t* h, void* alpha) {
  return guard([&] { MZ(h)->add_identity(real_scalar<Z>(alpha, "add_identity")); });
}
template <typename Z>
int axpy_z(void* a, hmat_matrix_t* x, hmat_matrix_t* y) {
  return guard([&] { MZ(y)->axpy_from(real_scalar<Z>(a, "axpy"), *MZ(x)); });
}
template <typename Z>
int gemm_z(char ta, char tb, const void* alpha, hmat_matrix_t* a, hmat_matrix_t* b,
           const void* beta, hmat_matrix_t* c) {
  return guard([&] {
    if (ta != 'N' || tb != 'N') throw std::runtime_error("gemm: only 'N', 'N' for complex matrices");
    MZ(c)->gemm_from('N', 'N', real_scalar<Z>(alpha, "gemm"), *MZ(a), *MZ(b), real_scalar<Z>(beta, "gemm"));
  });
}
int get_cluster_trees_z(hmat_matrix_t* h, const hmat_cluster_tree_t** r, const hmat_cluster_tree_t** c) {
  if (r) *r = reinterpret_cast<const hmat_cluster_tree_t*>(M(h)->zrows);
  if (c) *c = reinterpret_cast<const hmat_cluster_tree_t*>(M(h)->zcols);
  return 0;
}
int destroy_z(hmat_matrix_t* h) {
  Matrix* m = M(h);
  const hmx::ClusterTree* r = m->own_rows ? m->zrows : nullptr;
  const hmx::ClusterTree* c = m->own_cols ? m->zcols : nullptr;
  delete m;
  if (r) delete r;
  if (c && c != r) delete c;
  return 0;
}
// get_block / get_values of a complex matrix: entry (p, q) of the complex
// matrix (hmat numbering) is read from the real representation's 2 x 2 block
// at rows 2p, 2p + 1 and columns 2q, 2q + 1 (the average of the two places
// holding Re and of those holding Im, as the compression keeps them to eps)
template <typename Z>
void complex_entries(Matrix* m, const std::vector<int>& rp, const std::vector<int>& cp, Z* out) {
  if (m->factorization != -1)
    throw std::runtime_error("get_block/get_values: the factor of a complex matrix is held as the real "
                             "representation's LU, which has no complex entries");
  const int nr = (int)rp.size(), nc = (int)cp.size();
  std::vector<int> er(2 * (size_t)nr), ec(2 * (size_t)nc);
  for (int i = 0; i < nr; ++i) er[2 * i] = 2 * rp[i], er[2 * i + 1] = 2 * rp[i] + 1;
  for (int j = 0; j < nc; ++j) ec[2 * j] = 2 * cp[j], ec[2 * j + 1] = 2 * cp[j] + 1;
  std::vector<double> v(4 * (size_t)nr * nc);
  m->get_entries(er, ec, v.data(), 2 * nr);
  const size_t ld = 2 * (size_t)nr;
  for (int j = 0; j < nc; ++j)
    for (int i = 0; i < nr; ++i) {
      const double a = v[2 * i + 2 * j * ld], b = v[2 * i + 1 + 2 * j * ld];
      const double c = v[2 * i + (2 * j + 1) * ld], d = v[2 * i + 1 + (2 * j + 1) * ld];
      out[i + (size_t)j * nr] = Z(0.5 * (a + d), 0.5 * (b - c));
    }
}
template <typename Z>
int get_block_z(struct hmat_get_values_context_t* ctx) {
  return guard([&] {
    Matrix* m = MZ(ctx->matrix);
    const hmx::ClusterTree& R = *m->zrows;
    const hmx::ClusterTree& Cl = *m->zcols;
    if (ctx->row_offset < 0 || ctx->col_offset < 0 || ctx->row_offset + ctx->row_size > R.n ||
        ctx->col_offset + ctx->col_size > Cl.n)
      throw std::runtime_error("get_block: block out of range");
    std::vector<int> rp(ctx->row_size), cp(ctx->col_size);
    for (int i = 0; i < ctx->row_size; ++i) rp[i] = ctx->row_offset + i;
    for (int j = 0; j < ctx->col_size; ++j) cp[j] = ctx->col_offset + j;
    std::vector<Z> v((size_t)ctx->row_size * ctx->col_size);
    complex_entries(m, rp, cp, v.data());
    Z* out = static_cast<Z*>(ctx->values);
    if (ctx->renumber_rows) {
      if (ctx->row_offset != 0 || ctx->row_size != R.n) throw std::runtime_error("Cannot renumber");
      for (int j = 0; j < ctx->col_size; ++j)
        for (int i = 0; i < ctx->row_size; ++i)
          out[R.indices[i] + (size_t)j * ctx->row_size] = v[i + (size_t)j * ctx->row_size];
    } else {
      std::copy(v.begin(), v.end(), out);
    }
    ctx->row_indices = const_cast<int*>(R.indices.data()) + ctx->row_offset;
    ctx->col_indices = const_cast<int*>(Cl.indices.data()) + ctx->col_offset;
  });
}
template <typename Z>
int get_values_z(struct hmat_get_values_context_t* ctx) {
  return guard([&] {
    Matrix* m = MZ(ctx->matrix);
    const hmx::ClusterTree& R = *m->zrows;
    const hmx::ClusterTree& Cl = *m->zcols;
    std::vector<int> rp(ctx->row_size), cp(ctx->col_size);
    for (int i = 0; i < ctx->row_size; ++i) {
      const int u = ctx->row_indices[i] - 1;
      if (u < 0 || u >= R.n) throw std::runtime_error("get_values: row index out of range");
      rp[i] = R.indices_rev[u];
    }
    for (int j = 0; j < ctx->col_size; ++j) {
      const int u = ctx->col_indices[j] - 1;
      if (u < 0 || u >= Cl.n) throw std::runtime_error("get_values: column index out of range");
      cp[j] = Cl.indices_rev[u];
    }
    complex_entries(m, rp, cp, static_cast<Z*>(ctx->values));
  });
}
int u_logdet_z(hmat_matrix_t*, double*) { return unsupported("logdet of a complex matrix"); }
int u_diag_z(hmat_matrix_t*, void*, int) { return unsupported("extract_diagonal of a complex matrix"); }
int u_tri_z(hmat_matrix_t*, int, void*, int) {
  return unsupported("solve_lower_triangular of a complex matrix (LLt only)");
}

// A complex table: the real table of the component type, with the entries
// that see complex scalars, vectors or trees replaced
template <int VT, typename T, typename Z, int ZT>
void fill_complex(hmat_interface_t* i) {
  fill<VT, T>(i);
  i->value_type = (hmat_value_t)ZT;
  i->create_empty_hmatrix_admissibility = create_empty_z<VT>;
  i->factorize_generic = factorize_z;
  i->destroy = destroy_z;
  i->gemv = gemv_z<Z>;
  i->gemm_scalar = gemm_scalar_z<Z>;
  i->gemm_dense = gemm_dense_z<Z>;
  i->scale = scale_z<Z>;
  i->add_identity = add_identity_z<Z>;
  i->axpy = axpy_z<Z>;
  i->gemm = gemm_z<Z>;
  i->vector_reorder = vector_order<Z, false>;
  i->vector_restore = vector_order<Z, true>;
  i->get_cluster_trees = get_cluster_trees_z;
  i->logdet = u_logdet_z;
  i->extract_diagonal = u_diag_z;
  i->solve_lower_triangular = u_tri_z;
  i->solve_lower_triangular_dense = u_tri_z;
  i->get_block = get_block_z<Z>;
  i->get_values = get_values_z<Z>;
  // the stream holds the complex matrix (unfactorised); the engine reads it
  // into the real representation
  i->read_struct = read_struct<ZT>;
}

}  // namespace

extern "C" {

void hmat_init_mi355x_interface(hmat_interface_t* i, hmat_value_t type) {
  switch (type) {
    case HMAT_SIMPLE_PRECISION: fill<HMAT_SIMPLE_PRECISION, float>(i); break;
    case HMAT_DOUBLE_PRECISION: fill<HMAT_DOUBLE_PRECISION, double>(i); break;
    case HMAT_DOUBLE_COMPLEX:
      fill_complex<HMAT_DOUBLE_PRECISION, double, zd, HMAT_DOUBLE_COMPLEX>(i);
      break;
    case HMAT_SIMPLE_COMPLEX:
      fill_complex<HMAT_SIMPLE_PRECISION, float, zc, HMAT_SIMPLE_COMPLEX>(i);
      break;
  }
}

void hmat_mi355x_kernel_compute(void* ctx, int row, int col, void* result) {
  const auto* k = static_cast<const hmat_mi355x_kernel_t*>(ctx);
  double d[3] = {0, 0, 0};
  for (int q = 0; q < k->dimension && q < 3; ++q)
    d[q] = k->points[(size_t)row * k->dimension + q] - k->points[(size_t)col * k->dimension + q];
  hmx::KernelParams kp{k->kind, k->length, k->nugget,
                      k->kind == HMAT_MI355X_KERNEL_SHIFTED_EXPONENTIAL ? k->shift : 0.0,
                      k->wavenumber, k->regularization};
  if (k->kind == HMAT_MI355X_KERNEL_HELMHOLTZ) {
    // a complex value: (re, im)
    hmx::helmholtz_value(kp, d[0], d[1], d[2], row == col, (double*)result, (double*)result + 1);
    return;
  }
  *(double*)result = hmx::kernel_value(kp, d[0], d[1], d[2], row == col);
}

// ---- clustering / trees -------------------------------------------------------
hmat_clustering_algorithm_t* hmat_create_clustering_median(void) {
  return reinterpret_cast<hmat_clustering_algorithm_t*>(new ClusteringAlgo());
}
hmat_clustering_algorithm_t* hmat_create_clustering_max_dof(const hmat_clustering_algorithm_t* a, int max_dof) {
  auto* r = new ClusteringAlgo(*reinterpret_cast<const ClusteringAlgo*>(a));
  r->max_leaf = max_dof;
  return reinterpret_cast<hmat_clustering_algorithm_t*>(r);
}
void hmat_delete_clustering(hmat_clustering_algorithm_t* a) { delete reinterpret_cast<ClusteringAlgo*>(a); }
void hmat_set_clustering_divider(hmat_clustering_algorithm_t* a, int divider) {
  if (divider != 2) fprintf(stderr, "hmat_set_clustering_divider: only binary median trees are supported\n");
  reinterpret_cast<ClusteringAlgo*>(a)->divider = 2;
}

hmat_cluster_tree_t* hmat_create_cluster_tree(double* coord, int dimension, int size,
                                              hmat_clustering_algorithm_t* algo) {
  try {
    const auto* a = reinterpret_cast<const ClusteringAlgo*>(algo);
    const int leaf = a->max_leaf >= 0 ? a->max_leaf : g_settings.maxLeafSize;
    return reinterpret_cast<hmat_cluster_tree_t*>(hmx::build_cluster_tree(coord, dimension, size, leaf).release());
  } catch (const std::exception& e) {
    fail(e);
    return nullptr;
  }
}

struct hmat_cluster_tree_builder {
  ClusteringAlgo algo;
};
hmat_cluster_tree_builder_t* hmat_create_cluster_tree_builder(const hmat_clustering_algorithm_t* a) {
  return new hmat_cluster_tree_builder{*reinterpret_cast<const ClusteringAlgo*>(a)};
}
void hmat_delete_cluster_tree_builder(hmat_cluster_tree_builder_t* b) { delete b; }

hmat_cluster_tree_t* hmat_create_cluster_tree_generic(struct hmat_cluster_tree_create_context_t* ctx) {
  try {
    if (ctx->span_offsets || ctx->group_index)
      throw std::runtime_error("hmat_create_cluster_tree_generic: spans and groups are not supported");
    const int leaf = ctx->builder && ctx->builder->algo.max_leaf >= 0 ? ctx->builder->algo.max_leaf
                                                                      : g_settings.maxLeafSize;
    return reinterpret_cast<hmat_cluster_tree_t*>(
        hmx::build_cluster_tree(ctx->coordinates, (int)ctx->dimension, (int)ctx->number_of_dof, leaf).release());
  } catch (const std::exception& e) {
    fail(e);
    return nullptr;
  }
}

void hmat_delete_cluster_tree(const hmat_cluster_tree_t* t) {
  delete reinterpret_cast<const hmx::ClusterTree*>(t);
}
int hmat_tree_nodes_count(const hmat_cluster_tree_t* t) {
  return reinterpret_cast<const hmx::ClusterTree*>(t)->node_count();
}
int hmat_cluster_get_info(hmat_cluster_tree_t* t, hmat_cluster_info_t* info) {
  const auto* ct = reinterpret_cast<const hmx::ClusterTree*>(t);
  info->spatial_dimension = ct->dim;
  info->dimension = ct->n;
  info->nr_tree_nodes = ct->node_count();
  return 0;
}
const int* hmat_cluster_get_indices(const hmat_cluster_tree_t* t) {
  return reinterpret_cast<const hmx::ClusterTree*>(t)->indices.data();
}

// ---- admissibility / compression ---------------------------------------------
void hmat_init_admissibility_param(hmat_admissibility_param_t* p) {
  p->eta = 2;
  p->ratio = 0.0;
  p->max_width = (size_t)-1L;
  p->scale = -1;
}
hmat_admissibility_t* hmat_create_admissibility(hmat_admissibility_param_t* p) {
  auto* a = new Admissibility();
  a->p.eta = p->eta;
  a->p.ratio = p->ratio;
  a->p.max_width = p->max_width;
  return reinterpret_cast<hmat_admissibility_t*>(a);
}
void hmat_update_admissibility(hmat_admissibility_t* c, hmat_admissibility_param_t* p) {
  auto* a = reinterpret_cast<Admissibility*>(c);
  a->p.ratio = p->ratio;
  a->p.max_width = p->max_width;
}
hmat_admissibility_t* hmat_create_admissibility_standard(double eta) {
  auto* a = new Admissibility();
  a->p.eta = eta;
  return reinterpret_cast<hmat_admissibility_t*>(a);
}
// hmat_create_admissibility_hodlr c_default_interface.cpp:220-222
hmat_admissibility_t* hmat_create_admissibility_hodlr(void) {
  auto* a = new Admissibility();
  a->p.hodlr = true;
  return reinterpret_cast<hmat_admissibility_t*>(a);
}
void hmat_delete_admissibility(hmat_admissibility_t* c) { delete reinterpret_cast<Admissibility*>(c); }

hmat_compression_algorithm_t* hmat_create_compression_svd(double eps) {
  return reinterpret_cast<hmat_compression_algorithm_t*>(new Compression{hmat_compress_svd, eps});
}
hmat_compression_algorithm_t* hmat_create_compression_aca_partial(double eps) {
  return reinterpret_cast<hmat_compression_algorithm_t*>(new Compression{hmat_compress_aca_partial, eps});
}
hmat_compression_algorithm_t* hmat_create_compression_aca_plus(double eps) {
  return reinterpret_cast<hmat_compression_algorithm_t*>(new Compression{hmat_compress_aca_plus, eps});
}
void hmat_delete_compression(const hmat_compression_algorithm_t* c) {
  delete reinterpret_cast<const Compression*>(c);
}

// ---- contexts / settings -------------------------------------------------------
void hmat_assemble_context_init(hmat_assemble_context_t* c) {
  std::memset(c, 0, sizeof(*c));
  c->factorization = hmat_factorization_none;
  c->progress = nullptr;
}
void hmat_factorization_context_init(hmat_factorization_context_t* c) {
  c->factorization = hmat_factorization_lu;
  c->progress = nullptr;
}
void hmat_solve_context_init(struct hmat_solve_context_t* c) { std::memset(c, 0, sizeof(*c)); }

static hmat_progress_t g_progress = {0, 0, nullptr, nullptr};
hmat_progress_t* hmat_default_progress(void) { return &g_progress; }

void hmat_get_parameters(hmat_settings_t* s) { *s = g_settings; }
int hmat_set_parameters(hmat_settings_t* s) {
  if (s->compressionMinLeafSize != 100) {
    fprintf(stderr, "hmat_set_parameters: compressionMinLeafSize is fixed at 100 in this engine\n");
    return 1;
  }
  g_settings = *s;
  return 0;
}
const char* hmat_get_version(void) { return "hmat-mi355x 0.1"; }

// ---- flat helpers ----------------------------------------------------------------
const char* hmx_last_error(void) { return g_last_error.c_str(); }

int hmx_gemv_device(hmat_matrix_t* h, int precision, double alpha, const void* x, double beta,
                    void* y, int nrhs) {
  return guard([&] {
    Matrix* m = M(h);
    m->device_init();
    if (precision == 0)
      m->gemv_device<double>(alpha, (const double*)x, beta, (double*)y, nrhs, m->n_cols(), m->n_rows(), m->s64);
    else
      m->gemv_device<float>(alpha, (const float*)x, beta, (float*)y, nrhs, m->n_cols(), m->n_rows(), m->s32);
    if (!m->user_stream) HMX_CHECK(hipStreamSynchronize(m->stream));
  });
}

int hmx_gemv_device_noalloc(hmat_matrix_t* h, int precision, double alpha, const void* x,
                            double beta, void* y, int nrhs) {
  return guard([&] {
    Matrix* m = M(h);
    if (!m->assembled || !m->stream)
      throw std::runtime_error("hmx_gemv_device_noalloc: matrix not assembled");
    size_t elems = 0;
    double* scratch = hmx::noalloc_gemv_scratch(&elems);
    if (precision == 0)
      m->gemv_device_scratch<double>(alpha, (const double*)x, beta, (double*)y, nrhs, m->s64,
                                     scratch, elems);
    else
      m->gemv_device_scratch<float>(alpha, (const float*)x, beta, (float*)y, nrhs, m->s32,
                                    scratch, elems);
    if (!m->user_stream) HMX_CHECK(hipStreamSynchronize(m->stream));
  });
}

int hmx_solve_device(hmat_matrix_t* h, int precision, void* b, int nrhs) {
  return guard([&] {
    Matrix* m = M(h);
    if (precision == 0)
      m->solve_device<double>((double*)b, nrhs, m->n_rows(), m->s64);
    else
      m->solve_device<float>((float*)b, nrhs, m->n_rows(), m->s32);
    if (!m->user_stream) HMX_CHECK(hipStreamSynchronize(m->stream));
  });
}

int hmx_factorize_mixed(hmat_matrix_t* h, int algo) {
  return guard([&] { M(h)->factorize(algo, true); });
}

int hmx_solve_refined_device(hmat_matrix_t* h, double* b, int nrhs, double tol, int max_iter,
                             double* resid) {
  int it = -1;
  const int rc = guard([&] { it = M(h)->solve_refined(b, nrhs, M(h)->n_rows(), tol, max_iter, resid); });
  return rc ? -1 : it;
}

int hmx_run_batch(hmat_matrix_t** ms, hmat_assemble_context_t** ctx, int count, int algo,
                  int mixed, double tol, double** b, const int* nrhs, int threads,
                  double* seconds, int* status) {
  return guard([&] {
    if (count < 0 || threads < 1) throw std::runtime_error("hmx_run_batch: bad count or thread count");
    // several host threads submit at once: each system keeps to its own stream
    // and to its share of the device memory for temporaries
    size_t freeb = 0, total = 0;
    HMX_CHECK(hipMemGetInfo(&freeb, &total));
    const int workers = std::min(threads, std::max(count, 1));
    const size_t share = std::min((size_t)(0.5 * (double)freeb / workers), (size_t)96 << 30);
    for (int i = 0; i < count; ++i) {
      M(ms[i])->single_stream = true;
      M(ms[i])->temp_budget_limit = share;
    }
    std::atomic<int> next{0};
    auto work = [&]() {
      for (int i = next++; i < count; i = next++) {
        const auto t0 = std::chrono::steady_clock::now();
        int rc = assemble_generic(ms[i], ctx[i]);
        if (rc == 0)
          rc = guard([&] {
            Matrix* m = M(ms[i]);
            const int n = m->n_rows();
            if (!mixed) {
              m->factorize(algo, false);
              m->solve_host(b[i], nrhs[i], true);
              m->release_temporaries();
              return;
            }
            // FP32 factor, refinement against the FP64 matrix on device vectors
            m->factorize(algo, true);
            const size_t cnt = (size_t)n * nrhs[i];
            std::vector<double> hb(cnt);
            for (int r = 0; r < nrhs[i]; ++r)
              for (int q = 0; q < n; ++q) hb[q + (size_t)r * n] = b[i][m->rows->indices[q] + (size_t)r * n];
            hmx::DevBuf d;
            hmx::upload(d, hb.data(), cnt * sizeof(double), m->stream);
            double res = 0.0;
            m->solve_refined(d.as<double>(), nrhs[i], n, tol, 200, &res);
            if (!(res <= tol)) throw std::runtime_error("hmx_run_batch: refinement did not reach tol");
            HMX_CHECK(hipMemcpy(hb.data(), d.p, cnt * sizeof(double), hipMemcpyDeviceToHost));
            for (int r = 0; r < nrhs[i]; ++r)
              for (int q = 0; q < n; ++q) b[i][m->rows->indices[q] + (size_t)r * n] = hb[q + (size_t)r * n];
            m->release_temporaries();
          });
        seconds[i] = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        status[i] = rc;
      }
    };
    std::vector<std::thread> pool;
    for (int t = 1; t < threads && t < count; ++t) pool.emplace_back(work);
    work();
    for (auto& th : pool) th.join();
  });
}

int hmx_partition(hmat_matrix_t* h, int* out) {
  int count = 0;
  std::function<void(const hmx::Block*)> rec = [&](const hmx::Block* b) {
    if (out) {
      int* o = out + 5 * count;
      o[0] = b->rows->offset;
      o[1] = b->rows->size;
      o[2] = b->cols->offset;
      o[3] = b->cols->size;
      o[4] = (int)b->kind;
    }
    ++count;
    if (!b->is_leaf())
      for (int i = 0; i < b->nrow * b->ncol; ++i)
        if (b->child[i]) rec(b->child[i]);
  };
  rec(M(h)->bt->root);
  return count;
}

int hmx_leaf_count(hmat_matrix_t* h) { return (int)M(h)->bt->leaves.size(); }

int hmx_leaf_info(hmat_matrix_t* h, int leaf, int precision, int* info) {
  return guard([&] {
    Matrix* m = M(h);
    const hmx::LeafDesc& L = m->s64.hdesc.at(leaf);
    info[0] = L.m;
    info[1] = L.n;
    info[2] = L.r0;
    info[3] = L.c0;
    info[4] = L.kind;
    int r = -1;
    const hmx::DevBuf& rk = precision == 0 ? m->s64.ranks : m->s32.ranks;
    if (L.kind == 1 && rk.p)
      HMX_CHECK(hipMemcpy(&r, rk.as<int>() + leaf, sizeof(int), hipMemcpyDeviceToHost));
    info[5] = r;
  });
}

int hmx_leaf_table(hmat_matrix_t* h, int precision, int* out) {
  return guard([&] {
    Matrix* m = M(h);
    const auto& D = m->s64.hdesc;
    std::vector<int> rk(D.size(), -1);
    const hmx::DevBuf& rb = precision == 0 ? m->s64.ranks : m->s32.ranks;
    if (rb.p) HMX_CHECK(hipMemcpy(rk.data(), rb.p, rk.size() * sizeof(int), hipMemcpyDeviceToHost));
    for (size_t i = 0; i < D.size(); ++i) {
      int* o = out + 6 * i;
      o[0] = D[i].m;
      o[1] = D[i].n;
      o[2] = D[i].r0;
      o[3] = D[i].c0;
      o[4] = D[i].kind;
      o[5] = D[i].kind == 0 ? -1 : rk[i];
    }
  });
}

int hmx_leaf_data(hmat_matrix_t* h, int leaf, int precision, double* a, double* b) {
  return guard([&] {
    Matrix* m = M(h);
    const hmx::LeafDesc& L = m->s64.hdesc.at(leaf);
    auto fetch = [&](int64_t off, size_t count, double* dst) {
      if (precision == 0) {
        if (!m->s64.arena.p) throw std::runtime_error("no FP64 data");
        HMX_CHECK(hipMemcpy(dst, m->s64.data() + off, count * sizeof(double), hipMemcpyDeviceToHost));
      } else {
        if (!m->s32.arena.p) throw std::runtime_error("no FP32 data");
        std::vector<float> tmp(count);
        HMX_CHECK(hipMemcpy(tmp.data(), m->s32.data() + off, count * sizeof(float), hipMemcpyDeviceToHost));
        for (size_t i = 0; i < count; ++i) dst[i] = tmp[i];
      }
    };
    if (L.kind == 0) {
      fetch(L.a, (size_t)L.m * L.n, a);
    } else {
      int r = 0;
      const hmx::DevBuf& rk = precision == 0 ? m->s64.ranks : m->s32.ranks;
      HMX_CHECK(hipMemcpy(&r, rk.as<int>() + leaf, sizeof(int), hipMemcpyDeviceToHost));
      if (r > 0) {
        fetch(L.a, (size_t)L.m * r, a);
        fetch(L.b, (size_t)L.n * r, b);
      }
    }
  });
}

int hmx_stats(hmat_matrix_t* h, double* out) {
  const hmx::Stats& s = M(h)->stats;
  out[0] = s.t_assemble;
  out[1] = s.t_factor;
  out[2] = s.t_solve;
  out[3] = (double)s.tasks;
  out[4] = (double)s.waves;
  out[5] = (double)s.launches;
  out[6] = (double)s.segments;
  out[7] = (double)s.temp_bytes;
  return 0;
}

int hmx_set_stream(hmat_matrix_t* h, void* s) {
  return guard([&] {
    Matrix* m = M(h);
    m->device_init();
    HMX_CHECK(hipStreamSynchronize(m->stream));
    m->stream = s ? (hipStream_t)s : m->own_stream;
    m->user_stream = s != nullptr;
  });
}

void hmx_set_rank_capacity(int cap) { g_rank_cap = cap > 0 ? cap : 160; }

int hmx_autotune_solve(hmat_matrix_t* h, int nrhs, int* best_rows, double* best_ms) {
  return guard([&] {
    const double ms = M(h)->autotune_solve(nrhs, best_rows);
    if (best_ms) *best_ms = ms;
  });
}

int hmx_set_large_block_rows(hmat_matrix_t* h, int rows) {
  return guard([&] {
    if (rows < 1) throw std::runtime_error("hmx_set_large_block_rows: rows must be positive");
    Matrix* m = M(h);
    // complex rows; the real representation has twice as many
    m->giant_rows = m->embedded ? 2 * rows : rows;
  });
}

int hmx_set_solve_block_rows(hmat_matrix_t* h, int rows) {
  return guard([&] {
    if (rows < 1) throw std::runtime_error("hmx_set_solve_block_rows: rows must be positive");
    Matrix* m = M(h);
    m->inv_max_rows = rows;
    m->drop_solve_plans();
  });
}

int hmx_ftask_size(void) { return (int)sizeof(hmx::FTask); }

int hmx_run_tasks(int precision, int type, const void* tasks, int count, int* flags_out) {
  return guard([&] {
    if (count <= 0) return;
    if (type < 0 || type >= hmx::T_NTYPES) throw std::runtime_error("hmx_run_tasks: unknown task type");
    std::vector<hmx::FTask> t((const hmx::FTask*)tasks, (const hmx::FTask*)tasks + count);
    std::vector<hmx::GemmTile> tiles;
    int64_t ws = 0;
    for (int i = 0; i < count; ++i) {
      hmx::FTask& x = t[i];
      int64_t need = 0;
      if (type == hmx::T_FADD) {
        int kmax = x.tgt.cap;
        for (int q = 0; q < x.nparts; ++q) kmax += x.parts[q].rk.cap;
        need = hmx::fadd_workspace(x.tgt.m, x.tgt.n, kmax);
      } else if (type == hmx::T_FADD_DENSE) {
        need = hmx::fadd_dense_workspace(x.tgt.m, x.tgt.n);
      } else if (type == hmx::T_RKRK) {
        need = hmx::rkrk_workspace(x.parts[0].rk.cap, x.parts[1].rk.cap);
      } else if (type == hmx::T_HODLR) {
        need = hmx::hodlr_workspace(x.lda);
      } else if (type == hmx::T_GEMM) {
        for (int a = 0; a < (x.m.s + 63) / 64; ++a)
          for (int b = 0; b < (x.n.s + 63) / 64; ++b) tiles.push_back(hmx::GemmTile{i, a, b});
      }
      x.ws = ws;
      ws += need;
    }
    hipStream_t s = nullptr;
    HMX_CHECK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    struct StreamGuard {
      hipStream_t s;
      ~StreamGuard() { (void)hipStreamDestroy(s); }
    } sg{s};
    hmx::DevBuf dt, dtl, dws, dfl;
    hmx::upload(dt, t.data(), t.size() * sizeof(hmx::FTask), s);
    if (!tiles.empty()) hmx::upload(dtl, tiles.data(), tiles.size() * sizeof(hmx::GemmTile), s);
    dws.alloc(sizeof(double) * (size_t)std::max<int64_t>(ws, 1));
    dfl.alloc(2 * sizeof(int));
    HMX_CHECK(hipMemsetAsync(dfl.p, 0, 2 * sizeof(int), s));
    if (precision == 0)
      hmx::launch_task_kernel<double>(type, dt.as<hmx::FTask>(), count, dtl.as<hmx::GemmTile>(),
                                      (int)tiles.size(), dws.as<double>(), dfl.as<int>(), s);
    else
      hmx::launch_task_kernel<float>(type, dt.as<hmx::FTask>(), count, dtl.as<hmx::GemmTile>(),
                                     (int)tiles.size(), dws.as<double>(), dfl.as<int>(), s);
    HMX_CHECK(hipStreamSynchronize(s));
    if (flags_out) HMX_CHECK(hipMemcpy(flags_out, dfl.p, 2 * sizeof(int), hipMemcpyDeviceToHost));
  });
}

}  // extern "C"
