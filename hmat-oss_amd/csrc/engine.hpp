// The MI355X engine behind the C API: one Matrix object per hmat_matrix_t.
// The block tree lives on the host; every leaf lives in one HBM arena per
// precision; the hot operations run as batched HIP kernels on the matrix's
// stream.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "hmat/hmat.h"
#include "kernels.hpp"
#include "tree.hpp"

namespace hmx {

// Owning device allocation.
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() { release(); }
  void release();
  void alloc(size_t nbytes);          // discards old content
  void ensure(size_t nbytes);         // grows (discards) if too small
  template <typename T>
  T* as() const { return static_cast<T*>(p); }
};

void upload(DevBuf& b, const void* src, size_t bytes, hipStream_t s);

// Leaf storage of one precision.
template <typename T>
struct Store {
  DevBuf arena;
  DevBuf ranks;     // int per leaf (full: -1)
  DevBuf desc;      // LeafDesc per leaf
  std::vector<LeafDesc> hdesc;
  size_t elems = 0;
  bool valid = false;
  T* data() const { return arena.as<T>(); }
};

struct CompressionSpec {
  int method = 3;     // hmat_compress_t
  double eps = 1e-4;
};

// Entries from the caller's host functions (hmat_assemble_context_t,
// assembly.cpp:75-292): one of simple, block or advanced is set.
struct HostFunctions {
  hmat_interaction_func_t simple = nullptr;
  hmat_compute_func_t block = nullptr;
  void (*advanced)(struct hmat_block_compute_context_t*) = nullptr;
  hmat_prepare_func_t prepare = nullptr;
  void* user_context = nullptr;
};

struct KernelSpec {
  int kind = kExponential;
  double length = 1.0;
  double nugget = 0.0;
  double shift = 0.0;
  double wavenumber = 0.0;  // kHelmholtz
  double reg = 0.0;         // kHelmholtz
  int dim = 3;
  int n = 0;
  std::vector<double> points;  // user numbering
};

// H-gemv plan: item lists with offsets per right-hand side (built once after
// assembly), plus scratch buffers sized for the largest nrhs seen so far.
struct GemvPlan {
  bool built = false;
  int nrhs = 0;       // scratch buffers below hold this many right-hand sides
  DevBuf zitems, zred, yitems, contrib, slotlist;
  int nz = 0, nzred = 0, ny = 0;
  int max_rows = 0;   // rows of the largest destination leaf cluster
  int64_t zpart1 = 0, z1 = 0, slot1 = 0;  // scratch doubles per right-hand side
  DevBuf zpart, z, slots, acc;
};

struct Stats {
  double t_assemble = 0, t_factor = 0, t_solve = 0;
  long tasks = 0, waves = 0, launches = 0, segments = 0;
  size_t temp_bytes = 0;
};

class Matrix {
 public:
  Matrix(const ClusterTree* rows, const ClusterTree* cols, bool lower_sym,
         const AdmissibilityParams& adm, int value_type, int rank_cap);
  // A matrix over a given partition (read from a serialized structure);
  // cap_hint[l] raises the Rk capacity of leaf l (the rank it will hold).
  Matrix(const ClusterTree* rows, const ClusterTree* cols, bool lower_sym,
         std::unique_ptr<BlockTree> tree, int value_type, int rank_cap,
         const std::vector<int>& cap_hint);
  AdmissibilityParams adm;
  ~Matrix();

  // A complex (Z, or C with vt = HMAT_SIMPLE_PRECISION) matrix over the
  // caller's trees r x c: the engine stores the 2 x 2 real representation
  // [[Re, -Im], [Im, Re]] of every entry over the embedded trees
  // (embed_cluster_tree), so that every kernel runs in real arithmetic of the
  // type vt and a complex vector (interleaved re, im) is a real vector of the
  // embedded numbering.  er / ec: embedded trees to share (clones).
  static Matrix* create_complex(const ClusterTree* r, const ClusterTree* c,
                                const AdmissibilityParams& adm, int rank_cap,
                                std::shared_ptr<ClusterTree> er = nullptr,
                                std::shared_ptr<ClusterTree> ec = nullptr,
                                int vt = HMAT_DOUBLE_PRECISION);
  bool embedded = false;                    // a complex matrix (see create_complex)
  const ClusterTree* zrows = nullptr;       // the caller's complex trees
  const ClusterTree* zcols = nullptr;
  std::shared_ptr<ClusterTree> erows, ecols;  // rows / cols point to these

  int value_type;        // hmat_value_t
  const ClusterTree* rows;
  const ClusterTree* cols;
  bool own_rows = false, own_cols = false;
  bool lower_sym;
  double lowrank_eps = 1e-4;
  int rk_cap_max = 160;   // capacity limit of Rk panels
  std::unique_ptr<BlockTree> bt;
  hipStream_t stream = nullptr;       // stream every launch goes to
  hipStream_t own_stream = nullptr;   // the engine's own stream
  bool user_stream = false;
  // every launch on `stream` alone, no side streams or cross-stream events:
  // set for matrices driven from several host threads at once, whose streams
  // share the few hardware queues (a wait queued ahead of the event it waits
  // for, in another thread's stream on the same queue, would never end)
  bool single_stream = false;
  // upper bound on the factorisation's temporary memory (0: half the free
  // device memory, at most 96 GB); matrices factorised side by side share it
  size_t temp_budget_limit = 0;
  // frees the factorisation's temporaries and the recorded solve plans that
  // use them (the factor itself stays)
  void release_temporaries();

  Store<double> s64;
  Store<float> s32;
  bool assembled = false;
  int factorization = -1;    // hmat_factorization_t
  // progress of a factorisation (hmat_factorization_context_t.progress,
  // hmat_interface_t.set_progressbar): max = N, current = the end row of the
  // last diagonal leaf factorised, updated as task segments complete; of an
  // assembly (hmat_assemble_context_t.progress): max = the number of leaves,
  // current = the leaves assembled, updated between launches
  hmat_progress_t* progress = nullptr;
  void report_assembly(int done, int total);
  bool factor32 = false;      // factor stored in s32 (mixed precision)
  DevBuf pivots;              // LU: getrf interchanges of each diagonal leaf, at its rows
  DevBuf ldlt_diag;           // LDLt: D at the matrix's rows, in the factor's precision
  // HODLR (symmetric) factor: per 2 x 2 node, the r x r blocks X11 and KK of
  // the Woodbury form (hodlr.cpp:163-285) in the factor's precision, and the
  // log of det(M11) (FP64)
  // Non-symmetric HODLR factor (hodlr.cpp:45-116): per node the 2r x 2r block
  // KK = I + [[0, b01^T a10], [b10^T a01, 0]] (LU-factorised, interchanges in
  // hodlr_piv) and the panels a01, b01 of the transposed copy of the Rk block.
  struct HodlrNode {
    const Block* b;   // the 2 x 2 node; b->get(1, 0) is its Rk block
    int cap;          // capacity of that block's panels
    int64_t x11, kk;  // offsets (elements) into hodlr_buf (symmetric: r x r each;
                      // non-symmetric: no X11, KK of 2 cap x 2 cap)
    int64_t a01 = -1, b01 = -1;  // non-symmetric: n0 x cap and n1 x cap panels
    int64_t piv = -1;            // non-symmetric: 2 cap interchanges in hodlr_piv
  };
  int64_t hodlr_elems = 0, hodlr_pivs = 0;
  DevBuf hodlr_piv;
  std::vector<HodlrNode> hodlr_nodes;
  // dense inverses of the LLt / LDLt factor's diagonal blocks of at most
  // kInvMaxRows rows (built at the first solve, in the factor's precision):
  // the solves multiply by them instead of substituting leaf by leaf
  struct InvBlock {
    const Block* b;
    int64_t off;  // elements into inv_buf; m x m, ld m
  };
  std::vector<InvBlock> inv_blocks;
  int giant_rows = 8192;      // embedded rows and columns of a block compressed on the whole device
  int inv_max_rows = 2048;   // the largest diagonal blocks applied as inverses
  // times the solve (nrhs right-hand sides) with each candidate inv_max_rows
  // and keeps the fastest; returns its milliseconds per solve
  double autotune_solve(int nrhs, int* best_rows);
  void drop_solve_plans();   // recorded solve plans, their graphs and the inverses
  std::map<const Block*, int> inv_index;
  DevBuf inv_buf;
  std::map<const Block*, int> hodlr_index;
  DevBuf hodlr_buf, hodlr_logdet;
  int overflow = 0;
  Stats stats;

  // points of the kernel, hmat order (rows tree), SoA, plus user indices
  DevBuf pts_rows, user_rows;
  DevBuf pts_cols, user_cols;

  void device_init();   // stream + device scratch, created on first GPU use
  void assemble(const KernelSpec& k, const CompressionSpec& c);
  // Assembly from host functions: every leaf block is evaluated on the host
  // (staged in batches), then compressed on the device exactly as the
  // built-in kernels' blocks are.
  void assemble_host(const HostFunctions& f, const CompressionSpec& c);
  // y = alpha * A * x + beta * y, device pointers in hmat numbering, ld = N.
  template <typename T>
  void gemv_device(double alpha, const T* x, double beta, T* y, int nrhs, int ldx, int ldy,
                   const Store<T>& st);
  // The same product with every temporary in caller-provided device scratch
  // (scratch_elems doubles); allocates nothing.  x has ld = N cols, y ld = N rows.
  template <typename T>
  void gemv_device_scratch(double alpha, const T* x, double beta, T* y, int nrhs,
                           const Store<T>& st, double* scratch, size_t scratch_elems);
  size_t gemv_scratch_elems(int nrhs) const;
  void gemv_host(char trans, const void* alpha, void* x, const void* beta, void* y, int nrhs,
                 bool permute);

  // factorisation / solve (factor.cpp)
  void factorize(int algo, bool fp32);
  // part 0: the whole solve; 1: the lower-triangular sweep only (L x = b);
  // 2: the upper one only (L^T x = b) -- LLt and LDLt factors
  template <typename T>
  void solve_device(T* b, int nrhs, int ldb, const Store<T>& st, int part = 0);
  void solve_host(void* b, int nrhs, bool permute, int part = 0);
  // y <- op(F) x for a factorised matrix: F = L of an LLt factor, L W of a
  // symmetric HODLR factor (device pointers, hmat numbering, ld = N; y is
  // overwritten)
  template <typename T>
  void factor_gemv_device(char trans, const T* x, T* y, int nrhs, const Store<T>& st);
  double logdet();
  // Iterative refinement (FP32 factor, FP64 residual with the FP64 matrix).
  int solve_refined(double* b_dev, int nrhs, int ldb, double tol, int max_iter, double* resid_out);

  void info(hmat_info_t& out) const;

  // whole-matrix operations (ops.cpp; truncate_all / axpy_from in factor.cpp)
  Matrix* clone(bool structure_only);       // same trees and partition; null leaves if asked
  void transpose();                         // A <- A^T (unfactorised, one cluster tree)
  double norm();                            // Frobenius norm of the whole matrix
  void scale(double alpha);                 // A <- alpha A
  void add_identity(double alpha);          // A <- A + alpha I
  void extract_diagonal(double* diag);      // diagonal in hmat numbering (n_rows doubles)
  // entries (rpos[i], cpos[j]) in hmat numbering, mirrored for lower-symmetric
  // storage (a factor's stored blocks as they are), into out[i + j ld]
  void get_entries(const std::vector<int>& rpos, const std::vector<int>& cpos, double* out, int ld);
  template <typename T>
  void get_entries_t(const Store<T>& st, const std::vector<int>& rpos, const std::vector<int>& cpos,
                     double* out, int ld);
  void truncate_all();                      // recompress every Rk leaf at lowrank_eps
  void axpy_from(double alpha, Matrix& x);  // A <- A + alpha X, X of the same partition
  // this <- beta this + alpha op(a) op(b) (H-gemm), a and b over the same trees
  void gemm_from(char ta, char tb, double alpha, Matrix& a, Matrix& b, double beta);
  void hodlr_layout(bool sym);    // node list of a HODLR factor; checks the structure
  void hodlr_alloc(size_t elem);  // its r x r blocks and log-determinants
  // b <- this^{-1} b for a factorised matrix and an H-matrix b (solve_mat);
  // part 1 / 2: L^{-1} b / L^{-T} b only (LLt)
  void solve_mat(Matrix& b, int part = 0);

  int n_rows() const { return rows->n; }
  int n_cols() const { return cols->n; }

  template <typename T>
  Store<T>& store();

 private:
  void build_storage(const std::vector<int>* cap_hint = nullptr);
  void assemble_pass(const KernelSpec& k, const CompressionSpec& c);
  void assemble_host_pass(const HostFunctions& f, const CompressionSpec& c);
  bool assembly_overflow();
  void grow_rank_capacity();
  void svd_compress_dense(const std::vector<std::pair<int, int64_t>>& blocks, const double* dense,
                          double eps);
  // very large blocks of a complex matrix (at least giant_rows embedded rows
  // and columns): dense evaluation and a randomized SVD on the whole device
  // (giant.cpp); the panels are kept for the passes of one assembly, and a
  // rank above the leaf's capacity raises that leaf's capacity
  struct GiantPanels {
    DevBuf a, b;
    int rank = 0;
  };
  std::map<int, std::unique_ptr<GiantPanels>> giant_;
  std::vector<int> cap_hint_;   // per-leaf capacity floors (empty: none)
  void compress_giants(const std::vector<int>& list, PointSet rp, PointSet cp, KernelParams kp,
                       double eps);
  void finish_assembly(double t0);
  template <typename T>
  void init_store(Store<T>& st);
  void build_gemv_plan();
  template <typename T>
  void gemv_launch(double alpha, const T* x, double beta, T* y, int nrhs, int ldx, int ldy,
                   const Store<T>& st, double* zpart, double* z, double* slots, double* acc);
  GemvPlan plan_;
  DevBuf overflow_;
  std::shared_ptr<struct FactorState> fstate_;
  friend struct FactorState;
  friend struct Serializer;
};

// Serialization in the hmat-oss stream format (serialize.cpp).
void serialize_write_struct(Matrix& m, hmat_iostream f, void* user_data);
void serialize_write_data(Matrix& m, hmat_iostream f, void* user_data);
Matrix* serialize_read_struct(hmat_iostream f, void* user_data, int value_type, int rank_cap);
void serialize_read_data(Matrix& m, hmat_iostream f, void* user_data);

}  // namespace hmx
