// H-LLt factorisation and H-triangular solves on the GPU.
//
// The host walks the same recursion as hmat-oss (recursiveLltDecomposition
// recursion.cpp:349-390, recursiveSolveUpperTriangularRight recursion.cpp:86-134,
// HMatrix::gemm / recursiveGemm / leafGemm h_matrix.cpp:1375-1620,
// RkMatrix::gemmRk rk_matrix.cpp:1352-1415, HMatrix::axpy h_matrix.cpp:976-1073,
// the solves h_matrix.cpp:2297-2540) but executes nothing: every leaf-level
// operation becomes a task record.  Data-dependent quantities (Rk ranks) stay
// on the device; the host only knows capacities.  Tasks get a dependency wave
// (as soon as possible in program order, tracked per leaf buffer and per
// leaf-cluster row range of temporaries) and a segment of tasks runs as one
// batched kernel launch per (wave, task type).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <stdexcept>
#include <deque>
#include <functional>
#include <exception>
#include <thread>
#include <unordered_map>

#include "engine.hpp"
#include "factor_kernels.hpp"

// Built with -DHMX_DIAGNOSTICS=1, HMX_TRACE_LAUNCHES=<seconds> times every
// batched launch of the factorisation on the host (off in the library).
#ifndef HMX_DIAGNOSTICS
#define HMX_DIAGNOSTICS 0
#endif

namespace hmx {

namespace {

double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

constexpr int64_t kTempBase = int64_t(1) << 40;
// the LDLt diagonal (one vector at the matrix's rows), tracked per leaf cluster
constexpr int64_t kDiagBuf = kTempBase + (int64_t(1) << 31);
// leaf numbers from here on name scaled copies held by the executor (LDLt)
constexpr int kShadowLeaf = 1 << 28;
// the r x r blocks of the HODLR nodes (two per node)
constexpr int64_t kHodlrBase = int64_t(1) << 38;
// the dense inverses of the factor's diagonal blocks, one id per block
constexpr int64_t kInvBase = kTempBase + (int64_t(1) << 32);
// largest diagonal block of a factor that the solves apply through its
// dense inverse (at N = 200k, leaf 128: blocks of 1,562 rows, 128 per sweep
// instead of 2,048 leaf substitutions; measured against 1024: 0.1 s less
// per FP32 + refinement step)
constexpr int kInvMaxRows = 2048;  // default of Matrix::inv_max_rows

// Dense view; n may be dynamic (rank slot dn).
struct DV {
  void* p = nullptr;
  int ld = 0;
  int m = 0, n = 0;
  const int* dn = nullptr;
  int64_t buf = -1;
  int row0 = -1;  // absolute row of the first row (temporaries only)
  bool trans = false;  // logical view is the transpose of the storage
};

// Rk view: A (m x cap, ld la), B (n x cap, ld lb), rank slot.
struct RV {
  void* a = nullptr;
  void* b = nullptr;
  int la = 0, lb = 0;
  int m = 0, n = 0;
  int cap = 0;
  int* rank = nullptr;
  int64_t bufa = -1, bufb = -1;
  int row0 = -1, col0 = -1;
};

RV swap_rv(const RV& r) {
  RV s = r;
  std::swap(s.a, s.b);
  std::swap(s.la, s.lb);
  std::swap(s.m, s.n);
  std::swap(s.bufa, s.bufb);
  std::swap(s.row0, s.col0);
  return s;
}

struct Unit {
  int lw = 0;  // wave of the last writer
  int lr = 0;  // latest wave of a reader since
};

}  // namespace

struct SegLaunch {
  int type, first, count, tfirst, tcount;
  int wave;      // dependency wave: launches of one wave are independent
  bool overlap;  // may run beside the other launches of its wave (its workspace does not overlap theirs)
};

// One executed segment: its tasks and GEMM tiles in device memory and the
// ordered list of batched launches.  Solves keep their plans for replay.
struct SegmentPlan {
  DevBuf tasks, tiles;
  std::vector<SegLaunch> launches;
  int64_t ws = 0;
  int ranks_used = 0;
  long ntasks = 0;
  int waves = 0;
};
using PlanList = std::vector<std::shared_ptr<SegmentPlan>>;

struct FactorState {
  // temporaries of the current segment: chunks that grow on demand up to
  // 'temp_budget' bytes and are reused by the next segment
  std::vector<std::unique_ptr<DevBuf>> chunks;
  size_t chunk = 0, chunk_used = 0, temp_total = 0, temp_budget = 0, temp_used = 0;
  DevBuf ranks, ws, flags;
  int rank_cap = 0, rank_used = 0;
  int64_t next_temp = 0;
  // cached solve plans, keyed by (precision, nrhs), and their right-hand side
  std::map<std::pair<int, int>, PlanList> solve_plans;
  // replays of a recorded solve plan captured as one HIP graph; the captured
  // kernel arguments hold the buffer addresses below, so a graph is rebuilt
  // when any of them has moved
  struct SolveGraph {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    const void *ws = nullptr, *rhs = nullptr, *ranks = nullptr, *flags = nullptr;
    void release() {
      if (exec) (void)hipGraphExecDestroy(exec);
      if (graph) (void)hipGraphDestroy(graph);
      exec = nullptr;
      graph = nullptr;
    }
  };
  std::map<std::pair<int, int>, SolveGraph> solve_graphs;
  // right-hand sides of each recorded solve plan: the plan's tasks hold this
  // buffer's address, so every (precision, nrhs) keeps its own
  std::map<std::pair<int, int>, DevBuf> solve_rhs;
  DevBuf ir_x, ir_r, ir_norms, ir_p, ir_q, ir_z, ir_dots;
  void drop_graphs() {
    for (auto& g : solve_graphs) g.second.release();
    solve_graphs.clear();
  }
  // side streams (default priority) and events for running the independent
  // launches of one wave side by side
  static constexpr int kSideStreams = 2;
  hipStream_t side[kSideStreams] = {};
  hipEvent_t fork = nullptr, join[kSideStreams] = {};
  void ensure_side() {
    if (fork) return;
    for (int k = 0; k < kSideStreams; ++k) {
      HMX_CHECK(hipStreamCreateWithFlags(&side[k], hipStreamNonBlocking));
      HMX_CHECK(hipEventCreateWithFlags(&join[k], hipEventDisableTiming));
    }
    HMX_CHECK(hipEventCreateWithFlags(&fork, hipEventDisableTiming));
  }
  ~FactorState() {
    drop_graphs();
    if (fork) {
      (void)hipEventDestroy(fork);
      for (int k = 0; k < kSideStreams; ++k) {
        (void)hipEventDestroy(join[k]);
        (void)hipStreamDestroy(side[k]);
      }
    }
  }
};

namespace {
#if HMX_DIAGNOSTICS
// launch time per task type (diagnostics build only)
double g_type_seconds[32];
long g_type_launches[32];
#endif

template <typename T>
void execute_plan(const SegmentPlan& P, FactorState& fs, hipStream_t s, bool capturing = false,
                  bool single_stream = false) {
  const size_t ws_bytes = (size_t)std::max<int64_t>(P.ws, 1) * sizeof(double);
  // segments run asynchronously: a workspace still used by the previous one
  // is only replaced once the stream has drained
  if (ws_bytes > fs.ws.bytes) {
    // grown with headroom (at least twice the old size), so that the stream
    // drains for a reallocation a few times per factorisation, not per segment
    if (fs.ws.p && !capturing) HMX_CHECK(hipStreamSynchronize(s));
    fs.ws.ensure(capturing ? ws_bytes : std::max(ws_bytes, 2 * fs.ws.bytes));
  }
  // a fill kernel rather than a memset node, so that the plan can be captured
  if (P.ranks_used) launch_fill_int(fs.ranks.as<int>(), (size_t)P.ranks_used, 0, s);
  const FTask* dt = P.tasks.as<FTask>();
  const GemmTile* dtl = P.tiles.as<GemmTile>();
#if HMX_DIAGNOSTICS
  // HMX_TRACE_LAUNCHES=<seconds>: time every launch with events and report
  // the launches slower than that on stderr (not for capture)
  static const double trace_env = getenv("HMX_TRACE_LAUNCHES") ? atof(getenv("HMX_TRACE_LAUNCHES")) : -1.0;
  const double trace = capturing ? -1.0 : trace_env;
  std::vector<hipEvent_t> ev;
  if (trace >= 0) {
    ev.resize(P.launches.size() + 1);
    for (auto& e : ev) HMX_CHECK(hipEventCreate(&e));
    HMX_CHECK(hipEventRecord(ev[0], s));
  }
#else
  constexpr double trace = -1.0;
#endif
  const bool side_by_side = !capturing && trace < 0 && !single_stream;
  if (side_by_side) fs.ensure_side();
  for (size_t q = 0; q < P.launches.size();) {
    // the launches of one wave that may overlap go round-robin over the
    // matrix's stream and the side streams, joined before the next wave
    size_t e = q + 1;
    if (side_by_side && P.launches[q].overlap)
      while (e < P.launches.size() && P.launches[e].wave == P.launches[q].wave && P.launches[e].overlap) ++e;
    if (e - q == 1) {
      const SegLaunch& L = P.launches[q];
      launch_task_kernel<T>(L.type, dt + L.first, L.count, dtl + L.tfirst, L.tcount,
                            fs.ws.as<double>(), fs.flags.as<int>(), s);
#if HMX_DIAGNOSTICS
      if (trace >= 0) HMX_CHECK(hipEventRecord(ev[q + 1], s));
#endif
      q = e;
      continue;
    }
    const int used = (int)std::min<size_t>(e - q - 1, FactorState::kSideStreams);
    HMX_CHECK(hipEventRecord(fs.fork, s));
    for (int k = 0; k < used; ++k) HMX_CHECK(hipStreamWaitEvent(fs.side[k], fs.fork, 0));
    for (size_t j = q; j < e; ++j) {
      const SegLaunch& L = P.launches[j];
      const int slot = (int)((j - q) % (used + 1));
      hipStream_t st = slot == 0 ? s : fs.side[slot - 1];
      launch_task_kernel<T>(L.type, dt + L.first, L.count, dtl + L.tfirst, L.tcount,
                            fs.ws.as<double>(), fs.flags.as<int>(), st);
    }
    for (int k = 0; k < used; ++k) {
      HMX_CHECK(hipEventRecord(fs.join[k], fs.side[k]));
      HMX_CHECK(hipStreamWaitEvent(s, fs.join[k], 0));
    }
    q = e;
  }
#if HMX_DIAGNOSTICS
  if (trace >= 0) {
    HMX_CHECK(hipStreamSynchronize(s));
    for (size_t q = 0; q < P.launches.size(); ++q) {
      float ms = 0.f;
      HMX_CHECK(hipEventElapsedTime(&ms, ev[q], ev[q + 1]));
      const SegLaunch& L = P.launches[q];
      if (L.type >= 0 && L.type < 32) {
        g_type_seconds[L.type] += ms * 1e-3;
        ++g_type_launches[L.type];
      }
      if (ms * 1e-3 > trace)
        fprintf(stderr, "launch %zu type %d tasks %d tiles %d: %.4f s\n", q, L.type, L.count, L.tcount, ms * 1e-3);
    }
    for (auto& e : ev) (void)hipEventDestroy(e);
  }
#endif
}
}  // namespace

template <typename T>
class Exec {
 public:
  Exec(Matrix& M, Store<T>& st, FactorState& fs) : M_(M), st_(st), fs_(fs) {
    const auto& RL = M.rows->leaves;
    row2leaf_.assign(M.rows->n, 0);
    for (size_t q = 0; q < RL.size(); ++q)
      for (int i = RL[q]->offset; i < RL[q]->offset + RL[q]->size; ++i) row2leaf_[i] = (int)q;
    eps_ = M.lowrank_eps;
    cap_max_ = M.rk_cap_max;
    if (fs_.temp_budget == 0) {
      size_t freeb = 0, total = 0;
      HMX_CHECK(hipMemGetInfo(&freeb, &total));
      fs_.temp_budget = std::min((size_t)(0.5 * (double)freeb), (size_t)96 << 30);
      if (M.temp_budget_limit) fs_.temp_budget = std::min(fs_.temp_budget, M.temp_budget_limit);
    }
    if (fs_.rank_cap < (1 << 22)) {
      fs_.ranks.alloc(sizeof(int) << 22);
      fs_.rank_cap = 1 << 22;
    }
    fs_.flags.alloc(2 * sizeof(int));
    HMX_CHECK(hipMemsetAsync(fs_.flags.p, 0, 2 * sizeof(int), M.stream));
    reset_temps();
    fs_.rank_used = 0;
  }

  ~Exec() {
    if (submitter_.joinable()) submitter_.join();
    if (!inflight_.empty()) (void)hipStreamSynchronize(M_.stream);
  }

  // Waits until the launches of the segment handed to the submission thread
  // are all queued (the queue's depth holds the host back until the device has
  // run most of them); rethrows its error.
  void join_submitter() {
    if (!submitter_.joinable()) return;
    submitter_.join();
    if (submit_err_) {
      std::exception_ptr e = submit_err_;
      submit_err_ = nullptr;
      std::rethrow_exception(e);
    }
  }

  void reset_temps() {
    fs_.chunk = 0;
    fs_.chunk_used = 0;
    fs_.temp_used = 0;
  }

  // ---- operands from the tree ---------------------------------------------
  DV full_of(const Block* b) const {
    if (b->leaf >= kShadowLeaf) return shadow_.at(b->leaf - kShadowLeaf).d;
    const LeafDesc& L = st_.hdesc[b->leaf];
    DV d;
    d.p = st_.data() + L.a;
    d.ld = L.m;
    d.m = L.m;
    d.n = L.n;
    d.buf = b->leaf;
    return d;
  }
  RV rk_of(const Block* b) const {
    if (b->leaf >= kShadowLeaf) return shadow_.at(b->leaf - kShadowLeaf).r;
    const LeafDesc& L = st_.hdesc[b->leaf];
    RV r;
    r.a = st_.data() + L.a;
    r.b = st_.data() + L.b;
    r.la = L.m;
    r.lb = L.n;
    r.m = L.m;
    r.n = L.n;
    r.cap = L.cap;
    r.rank = st_.ranks.template as<int>() + b->leaf;
    r.bufa = r.bufb = b->leaf;
    r.row0 = b->rows->offset;
    r.col0 = b->cols->offset;
    return r;
  }

  // ---- temporaries ------------------------------------------------------------
  void* talloc(size_t elems) {
    const size_t bytes = (elems * sizeof(T) + 255) / 256 * 256;
    for (;;) {
      if (fs_.chunk < fs_.chunks.size()) {
        DevBuf& c = *fs_.chunks[fs_.chunk];
        if (fs_.chunk_used + bytes <= c.bytes) {
          void* p = (char*)c.p + fs_.chunk_used;
          fs_.chunk_used += bytes;
          fs_.temp_used += bytes;
          return p;
        }
        ++fs_.chunk;
        fs_.chunk_used = 0;
        continue;
      }
      // no chunk left: grow, or run what is queued and start over
      const size_t want = std::max(bytes, fs_.chunks.empty() ? ((size_t)256 << 20)
                                                             : fs_.chunks.back()->bytes * 2);
      if (fs_.temp_total + want <= fs_.temp_budget || fs_.chunks.empty()) {
        auto b = std::make_unique<DevBuf>();
        const double ta0 = now_s();
        b->alloc(want);
        t_alloc_ += now_s() - ta0;
        fs_.temp_total += want;
        fs_.chunks.push_back(std::move(b));
        continue;
      }
      flush(true);  // throws if temporaries are live
      if (bytes > fs_.chunks.back()->bytes) {
        auto b = std::make_unique<DevBuf>();
        const double ta1 = now_s();
        b->alloc(bytes);
        t_alloc_ += now_s() - ta1;
        fs_.temp_total += bytes;
        fs_.chunks.push_back(std::move(b));
      }
    }
  }
  int* ralloc() {
    if (fs_.rank_used >= fs_.rank_cap) flush(true);
    return fs_.ranks.as<int>() + fs_.rank_used++;
  }
  DV temp_dense(int m, int n, int row0) {
    DV d;
    d.p = talloc((size_t)m * n);
    d.ld = m;
    d.m = m;
    d.n = n;
    d.buf = kTempBase + fs_.next_temp++;
    d.row0 = row0;
    ++live_;
    return d;
  }
  RV temp_rk(int m, int n, int cap, int row0, int col0) {
    RV r;
    r.a = talloc((size_t)m * cap);
    r.b = talloc((size_t)n * cap);
    r.la = m;
    r.lb = n;
    r.m = m;
    r.n = n;
    r.cap = cap;
    r.rank = ralloc();
    r.bufa = kTempBase + fs_.next_temp++;
    r.bufb = kTempBase + fs_.next_temp++;
    r.row0 = row0;
    r.col0 = col0;
    ++live_;
    return r;
  }
  void release() { --live_; }

  // ---- dependency keys --------------------------------------------------------
  void keys(int64_t buf, int row0, int m, std::vector<uint64_t>& out) const {
    if (buf < kTempBase || row0 < 0) {
      out.push_back((uint64_t)buf << 20 | 0xFFFFF);
      return;
    }
    if (m <= 0) return;
    if (row0 + m > (int)row2leaf_.size())
      throw std::runtime_error("dependency keys: rows [" + std::to_string(row0) + ", " +
                               std::to_string(row0 + m) + ") outside the matrix");
    const int q0 = row2leaf_[row0], q1 = row2leaf_[row0 + m - 1];
    for (int q = q0; q <= q1; ++q) out.push_back((uint64_t)buf << 20 | (uint64_t)q);
  }
  // row0 numbers the storage rows of a temporary (a transposed view selects
  // storage columns and keeps row0), so the keys always span the storage rows
  void rd(const DV& d) { keys(d.buf, d.row0, d.m, r_); }
  void wr(const DV& d) { keys(d.buf, d.row0, d.m, w_); }
  void rd(const RV& r) {
    keys(r.bufa, r.row0, r.m, r_);
    keys(r.bufb, r.col0, r.n, r_);
  }
  void wr(const RV& r) {
    keys(r.bufa, r.row0, r.m, w_);
    keys(r.bufb, r.col0, r.n, w_);
  }

  void emit(FTask& t, int64_t ws_doubles = 0) {
    int wave = 1;
    for (uint64_t k : r_) {
      auto it = units_.find(k);
      if (it != units_.end()) wave = std::max(wave, it->second.lw + 1);
    }
    for (uint64_t k : w_) {
      auto it = units_.find(k);
      if (it != units_.end()) wave = std::max(wave, std::max(it->second.lw, it->second.lr) + 1);
    }
    for (uint64_t k : r_) {
      Unit& u = units_[k];
      u.lr = std::max(u.lr, wave);
    }
    for (uint64_t k : w_) {
      Unit& u = units_[k];
      u.lw = wave;
      u.lr = wave;
    }
    r_.clear();
    w_.clear();
    t.ws = ws_doubles;
    tasks_.push_back(t);
    waves_.push_back(wave);
    max_wave_ = std::max(max_wave_, wave);
  }

  static Dim dim(int s, const int* d = nullptr) { return Dim{s, d}; }

  // C (+)= alpha op(A) op(B); A, B dense views (op from their trans flag).
  // GEMMs with a long static inner dimension and a small output (the z = B^T X
  // of a large Rk panel) are split along the inner dimension: independent
  // partial products into a temporary, then a fixed-order sum.
  // Small outputs (at most 256 x 64 by capacity: the right-hand sides of a
  // solve, a rank-sized core) split from 512 on, into slices of 256.
  static constexpr int kSplitP = 4096, kChunkP = 2048;
  static constexpr int kSplitSmallP = 512, kChunkSmallP = 256;

  void gemm(const DV& C, bool accumulate, double alpha, const DV& A, const DV& B, Dim m, Dim n,
            Dim p) {
    if (m.s <= 0 || n.s <= 0) return;
    const bool small_out = (int64_t)m.s * n.s <= 256 * 64;
    const int split_p = small_out ? kSplitSmallP : kSplitP, chunk_p = small_out ? kChunkSmallP : kChunkP;
    if (p.d == nullptr && p.s > split_p && (int64_t)m.s * n.s <= 256 * 256) {
      const int kChunkP = chunk_p;
      const int nch = (p.s + kChunkP - 1) / kChunkP;
      DV tmp = temp_dense(m.s, n.s * nch, -1);
      std::vector<DV> slice(nch);
      for (int q = 0; q < nch; ++q) {
        const int p0 = q * kChunkP, pc = std::min(kChunkP, p.s - p0);
        DV As = A, Bs = B;
        As.p = (T*)A.p + (A.trans ? (size_t)p0 : (size_t)p0 * A.ld);
        Bs.p = (T*)B.p + (B.trans ? (size_t)p0 * B.ld : (size_t)p0);
        // each slice is its own dependency unit, so the partial products share a wave
        slice[q] = tmp;
        slice[q].p = (T*)tmp.p + (size_t)q * m.s * n.s;
        slice[q].n = n.s;
        slice[q].buf = kTempBase + fs_.next_temp++;
        gemm_task(slice[q], false, alpha, As, Bs, m, n, dim(pc), A, B, slice[q]);
      }
      FTask t{};
      t.type = T_SUMK;
      t.mode = accumulate ? 1 : 0;
      t.m = m;
      t.n = n;
      t.p = dim(nch);
      t.A = tmp.p;
      t.lda = m.s;
      t.ldb = m.s * n.s;
      t.C = C.p;
      t.ldc = C.ld;
      for (const DV& d : slice) rd(d);
      if (accumulate) rd(C);
      wr(C);
      emit(t);
      release();
      return;
    }
    gemm_task(C, accumulate, alpha, A, B, m, n, p, A, B, C);
  }

  // One GEMM task; rA, rB, wC name the buffers whose dependency keys it uses.
  void gemm_task(const DV& C, bool accumulate, double alpha, const DV& A, const DV& B, Dim m, Dim n,
                 Dim p, const DV& rA, const DV& rB, const DV& wC) {
    FTask t{};
    t.type = T_GEMM;
    t.mode = accumulate ? 1 : 0;
    t.ta = A.trans;
    t.tb = B.trans;
    t.m = m;
    t.n = n;
    t.p = p;
    t.alpha = alpha;
    t.A = A.p;
    t.B = B.p;
    t.C = C.p;
    t.lda = A.ld;
    t.ldb = B.ld;
    t.ldc = C.ld;
    rd(rA);
    rd(rB);
    if (accumulate) rd(wC);
    wr(wC);
    emit(t);
  }

  void zero(const DV& C) {
    FTask t{};
    t.type = T_ZERO;
    t.m = dim(C.m);
    t.n = dim(C.n);
    t.C = C.p;
    t.ldc = C.ld;
    wr(C);
    emit(t);
  }

  // ---- leaf operations ---------------------------------------------------------
  void potrf(const Block* b) {
    prog_row_ = std::max(prog_row_, b->rows->offset + b->m());
    DV d = full_of(b);
    FTask t{};
    t.type = T_POTRF;
    t.m = dim(d.m);
    t.A = d.p;
    t.lda = d.ld;
    rd(d);
    wr(d);
    emit(t);
  }

  // HMatrix::luDecomposition at a leaf (full_matrix.cpp:230-237): getrf in
  // place, row interchanges into the matrix's pivot array at the leaf's rows.
  void getrf(const Block* b) {
    prog_row_ = std::max(prog_row_, b->rows->offset + b->m());
    DV d = full_of(b);
    FTask t{};
    t.type = T_GETRF;
    t.m = dim(d.m);
    t.A = d.p;
    t.lda = d.ld;
    t.B = piv_of(b);
    rd(d);
    wr(d);
    emit(t);
  }

  int* piv_of(const Block* b) const {
    if (!M_.pivots.p) throw std::runtime_error("LU pivots not allocated");
    return M_.pivots.as<int>() + b->rows->offset;
  }

  // mode 0: X <- X L^{-T}; mode 1: X <- L^{-1} X; mode 2: X <- L^{-T} X, with
  // the kTrsm* flags of ftask.hpp for the LU leaves.
  // 'other' is the free dimension of X (rows for mode 0, columns otherwise).
  void trsm(int mode, const Block* Lb, const DV& X, Dim other) {
    DV L = full_of(Lb);
    FTask t{};
    t.type = T_TRSM;
    t.mode = mode;
    t.n = dim(L.m);
    t.m = other;
    t.A = L.p;
    t.lda = L.ld;
    if (mode & kTrsmPiv) t.B = piv_of(Lb);
    t.C = X.p;
    t.ldc = X.ld;
    rd(L);
    rd(X);
    wr(X);
    emit(t);
  }

  // Rk target += sum alpha_i parts_i (parts at row/col offsets), truncate.
  void fadd(const RV& tgt, const std::vector<std::pair<double, RV>>& parts) {
    size_t q = 0;
    while (q < parts.size()) {
      FTask t{};
      t.type = T_FADD;
      t.eps = eps_;
      t.tgt = ref(tgt);
      int kmax = tgt.cap;
      rd(tgt);
      int np = 0;
      for (; q < parts.size() && np < 4; ++q, ++np) {
        const RV& p = parts[q].second;
        Part& P = t.parts[np];
        P.rk = ref(p);
        P.roff = p.row0 - tgt.row0;
        P.coff = p.col0 - tgt.col0;
        if (P.roff < 0 || P.coff < 0 || P.roff + p.m > tgt.m || P.coff + p.n > tgt.n)
          throw std::runtime_error("formatted add: part outside its target");
        P.alpha = parts[q].first;
        kmax += p.cap;
        rd(p);
      }
      t.nparts = np;
      wr(tgt);
      emit(t, fadd_workspace(tgt.m, tgt.n, kmax));
    }
  }

  // Rk target += alpha * dense part (at offsets), eval + truncated SVD.
  void fadd_dense(const RV& tgt, double alpha, const DV& part, int roff, int coff, bool replace = false) {
    FTask t{};
    t.type = T_FADD_DENSE;
    t.mode = replace ? 1 : 0;
    t.eps = eps_;
    t.tgt = ref(tgt);
    t.alpha = alpha;
    t.A = part.p;
    t.lda = part.ld;
    t.m = dim(part.m);
    t.n = dim(part.n);
    t.parts[0].roff = roff;
    t.parts[0].coff = coff;
    rd(tgt);
    rd(part);
    wr(tgt);
    emit(t, fadd_dense_workspace(tgt.m, tgt.n));
  }

  static RkRef ref(const RV& r) {
    RkRef x;
    x.a = r.a;
    x.b = r.b;
    x.la = r.la;
    x.lb = r.lb;
    x.m = r.m;
    x.n = r.n;
    x.cap = r.cap;
    x.rank = r.rank;
    return x;
  }

  // ---- H-gemv with a dense multi-column operand: Y += alpha op(h) X ----------
  // X has rows aligned with op(h)'s columns; its column count may be dynamic.
  void hgemv(const Block* h, char tr, double alpha, const DV& X, const DV& Y, Dim ncol) {
    if (!h->is_leaf()) {
      const int ni = tr == 'N' ? h->nrow : h->ncol;
      const int nj = tr == 'N' ? h->ncol : h->nrow;
      for (int i = 0; i < ni; ++i)
        for (int j = 0; j < nj; ++j) {
          char t = tr;
          const Block* ch = child_for_gemm(h, t, i, j);
          if (!ch) continue;
          int co = ch->cols->offset - h->cols->offset, cs = ch->n();
          int ro = ch->rows->offset - h->rows->offset, rs = ch->m();
          if (t != 'N') {
            std::swap(co, ro);
            std::swap(cs, rs);
          }
          hgemv(ch, t, alpha, rows_of(X, co, cs), rows_of(Y, ro, rs), ncol);
        }
      return;
    }
    if (h->kind == Kind::Full) {
      DV M = full_of(h);
      M.trans = tr != 'N';
      const int mo = tr == 'N' ? h->m() : h->n();
      const int po = tr == 'N' ? h->n() : h->m();
      gemm(Y, true, alpha, M, X, dim(mo), ncol, dim(po));
      return;
    }
    RV r = rk_of(h);
    if (tr != 'N') r = swap_rv(r);
    // z = B^T X (k x ncol), Y += alpha A z
    DV z = temp_dense(r.cap, std::max(ncol.s, 1), -1);
    DV Bt = dv_panel(r.b, r.lb, r.n, r.cap, r.bufb, r.col0);
    Bt.trans = true;
    gemm(z, false, 1.0, Bt, X, dim(r.cap, r.rank), ncol, dim(r.n));
    DV Ap = dv_panel(r.a, r.la, r.m, r.cap, r.bufa, r.row0);
    gemm(Y, true, alpha, Ap, z, dim(r.m), ncol, dim(r.cap, r.rank));
    release();
  }

  static DV dv_panel(void* p, int ld, int m, int n, int64_t buf, int row0) {
    DV d;
    d.p = p;
    d.ld = ld;
    d.m = m;
    d.n = n;
    d.buf = buf;
    d.row0 = row0;
    return d;
  }
  DV rows_of(const DV& X, int off, int len) const {
    DV s = X;
    if (X.trans) {
      s.p = (char*)X.p + sizeof(T) * (size_t)off * X.ld;
      s.n = len;
    } else {
      s.p = (char*)X.p + sizeof(T) * (size_t)off;
      s.m = len;
      if (X.row0 >= 0) s.row0 = X.row0 + off;
    }
    return s;
  }

  static const Block* child_for_gemm(const Block* h, char& t, int i, int j) {
    // getChildForGEMM h_matrix.cpp:772-788
    if (t != 'N') std::swap(i, j);
    if (h->sym && j > i) {
      t = t == 'N' ? 'T' : 'N';
      return h->get(j, i);
    }
    return h->get(i, j);
  }

  // ---- products (HMatrix::multiplyRkMatrix h_matrix.cpp:1874-1917) ------------
  static RV op_rk(const RV& r, char t) { return t == 'N' ? r : swap_rv(r); }

  // Returns an Rk spanning op(a).rows x op(b).cols; sets 'owned' if it is a temp
  // (aliased panels of the operands are never released).
  RV multiply_rk(char ta, char tb, const Block* a, const Block* b, int& owned) {
    const Cluster* ar = ta == 'N' ? a->rows : a->cols;
    const Cluster* bc = tb == 'N' ? b->cols : b->rows;
    owned = 0;
    if (a->kind == Kind::Rk && !b->is_leaf()) {
      // multiplyRkH rk_matrix.cpp:1127-1189: newA = A, newB = op(H)^T B
      RV r = op_rk(rk_of(a), ta);
      DV nb = temp_dense(bc->size, r.cap, bc->offset);
      zero(nb);
      DV X = dv_panel(r.b, r.lb, r.n, r.cap, r.bufb, r.col0);
      hgemv(b, tb == 'N' ? 'T' : 'N', 1.0, X, nb, dim(r.cap, r.rank));
      RV out = r;
      out.b = nb.p;
      out.lb = nb.ld;
      out.n = bc->size;
      out.bufb = nb.buf;
      out.row0 = ar->offset;
      out.col0 = bc->offset;
      owned = 1;
      return out;
    }
    if (!a->is_leaf() && b->kind == Kind::Rk) {
      // multiplyHRk rk_matrix.cpp:1192-1243: newA = op(H) A, newB = B
      RV r = op_rk(rk_of(b), tb);
      DV na = temp_dense(ar->size, r.cap, ar->offset);
      zero(na);
      DV X = dv_panel(r.a, r.la, r.m, r.cap, r.bufa, r.row0);
      hgemv(a, ta, 1.0, X, na, dim(r.cap, r.rank));
      RV out = r;
      out.a = na.p;
      out.la = na.ld;
      out.m = ar->size;
      out.bufa = na.buf;
      out.row0 = ar->offset;
      out.col0 = bc->offset;
      owned = 1;
      return out;
    }
    if (a->kind == Kind::Rk && b->kind == Kind::Rk) {
      // multiplyRkRk rk_matrix.cpp:1245-1333
      RV r1 = op_rk(rk_of(a), ta), r2 = op_rk(rk_of(b), tb);
      RV out = temp_rk(ar->size, bc->size, std::min(r1.cap, r2.cap), ar->offset, bc->offset);
      FTask t{};
      t.type = T_RKRK;
      t.eps = eps_;
      t.parts[0].rk = ref(r1);
      t.parts[1].rk = ref(r2);
      t.tgt = ref(out);
      rd(r1);
      rd(r2);
      wr(out);
      emit(t, rkrk_workspace(r1.cap, r2.cap));
      owned = 2;
      return out;
    }
    if (a->kind == Kind::Rk && b->kind == Kind::Full) {
      // multiplyRkFull rk_matrix.cpp:1013-1077: newA = A, newB = op(M)^T B
      RV r = op_rk(rk_of(a), ta);
      DV M = full_of(b);
      M.trans = (tb == 'N');  // op(M)^T
      DV nb = temp_dense(bc->size, r.cap, bc->offset);
      DV X = dv_panel(r.b, r.lb, r.n, r.cap, r.bufb, r.col0);
      gemm(nb, false, 1.0, M, X, dim(bc->size), dim(r.cap, r.rank), dim(r.n));
      RV out = r;
      out.b = nb.p;
      out.lb = nb.ld;
      out.n = bc->size;
      out.bufb = nb.buf;
      out.row0 = ar->offset;
      out.col0 = bc->offset;
      owned = 1;
      return out;
    }
    if (a->kind == Kind::Full && b->kind == Kind::Rk) {
      // multiplyFullRk rk_matrix.cpp:1080-1124: newA = op(M) A, newB = B
      RV r = op_rk(rk_of(b), tb);
      DV M = full_of(a);
      M.trans = (ta != 'N');
      DV na = temp_dense(ar->size, r.cap, ar->offset);
      DV X = dv_panel(r.a, r.la, r.m, r.cap, r.bufa, r.row0);
      gemm(na, false, 1.0, M, X, dim(ar->size), dim(r.cap, r.rank), dim(r.m));
      RV out = r;
      out.a = na.p;
      out.la = na.ld;
      out.m = ar->size;
      out.bufa = na.buf;
      out.row0 = ar->offset;
      out.col0 = bc->offset;
      owned = 1;
      return out;
    }
    throw std::runtime_error("multiply_rk: unsupported operand combination");
  }

  // Dense product of leaves / H with a full leaf (multiplyFullMatrix h_matrix.cpp:1920-1950).
  DV multiply_full(char ta, char tb, const Block* a, const Block* b) {
    const Cluster* ar = ta == 'N' ? a->rows : a->cols;
    const Cluster* bc = tb == 'N' ? b->cols : b->rows;
    if (a->kind == Kind::Full && b->kind == Kind::Full) {
      DV out = temp_dense(ar->size, bc->size, ar->offset);
      DV A = full_of(a), B = full_of(b);
      A.trans = ta != 'N';
      B.trans = tb != 'N';
      const int p = ta == 'N' ? a->n() : a->m();
      gemm(out, false, 1.0, A, B, dim(ar->size), dim(bc->size), dim(p));
      return out;
    }
    if (!a->is_leaf() && b->kind == Kind::Full) {
      DV out = temp_dense(ar->size, bc->size, ar->offset);
      zero(out);
      DV X = full_of(b);
      X.trans = tb != 'N';
      hgemv(a, ta, 1.0, X, out, dim(bc->size));
      return out;
    }
    if (a->kind == Kind::Full && !b->is_leaf()) {
      // multiplyFullH h_matrix.cpp:1623-1646: op(M) op(H) = (op(H)^T op(M)^T)^T,
      // the transposed product by H-gemv, then one transposing copy
      DV outT = temp_dense(bc->size, ar->size, bc->offset);
      zero(outT);
      DV X = full_of(a);
      X.trans = (ta == 'N');  // op(M)^T
      hgemv(b, tb == 'N' ? 'T' : 'N', 1.0, X, outT, dim(ar->size));
      DV out = temp_dense(ar->size, bc->size, ar->offset);
      FTask t{};
      t.type = T_AXPY;
      t.mode = 1;
      t.ta = 1;
      t.m = dim(ar->size);
      t.n = dim(bc->size);
      t.alpha = 1.0;
      t.A = outT.p;
      t.lda = outT.ld;
      t.C = out.p;
      t.ldc = out.ld;
      rd(outT);
      wr(out);
      emit(t);
      release();  // outT
      return out;
    }
    throw std::runtime_error("multiply_full: unsupported operand combination");
  }

  // ---- axpy into the tree (HMatrix::axpy h_matrix.cpp:976-1073) ---------------
  // rk spans clusters starting at (rk.row0, rk.col0).
  void axpy_rk(const Block* c, double alpha, const RV& rk) {
    if (!c->is_leaf()) {
      for (int q = 0; q < c->nrow * c->ncol; ++q)
        if (c->child[q]) axpy_rk(c->child[q], alpha, subset(rk, c->child[q]));
      return;
    }
    RV s = subset(rk, c);
    if (c->kind == Kind::Rk) {
      RV tg = rk_of(c);
      tg.row0 = c->rows->offset;
      tg.col0 = c->cols->offset;
      fadd(tg, {{alpha, s}});
    } else {
      DV C = full_of(c);
      DV A = dv_panel(s.a, s.la, s.m, s.cap, s.bufa, s.row0);
      DV B = dv_panel(s.b, s.lb, s.n, s.cap, s.bufb, s.col0);
      B.trans = true;
      gemm(C, true, alpha, A, B, dim(c->m()), dim(c->n()), dim(s.cap, s.rank));
    }
  }

  RV subset(const RV& r, const Block* c) const {
    RV s = r;
    const int ro = c->rows->offset - r.row0, co = c->cols->offset - r.col0;
    if (ro < 0 || co < 0 || ro + c->m() > r.m || co + c->n() > r.n)
      throw std::runtime_error("axpy: block outside the Rk operand");
    s.a = (char*)r.a + sizeof(T) * (size_t)ro;
    s.b = (char*)r.b + sizeof(T) * (size_t)co;
    s.m = c->m();
    s.n = c->n();
    s.row0 = c->rows->offset;
    s.col0 = c->cols->offset;
    return s;
  }

  void axpy_full(const Block* c, double alpha, const DV& f, int frow0, int fcol0) {
    if (!c->is_leaf()) {
      for (int q = 0; q < c->nrow * c->ncol; ++q)
        if (c->child[q]) axpy_full(c->child[q], alpha, f, frow0, fcol0);
      return;
    }
    const int ro = c->rows->offset - frow0, co = c->cols->offset - fcol0;
    DV sub = f;
    sub.p = (char*)f.p + sizeof(T) * ((size_t)ro + (size_t)co * f.ld);
    sub.m = c->m();
    sub.n = c->n();
    sub.row0 = f.row0 >= 0 ? f.row0 + ro : -1;
    if (c->kind == Kind::Rk) {
      RV tg = rk_of(c);
      fadd_dense(tg, alpha, sub, 0, 0);
    } else {
      DV C = full_of(c);
      FTask t{};
      t.type = T_AXPY;
      t.m = dim(c->m());
      t.n = dim(c->n());
      t.alpha = alpha;
      t.A = sub.p;
      t.lda = sub.ld;
      t.C = C.p;
      t.ldc = C.ld;
      rd(sub);
      rd(C);
      wr(C);
      emit(t);
    }
  }

  // ---- dense evaluation of an H block and its write-back ----------------------------
  DV sub_of(const DV& D, int ro, int co, int m, int n) const {
    DV s = D;
    s.p = (char*)D.p + sizeof(T) * ((size_t)ro + (size_t)co * D.ld);
    s.m = m;
    s.n = n;
    if (D.row0 >= 0) s.row0 = D.row0 + ro;
    return s;
  }
  template <typename F>
  static void for_each_leaf(const Block* b, F&& f) {
    if (b->is_leaf()) {
      f(b);
      return;
    }
    for (int q = 0; q < b->nrow * b->ncol; ++q)
      if (b->child[q]) for_each_leaf(b->child[q], f);
  }
  // HMatrix::evalPart: a dense temporary holding every leaf of b (released by the caller)
  DV eval_block(const Block* b) {
    DV D = temp_dense(b->m(), b->n(), b->rows->offset);
    zero(D);
    for_each_leaf(b, [&](const Block* lf) {
      DV sub = sub_of(D, lf->rows->offset - b->rows->offset, lf->cols->offset - b->cols->offset,
                      lf->m(), lf->n());
      if (lf->kind == Kind::Full) {
        DV F = full_of(lf);
        FTask t{};
        t.type = T_AXPY;
        t.mode = 1;
        t.m = dim(lf->m());
        t.n = dim(lf->n());
        t.alpha = 1.0;
        t.A = F.p;
        t.lda = F.ld;
        t.C = sub.p;
        t.ldc = sub.ld;
        rd(F);
        wr(sub);
        emit(t);
      } else {
        RV r = rk_of(lf);
        DV A = dv_panel(r.a, r.la, r.m, r.cap, r.bufa, r.row0);
        DV B = dv_panel(r.b, r.lb, r.n, r.cap, r.bufb, r.col0);
        B.trans = true;
        gemm(sub, true, 1.0, A, B, dim(lf->m()), dim(lf->n()), dim(r.cap, r.rank));
      }
    });
    return D;
  }
  // HMatrix::clear + HMatrix::axpy(1, FullMatrix) h_matrix.cpp:2201-2215, 1034-1073:
  // full leaves take their part of D, Rk leaves the truncated SVD of it
  // (RkMatrix::formattedAddParts with a full part on an empty Rk, rk_matrix.cpp:984-1010).
  void clear_and_set(const Block* b, const DV& D) {
    for_each_leaf(b, [&](const Block* lf) {
      DV sub = sub_of(D, lf->rows->offset - b->rows->offset, lf->cols->offset - b->cols->offset,
                      lf->m(), lf->n());
      if (lf->kind == Kind::Full) {
        DV C = full_of(lf);
        FTask t{};
        t.type = T_AXPY;
        t.mode = 1;
        t.m = dim(lf->m());
        t.n = dim(lf->n());
        t.alpha = 1.0;
        t.A = sub.p;
        t.lda = sub.ld;
        t.C = C.p;
        t.ldc = C.ld;
        rd(sub);
        wr(C);
        emit(t);
      } else {
        fadd_dense(rk_of(lf), 1.0, sub, 0, 0, true);
      }
    });
  }

  // ---- RkMatrix::gemmRk rk_matrix.cpp:1352-1415 -------------------------------
  void gemm_rk(const RV& tgt, char ta, char tb, double alpha, const Block* a, const Block* b) {
    if (!a->is_leaf() && !b->is_leaf()) {
      const int nbr = ta == 'N' ? a->nrow : a->ncol;
      const int nbc = tb == 'N' ? b->ncol : b->nrow;
      const int nbk = ta == 'N' ? a->ncol : a->nrow;
      std::vector<RV> subs(nbr * nbc);
      std::vector<char> made(nbr * nbc, 0);
      for (int i = 0; i < nbr; ++i)
        for (int j = 0; j < nbc; ++j) {
          const int p = i + j * nbr;
          for (int k = 0; k < nbk; ++k) {
            const Block* aik = ta == 'N' ? a->get(i, k) : a->get(k, i);
            const Block* bkj = tb == 'N' ? b->get(k, j) : b->get(j, k);
            if (!aik || !bkj) continue;
            if (!made[p]) {
              const Cluster* sr = ta == 'N' ? aik->rows : aik->cols;
              const Cluster* sc = tb == 'N' ? bkj->cols : bkj->rows;
              subs[p] = temp_rk(sr->size, sc->size, std::min(std::min(sr->size, sc->size), cap_max_),
                                sr->offset, sc->offset);
              made[p] = 1;
            }
            gemm_rk(subs[p], ta, tb, alpha, aik, bkj);
          }
        }
      std::vector<std::pair<double, RV>> parts;
      for (size_t p = 0; p < subs.size(); ++p)
        if (made[p]) parts.push_back({1.0, subs[p]});
      if (!parts.empty()) fadd(tgt, parts);
      for (size_t p = 0; p < subs.size(); ++p)
        if (made[p]) release();
      return;
    }
    if (a->kind == Kind::Rk || b->kind == Kind::Rk) {
      int owned = 0;
      RV r = multiply_rk(ta, tb, a, b, owned);
      fadd(tgt, {{alpha, r}});
      if (owned) release();
      return;
    }
    // full x full (or H x full): dense product, acaFull, add
    DV f = multiply_full(ta, tb, a, b);
    const int m = f.m, n = f.n;
    RV r = temp_rk(m, n, std::min(m, n), f.row0, (tb == 'N' ? b->cols : b->rows)->offset);
    FTask t{};
    t.type = T_ACAFULL;
    t.eps = eps_;
    t.A = f.p;
    t.lda = f.ld;
    t.m = dim(m);
    t.n = dim(n);
    t.tgt = ref(r);
    rd(f);
    wr(f);
    wr(r);
    emit(t);
    release();  // f
    fadd(tgt, {{alpha, r}});
    release();  // r
  }

  // ---- H-gemm (HMatrix::gemm + recursiveGemm + leafGemm) ------------------------
  static bool intersects(const Cluster* x, const Cluster* y) {
    return x->offset < y->offset + y->size && y->offset < x->offset + x->size;
  }

  void hgemm(const Block* c, char ta, char tb, double alpha, const Block* a, const Block* b) {
    if (!c->is_leaf() && !a->is_leaf() && !b->is_leaf()) {
      const int row_a = ta == 'N' ? a->nrow : a->ncol;
      const int col_a = ta == 'N' ? a->ncol : a->nrow;
      const int row_b = tb == 'N' ? b->nrow : b->ncol;
      const int col_b = tb == 'N' ? b->ncol : b->nrow;
      for (int i = 0; i < c->nrow; ++i)
        for (int j = 0; j < c->ncol; ++j) {
          const Block* ch = c->get(i, j);
          if (!ch) continue;
          for (int ia = 0; ia < row_a; ++ia)
            for (int jb = 0; jb < col_b; ++jb)
              for (int k = 0; k < col_a; ++k) {
                char tA = ta;
                const Block* ca = child_for_gemm(a, tA, ia, k);
                if (!ca) continue;
                const Cluster* car = tA == 'N' ? ca->rows : ca->cols;
                const Cluster* cac = tA == 'N' ? ca->cols : ca->rows;
                if (!intersects(car, ch->rows)) continue;
                for (int l = 0; l < row_b; ++l) {
                  char tB = tb;
                  const Block* cb = child_for_gemm(b, tB, l, jb);
                  if (!cb) continue;
                  const Cluster* cbr = tB == 'N' ? cb->rows : cb->cols;
                  const Cluster* cbc = tB == 'N' ? cb->cols : cb->rows;
                  if (!intersects(cbc, ch->cols) || !intersects(cac, cbr)) continue;
                  hgemm(ch, tA, tB, alpha, ca, cb);
                }
              }
          maybe_flush();
        }
      return;
    }
    const Cluster* ar = ta == 'N' ? a->rows : a->cols;
    const Cluster* ac = ta == 'N' ? a->cols : a->rows;
    const Cluster* br = tb == 'N' ? b->rows : b->cols;
    const Cluster* bcl = tb == 'N' ? b->cols : b->rows;
    if (ar != c->rows || bcl != c->cols || ac != br)
      throw std::runtime_error("H-gemm: operands not aligned with the cluster tree");
    leaf_gemm(c, ta, tb, alpha, a, b);
  }

  void leaf_gemm(const Block* c, char ta, char tb, double alpha, const Block* a, const Block* b) {
    if (!c->is_leaf()) {
      if (a->kind == Kind::Rk || b->kind == Kind::Rk) {
        int owned = 0;
        RV r = multiply_rk(ta, tb, a, b, owned);
        axpy_rk(c, alpha, r);
        if (owned) release();
      } else {
        DV f = multiply_full(ta, tb, a, b);
        axpy_full(c, alpha, f, (ta == 'N' ? a->rows : a->cols)->offset,
                  (tb == 'N' ? b->cols : b->rows)->offset);
        release();
      }
      return;
    }
    if (c->kind == Kind::Rk) {
      RV tg = rk_of(c);
      tg.row0 = c->rows->offset;
      tg.col0 = c->cols->offset;
      gemm_rk(tg, ta, tb, alpha, a, b);
      return;
    }
    if (!a->is_leaf() && !b->is_leaf()) {
      full_hh_gemm(c, ta, tb, alpha, a, b);
      return;
    }
    DV C = full_of(c);
    if (a->kind == Kind::Rk || b->kind == Kind::Rk) {
      int owned = 0;
      RV r = multiply_rk(ta, tb, a, b, owned);
      DV A = dv_panel(r.a, r.la, r.m, r.cap, r.bufa, r.row0);
      DV B = dv_panel(r.b, r.lb, r.n, r.cap, r.bufb, r.col0);
      B.trans = true;
      gemm(C, true, alpha, A, B, dim(c->m()), dim(c->n()), dim(r.cap, r.rank));
      if (owned) release();
      return;
    }
    if (a->is_leaf() && b->is_leaf()) {
      DV A = full_of(a), B = full_of(b);
      A.trans = ta != 'N';
      B.trans = tb != 'N';
      const int p = ta == 'N' ? a->n() : a->m();
      gemm(C, true, alpha, A, B, dim(c->m()), dim(c->n()), dim(p));
      return;
    }
    DV f = multiply_full(ta, tb, a, b);
    FTask t{};
    t.type = T_AXPY;
    t.m = dim(c->m());
    t.n = dim(c->n());
    t.alpha = alpha;
    t.A = f.p;
    t.lda = f.ld;
    t.C = C.p;
    t.ldc = C.ld;
    rd(f);
    rd(C);
    wr(C);
    emit(t);
    release();
  }

  // fullHHGemm h_matrix.cpp:1449-1472: a full leaf target and two H operands;
  // recurse on the operands' children (the target's clusters are leaves, so
  // the children keep its rows / columns) until one operand is a leaf.
  void full_hh_gemm(const Block* c, char ta, char tb, double alpha, const Block* a, const Block* b) {
    if (!a->is_leaf() && !b->is_leaf()) {
      const int ni = ta == 'N' ? a->nrow : a->ncol;
      const int nj = tb == 'N' ? b->ncol : b->nrow;
      const int nk = ta == 'N' ? a->ncol : a->nrow;
      for (int i = 0; i < ni; ++i)
        for (int j = 0; j < nj; ++j)
          for (int k = 0; k < nk; ++k) {
            char tA = ta, tB = tb;
            const Block* ca = child_for_gemm(a, tA, i, k);
            const Block* cb = child_for_gemm(b, tB, k, j);
            if (ca && cb) full_hh_gemm(c, tA, tB, alpha, ca, cb);
          }
      return;
    }
    hgemm(c, ta, tb, alpha, a, b);
  }

  // ---- triangular solves -----------------------------------------------------------
  // ---- dense inverses of diagonal blocks (solves) ------------------------------------
  DV inv_dv(int q) const {
    const Matrix::InvBlock& ib = M_.inv_blocks[q];
    const int m = ib.b->m();
    DV d;
    d.p = M_.inv_buf.template as<T>() + ib.off;
    d.ld = std::max(m, 1);
    d.m = m;
    d.n = m;
    d.buf = kInvBase + q;
    d.row0 = ib.b->rows->offset;
    return d;
  }
  // L_B^{-1} of every listed block B: the block's own forward substitution
  // on the identity (blocks are independent, so their tasks share waves)
  void build_inverses() {
    use_inv_ = false;
    for (size_t q = 0; q < M_.inv_blocks.size(); ++q) {
      const Block* b = M_.inv_blocks[q].b;
      DV I = inv_dv((int)q);
      FTask t{};
      t.type = T_ZERO;
      t.mode = 1;
      t.m = dim(I.m);
      t.n = dim(I.n);
      t.C = I.p;
      t.ldc = I.ld;
      wr(I);
      emit(t);
      solve_lower_left(b, I, dim(b->m()));
      maybe_flush();
    }
  }
  void use_inverses(bool v) { use_inv_ = v; }
  // X <- L_B^{-1} X (trans: L_B^{-T} X) through the dense inverse: partial
  // products over 256-column slices of the inverse into a temporary, summed
  // into X (the sum runs after every product has read X)
  void inverse_apply(int q, const DV& X, Dim ncol, bool trans) {
    DV A = inv_dv(q);
    A.trans = trans;
    const int m = A.m, nc = std::max(ncol.s, 1);
    constexpr int kSlice = 256;
    const int nch = (m + kSlice - 1) / kSlice;
    DV tmp = temp_dense(m, nc * nch, -1);
    std::vector<DV> slice(nch);
    for (int k = 0; k < nch; ++k) {
      const int p0 = k * kSlice, pc = std::min(kSlice, m - p0);
      DV As = A;
      As.p = (T*)A.p + (trans ? (size_t)p0 : (size_t)p0 * A.ld);
      DV Xs = rows_of(X, p0, pc);
      slice[k] = tmp;
      slice[k].p = (T*)tmp.p + (size_t)k * m * nc;
      slice[k].n = nc;
      slice[k].buf = kTempBase + fs_.next_temp++;
      gemm_task(slice[k], false, 1.0, As, Xs, dim(m), ncol, dim(pc), A, Xs, slice[k]);
    }
    FTask t{};
    t.type = T_SUMK;
    t.mode = 0;
    t.m = dim(m);
    t.n = ncol;
    t.p = dim(nch);
    t.A = tmp.p;
    t.lda = m;
    t.ldb = m * nc;
    t.C = X.p;
    t.ldc = X.ld;
    for (const DV& d : slice) rd(d);
    rd(X);
    wr(X);
    emit(t);
    release();
  }
  int inv_index(const Block* b) const {
    if (!use_inv_) return -1;
    auto it = M_.inv_index.find(b);
    return it == M_.inv_index.end() ? -1 : it->second;
  }

  // HMatrix::solveLowerTriangularLeft(ScalarArray) h_matrix.cpp:2338-2372
  void solve_lower_left(const Block* L, const DV& X, Dim ncol) {
    const int q = inv_index(L);
    if (q >= 0) {
      inverse_apply(q, X, ncol, false);
      return;
    }
    if (L->is_leaf()) {
      trsm(1, L, X, ncol);
      return;
    }
    int off = 0;
    std::vector<std::pair<int, int>> subs;
    for (int i = 0; i < L->nrow; ++i) {
      const Block* d = L->get(i, i);
      subs.push_back({off, d->n()});
      for (int j = 0; j < i; ++j) {
        const Block* lij = L->get(i, j);
        if (lij)
          hgemv(lij, 'N', -1.0, rows_of(X, subs[j].first, subs[j].second),
                rows_of(X, subs[i].first, subs[i].second), ncol);
      }
      solve_lower_left(d, rows_of(X, off, d->n()), ncol);
      off += d->n();
    }
  }

  // HMatrix::solveUpperTriangularLeft(ScalarArray) h_matrix.cpp:2504-2540 (Uplo LOWER)
  void solve_upper_left(const Block* L, const DV& X, Dim ncol) {
    const int q = inv_index(L);
    if (q >= 0) {
      inverse_apply(q, X, ncol, true);
      return;
    }
    if (L->is_leaf()) {
      trsm(2, L, X, ncol);
      return;
    }
    std::vector<std::pair<int, int>> subs;
    int off = 0;
    for (int i = 0; i < L->nrow; ++i) {
      subs.push_back({off, L->get(i, i)->n()});
      off += L->get(i, i)->n();
    }
    for (int i = L->nrow - 1; i >= 0; --i) {
      solve_upper_left(L->get(i, i), rows_of(X, subs[i].first, subs[i].second), ncol);
      for (int j = 0; j < i; ++j) {
        const Block* lij = L->get(i, j);
        if (lij)
          hgemv(lij, 'T', -1.0, rows_of(X, subs[i].first, subs[i].second),
                rows_of(X, subs[j].first, subs[j].second), ncol);
      }
    }
  }

  // HMatrix::solveUpperTriangularRight(HMatrix) h_matrix.cpp:2380-2415 +
  // recursiveSolveUpperTriangularRight recursion.cpp:86-134 (LLt, Uplo LOWER)
  void solve_upper_right_h(const Block* L, const Block* b) {
    if (!b) return;
    if (!L->is_leaf() && !b->is_leaf()) {
      if (L->nrow == b->ncol) {
        for (int k = 0; k < b->nrow; ++k)
          for (int i = 0; i < L->nrow; ++i) {
            const Block* bki = b->get(k, i);
            if (!bki) continue;
            for (int j = 0; j < i; ++j) {
              const Block* uji = L->get(i, j);
              const Block* bkj = b->get(k, j);
              if (bkj && uji) hgemm(bki, 'N', 'T', -1.0, bkj, uji);
            }
            solve_upper_right_h(L->get(i, i), bki);
          }
      } else if (L->nrow > 1 && b->ncol == 1 && b->nrow > 1) {
        for (int k = 0; k < b->nrow; ++k) solve_upper_right_h(L, b->get(k, 0));
      } else {
        throw std::runtime_error("solveUpperTriangularRight: case not handled");
      }
      return;
    }
    if (!b->is_leaf()) {
      // h_matrix.cpp:2386-2397: evaluate b, solve the dense block, clear b and add it back
      DV D = eval_block(b);
      trsm(0, L, D, dim(D.m));
      clear_and_set(b, D);
      release();
      return;
    }
    if (b->kind == Kind::Full) {
      if (!L->is_leaf()) throw std::runtime_error("solveUpperTriangularRight: full block with an H diagonal");
      DV X = full_of(b);
      trsm(0, L, X, dim(b->m()));
    } else {
      // Xb^T L^T = Bb^T  <=>  L Xb = Bb  (solve on the B panel only)
      RV r = rk_of(b);
      DV P = dv_panel(r.b, r.lb, r.n, r.cap, r.bufb, r.col0);
      solve_lower_left(L, P, dim(r.cap, r.rank));
    }
  }

  // ---- H-LU (non-symmetric storage) ------------------------------------------------
  // HMatrix::solveLowerTriangularLeft(ScalarArray) h_matrix.cpp:2338-2372 with
  // LU, UNIT, LOWER: leaves apply their interchanges then a unit lower solve.
  void lu_solve_lower_left(const Block* L, const DV& X, Dim ncol) {
    if (L->is_leaf()) {
      trsm(1 | kTrsmUnit | kTrsmPiv, L, X, ncol);
      return;
    }
    int off = 0;
    std::vector<std::pair<int, int>> subs;
    for (int i = 0; i < L->nrow; ++i) {
      const Block* d = L->get(i, i);
      subs.push_back({off, d->n()});
      for (int j = 0; j < i; ++j) {
        const Block* lij = L->get(i, j);
        if (lij)
          hgemv(lij, 'N', -1.0, rows_of(X, subs[j].first, subs[j].second),
                rows_of(X, subs[i].first, subs[i].second), ncol);
      }
      lu_solve_lower_left(d, rows_of(X, off, d->n()), ncol);
      off += d->n();
    }
  }

  // HMatrix::solveUpperTriangularLeft(ScalarArray) h_matrix.cpp:2504-2540
  // with Uplo UPPER: X <- U^{-1} X (u_ji = get(j, i), applied 'N').
  void lu_solve_upper_left(const Block* U, const DV& X, Dim ncol) {
    if (U->is_leaf()) {
      trsm(2 | kTrsmTransA, U, X, ncol);
      return;
    }
    std::vector<std::pair<int, int>> subs;
    int off = 0;
    for (int i = 0; i < U->nrow; ++i) {
      subs.push_back({off, U->get(i, i)->n()});
      off += U->get(i, i)->n();
    }
    for (int i = U->nrow - 1; i >= 0; --i) {
      lu_solve_upper_left(U->get(i, i), rows_of(X, subs[i].first, subs[i].second), ncol);
      for (int j = 0; j < i; ++j) {
        const Block* uji = U->get(j, i);
        if (uji)
          hgemv(uji, 'N', -1.0, rows_of(X, subs[i].first, subs[i].second),
                rows_of(X, subs[j].first, subs[j].second), ncol);
      }
    }
  }

  // solveLowerTriangularLeft(ScalarArray) with Uplo UPPER (no interchanges):
  // X <- U^{-T} X (u_ji = get(j, i), applied 'T').
  void lu_solve_upper_t_left(const Block* U, const DV& X, Dim ncol) {
    if (U->is_leaf()) {
      trsm(1 | kTrsmTransA, U, X, ncol);
      return;
    }
    int off = 0;
    std::vector<std::pair<int, int>> subs;
    for (int i = 0; i < U->nrow; ++i) {
      const Block* d = U->get(i, i);
      subs.push_back({off, d->n()});
      for (int j = 0; j < i; ++j) {
        const Block* uji = U->get(j, i);
        if (uji)
          hgemv(uji, 'T', -1.0, rows_of(X, subs[j].first, subs[j].second),
                rows_of(X, subs[i].first, subs[i].second), ncol);
      }
      lu_solve_upper_t_left(d, rows_of(X, off, d->n()), ncol);
      off += d->n();
    }
  }

  // HMatrix::solveLowerTriangularLeft(HMatrix) h_matrix.cpp:2297-2336 +
  // recursiveSolveLowerTriangularLeft recursion.cpp:191-233 (LU, UNIT, LOWER)
  void lu_solve_lower_left_h(const Block* L, const Block* b) {
    if (!b) return;
    if (!L->is_leaf() && !b->is_leaf()) {
      if (L->ncol == b->nrow) {
        for (int k = 0; k < b->ncol; ++k)
          for (int i = 0; i < L->nrow; ++i) {
            const Block* bik = b->get(i, k);
            if (!bik) continue;
            for (int j = 0; j < i; ++j) {
              const Block* lij = L->get(i, j);
              const Block* bjk = b->get(j, k);
              if (lij && bjk) hgemm(bik, 'N', 'N', -1.0, lij, bjk);
            }
            lu_solve_lower_left_h(L->get(i, i), bik);
          }
      } else if (L->ncol > 1 && b->nrow == 1 && b->ncol > 1) {
        for (int k = 0; k < b->ncol; ++k) lu_solve_lower_left_h(L, b->get(0, k));
      } else {
        throw std::runtime_error("solveLowerTriangularLeft: case not handled");
      }
      return;
    }
    if (!b->is_leaf()) {
      if (b->nrow == 1) {
        // h_matrix.cpp:2305-2311: one block row, solve column block by column block
        for (int k = 0; k < b->ncol; ++k) lu_solve_lower_left_h(L, b->get(0, k));
        return;
      }
      // h_matrix.cpp:2313-2320: evaluate b, solve, clear b and add it back
      DV D = eval_block(b);
      lu_solve_lower_left(L, D, dim(D.n));
      clear_and_set(b, D);
      release();
      return;
    }
    if (b->kind == Kind::Full) {
      lu_solve_lower_left(L, full_of(b), dim(b->n()));
    } else {
      RV r = rk_of(b);
      lu_solve_lower_left(L, dv_panel(r.a, r.la, r.m, r.cap, r.bufa, r.row0), dim(r.cap, r.rank));
    }
  }

  // HMatrix::solveUpperTriangularRight(HMatrix) h_matrix.cpp:2380-2415 +
  // recursiveSolveUpperTriangularRight recursion.cpp:86-134 (LU, NONUNIT, UPPER)
  void lu_solve_upper_right_h(const Block* U, const Block* b) {
    if (!b) return;
    if (!U->is_leaf() && !b->is_leaf()) {
      if (U->nrow == b->ncol) {
        for (int k = 0; k < b->nrow; ++k)
          for (int i = 0; i < U->nrow; ++i) {
            const Block* bki = b->get(k, i);
            if (!bki) continue;
            for (int j = 0; j < i; ++j) {
              const Block* uji = U->get(j, i);
              const Block* bkj = b->get(k, j);
              if (bkj && uji) hgemm(bki, 'N', 'N', -1.0, bkj, uji);
            }
            lu_solve_upper_right_h(U->get(i, i), bki);
          }
      } else if (U->nrow > 1 && b->ncol == 1 && b->nrow > 1) {
        for (int k = 0; k < b->nrow; ++k) lu_solve_upper_right_h(U, b->get(k, 0));
      } else {
        throw std::runtime_error("solveUpperTriangularRight: case not handled");
      }
      return;
    }
    if (!b->is_leaf()) {
      // h_matrix.cpp:2386-2397: evaluate b, solve, clear b and add it back
      DV D = eval_block(b);
      trsm(0 | kTrsmTransA, U, D, dim(D.m));
      clear_and_set(b, D);
      release();
      return;
    }
    if (b->kind == Kind::Full) {
      if (!U->is_leaf()) throw std::runtime_error("solveUpperTriangularRight: full block with an H diagonal");
      trsm(0 | kTrsmTransA, U, full_of(b), dim(b->m()));
    } else {
      // Xb^T U = Bb^T  <=>  U^T Xb = Bb (solve on the B panel only)
      RV r = rk_of(b);
      lu_solve_upper_t_left(U, dv_panel(r.b, r.lb, r.n, r.cap, r.bufb, r.col0), dim(r.cap, r.rank));
    }
  }

  // HMatrix::luDecomposition h_matrix.cpp:2567 + recursiveLuDecomposition
  // recursion.cpp:238-287
  void lu(const Block* h) {
    if (h->is_leaf()) {
      if (h->kind != Kind::Full) throw std::runtime_error("lu: Rk diagonal block");
      getrf(h);
      return;
    }
    if (h->nrow != h->ncol) throw std::runtime_error("lu: non-square block grid");
    const int n = h->nrow;
    for (int k = 0; k < n; ++k) {
      const Block* hkk = h->get(k, k);
      if (!hkk) continue;
      lu(hkk);
      for (int i = k + 1; i < n; ++i)
        if (h->get(k, i)) lu_solve_lower_left_h(hkk, h->get(k, i));
      for (int i = k + 1; i < n; ++i)
        if (h->get(i, k)) lu_solve_upper_right_h(hkk, h->get(i, k));
      maybe_flush();
      for (int i = k + 1; i < n; ++i) {
        if (!h->get(i, k)) continue;
        for (int j = k + 1; j < n; ++j)
          if (h->get(i, j) && h->get(k, j)) {
            hgemm(h->get(i, j), 'N', 'N', -1.0, h->get(i, k), h->get(k, j));
            maybe_flush();
          }
      }
    }
  }

  // HMatrix::lltDecomposition h_matrix.cpp:2547-2564 + recursiveLltDecomposition
  void llt(const Block* h) {
    if (h->is_leaf()) {
      if (h->kind != Kind::Full) throw std::runtime_error("llt: Rk diagonal block");
      potrf(h);
      return;
    }
    if (h->nrow != h->ncol) throw std::runtime_error("llt: non-square block grid");
    const int n = h->nrow;
    for (int k = 0; k < n; ++k) {
      llt(h->get(k, k));
      for (int i = k + 1; i < n; ++i)
        if (h->get(i, k)) solve_upper_right_h(h->get(k, k), h->get(i, k));
      maybe_flush();
      for (int i = k + 1; i < n; ++i) {
        if (!h->get(i, k)) continue;
        for (int j = k + 1; j <= i; ++j)
          if (h->get(i, j) && h->get(j, k)) {
            hgemm(h->get(i, j), 'N', 'T', -1.0, h->get(i, k), h->get(j, k));
            maybe_flush();
          }
      }
    }
  }

  // ---- H-LDLt (HMatrix::ldltDecomposition h_matrix.cpp:2747-2769 +
  // recursiveLdltDecomposition recursion.cpp:37-83) ----------------------------------
  // D lives in one vector at the matrix's rows (the reference keeps it per
  // diagonal leaf, full_matrix.cpp:205-216); the leaves hold the unit lower L.
  T* diag_ptr() const {
    if (!M_.ldlt_diag.p) throw std::runtime_error("LDLt diagonal not allocated");
    return M_.ldlt_diag.template as<T>();
  }
  DV diag_of(int row0, int len) const {
    DV d;
    d.p = diag_ptr() + row0;
    d.ld = std::max(len, 1);
    d.m = len;
    d.n = 1;
    d.buf = kDiagBuf;
    d.row0 = row0;
    return d;
  }

  void ldlt_leaf(const Block* b) {
    prog_row_ = std::max(prog_row_, b->rows->offset + b->m());
    DV a = full_of(b);
    DV dg = diag_of(b->rows->offset, b->m());
    FTask t{};
    t.type = T_LDLT;
    t.m = dim(a.m);
    t.A = a.p;
    t.lda = a.ld;
    t.B = dg.p;
    rd(a);
    wr(a);
    wr(dg);
    emit(t);
  }

  // C = A diag(d) (cols) or diag(d) A, or with the inverse; d starts at the
  // matrix row / column d0.  C may be A.
  void dscale(const DV& C, const DV& A, bool cols, bool inverse, int d0, Dim m, Dim n) {
    DV dg = diag_of(d0, cols ? n.s : m.s);
    FTask t{};
    t.type = T_DSCALE;
    t.mode = (cols ? kDscaleCols : 0) | (inverse ? kDscaleInv : 0);
    t.m = m;
    t.n = n;
    t.A = A.p;
    t.lda = A.ld;
    t.B = dg.p;
    t.C = C.p;
    t.ldc = C.ld;
    rd(A);
    rd(dg);
    if (C.p == A.p) rd(C);
    wr(C);
    emit(t);
  }

  // HMatrix::multiplyWithDiag(D, side, inverse) h_matrix.cpp:1953-1998: every
  // leaf times the diagonal at its columns (Side::RIGHT; an Rk leaf: the rows
  // of B) or at its rows (left; an Rk leaf: the rows of A)
  void scale_by_diag(const Block* b, bool inverse, bool left = false) {
    if (!b) return;
    if (!b->is_leaf()) {
      for (Block* c : b->child)
        if (c) scale_by_diag(c, inverse, left);
      return;
    }
    if (b->kind == Kind::Full) {
      DV f = full_of(b);
      dscale(f, f, !left, inverse, left ? b->rows->offset : b->cols->offset, dim(f.m), dim(f.n));
    } else if (left) {
      RV r = rk_of(b);
      DV A = dv_panel(r.a, r.la, r.m, r.cap, r.bufa, r.row0);
      dscale(A, A, false, inverse, b->rows->offset, dim(r.m), dim(r.cap, r.rank));
    } else {
      RV r = rk_of(b);
      DV B = dv_panel(r.b, r.lb, r.n, r.cap, r.bufb, r.col0);
      dscale(B, B, false, inverse, b->cols->offset, dim(r.n), dim(r.cap, r.rank));
    }
  }

  // A copy of b times the diagonal at its columns, the operand M D of
  // mdntProduct / mdmtProduct (h_matrix.cpp:2585-2690: copy, multiplyWithDiag,
  // product): full leaves into temporaries, Rk leaves share A and the rank slot
  // and get a scaled copy of B.  Readable through full_of / rk_of until
  // release_twins().
  const Block* scaled_twin(const Block* b) {
    twins_.push_back(*b);
    Block* t = &twins_.back();
    if (!b->is_leaf()) {
      for (int q = 0; q < 4; ++q)
        if (b->child[q]) t->child[q] = const_cast<Block*>(scaled_twin(b->child[q]));
      return t;
    }
    Shadow sh;
    if (b->kind == Kind::Full) {
      DV f = full_of(b);
      sh.d = temp_dense(f.m, f.n, b->rows->offset);
      dscale(sh.d, f, true, false, b->cols->offset, dim(f.m), dim(f.n));
    } else {
      RV r = rk_of(b);
      sh.r = r;
      sh.r.b = talloc((size_t)r.n * r.cap);
      sh.r.lb = r.n;
      sh.r.bufb = kTempBase + fs_.next_temp++;
      ++live_;
      DV src = dv_panel(r.b, r.lb, r.n, r.cap, r.bufb, r.col0);
      DV dst = dv_panel(sh.r.b, sh.r.lb, r.n, r.cap, sh.r.bufb, r.col0);
      dscale(dst, src, false, false, b->cols->offset, dim(r.n), dim(r.cap, r.rank));
    }
    t->leaf = kShadowLeaf + (int)shadow_.size();
    shadow_.push_back(sh);
    ++twin_live_;
    return t;
  }
  void release_twins() {
    live_ -= twin_live_;
    twin_live_ = 0;
    shadow_.clear();
    twins_.clear();
  }

  void ldlt(const Block* h) {
    if (h->is_leaf()) {
      if (h->kind != Kind::Full) throw std::runtime_error("ldlt: Rk diagonal block");
      ldlt_leaf(h);
      return;
    }
    if (h->nrow != h->ncol) throw std::runtime_error("ldlt: non-square block grid");
    const int n = h->nrow;
    for (int k = 0; k < n; ++k) {
      ldlt(h->get(k, k));
      // L_ik = H_ik L_kk^{-T} D_k^{-1}
      for (int i = k + 1; i < n; ++i)
        if (h->get(i, k)) {
          solve_upper_right_h(h->get(k, k), h->get(i, k));
          scale_by_diag(h->get(i, k), true);
        }
      maybe_flush();
      // H_ij -= (L_ik D_k) L_jk^T for k < j <= i
      for (int i = k + 1; i < n; ++i) {
        if (!h->get(i, k)) continue;
        const Block* w = scaled_twin(h->get(i, k));
        for (int j = k + 1; j <= i; ++j)
          if (h->get(i, j) && h->get(j, k)) hgemm(h->get(i, j), 'N', 'T', -1.0, w, h->get(j, k));
        release_twins();
        maybe_flush();
      }
    }
  }

  // A view of another matrix's leaves (the same cluster trees): a twin of its
  // block tree whose leaves alias its store under buffer ids base + leaf.
  const Block* foreign_view(const Block* b, Store<T>& xs, int64_t base) {
    twins_.push_back(*b);
    Block* t = &twins_.back();
    if (!b->is_leaf()) {
      for (int q = 0; q < 4; ++q)
        if (b->child[q]) t->child[q] = const_cast<Block*>(foreign_view(b->child[q], xs, base));
      return t;
    }
    const LeafDesc& L = xs.hdesc[b->leaf];
    Shadow sh;
    sh.d.p = xs.data() + L.a;
    sh.d.ld = L.m;
    sh.d.m = L.m;
    sh.d.n = L.n;
    sh.d.buf = base + b->leaf;
    sh.r.a = xs.data() + L.a;
    sh.r.b = xs.data() + L.b;
    sh.r.la = L.m;
    sh.r.lb = L.n;
    sh.r.m = L.m;
    sh.r.n = L.n;
    sh.r.cap = L.cap;
    sh.r.rank = xs.ranks.template as<int>() + b->leaf;
    sh.r.bufa = sh.r.bufb = base + b->leaf;
    sh.r.row0 = b->rows->offset;
    sh.r.col0 = b->cols->offset;
    t->leaf = kShadowLeaf + (int)shadow_.size();
    shadow_.push_back(sh);
    return t;
  }

  // HMatrix::gemm h_matrix.cpp:1570-1620 (after the scaling by beta): this
  // matrix += alpha op(A) op(B), A and B other matrices over the same trees
  void gemm_matrices(char ta, char tb, double alpha, Matrix& A, Store<T>& as, Matrix& B,
                     Store<T>& bs) {
    constexpr int64_t kABase = int64_t(1) << 36, kBBase = kABase + (int64_t(1) << 30);
    const Block* va = foreign_view(A.bt->root, as, kABase);
    const Block* vb = foreign_view(B.bt->root, bs, kBBase);
    hgemm(M_.bt->root, ta, tb, alpha, va, vb);
    maybe_flush();
  }

  // Left triangular solves with an H-matrix right-hand side b (in place):
  // HMatrix::solveLowerTriangularLeft(HMatrix) h_matrix.cpp:2297-2335 +
  // recursion.cpp:191-233 and solveUpperTriangularLeft(HMatrix) :2463-2500 +
  // recursion.cpp:393-444.  kind 0: L X = B (LLt / LDLt factor), 1: L^T X = B,
  // 2: L X = B (LU, unit L with the leaves' interchanges), 3: U X = B (LU).
  void dense_left_solve(int kind, const Block* F, const DV& X, Dim ncol) {
    switch (kind) {
      case 0: solve_lower_left(F, X, ncol); break;
      case 1: solve_upper_left(F, X, ncol); break;
      case 2: lu_solve_lower_left(F, X, ncol); break;
      default: lu_solve_upper_left(F, X, ncol); break;
    }
  }
  void solve_left_h(int kind, const Block* F, const Block* b) {
    if (!b) return;
    const bool lower = kind == 0 || kind == 2;
    if (!F->is_leaf() && !b->is_leaf()) {
      if (F->ncol == b->nrow) {
        for (int k = 0; k < b->ncol; ++k) {
          if (lower) {
            for (int i = 0; i < F->nrow; ++i) {
              const Block* bik = b->get(i, k);
              if (!bik) continue;
              for (int j = 0; j < i; ++j)
                if (F->get(i, j) && b->get(j, k)) hgemm(bik, 'N', 'N', -1.0, F->get(i, j), b->get(j, k));
              solve_left_h(kind, F->get(i, i), bik);
            }
          } else {
            for (int i = F->nrow - 1; i >= 0; --i) {
              const Block* bik = b->get(i, k);
              if (!bik) continue;
              solve_left_h(kind, F->get(i, i), bik);
              for (int j = 0; j < i; ++j) {
                const Block* bjk = b->get(j, k);
                if (!bjk) continue;
                if (kind == 1 && F->get(i, j)) hgemm(bjk, 'T', 'N', -1.0, F->get(i, j), bik);
                if (kind == 3 && F->get(j, i)) hgemm(bjk, 'N', 'N', -1.0, F->get(j, i), bik);
              }
            }
          }
          maybe_flush();
        }
      } else if (F->ncol > 1 && b->nrow == 1 && b->ncol > 1) {
        for (int k = 0; k < b->ncol; ++k) solve_left_h(kind, F, b->get(0, k));
      } else {
        throw std::runtime_error("solve_mat: triangular solve case not handled");
      }
      return;
    }
    if (!b->is_leaf()) {
      if (b->nrow == 1) {
        for (int k = 0; k < b->ncol; ++k) solve_left_h(kind, F, b->get(0, k));
        return;
      }
      DV D = eval_block(b);
      dense_left_solve(kind, F, D, dim(D.n));
      clear_and_set(b, D);
      release();
      return;
    }
    if (b->kind == Kind::Full) {
      dense_left_solve(kind, F, full_of(b), dim(b->n()));
    } else {
      RV r = rk_of(b);
      dense_left_solve(kind, F, dv_panel(r.a, r.la, r.m, r.cap, r.bufa, r.row0), dim(r.cap, r.rank));
    }
  }

  // HMatrix::solve(HMatrix* b, algo) h_matrix.cpp:2937-2966: B <- A^{-1} B for
  // another matrix B whose rows use this matrix's column tree
  // (part 1: L X = B only, part 2: L^T X = B only, as DefaultEngine::solveLower
  // default_engine.cpp:243-253 for LLt factors)
  void solve_matrix(Matrix& B, Store<T>& bs, int part) {
    constexpr int64_t kBBase = (int64_t(1) << 36) + (int64_t(1) << 30);
    const Block* vb = foreign_view(B.bt->root, bs, kBBase);
    const Block* F = M_.bt->root;
    if (part != 0) {
      solve_left_h(part == 1 ? 0 : 1, F, vb);
    } else if (M_.factorization == hmat_factorization_lu) {
      solve_left_h(2, F, vb);
      solve_left_h(3, F, vb);
    } else {
      solve_left_h(0, F, vb);
      if (M_.factorization == hmat_factorization_ldlt) scale_by_diag(vb, true, true);
      solve_left_h(1, F, vb);
    }
  }

  // ---- symmetric HODLR (hodlr.cpp) -------------------------------------------------
  // A 2 x 2 node m = [m00, .; a b^T, m11] is factored as
  // P W W^T P^T with P = diag(L00, L11) (the children's factors) and
  // W = I + U X U^T, U = diag(b~, a~) the Rk panels after the children's
  // lower solves; a node keeps X11 and KK = (I + X11 a~^T a~)^{-1} X11.
  DV hodlr_block(int q, int which) const {
    const Matrix::HodlrNode& nd = M_.hodlr_nodes[q];
    DV d;
    d.p = M_.hodlr_buf.template as<T>() + (which == 0 ? nd.x11 : nd.kk);
    d.ld = std::max(nd.cap, 1);
    d.m = nd.cap;
    d.n = nd.cap;
    d.buf = kHodlrBase + 2 * (int64_t)q + which;
    return d;
  }

  // HODLR<T>::factorizeSym / factorizeSym hodlr.cpp:287-312
  void hodlr_factor(const Block* a) {
    const Block *a00 = a->get(0, 0), *a11 = a->get(1, 1), *a10 = a->get(1, 0);
    if (a00->is_leaf()) potrf(a00); else hodlr_factor(a00);
    if (a11->is_leaf()) potrf(a11); else hodlr_factor(a11);
    RV r = rk_of(a10);
    const Dim rk = dim(r.cap, r.rank);
    DV A = dv_panel(r.a, r.la, r.m, r.cap, r.bufa, r.row0);
    DV B = dv_panel(r.b, r.lb, r.n, r.cap, r.bufb, r.col0);
    hodlr_lower(a00, B, rk);  // b <- L00^-1 b
    hodlr_lower(a11, A, rk);  // a <- L11^-1 a
    const int q = M_.hodlr_index.at(a);
    DV G = temp_dense(r.cap, r.cap, -1), P = temp_dense(r.cap, r.cap, -1);
    DV At = A, Bt = B;
    At.trans = Bt.trans = true;
    gemm(G, false, 1.0, At, A, rk, rk, dim(r.m));
    gemm(P, false, 1.0, Bt, B, rk, rk, dim(r.n));
    DV X = hodlr_block(q, 0), K = hodlr_block(q, 1);
    FTask t{};
    t.type = T_HODLR;
    t.m = rk;
    t.A = G.p;
    t.B = P.p;
    t.lda = r.cap;
    t.C = X.p;
    t.tgt.a = K.p;
    t.tgt.b = M_.hodlr_logdet.template as<double>() + q;
    rd(G);
    rd(P);
    wr(X);
    wr(K);
    emit(t, hodlr_workspace(r.cap));
    release();
    release();
    maybe_flush();
  }

  // X <- L^{-1} X for the HODLR factor of m (solveLowerTriangularLeft
  // hodlr.cpp:138-156): the children, then x1 -= a (b^T x0), then the
  // Woodbury correction x1 -= a KK a^T x1.
  void hodlr_lower(const Block* m, const DV& X, Dim ncol) {
    if (m->is_leaf()) {
      trsm(1, m, X, ncol);
      return;
    }
    const int n0 = m->get(0, 0)->m(), n1 = m->get(1, 1)->m();
    DV X0 = rows_of(X, 0, n0), X1 = rows_of(X, n0, n1);
    hodlr_lower(m->get(0, 0), X0, ncol);
    hodlr_lower(m->get(1, 1), X1, ncol);
    RV r = rk_of(m->get(1, 0));
    const Dim rk = dim(r.cap, r.rank);
    DV A = dv_panel(r.a, r.la, r.m, r.cap, r.bufa, r.row0);
    DV B = dv_panel(r.b, r.lb, r.n, r.cap, r.bufb, r.col0);
    DV At = A, Bt = B;
    At.trans = Bt.trans = true;
    DV z = temp_dense(r.cap, std::max(ncol.s, 1), -1), u = temp_dense(r.cap, std::max(ncol.s, 1), -1);
    gemm(z, false, 1.0, Bt, X0, rk, ncol, dim(n0));
    gemm(X1, true, -1.0, A, z, dim(n1), ncol, rk);
    gemm(z, false, 1.0, At, X1, rk, ncol, dim(n1));
    gemm(u, false, 1.0, hodlr_block(M_.hodlr_index.at(m), 1), z, rk, ncol, rk);
    gemm(X1, true, -1.0, A, u, dim(n1), ncol, rk);
    release();
    release();
  }

  // X <- L^{-T} X (solveUpperTriangularLeft hodlr.cpp:113-136): the Woodbury
  // correction x1 -= a KK^T a^T x1, then x0 -= b a^T x1, then the children.
  void hodlr_upper(const Block* m, const DV& X, Dim ncol) {
    if (m->is_leaf()) {
      trsm(2, m, X, ncol);
      return;
    }
    const int n0 = m->get(0, 0)->m(), n1 = m->get(1, 1)->m();
    DV X0 = rows_of(X, 0, n0), X1 = rows_of(X, n0, n1);
    RV r = rk_of(m->get(1, 0));
    const Dim rk = dim(r.cap, r.rank);
    DV A = dv_panel(r.a, r.la, r.m, r.cap, r.bufa, r.row0);
    DV B = dv_panel(r.b, r.lb, r.n, r.cap, r.bufb, r.col0);
    DV At = A;
    At.trans = true;
    DV Kt = hodlr_block(M_.hodlr_index.at(m), 1);
    Kt.trans = true;
    DV z = temp_dense(r.cap, std::max(ncol.s, 1), -1), u = temp_dense(r.cap, std::max(ncol.s, 1), -1);
    gemm(z, false, 1.0, At, X1, rk, ncol, dim(n1));
    gemm(u, false, 1.0, Kt, z, rk, ncol, rk);
    gemm(X1, true, -1.0, A, u, dim(n1), ncol, rk);
    gemm(z, false, 1.0, At, X1, rk, ncol, dim(n1));
    gemm(X0, true, -1.0, B, z, dim(n0), ncol, rk);
    release();
    release();
    hodlr_upper(m->get(0, 0), X0, ncol);
    hodlr_upper(m->get(1, 1), X1, ncol);
  }

  // C (+)= alpha A, dense views of equal shape (a T_AXPY task; set: C = alpha A)
  void dense_axpy(const DV& C, double alpha, const DV& A, Dim m, Dim n, bool set) {
    FTask t{};
    t.type = T_AXPY;
    t.mode = set ? 1 : 0;
    t.m = m;
    t.n = n;
    t.alpha = alpha;
    t.A = A.p;
    t.lda = A.ld;
    t.C = C.p;
    t.ldc = C.ld;
    rd(A);
    if (!set) rd(C);
    wr(C);
    emit(t);
  }

  // Y += op(L) X for the stored blocks of a factorised matrix, without the
  // mirrored upper blocks of symmetric storage: HMatrix::gemv of an LLt factor
  // (isTriLower, isLower off: getChildForGEMM returns the absent upper children
  // as null, h_matrix.cpp:771-788, 790-857); the diagonal leaves hold L with
  // their upper triangle zeroed (scalar_array.cpp:822-826, as POTRF leaves it).
  void tri_gemv(const Block* h, char tr, const DV& X, const DV& Y, Dim ncol) {
    if (h->is_leaf()) {
      hgemv(h, tr, 1.0, X, Y, ncol);
      return;
    }
    for (int i = 0; i < h->nrow; ++i)
      for (int j = 0; j < h->ncol; ++j) {
        const Block* ch = h->get(i, j);
        if (!ch) continue;
        int co = ch->cols->offset - h->cols->offset, cs = ch->n();
        int ro = ch->rows->offset - h->rows->offset, rs = ch->m();
        if (tr != 'N') {
          std::swap(co, ro);
          std::swap(cs, rs);
        }
        tri_gemv(ch, tr, rows_of(X, co, cs), rows_of(Y, ro, rs), ncol);
      }
  }

  // Y += op(L W) X for a symmetric HODLR factor (HODLR<T>::gemv, hodlr.cpp:
  // 325-378): per node, with a10 = a b^T, W11 = I + a X11 a^T,
  //   'N': y1 += L1 (x1 + a (b^T x0 + X11 a^T x1)),  y0 += L0 x0;
  //   'T': y0 += L0^T x0 + b a^T L1^T x1,  y1 += L1^T x1 + a X11^T a^T L1^T x1.
  void hodlr_gemv(const Block* m, char tr, const DV& X, const DV& Y, Dim ncol) {
    if (m->is_leaf()) {
      hgemv(m, tr, 1.0, X, Y, ncol);
      return;
    }
    const int n0 = m->get(0, 0)->m(), n1 = m->get(1, 1)->m();
    DV X0 = rows_of(X, 0, n0), X1 = rows_of(X, n0, n1);
    DV Y0 = rows_of(Y, 0, n0), Y1 = rows_of(Y, n0, n1);
    RV r = rk_of(m->get(1, 0));
    const Dim rk = dim(r.cap, r.rank);
    const int w = std::max(ncol.s, 1);
    DV A = dv_panel(r.a, r.la, r.m, r.cap, r.bufa, r.row0);
    DV B = dv_panel(r.b, r.lb, r.n, r.cap, r.bufb, r.col0);
    DV At = A, Bt = B;
    At.trans = Bt.trans = true;
    DV X11 = hodlr_block(M_.hodlr_index.at(m), 0);
    if (tr == 'N') {
      DV t = temp_dense(r.cap, w, -1), u = temp_dense(r.cap, w, -1), x1b = temp_dense(n1, w, -1);
      gemm(t, false, 1.0, At, X1, rk, ncol, dim(n1));
      gemm(u, false, 1.0, X11, t, rk, ncol, rk);
      gemm(u, true, 1.0, Bt, X0, rk, ncol, dim(n0));
      dense_axpy(x1b, 1.0, X1, dim(n1), ncol, true);
      gemm(x1b, true, 1.0, A, u, dim(n1), ncol, rk);
      hodlr_gemv(m->get(1, 1), tr, x1b, Y1, ncol);
      hodlr_gemv(m->get(0, 0), tr, X0, Y0, ncol);
      release();
      release();
      release();
      return;
    }
    hodlr_gemv(m->get(0, 0), tr, X0, Y0, ncol);
    DV l = temp_dense(n1, w, -1), t = temp_dense(r.cap, w, -1), u = temp_dense(r.cap, w, -1);
    zero(l);
    hodlr_gemv(m->get(1, 1), tr, X1, l, ncol);
    gemm(t, false, 1.0, At, l, rk, ncol, dim(n1));
    gemm(Y0, true, 1.0, B, t, dim(n0), ncol, rk);
    DV X11t = X11;
    X11t.trans = true;
    gemm(u, false, 1.0, X11t, t, rk, ncol, rk);
    gemm(Y1, true, 1.0, A, u, dim(n1), ncol, rk);
    dense_axpy(Y1, 1.0, l, dim(n1), ncol, false);
    release();
    release();
    release();
  }

  // ---- non-symmetric HODLR (hodlr.cpp:45-116) --------------------------------------
  // The n0 x cap (which 2: a01) or n1 x cap (3: b01) panel of node q.
  DV hodlr_panel(int q, int which) const {
    const Matrix::HodlrNode& nd = M_.hodlr_nodes[q];
    const Block* b10 = nd.b->get(1, 0);
    DV d;
    d.p = M_.hodlr_buf.template as<T>() + (which == 2 ? nd.a01 : nd.b01);
    d.ld = which == 2 ? b10->n() : b10->m();
    d.m = d.ld;
    d.n = nd.cap;
    d.buf = kHodlrBase + (int64_t(1) << 32) + 2 * (int64_t)q + (which - 2);
    return d;
  }
  // KK of node q (2 cap x 2 cap, ld 2 cap)
  DV hodlr_kk(int q) const {
    const Matrix::HodlrNode& nd = M_.hodlr_nodes[q];
    DV d;
    d.p = M_.hodlr_buf.template as<T>() + nd.kk;
    d.ld = d.m = d.n = 2 * nd.cap;
    d.buf = kHodlrBase + 2 * (int64_t)q + 1;
    return d;
  }
  // C <- I (a T_ZERO task, mode 1)
  void identity(const DV& C) {
    FTask t{};
    t.type = T_ZERO;
    t.mode = 1;
    t.m = dim(C.m);
    t.n = dim(C.n);
    t.C = C.p;
    t.ldc = C.ld;
    wr(C);
    emit(t);
  }
  // TRSM (modes and flags as trsm) with a dense triangular operand and its
  // interchanges
  void trsm_dense(int mode, const DV& L, int* piv, const DV& X, Dim other) {
    FTask t{};
    t.type = T_TRSM;
    t.mode = mode;
    t.n = dim(L.m);
    t.m = other;
    t.A = L.p;
    t.lda = L.ld;
    if (mode & kTrsmPiv) t.B = piv;
    t.C = X.p;
    t.ldc = X.ld;
    rd(L);
    rd(X);
    wr(X);
    emit(t);
  }

  // HODLR<T>::factorize / factorize hodlr.cpp:81-116: the children (leaves
  // by LLt), a01 b01^T = (a10 b10^T)^T (desymmetrize), a01 <- M00^{-1} a01,
  // a10 <- M11^{-1} a10, KK = I + [[0, b01^T a10], [b10^T a01, 0]], LU of KK.
  void hodlr_factor_ns(const Block* a) {
    const Block *a00 = a->get(0, 0), *a11 = a->get(1, 1), *a10 = a->get(1, 0);
    if (a00->is_leaf()) potrf(a00); else hodlr_factor_ns(a00);
    if (a11->is_leaf()) potrf(a11); else hodlr_factor_ns(a11);
    const int q = M_.hodlr_index.at(a);
    const int n0 = a00->m(), n1 = a11->m();
    RV r = rk_of(a10);
    const Dim rk = dim(r.cap, r.rank);
    DV A10 = dv_panel(r.a, r.la, r.m, r.cap, r.bufa, r.row0);
    DV B10 = dv_panel(r.b, r.lb, r.n, r.cap, r.bufb, r.col0);
    DV A01 = hodlr_panel(q, 2), B01 = hodlr_panel(q, 3);
    dense_axpy(A01, 1.0, B10, dim(n0), rk, true);
    dense_axpy(B01, 1.0, A10, dim(n1), rk, true);
    hodlr_solve_ns(a00, A01, rk);
    hodlr_solve_ns(a11, A10, rk);
    DV K = hodlr_kk(q);
    identity(K);
    const int cap = r.cap;
    DV K01 = K, K10 = K;
    K01.p = (char*)K.p + sizeof(T) * (size_t)cap * K.ld;
    K10.p = (char*)K.p + sizeof(T) * (size_t)cap;
    K01.m = K01.n = K10.m = K10.n = cap;
    DV B10t = B10, B01t = B01;
    B10t.trans = B01t.trans = true;
    gemm(K01, true, 1.0, B01t, A10, rk, rk, dim(n1));
    gemm(K10, true, 1.0, B10t, A01, rk, rk, dim(n0));
    FTask t{};
    t.type = T_GETRF;
    t.m = dim(K.m);
    t.A = K.p;
    t.lda = K.ld;
    t.B = M_.hodlr_piv.template as<int>() + M_.hodlr_nodes[q].piv;
    rd(K);
    wr(K);
    emit(t);
    maybe_flush();
  }

  // X <- M^{-1} X with the non-symmetric HODLR factor of m (solve hodlr.cpp:
  // 45-74): the leaves' LLt solve, the children, then the Woodbury step
  // X -= U KK^{-1} V^T X with U = diag(a01, a10), V = [b01; b10] crosswise.
  void hodlr_solve_ns(const Block* m, const DV& X, Dim ncol) {
    if (m->is_leaf()) {
      trsm(1, m, X, ncol);
      trsm(2, m, X, ncol);
      return;
    }
    const int n0 = m->get(0, 0)->m(), n1 = m->get(1, 1)->m();
    DV X0 = rows_of(X, 0, n0), X1 = rows_of(X, n0, n1);
    hodlr_solve_ns(m->get(0, 0), X0, ncol);
    hodlr_solve_ns(m->get(1, 1), X1, ncol);
    const int q = M_.hodlr_index.at(m);
    RV r = rk_of(m->get(1, 0));
    const Dim rk = dim(r.cap, r.rank);
    const int cap = r.cap, w = std::max(ncol.s, 1);
    DV A10 = dv_panel(r.a, r.la, r.m, r.cap, r.bufa, r.row0);
    DV B10 = dv_panel(r.b, r.lb, r.n, r.cap, r.bufb, r.col0);
    DV A01 = hodlr_panel(q, 2), B01 = hodlr_panel(q, 3);
    DV B10t = B10, B01t = B01;
    B10t.trans = B01t.trans = true;
    DV tv = temp_dense(2 * cap, w, -1);
    zero(tv);
    DV tv0 = rows_of(tv, 0, cap), tv1 = rows_of(tv, cap, cap);
    gemm(tv0, false, 1.0, B01t, X1, rk, ncol, dim(n1));
    gemm(tv1, false, 1.0, B10t, X0, rk, ncol, dim(n0));
    DV K = hodlr_kk(q);
    int* piv = M_.hodlr_piv.template as<int>() + M_.hodlr_nodes[q].piv;
    trsm_dense(1 | kTrsmUnit | kTrsmPiv, K, piv, tv, ncol);
    trsm_dense(2 | kTrsmTransA, K, nullptr, tv, ncol);
    gemm(X0, true, -1.0, A01, tv0, dim(n0), ncol, rk);
    gemm(X1, true, -1.0, A10, tv1, dim(n1), ncol, rk);
    release();
  }

  // HMatrix::solveDiagonal h_matrix.cpp:2969-2983 (LDLt): X <- D^{-1} X, one
  // task per diagonal leaf
  void solve_diag(const DV& X, Dim ncol) {
    for (const Block* b : M_.bt->leaves) {
      if (b->rows != b->cols) continue;
      DV xs = rows_of(X, b->rows->offset, b->m());
      dscale(xs, xs, false, true, b->rows->offset, dim(b->m()), ncol);
    }
  }

  // ---- whole-matrix updates -----------------------------------------------------------
  // RkMatrix::truncate (rk_matrix.cpp:490-585) of every Rk leaf at eps: a
  // formatted-add task whose target is recompressed alone (mode 1)
  void truncate_leaves() {
    for (const Block* b : M_.bt->leaves) {
      if (b->kind != Kind::Rk) continue;
      RV tg = rk_of(b);
      FTask t{};
      t.type = T_FADD;
      t.mode = 1;
      t.eps = eps_;
      t.tgt = ref(tg);
      t.nparts = 0;
      rd(tg);
      wr(tg);
      emit(t, fadd_workspace(tg.m, tg.n, tg.cap));
      maybe_flush();
    }
  }

  // HMatrix::axpy(alpha, X) (h_matrix.cpp:930-1073) for X of the same
  // partition, stored in xs: Rk leaves by formatted addition + truncation,
  // full leaves by a dense axpy.  X's leaves are read-only dependency units.
  void axpy_matrix(double alpha, Store<T>& xs) {
    constexpr int64_t kXBase = int64_t(1) << 36;
    const int* xr = xs.ranks.template as<int>();
    for (size_t l = 0; l < M_.bt->leaves.size(); ++l) {
      const Block* b = M_.bt->leaves[l];
      const LeafDesc& L = xs.hdesc[l];
      if (L.kind == 0) {
        DV f;
        f.p = xs.data() + L.a;
        f.ld = L.m;
        f.m = L.m;
        f.n = L.n;
        f.buf = kXBase + (int64_t)l;
        axpy_full(b, alpha, f, b->rows->offset, b->cols->offset);
      } else {
        RV r;
        r.a = xs.data() + L.a;
        r.b = xs.data() + L.b;
        r.la = L.m;
        r.lb = L.n;
        r.m = L.m;
        r.n = L.n;
        r.cap = L.cap;
        r.rank = const_cast<int*>(xr) + l;
        r.bufa = r.bufb = kXBase + (int64_t)l;
        r.row0 = b->rows->offset;
        r.col0 = b->cols->offset;
        axpy_rk(b, alpha, r);
      }
      maybe_flush();
    }
  }

  // ---- execution -----------------------------------------------------------------------
  int prog_row_ = 0;  // end row of the last diagonal leaf factorisation emitted

  void maybe_flush() {
    if (live_ != 0) return;
    // with a progress reporter, segments of at most kMaxTasksProgress tasks so
    // that reports come at least every few seconds on large factorisations
    const size_t limit = M_.progress && M_.progress->update ? kMaxTasksProgress : kMaxTasks;
    if (tasks_.size() >= limit || fs_.temp_used > fs_.temp_budget / 2) flush(false);
  }

  void flush(bool forced) {
    if (forced && live_ != 0)
      throw std::runtime_error("task segment limit reached while temporaries are live");
    if (tasks_.empty()) return;
    run_segment();
    if (M_.progress && M_.progress->update) {
      // hmat_progress_t (h_matrix.cpp:2554-2557): the end row of the last
      // diagonal leaf factorised, once the segment holding it is done
      join_submitter();
      HMX_CHECK(hipStreamSynchronize(M_.stream));
      M_.progress->current = M_.embedded ? prog_row_ / 2 : prog_row_;
      M_.progress->update(M_.progress);
    }
    tasks_.clear();
    waves_.clear();
    units_.clear();
    max_wave_ = 0;
    reset_temps();
    fs_.rank_used = 0;
  }

  void run_segment() {
    const double tseg0 = now_s();
    struct SegTimer {
      double& acc;
      double t0;
      ~SegTimer() { acc += now_s() - t0; }
    } seg_timer{t_seg_, tseg0};
    hipStream_t s = M_.stream;
    const int W = max_wave_;
    auto plan = std::make_shared<SegmentPlan>();
    const size_t n = tasks_.size();
    // order: wave, then type
    std::vector<int> order(n);
    for (size_t i = 0; i < n; ++i) order[i] = (int)i;
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) {
      if (waves_[x] != waves_[y]) return waves_[x] < waves_[y];
      return tasks_[x].type < tasks_[y].type;
    });
    std::vector<FTask> sorted(n);
    std::vector<GemmTile> tiles;
    std::vector<SegLaunch>& launches = plan->launches;
    int64_t ws_max = 0;
    const int64_t ws_budget = (int64_t)(4ll << 30) / 8;  // doubles per launch
    int64_t wave_ws = 0;  // workspace taken by the overlapping launches of the current wave
    size_t i = 0;
    while (i < n) {
      const int w = waves_[order[i]], ty = tasks_[order[i]].type;
      SegLaunch L{ty, (int)i, 0, (int)tiles.size(), 0, w, false};
      int64_t ws = 0;
      while (i < n && waves_[order[i]] == w && tasks_[order[i]].type == ty) {
        FTask t = tasks_[order[i]];
        const int64_t need = t.ws;
        if (L.count > 0 && ws + need > ws_budget) break;  // split oversized launches
        t.ws = ws;
        ws += need;
        sorted[i] = t;
        if (ty == T_GEMM) {
          const int tm = (t.m.s + 63) / 64, tn = (t.n.s + 63) / 64;
          for (int a = 0; a < tm; ++a)
            for (int b = 0; b < tn; ++b) tiles.push_back(GemmTile{(int)i - L.first, a, b});
        }
        ++L.count;
        ++i;
      }
      L.tcount = (int)tiles.size() - L.tfirst;
      // workspace after the other launches of the wave, so that they can run
      // side by side; a launch that would exceed the budget runs alone
      int64_t base = 0;
      if (!launches.empty() && launches.back().wave == w && launches.back().overlap) base = wave_ws;
      if (base + ws <= ws_budget) {
        L.overlap = true;
        if (base > 0)
          for (int q = L.first; q < L.first + L.count; ++q) sorted[q].ws += base;
        wave_ws = base + ws;
      } else {
        wave_ws = 0;
      }
      ws_max = std::max(ws_max, L.overlap ? wave_ws : ws);
      launches.push_back(L);
    }
    // the task lists go up with a synchronous copy on the null stream, which
    // does not wait for the matrix's (non-blocking) stream: the host builds the
    // next segment while the device still runs this one
    auto put = [](DevBuf& b, const void* src, size_t bytes) {
      b.ensure(bytes);
      HMX_CHECK(hipMemcpy(b.p, src, bytes, hipMemcpyHostToDevice));
    };
    put(plan->tasks, sorted.data(), n * sizeof(FTask));
    if (!tiles.empty()) {
      put(plan->tiles, tiles.data(), tiles.size() * sizeof(GemmTile));
    } else {
      plan->tiles.ensure(sizeof(GemmTile));
      HMX_CHECK(hipMemsetAsync(plan->tiles.p, 0, sizeof(GemmTile), s));
    }
    plan->ws = ws_max;
    plan->ranks_used = fs_.rank_used;
    plan->ntasks = (long)n;
    plan->waves = W;
    // the previous segment's launches are queued before this one's
    join_submitter();
    if (record_) {
      execute_plan<T>(*plan, fs_, s, false, M_.single_stream);
      HMX_CHECK(hipStreamSynchronize(s));
      record_->push_back(plan);
    } else {
      inflight_.push_back(plan);  // kept until check_flags() has waited for the stream
      // a thread queues the launches, so that the host builds the next
      // segment meanwhile: queuing ~10k launches waits on the device (a full
      // queue), and building the next segment after it left the device idle
      const bool one = M_.single_stream;
      submitter_ = std::thread([this, plan, s, one] {
        try {
          execute_plan<T>(*plan, fs_, s, false, one);
        } catch (...) {
          submit_err_ = std::current_exception();
        }
      });
    }
    M_.stats.tasks += (long)n;
    M_.stats.waves += W;
    M_.stats.launches += (long)launches.size();
    M_.stats.segments += 1;
    M_.stats.temp_bytes = std::max(M_.stats.temp_bytes, fs_.temp_used);
  }

  // Waits for the segments in flight; a failing pivot (potrf's info > 0) or a
  // capacity overflow seen by any of them raises here.  Launches after a
  // failing pivot return at once (the failure flag), as the reference stops.
  void check_flags() {
    join_submitter();
    HMX_CHECK(hipStreamSynchronize(M_.stream));
    inflight_.clear();
    int f[2] = {0, 0};
    HMX_CHECK(hipMemcpy(f, fs_.flags.p, sizeof(f), hipMemcpyDeviceToHost));
    if (f[1] && lu_mode_) throw std::runtime_error("luDecomposition: getrf found a zero pivot (singular matrix)");
    if (f[1] && ldlt_mode_)
      throw std::runtime_error("ldltDecomposition: zero pivot (no LDLt factorisation without pivoting)");
    if (f[1]) throw std::runtime_error("lltDecomposition: matrix is not positive definite");
    if (f[0]) throw std::runtime_error("an Rk rank exceeded the panel capacity during the factorisation");
  }

  int live() const { return live_; }
  void record_into(PlanList* p) { record_ = p; }
  void set_lu(bool v) { lu_mode_ = v; }
  void set_ldlt(bool v) { ldlt_mode_ = v; }
  void report(double total) const {
#if HMX_DIAGNOSTICS
    if (getenv("HMX_TRACE_LAUNCHES")) {
      fprintf(stderr, "factorisation: total %.4f s, segments %.4f s, temp allocations %.4f s, %zu units\n",
              total, t_seg_, t_alloc_, units_.size());
      for (int q = 0; q < 32; ++q)
        if (g_type_launches[q])
          fprintf(stderr, "  task type %d: %ld launches, %.4f s\n", q, g_type_launches[q], g_type_seconds[q]);
    }
#else
    (void)total;
#endif
  }

 private:
  // segments of at most this many tasks, so that the device starts early and
  // the host builds the next segment while the device runs the current one
  static constexpr size_t kMaxTasks = 250000;
  static constexpr size_t kMaxTasksProgress = 20000;
  std::vector<std::shared_ptr<SegmentPlan>> inflight_;
  std::thread submitter_;          // queues the launches of the last segment
  std::exception_ptr submit_err_;
  Matrix& M_;
  Store<T>& st_;
  FactorState& fs_;
  std::vector<int> row2leaf_;
  double eps_ = 1e-4;
  int cap_max_ = 160;
  int live_ = 0;
  std::vector<FTask> tasks_;
  std::vector<int> waves_;
  std::unordered_map<uint64_t, Unit> units_;
  std::vector<uint64_t> r_, w_;
  int max_wave_ = 0;
  PlanList* record_ = nullptr;
  bool lu_mode_ = false, ldlt_mode_ = false;
  bool use_inv_ = false;
  double t_alloc_ = 0.0, t_seg_ = 0.0;
  struct Shadow {
    DV d;
    RV r;
  };
  std::deque<Block> twins_;
  std::vector<Shadow> shadow_;
  int twin_live_ = 0;
};

// ---------------------------------------------------------------------------
// Matrix entry points
// ---------------------------------------------------------------------------
void Matrix::factorize(int algo, bool fp32) {
  if (!assembled) throw std::runtime_error("factorize: matrix not assembled");
  if (factorization != -1) throw std::runtime_error("factorize: matrix already factorised");
  if (embedded && algo != hmat_factorization_lu)
    throw std::runtime_error("factorize: complex matrices are factorised by hmat_factorization_lu");
  if (algo != hmat_factorization_llt && algo != hmat_factorization_lu && algo != hmat_factorization_ldlt &&
      algo != hmat_factorization_hodlrsym && algo != hmat_factorization_hodlr)
    throw std::runtime_error("factorize: only hmat_factorization_llt, hmat_factorization_ldlt, "
                             "hmat_factorization_hodlr, hmat_factorization_hodlrsym and "
                             "hmat_factorization_lu are provided");
  const bool hodlr_any = algo == hmat_factorization_hodlrsym || algo == hmat_factorization_hodlr;
  if ((algo == hmat_factorization_llt || algo == hmat_factorization_ldlt || hodlr_any) && !lower_sym)
    throw std::runtime_error("factorize: LLt, LDLt and HODLR need lower-symmetric storage");
  if (hodlr_any) hodlr_layout(algo == hmat_factorization_hodlrsym);
  inv_blocks.clear();
  inv_index.clear();
  inv_buf.release();
  if (algo == hmat_factorization_lu && lower_sym)
    throw std::runtime_error("factorize: LU needs non-symmetric storage");
  if (algo == hmat_factorization_lu && rows->n != cols->n)
    throw std::runtime_error("factorize: LU needs a square matrix");
  device_init();
  const double t0 = now_s();
  if (progress) {
    progress->max = embedded ? rows->n / 2 : rows->n;
    progress->current = 0;
  }
  if (!fstate_) fstate_ = std::make_shared<FactorState>();
  fstate_->solve_plans.clear();
  fstate_->drop_graphs();
  stats.tasks = stats.waves = stats.launches = stats.segments = 0;
  if (fp32 || value_type == HMAT_SIMPLE_PRECISION) {
    if (value_type == HMAT_DOUBLE_PRECISION) {
      init_store(s32);
      launch_convert<double, float>(s64.data(), s32.data(), s64.elems, stream);
      HMX_CHECK(hipMemcpyAsync(s32.ranks.p, s64.ranks.p, s64.hdesc.size() * sizeof(int),
                               hipMemcpyDeviceToDevice, stream));
      s32.valid = true;
    }
    Exec<float> ex(*this, s32, *fstate_);
    ex.set_lu(algo == hmat_factorization_lu);
    ex.set_ldlt(algo == hmat_factorization_ldlt);
    if (algo == hmat_factorization_lu) {
      pivots.alloc(sizeof(int) * (size_t)std::max(rows->n, 1));
      ex.lu(bt->root);
    } else if (algo == hmat_factorization_ldlt) {
      ldlt_diag.alloc(sizeof(float) * (size_t)std::max(rows->n, 1));
      ex.ldlt(bt->root);
    } else if (algo == hmat_factorization_hodlrsym) {
      hodlr_alloc(sizeof(float));
      ex.hodlr_factor(bt->root);
    } else if (algo == hmat_factorization_hodlr) {
      hodlr_alloc(sizeof(float));
      ex.hodlr_factor_ns(bt->root);
    } else {
      ex.llt(bt->root);
    }
    ex.flush(false);
    ex.check_flags();
    ex.report(now_s() - t0);
    factor32 = true;
  } else {
    Exec<double> ex(*this, s64, *fstate_);
    ex.set_lu(algo == hmat_factorization_lu);
    ex.set_ldlt(algo == hmat_factorization_ldlt);
    if (algo == hmat_factorization_lu) {
      pivots.alloc(sizeof(int) * (size_t)std::max(rows->n, 1));
      ex.lu(bt->root);
    } else if (algo == hmat_factorization_ldlt) {
      ldlt_diag.alloc(sizeof(double) * (size_t)std::max(rows->n, 1));
      ex.ldlt(bt->root);
    } else if (algo == hmat_factorization_hodlrsym) {
      hodlr_alloc(sizeof(double));
      ex.hodlr_factor(bt->root);
    } else if (algo == hmat_factorization_hodlr) {
      hodlr_alloc(sizeof(double));
      ex.hodlr_factor_ns(bt->root);
    } else {
      ex.llt(bt->root);
    }
    ex.flush(false);
    ex.check_flags();
    ex.report(now_s() - t0);
    factor32 = false;
  }
  factorization = algo;
  stats.t_factor = now_s() - t0;
}

template <typename T>
void Matrix::solve_device(T* b, int nrhs, int ldb, const Store<T>& st, int part) {
  if (factorization == -1) throw std::runtime_error("solve: matrix not factorised");
  if ((factor32 ? 1 : 0) != (sizeof(T) == 4 ? 1 : 0))
    throw std::runtime_error("solve: the factor is stored in the other precision");
  device_init();
  const double t0 = now_s();
  if (!fstate_) fstate_ = std::make_shared<FactorState>();
  FactorState& fs = *fstate_;
  const int n = rows->n;
  // The persistent-sweep solve (sweep.cpp, k_sweep.hip) did not complete on the
  // device and its cause is not found: it is switched off, and asking for it
  // (HMX_SOLVE_SWEEP=1) is an error.  The recorded task plans below solve.
  const char* sweep_env = getenv("HMX_SOLVE_SWEEP");
  if (sweep_env && atoi(sweep_env))
    throw std::runtime_error("solve: the persistent-sweep solve is disabled (it did not complete "
                             "on the device; cause not found)");
  if (part != 0 && (factorization == hmat_factorization_lu || factorization == hmat_factorization_hodlr))
    throw std::runtime_error("solve_lower_triangular: provided for LLt, LDLt and symmetric HODLR factors only");
  const auto key = std::make_pair((int)sizeof(T) + 16 * part, nrhs);
  const size_t bytes = (size_t)n * nrhs * sizeof(T);
  auto it = fs.solve_plans.find(key);
  if (it == fs.solve_plans.end()) fs.solve_rhs.erase(key);
  DevBuf& rhs = fs.solve_rhs[key];
  const bool inverses = factorization == hmat_factorization_llt || factorization == hmat_factorization_ldlt;
  if (it == fs.solve_plans.end() && inverses && inv_blocks.empty()) {
    // the dense inverses of the diagonal blocks of at most inv_max_rows rows
    std::function<void(const Block*)> pick = [&](const Block* b) {
      if (b->m() <= inv_max_rows || b->is_leaf()) {
        inv_index[b] = (int)inv_blocks.size();
        inv_blocks.push_back(InvBlock{b, 0});
        return;
      }
      for (int i = 0; i < b->nrow; ++i) pick(b->get(i, i));
    };
    pick(bt->root);
    int64_t off = 0;
    for (InvBlock& ib : inv_blocks) {
      ib.off = off;
      off += (int64_t)ib.b->m() * ib.b->m();
    }
    inv_buf.alloc(sizeof(T) * (size_t)std::max<int64_t>(off, 1));
    Exec<T> bx(*this, const_cast<Store<T>&>(st), fs);
    bx.build_inverses();
    bx.flush(false);
    bx.check_flags();
  }
  if (it == fs.solve_plans.end()) {
    // record the forward + backward substitution once for this nrhs
    fs.solve_rhs[key].ensure(bytes);
    PlanList plans;
    Exec<T> ex(*this, const_cast<Store<T>&>(st), fs);
    ex.record_into(&plans);
    ex.use_inverses(inverses);
    DV X;
    X.p = fs.solve_rhs[key].p;
    X.ld = n;
    X.m = n;
    X.n = nrhs;
    X.buf = kTempBase + (int64_t(1) << 30);  // tracked per leaf cluster of rows
    X.row0 = 0;
    HMX_CHECK(hipMemcpy2DAsync(rhs.p, (size_t)n * sizeof(T), b, (size_t)ldb * sizeof(T),
                               (size_t)n * sizeof(T), nrhs, hipMemcpyDeviceToDevice, stream));
    if (factorization == hmat_factorization_lu) {
      // HMatrix::solve (LU) h_matrix.cpp:2772-2779
      ex.lu_solve_lower_left(bt->root, X, Exec<T>::dim(nrhs));
      ex.lu_solve_upper_left(bt->root, X, Exec<T>::dim(nrhs));
    } else if (factorization == hmat_factorization_hodlr) {
      // HODLR<T>::solve (DefaultEngine::solve default_engine.cpp:203-204,
      // hodlr.cpp:434-437)
      ex.hodlr_solve_ns(bt->root, X, Exec<T>::dim(nrhs));
    } else if (factorization == hmat_factorization_hodlrsym) {
      // HODLR<T>::solveSymLower then solveSymUpper (DefaultEngine::solve
      // default_engine.cpp:205-208, hodlr.cpp:455-470)
      if (part != 2) ex.hodlr_lower(bt->root, X, Exec<T>::dim(nrhs));
      if (part != 1) ex.hodlr_upper(bt->root, X, Exec<T>::dim(nrhs));
    } else {
      // HMatrix::solveLlt h_matrix.cpp:3010-3018 (part 1: L x = b only, part 2:
      // L^T x = b only, as solveLower(b, transpose) of the solve_lower_triangular slot)
      // (LDLt, HMatrix::solveLdlt h_matrix.cpp:2990-3002: the unit L sweeps
      // with X <- D^{-1} X between them)
      if (part != 2) ex.solve_lower_left(bt->root, X, Exec<T>::dim(nrhs));
      if (part == 0 && factorization == hmat_factorization_ldlt) ex.solve_diag(X, Exec<T>::dim(nrhs));
      if (part != 1) ex.solve_upper_left(bt->root, X, Exec<T>::dim(nrhs));
    }
    ex.flush(false);
    it = fs.solve_plans.emplace(key, std::move(plans)).first;
  } else {
    HMX_CHECK(hipMemcpy2DAsync(rhs.p, (size_t)n * sizeof(T), b, (size_t)ldb * sizeof(T),
                               (size_t)n * sizeof(T), nrhs, hipMemcpyDeviceToDevice, stream));
    // replays run as one HIP graph: the substitution is a long chain of small
    // batched launches whose cost is mostly launch overhead.  Matrices driven
    // from several host threads replay eagerly (a capture in one thread makes
    // other threads' synchronous calls fail).
    int64_t wsmax = 1;
    for (const auto& p : it->second) wsmax = std::max(wsmax, p->ws);
    fs.ws.ensure((size_t)wsmax * sizeof(double));
    if (single_stream) {
      for (const auto& p : it->second) execute_plan<T>(*p, fs, stream, false, true);
      HMX_CHECK(hipMemcpy2DAsync(b, (size_t)ldb * sizeof(T), rhs.p, (size_t)n * sizeof(T),
                                 (size_t)n * sizeof(T), nrhs, hipMemcpyDeviceToDevice, stream));
      if (!user_stream) HMX_CHECK(hipStreamSynchronize(stream));
      stats.t_solve = now_s() - t0;
      return;
    }
    FactorState::SolveGraph& g = fs.solve_graphs[key];
    if (!g.exec || g.ws != fs.ws.p || g.rhs != rhs.p || g.ranks != fs.ranks.p ||
        g.flags != fs.flags.p) {
      g.release();
      HMX_CHECK(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
      try {
        for (const auto& p : it->second) execute_plan<T>(*p, fs, stream, true);
      } catch (...) {
        hipGraph_t broken = nullptr;
        (void)hipStreamEndCapture(stream, &broken);
        if (broken) (void)hipGraphDestroy(broken);
        throw;
      }
      HMX_CHECK(hipStreamEndCapture(stream, &g.graph));
      HMX_CHECK(hipGraphInstantiate(&g.exec, g.graph, nullptr, nullptr, 0));
      g.ws = fs.ws.p;
      g.rhs = rhs.p;
      g.ranks = fs.ranks.p;
      g.flags = fs.flags.p;
    }
    HMX_CHECK(hipGraphLaunch(g.exec, stream));
  }
  HMX_CHECK(hipMemcpy2DAsync(b, (size_t)ldb * sizeof(T), rhs.p, (size_t)n * sizeof(T),
                             (size_t)n * sizeof(T), nrhs, hipMemcpyDeviceToDevice, stream));
  if (!user_stream) HMX_CHECK(hipStreamSynchronize(stream));
  stats.t_solve = now_s() - t0;
}
template void Matrix::solve_device<double>(double*, int, int, const Store<double>&, int);
template void Matrix::solve_device<float>(float*, int, int, const Store<float>&, int);

template <typename T>
void Matrix::factor_gemv_device(char trans, const T* x, T* y, int nrhs, const Store<T>& st) {
  if (factorization != hmat_factorization_llt && factorization != hmat_factorization_hodlrsym)
    throw std::runtime_error("gemv of a factorised matrix: provided for LLt and HODLR factors "
                             "(the product with L, or with L W)");
  if ((factor32 ? 1 : 0) != (sizeof(T) == 4 ? 1 : 0))
    throw std::runtime_error("gemv: the factor is stored in the other precision");
  if (trans != 'N' && trans != 'T' && trans != 'C') throw std::runtime_error("gemv: bad trans");
  if (nrhs < 1) throw std::runtime_error("gemv: nrhs must be positive");
  device_init();
  if (!fstate_) fstate_ = std::make_shared<FactorState>();
  const int n = rows->n;
  DevBuf bx, by;
  bx.alloc(sizeof(T) * (size_t)n * nrhs);
  by.alloc(sizeof(T) * (size_t)n * nrhs);
  HMX_CHECK(hipMemcpyAsync(bx.p, x, sizeof(T) * (size_t)n * nrhs, hipMemcpyDeviceToDevice, stream));
  HMX_CHECK(hipMemsetAsync(by.p, 0, sizeof(T) * (size_t)n * nrhs, stream));
  HMX_CHECK(hipStreamSynchronize(stream));
  {
    Exec<T> ex(*this, const_cast<Store<T>&>(st), *fstate_);
    DV X, Y;
    X.p = bx.p;
    Y.p = by.p;
    X.ld = Y.ld = X.m = Y.m = n;
    X.n = Y.n = nrhs;
    X.buf = kTempBase + (int64_t(1) << 30) + 1;  // tracked per leaf cluster of rows
    Y.buf = kTempBase + (int64_t(1) << 30) + 2;
    X.row0 = Y.row0 = 0;
    const char t = trans == 'N' ? 'N' : 'T';
    if (factorization == hmat_factorization_hodlrsym)
      ex.hodlr_gemv(bt->root, t, X, Y, Exec<T>::dim(nrhs));
    else
      ex.tri_gemv(bt->root, t, X, Y, Exec<T>::dim(nrhs));
    ex.flush(false);
    ex.check_flags();
  }
  HMX_CHECK(hipMemcpyAsync(y, by.p, sizeof(T) * (size_t)n * nrhs, hipMemcpyDeviceToDevice, stream));
  HMX_CHECK(hipStreamSynchronize(stream));
}
template void Matrix::factor_gemv_device<double>(char, const double*, double*, int, const Store<double>&);

void Matrix::drop_solve_plans() {
  if (fstate_) {
    HMX_CHECK(hipStreamSynchronize(stream));
    fstate_->drop_graphs();
    fstate_->solve_graphs.clear();
    fstate_->solve_plans.clear();
    fstate_->solve_rhs.clear();
  }
  inv_blocks.clear();
  inv_index.clear();
  inv_buf.release();
}

// The solve's schedule picked by measurement: each candidate size of the
// diagonal blocks the solves apply as dense inverses is built, recorded and
// replayed on a random right-hand side block; the fastest replay is kept.
double Matrix::autotune_solve(int nrhs, int* best_rows) {
  if (factorization != hmat_factorization_llt && factorization != hmat_factorization_ldlt)
    throw std::runtime_error("autotune_solve: provided for LLt and LDLt factors");
  if (nrhs < 1) throw std::runtime_error("autotune_solve: nrhs must be positive");
  device_init();
  const int n = rows->n;
  const size_t esz = factor32 ? sizeof(float) : sizeof(double);
  DevBuf b;
  b.alloc(esz * (size_t)n * nrhs);
  HMX_CHECK(hipMemsetAsync(b.p, 0, b.bytes, stream));
  hipEvent_t e0, e1;
  HMX_CHECK(hipEventCreate(&e0));
  HMX_CHECK(hipEventCreate(&e1));
  struct Events {
    hipEvent_t a, b;
    ~Events() {
      (void)hipEventDestroy(a);
      (void)hipEventDestroy(b);
    }
  } ev{e0, e1};
  struct KeepUserStream {
    bool& flag;
    bool saved;
    ~KeepUserStream() { flag = saved; }
  } keep{user_stream, user_stream};
  user_stream = true;  // the solves below do not wait on their own
  double best = -1.0;
  int best_r = inv_max_rows;
  for (int r : {512, 1024, 2048, 4096}) {
    inv_max_rows = r;
    drop_solve_plans();
    auto solve = [&] {
      if (factor32)
        solve_device<float>(b.as<float>(), nrhs, n, s32);
      else
        solve_device<double>(b.as<double>(), nrhs, n, s64);
    };
    solve();  // inverses, recording
    solve();  // graph capture
    HMX_CHECK(hipEventRecord(e0, stream));
    for (int q = 0; q < 3; ++q) solve();
    HMX_CHECK(hipEventRecord(e1, stream));
    HMX_CHECK(hipEventSynchronize(e1));
    float ms = 0.f;
    HMX_CHECK(hipEventElapsedTime(&ms, e0, e1));
    if (best < 0 || ms / 3 < best) {
      best = ms / 3;
      best_r = r;
    }
  }
  inv_max_rows = best_r;
  drop_solve_plans();
  if (best_rows) *best_rows = best_r;
  return best;
}
template void Matrix::factor_gemv_device<float>(char, const float*, float*, int, const Store<float>&);

void Matrix::solve_host(void* b, int nrhs, bool permute, int part) {
  const int n = rows->n;
  auto run = [&](auto* tag) {
    using T = std::remove_pointer_t<decltype(tag)>;
    std::vector<T> h((size_t)n * nrhs);
    T* ub = (T*)b;
    if (permute) {
      for (int r = 0; r < nrhs; ++r)
        for (int i = 0; i < n; ++i) h[i + (size_t)r * n] = ub[rows->indices[i] + (size_t)r * n];
    } else {
      std::memcpy(h.data(), ub, h.size() * sizeof(T));
    }
    DevBuf d;
    upload(d, h.data(), h.size() * sizeof(T), stream);
    solve_device<T>(d.as<T>(), nrhs, n, store<T>(), part);
    HMX_CHECK(hipMemcpyAsync(h.data(), d.p, h.size() * sizeof(T), hipMemcpyDeviceToHost, stream));
    HMX_CHECK(hipStreamSynchronize(stream));
    if (permute) {
      for (int r = 0; r < nrhs; ++r)
        for (int i = 0; i < n; ++i) ub[rows->indices[i] + (size_t)r * n] = h[i + (size_t)r * n];
    } else {
      std::memcpy(ub, h.data(), h.size() * sizeof(T));
    }
  };
  device_init();
  if (value_type == HMAT_DOUBLE_PRECISION) {
    if (factor32) throw std::runtime_error("solve: the matrix holds a mixed-precision factor; use the refined solve");
    run((double*)nullptr);
  } else if (value_type == HMAT_SIMPLE_PRECISION) {
    run((float*)nullptr);
  } else {
    // complex matrices are stored as their real representation (double)
    throw std::runtime_error("solve: unsupported value type");
  }
}

void Matrix::release_temporaries() {
  if (!fstate_) return;
  HMX_CHECK(hipStreamSynchronize(stream));
  FactorState& fs = *fstate_;
  fs.drop_graphs();
  fs.solve_plans.clear();
  fs.solve_rhs.clear();
  fs.chunks.clear();
  fs.chunk = fs.chunk_used = fs.temp_total = fs.temp_used = 0;
  fs.temp_budget = 0;
  fs.ws.release();
}

void Matrix::truncate_all() {
  if (!assembled) throw std::runtime_error("truncate: matrix not assembled");
  if (factorization != -1) throw std::runtime_error("truncate: matrix has been factorised");
  device_init();
  if (!fstate_) fstate_ = std::make_shared<FactorState>();
  if (s64.valid) {
    Exec<double> ex(*this, s64, *fstate_);
    ex.truncate_leaves();
    ex.flush(false);
    ex.check_flags();
  } else if (s32.valid) {
    Exec<float> ex(*this, s32, *fstate_);
    ex.truncate_leaves();
    ex.flush(false);
    ex.check_flags();
  }
}

void Matrix::axpy_from(double alpha, Matrix& x) {
  if (!assembled || !x.assembled) throw std::runtime_error("axpy: matrix not assembled");
  if (factorization != -1 || x.factorization != -1)
    throw std::runtime_error("axpy: factorised matrices are not supported");
  if (x.rows != rows || x.cols != cols || x.lower_sym != lower_sym || x.value_type != value_type ||
      x.s64.hdesc.size() != s64.hdesc.size())
    throw std::runtime_error("axpy: the matrices must share cluster trees, storage and partition");
  for (size_t l = 0; l < s64.hdesc.size(); ++l) {
    const LeafDesc &a = s64.hdesc[l], &b = x.s64.hdesc[l];
    if (a.m != b.m || a.n != b.n || a.r0 != b.r0 || a.c0 != b.c0 || a.kind != b.kind || a.cap != b.cap)
      throw std::runtime_error("axpy: the matrices must share cluster trees, storage and partition");
  }
  device_init();
  x.device_init();
  HMX_CHECK(hipStreamSynchronize(x.stream));
  if (!fstate_) fstate_ = std::make_shared<FactorState>();
  if (s64.valid && x.s64.valid) {
    Exec<double> ex(*this, s64, *fstate_);
    ex.axpy_matrix(alpha, x.s64);
    ex.flush(false);
    ex.check_flags();
  } else if (s32.valid && x.s32.valid) {
    Exec<float> ex(*this, s32, *fstate_);
    ex.axpy_matrix(alpha, x.s32);
    ex.flush(false);
    ex.check_flags();
  } else {
    throw std::runtime_error("axpy: no data in a common precision");
  }
}

// The node list of a symmetric HODLR factor (HODLRNode::create hodlr.cpp:
// 394-404); the matrix must have the HODLR structure: every inner block a
// lower-symmetric 2 x 2 node whose (1, 0) child is an Rk leaf, every leaf
// below the diagonal Rk, the diagonal leaves full.
void Matrix::hodlr_layout(bool sym) {
  hodlr_nodes.clear();
  hodlr_index.clear();
  int64_t off = 0, poff = 0;
  std::function<void(const Block*)> walk = [&](const Block* b) {
    if (b->is_leaf()) {
      if (b->kind != Kind::Full) throw std::runtime_error("Not HODLR matrix");
      return;
    }
    const Block* b10 = b->nrow == 2 && b->ncol == 2 ? b->get(1, 0) : nullptr;
    if (!b->sym || !b10 || b->get(0, 1) || !b10->is_leaf() || b10->kind != Kind::Rk)
      throw std::runtime_error("Not HODLR matrix");
    const int cap = s64.hdesc.at(b10->leaf).cap;
    hodlr_index[b] = (int)hodlr_nodes.size();
    if (sym) {
      hodlr_nodes.push_back(HodlrNode{b, cap, off, off + (int64_t)cap * cap});
      off += 2 * (int64_t)cap * cap;
    } else {
      HodlrNode nd{b, cap, -1, off};
      off += 4 * (int64_t)cap * cap;
      nd.a01 = off;
      off += (int64_t)b10->n() * cap;
      nd.b01 = off;
      off += (int64_t)b10->m() * cap;
      nd.piv = poff;
      poff += 2 * (int64_t)cap;
      hodlr_nodes.push_back(nd);
    }
    walk(b->get(0, 0));
    walk(b->get(1, 1));
  };
  if (bt->root->is_leaf()) throw std::runtime_error("Not HODLR matrix");
  walk(bt->root);
  hodlr_elems = off;
  hodlr_pivs = poff;
}

void Matrix::hodlr_alloc(size_t elem) {
  hodlr_buf.alloc(elem * (size_t)std::max<int64_t>(hodlr_elems, 1));
  hodlr_piv.alloc(sizeof(int) * (size_t)std::max<int64_t>(hodlr_pivs, 1));
  hodlr_logdet.alloc(sizeof(double) * std::max<size_t>(hodlr_nodes.size(), 1));
  HMX_CHECK(hipMemsetAsync(hodlr_logdet.p, 0, hodlr_logdet.bytes, stream));
}

void Matrix::gemm_from(char ta, char tb, double alpha, Matrix& a, Matrix& b, double beta) {
  auto norm_t = [](char t) {
    if (t == 'N' || t == 'n') return 'N';
    if (t == 'T' || t == 't' || t == 'C' || t == 'c') return 'T';
    throw std::runtime_error(std::string("gemm: bad transpose flag ") + t);
  };
  ta = norm_t(ta);
  tb = norm_t(tb);
  if (&a == this || &b == this) throw std::runtime_error("gemm: C must differ from A and B");
  if (!assembled || !a.assembled || !b.assembled) throw std::runtime_error("gemm: matrix not assembled");
  if (factorization != -1 || a.factorization != -1 || b.factorization != -1)
    throw std::runtime_error("gemm: factorised matrices are not supported");
  if (a.value_type != value_type || b.value_type != value_type)
    throw std::runtime_error("gemm: the matrices must have the same scalar type");
  const ClusterTree* ar = ta == 'N' ? a.rows : a.cols;
  const ClusterTree* ac = ta == 'N' ? a.cols : a.rows;
  const ClusterTree* br = tb == 'N' ? b.rows : b.cols;
  const ClusterTree* bc = tb == 'N' ? b.cols : b.rows;
  if (ar != rows || bc != cols || ac != br)
    throw std::runtime_error("gemm: the matrices must share their cluster trees (op(A) rows = C rows, "
                             "op(B) cols = C cols, op(A) cols = op(B) rows)");
  if (beta != 1.0) scale(beta);
  device_init();
  a.device_init();
  b.device_init();
  HMX_CHECK(hipStreamSynchronize(a.stream));
  HMX_CHECK(hipStreamSynchronize(b.stream));
  if (!fstate_) fstate_ = std::make_shared<FactorState>();
  if (s64.valid && a.s64.valid && b.s64.valid) {
    Exec<double> ex(*this, s64, *fstate_);
    ex.gemm_matrices(ta, tb, alpha, a, a.s64, b, b.s64);
    ex.flush(false);
    ex.check_flags();
  } else if (s32.valid && a.s32.valid && b.s32.valid) {
    Exec<float> ex(*this, s32, *fstate_);
    ex.gemm_matrices(ta, tb, alpha, a, a.s32, b, b.s32);
    ex.flush(false);
    ex.check_flags();
  } else {
    throw std::runtime_error("gemm: no data in a common precision");
  }
}

void Matrix::solve_mat(Matrix& b, int part) {
  if (factorization == -1) throw std::runtime_error("solve_mat: matrix not factorised");
  if (factorization == hmat_factorization_hodlrsym || factorization == hmat_factorization_hodlr)
    throw std::runtime_error("solve_mat: H-matrix right-hand sides are not supported with HODLR factors");
  if (part != 0 && factorization != hmat_factorization_llt)
    throw std::runtime_error("solveLower only supported by LLt factorization.");
  if (&b == this) throw std::runtime_error("solve_mat: B must differ from the factorised matrix");
  if (!b.assembled) throw std::runtime_error("solve_mat: B not assembled");
  if (b.factorization != -1) throw std::runtime_error("solve_mat: B has been factorised");
  if (b.lower_sym) throw std::runtime_error("solve_mat: B must be stored non-symmetric");
  if (b.rows != cols) throw std::runtime_error("solve_mat: B's rows must use the matrix's column tree");
  if (b.value_type != value_type) throw std::runtime_error("solve_mat: B has another scalar type");
  if (factor32 && value_type == HMAT_DOUBLE_PRECISION)
    throw std::runtime_error("solve_mat: mixed-precision factors are not supported");
  device_init();
  b.device_init();
  HMX_CHECK(hipStreamSynchronize(b.stream));
  if (!fstate_) fstate_ = std::make_shared<FactorState>();
  if (factor32) {
    if (!b.s32.valid) throw std::runtime_error("solve_mat: B has no single-precision data");
    Exec<float> ex(*this, s32, *fstate_);
    ex.solve_matrix(b, b.s32, part);
    ex.flush(false);
    ex.check_flags();
  } else {
    if (!b.s64.valid) throw std::runtime_error("solve_mat: B has no double-precision data");
    Exec<double> ex(*this, s64, *fstate_);
    ex.solve_matrix(b, b.s64, part);
    ex.flush(false);
    ex.check_flags();
  }
}

double Matrix::logdet() {
  // HMatrix::logdet h_matrix.cpp:2925-2936: sum of log of the diagonal of L
  if (factorization != hmat_factorization_llt && factorization != hmat_factorization_hodlrsym)
    throw std::runtime_error("logdet is only supported for LLt or HODLR factorized matrices.");
  double r = 0.0;
  if (factorization == hmat_factorization_hodlrsym) {
    // logdet hodlr.cpp:158-166: the leaves' log det L plus log det M11 per node
    std::vector<double> ld(hodlr_nodes.size());
    HMX_CHECK(hipMemcpy(ld.data(), hodlr_logdet.p, ld.size() * sizeof(double), hipMemcpyDeviceToHost));
    for (double v : ld) r += v;
  }
  for (const Block* b : bt->leaves) {
    if (b->kind != Kind::Full || b->rows != b->cols) continue;
    const LeafDesc& L = s64.hdesc[b->leaf];
    const int n = L.m;
    if (factor32) {
      std::vector<float> d((size_t)n * n);
      HMX_CHECK(hipMemcpy(d.data(), s32.data() + L.a, d.size() * sizeof(float), hipMemcpyDeviceToHost));
      for (int i = 0; i < n; ++i) r += std::log((double)d[i + (size_t)i * n]);
    } else {
      std::vector<double> d((size_t)n * n);
      HMX_CHECK(hipMemcpy(d.data(), s64.data() + L.a, d.size() * sizeof(double), hipMemcpyDeviceToHost));
      for (int i = 0; i < n; ++i) r += std::log(d[i + (size_t)i * n]);
    }
  }
  return r;
}

// Iterative refinement: x_{k+1} = x_k + (L L^T)^{-1} (b - A x_k) with the
// FP32 factor for the correction and the FP64 H-matrix for the residual.
// Refinement of the FP32 factor's solution against the FP64 H-matrix.  For
// an LU factor (A not symmetric) plain sweeps x += M (b - A x); for LLt,
// accelerated by conjugate gradients: per right-hand side, CG on A x = b
// with the FP32 H-LLt solve as preconditioner (z = (L L^T)^{-1} r in FP32),
// A p by the FP64 H-gemv.  Each iteration costs one FP32 solve and one FP64
// gemv, as a plain refinement sweep, and converges at the rate of the
// preconditioned CG instead of the contraction of the sweep.  When the
// recursive residual meets tol, the true residual b - A x is checked; if it
// does not, CG restarts from it.  Returns the iterations (FP32 solves) used.
int Matrix::solve_refined(double* b, int nrhs, int ldb, double tol, int max_iter, double* resid) {
  if (!factor32 || factorization == -1) throw std::runtime_error("refined solve needs hmx_factorize_mixed first");
  if (!s64.valid) throw std::runtime_error("refined solve needs the FP64 matrix");
  device_init();
  if (!fstate_) fstate_ = std::make_shared<FactorState>();
  FactorState& fs = *fstate_;
  const int n = rows->n;
  const size_t cnt = (size_t)n * nrhs;
  for (DevBuf* d : {&fs.ir_x, &fs.ir_r, &fs.ir_p, &fs.ir_q, &fs.ir_z}) d->ensure(cnt * sizeof(double));
  fs.ir_norms.ensure(2 * nrhs * sizeof(double));
  fs.ir_dots.ensure(4 * nrhs * sizeof(double));
  DevBuf d32;
  d32.alloc(cnt * sizeof(float));
  double* x = fs.ir_x.as<double>();
  double* r = fs.ir_r.as<double>();
  double* p = fs.ir_p.as<double>();
  double* q = fs.ir_q.as<double>();
  double* z = fs.ir_z.as<double>();
  double* rz[2] = {fs.ir_dots.as<double>(), fs.ir_dots.as<double>() + nrhs};
  double* pq = fs.ir_dots.as<double>() + 2 * nrhs;
  std::vector<double> bn(nrhs), rn(nrhs);
  auto residual_rel = [&](const double* v) {
    launch_col_sumsq(v, n, nrhs, fs.ir_norms.as<double>(), stream);
    HMX_CHECK(hipMemcpyAsync(rn.data(), fs.ir_norms.p, nrhs * sizeof(double), hipMemcpyDeviceToHost, stream));
    HMX_CHECK(hipStreamSynchronize(stream));
    double rel = 0.0;
    for (int j = 0; j < nrhs; ++j) rel = std::max(rel, bn[j] > 0 ? std::sqrt(rn[j] / bn[j]) : 0.0);
    return rel;
  };
  HMX_CHECK(hipMemcpy2DAsync(r, (size_t)n * sizeof(double), b, (size_t)ldb * sizeof(double),
                             (size_t)n * sizeof(double), nrhs, hipMemcpyDeviceToDevice, stream));
  HMX_CHECK(hipMemsetAsync(x, 0, cnt * sizeof(double), stream));
  launch_col_sumsq(r, n, nrhs, fs.ir_norms.as<double>(), stream);
  HMX_CHECK(hipMemcpyAsync(bn.data(), fs.ir_norms.p, nrhs * sizeof(double), hipMemcpyDeviceToHost, stream));
  HMX_CHECK(hipStreamSynchronize(stream));
  int it = 0, cur = 0;
  double rel = 1.0;
  const bool saved = user_stream;
  user_stream = true;  // keep the inner calls stream-ordered
  // z = M r (FP32 solve), rz[cur] = r . z
  auto precondition = [&]() {
    launch_convert<double, float>(r, d32.as<float>(), cnt, stream);
    solve_device<float>(d32.as<float>(), nrhs, n, s32);
    launch_convert<float, double>(d32.as<float>(), z, cnt, stream);
    ++it;
  };
  try {
    if (factorization != hmat_factorization_llt) {
      // LU (non-symmetric A) and LDLt (A possibly indefinite): plain
      // refinement sweeps x += M r, r = b - A x
      while (it < max_iter) {
        precondition();
        launch_vec_add(x, z, cnt, stream);
        HMX_CHECK(hipMemcpy2DAsync(r, (size_t)n * sizeof(double), b, (size_t)ldb * sizeof(double),
                                   (size_t)n * sizeof(double), nrhs, hipMemcpyDeviceToDevice, stream));
        gemv_device<double>(-1.0, x, 1.0, r, nrhs, n, n, s64);
        rel = residual_rel(r);
        if (rel <= tol) break;
      }
    }
    bool fresh = true;  // p must restart from z
    while (factorization == hmat_factorization_llt && it < max_iter) {
      precondition();
      const int nxt = cur ^ 1;
      launch_col_dot(r, z, n, nrhs, rz[nxt], stream);
      if (fresh) {
        HMX_CHECK(hipMemcpyAsync(p, z, cnt * sizeof(double), hipMemcpyDeviceToDevice, stream));
        fresh = false;
      } else {
        launch_cg_p(p, z, rz[nxt], rz[cur], n, nrhs, stream);
      }
      cur = nxt;
      gemv_device<double>(1.0, p, 0.0, q, nrhs, n, n, s64);
      launch_col_dot(p, q, n, nrhs, pq, stream);
      launch_cg_xr(x, r, p, q, rz[cur], pq, n, nrhs, stream);
      rel = residual_rel(r);
      if (rel <= tol) {
        // confirm with the true residual; restart from it if the recursion drifted
        HMX_CHECK(hipMemcpy2DAsync(r, (size_t)n * sizeof(double), b, (size_t)ldb * sizeof(double),
                                   (size_t)n * sizeof(double), nrhs, hipMemcpyDeviceToDevice, stream));
        gemv_device<double>(-1.0, x, 1.0, r, nrhs, n, n, s64);
        rel = residual_rel(r);
        if (rel <= tol) break;
        fresh = true;
      }
    }
  } catch (...) {
    user_stream = saved;
    throw;
  }
  user_stream = saved;
  HMX_CHECK(hipMemcpy2DAsync(b, (size_t)ldb * sizeof(double), x, (size_t)n * sizeof(double),
                             (size_t)n * sizeof(double), nrhs, hipMemcpyDeviceToDevice, stream));
  HMX_CHECK(hipStreamSynchronize(stream));
  if (resid) *resid = rel;
  return it;
}

}  // namespace hmx
