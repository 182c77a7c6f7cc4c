// Item lists of the persistent H-LLt solve sweeps (the layout and item kinds
// are described in sweep.hpp) and the launches.
//
// Forward sweep, L y = b (HMatrix::solveLowerTriangularLeft h_matrix.cpp:2338-2372):
// block row q of L holds full leaves (q, j < q), applied as L_qj y_j, and rows
// of Rk leaves (I, J) with q in I, applied as A[q] (B^T y_J).  Backward sweep,
// L^T x = y (HMatrix::solveUpperTriangularLeft h_matrix.cpp:2504-2540): block
// column q of L, applied transposed: L_iq^T x_i for full leaves (i > q, q),
// B[q] (A^T x_I) for Rk leaves (I, J) with q in J.  The diagonal leaf is
// applied through its inverse, formed once per factorisation by the TRSM task
// kernel (X <- L^{-1} X on a unit matrix).
#include "sweep_host.hpp"

#include <algorithm>
#include <stdexcept>

#include "factor_kernels.hpp"

namespace hmx {

namespace {
constexpr int kMaxRows = kSweepChunk;  // leaf clusters and Rk capacities the kernel stages
constexpr double kWaitSeconds = 2.0;    // a wait longer than this is a failure
constexpr int kMaxGroups = 1024;       // 4 workgroups per compute unit

template <typename V>
void put(DevBuf& d, const std::vector<V>& v, hipStream_t s) {
  upload(d, v.data(), v.size() * sizeof(V), s);
}
}  // namespace

std::unique_ptr<SweepSolver> SweepSolver::build(const Matrix& M, std::string* why) {
  auto fail = [&](const std::string& w) -> std::unique_ptr<SweepSolver> {
    if (why) *why = w;
    return nullptr;
  };
  if (!M.lower_sym || M.rows != M.cols) return fail("not lower-symmetric storage");
  const ClusterTree& ct = *M.rows;
  const auto& RL = ct.leaves;
  const int Q = (int)RL.size();
  for (const Cluster* c : RL)
    if (c->size > kMaxRows || c->size < 1) return fail("leaf cluster outside 1..256 rows");
  std::vector<int> row2leaf(ct.n);
  for (int q = 0; q < Q; ++q)
    for (int i = RL[q]->offset; i < RL[q]->offset + RL[q]->size; ++i) row2leaf[i] = q;
  const auto& D = M.s64.hdesc;
  auto sv = std::make_unique<SweepSolver>();
  SweepSolver& S = *sv;
  S.nrows_ = Q;
  S.diag_leaf_.assign(Q, -1);

  struct Keyed {
    int key;
    SwContrib c;
  };
  std::vector<std::vector<Keyed>> per[2];
  per[0].resize(Q);
  per[1].resize(Q);
  std::vector<std::vector<SwChunk>> chunk_at[2];  // chunks bucketed by dep position
  chunk_at[0].resize(Q);
  chunk_at[1].resize(Q);
  int64_t zoff = 0, zpoff = 0;
  std::vector<int64_t> slot_z;  // z offset of each Rk slot
  const auto& leaves = M.bt->leaves;
  for (size_t l = 0; l < leaves.size(); ++l) {
    const Block* b = leaves[l];
    const LeafDesc& L = D[l];
    if (b->kind == Kind::Full) {
      if (!b->rows->is_leaf() || !b->cols->is_leaf()) return fail("full leaf spanning several clusters");
      const int qr = b->rows->leaf_index, qc = b->cols->leaf_index;
      if (qr == qc) {
        S.diag_leaf_[qr] = (int)l;
        continue;
      }
      if (qr < qc) return fail("full leaf above the diagonal");
      per[0][qr].push_back({qc, SwContrib{SC_FULL, (int)l, RL[qc]->offset, qc, 0, 0, 0}});
      per[1][qc].push_back({Q - 1 - qr, SwContrib{SC_FULL_T, (int)l, RL[qr]->offset, Q - 1 - qr, 0, 0, 0}});
      continue;
    }
    if (L.cap > kMaxRows) return fail("Rk capacity above 256");
    const Cluster* I = b->rows;
    const Cluster* J = b->cols;
    const int i0 = I->first_leaf, i1 = I->first_leaf + I->nleaves - 1;
    const int j0 = J->first_leaf, j1 = J->first_leaf + J->nleaves - 1;
    if (j1 >= i0) return fail("Rk leaf touching the diagonal");
    const int slot = S.nslots_++;
    const int64_t z = zoff;
    slot_z.push_back(z);
    zoff += (int64_t)L.cap * kSweepNR;
    // forward: z = B^T y_J over chunks of J's rows; complete when the last row is done
    // backward: w = A^T x_I over chunks of I's rows; complete when the first row is done
    for (int d = 0; d < 2; ++d) {
      const int len = d == 0 ? L.n : L.m;
      const int base = d == 0 ? J->offset : I->offset;
      const int nch = (len + kSweepChunk - 1) / kSweepChunk;
      for (int r0 = 0; r0 < len; r0 += kSweepChunk) {
        SwChunk c{};
        c.leaf = (int)l;
        c.slot = slot;
        c.row0 = r0;
        c.rows = std::min(kSweepChunk, len - r0);
        c.dep = d == 0 ? row2leaf[base + r0 + c.rows - 1] : Q - 1 - row2leaf[base + r0];
        c.single = nch == 1;
        if (nch == 1) {
          c.out = z;
        } else {
          c.out = zpoff;
          zpoff += (int64_t)L.cap * kSweepNR;
        }
        chunk_at[d][c.dep].push_back(c);
      }
    }
    for (int q = i0; q <= i1; ++q)
      per[0][q].push_back({j1, SwContrib{SC_RK, (int)l, slot, 0, RL[q]->offset - I->offset, 0, z}});
    for (int q = j0; q <= j1; ++q)
      per[1][q].push_back({Q - 1 - i0, SwContrib{SC_RK, (int)l, slot, 0, RL[q]->offset - J->offset, 0, z}});
  }
  // inverses of the diagonal leaves, in hmat order
  S.hdiag_.resize(Q);
  for (int q = 0; q < Q; ++q) {
    const int l = S.diag_leaf_[q];
    if (l < 0) return fail("cluster without a diagonal full leaf");
    if (D[l].m != RL[q]->size || D[l].n != RL[q]->size) return fail("diagonal leaf size");
    S.hdiag_[q] = SwRow{RL[q]->offset, RL[q]->size, 0, 0, S.inv_elems_};
    S.inv_elems_ += ((int64_t)RL[q]->size * RL[q]->size + 63) / 64 * 64;
  }
  const hipStream_t st = M.stream;
  for (int d = 0; d < 2; ++d) {
    // ticket order: row p, then the chunks its completion completes
    std::vector<SwItem> order;
    std::vector<SwRow> rows(Q);
    std::vector<SwContrib> contrib;
    std::vector<SwChunk> pending;  // chunk records in ticket order
    for (int pos = 0; pos < Q; ++pos) {
      const int q = d == 0 ? pos : Q - 1 - pos;
      auto& list = per[d][q];
      std::stable_sort(list.begin(), list.end(),
                       [](const Keyed& a, const Keyed& b) { return a.key < b.key; });
      SwRow R = S.hdiag_[q];
      R.cbeg = (int)contrib.size();
      for (const Keyed& k : list) contrib.push_back(k.c);
      R.cend = (int)contrib.size();
      rows[pos] = R;
      order.push_back(SwItem{SW_ROW, pos});
      for (const SwChunk& c : chunk_at[d][pos]) {
        order.push_back(SwItem{SW_ZPART, (int)pending.size()});
        pending.push_back(c);
      }
    }
    // chunk records grouped by leaf (a reduction reads a contiguous range);
    // the reduction of a multi-chunk leaf follows its last chunk in ticket order
    std::vector<std::vector<int>> of(S.nslots_);
    for (size_t c = 0; c < pending.size(); ++c) of[pending[c].slot].push_back((int)c);
    std::vector<int> where(pending.size()), beg(S.nslots_);
    std::vector<SwChunk> chunks;
    for (int sl = 0; sl < S.nslots_; ++sl) {
      beg[sl] = (int)chunks.size();
      for (int c : of[sl]) {
        where[c] = (int)chunks.size();
        chunks.push_back(pending[c]);
      }
    }
    std::vector<SwItem> items;
    std::vector<SwRed> reds;
    std::vector<int> seen(S.nslots_, 0);
    for (const SwItem& it : order) {
      if (it.type != SW_ZPART) {
        items.push_back(it);
        continue;
      }
      const SwChunk& c = pending[it.idx];
      items.push_back(SwItem{SW_ZPART, where[it.idx]});
      const int cnt = (int)of[c.slot].size();
      if (++seen[c.slot] == cnt && cnt > 1) {
        items.push_back(SwItem{SW_ZRED, (int)reds.size()});
        reds.push_back(SwRed{c.leaf, c.slot, beg[c.slot], beg[c.slot] + cnt, slot_z[c.slot]});
      }
    }
    Dir& Dd = S.dir_[d];
    Dd.nitems = (int)items.size();
    put(Dd.items, items, st);
    put(Dd.rows, rows, st);
    put(Dd.contrib, contrib, st);
    put(Dd.chunks, chunks, st);
    put(Dd.reds, reds, st);
  }
  S.z_elems_ = std::max<int64_t>(zoff, 1);
  S.zpart_elems_ = std::max<int64_t>(zpoff, 1);
  S.z_.alloc(S.z_elems_ * sizeof(double));
  S.zpart_.alloc(S.zpart_elems_ * sizeof(double));
  S.sync_.alloc((size_t)sweep_sync_ints(Q, S.nslots_) * sizeof(int));
  S.fail_.alloc(kSwFailInts * sizeof(int));
  HMX_CHECK(hipMemsetAsync(S.fail_.p, 0, kSwFailInts * sizeof(int), st));
  int dev = 0, khz = 0;
  HMX_CHECK(hipGetDevice(&dev));
  HMX_CHECK(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev));
  if (khz <= 0) return fail("no wall clock rate");
  S.wait_ticks_ = (int64_t)(kWaitSeconds * 1e3 * khz);
  S.flags_.alloc(2 * sizeof(int));
  HMX_CHECK(hipMemsetAsync(S.flags_.p, 0, 2 * sizeof(int), st));
  put(S.diag_rows_, S.hdiag_, st);
  return sv;
}

template <typename T>
void SweepSolver::ensure_inverse(const Store<T>& st, hipStream_t s) {
  const int p = sizeof(T) == 8 ? 0 : 1;
  if (inv_ready_[p]) return;
  DevBuf& inv = p == 0 ? inv64_ : inv32_;
  inv.alloc((size_t)inv_elems_ * sizeof(T));
  launch_sweep_identity<T>(inv.as<T>(), diag_rows_.as<SwRow>(), nrows_, s);
  std::vector<FTask> tasks(nrows_);
  for (int q = 0; q < nrows_; ++q) {
    const LeafDesc& L = st.hdesc[diag_leaf_[q]];
    FTask t{};
    t.type = T_TRSM;
    t.mode = 1;  // X <- L^{-1} X
    t.n = Dim{L.m, nullptr};
    t.m = Dim{L.m, nullptr};
    t.A = st.data() + L.a;
    t.lda = L.m;
    t.C = inv.as<T>() + hdiag_[q].inv;
    t.ldc = L.m;
    tasks[q] = t;
  }
  put(trsm_tasks_, tasks, s);
  launch_task_kernel<T>(T_TRSM, trsm_tasks_.as<FTask>(), nrows_, nullptr, 0, nullptr,
                        flags_.as<int>(), s);
  inv_ready_[p] = true;
}

template <typename T>
void SweepSolver::solve(Matrix& M, const Store<T>& st, T* x, int nrhs, int ldx, hipStream_t s) {
  ensure_inverse<T>(st, s);
  const T* inv = (sizeof(T) == 8 ? inv64_ : inv32_).template as<T>();
  for (int c0 = 0; c0 < nrhs; c0 += kSweepNR) {
    const int nr = std::min(kSweepNR, nrhs - c0);
    for (int d = 0; d < 2; ++d) {
      const Dir& Dd = dir_[d];
      SweepLaunch L{};
      L.items = Dd.items.as<SwItem>();
      L.rows = Dd.rows.as<SwRow>();
      L.contrib = Dd.contrib.as<SwContrib>();
      L.chunks = Dd.chunks.as<SwChunk>();
      L.reds = Dd.reds.as<SwRed>();
      L.desc = st.desc.template as<LeafDesc>();
      L.ranks = st.ranks.template as<int>();
      L.z = z_.as<double>();
      L.zpart = zpart_.as<double>();
      L.sync = sync_.as<int>();
      L.fail = fail_.as<int>();
      L.wait_ticks = wait_ticks_;
      L.nitems = Dd.nitems;
      L.nslots = nslots_;
      L.nrows = nrows_;
      L.fwd = d == 0;
      L.nrhs = nr;
      L.ldx = ldx;
      L.max_groups = kMaxGroups;
      launch_sweep<T>(L, st.data(), inv, x + (size_t)c0 * ldx, s);
    }
  }
  (void)M;
}

void SweepSolver::check(hipStream_t s) {
  HMX_CHECK(hipStreamSynchronize(s));
  int f[kSwFailInts] = {};
  HMX_CHECK(hipMemcpy(f, fail_.p, sizeof(f), hipMemcpyDeviceToHost));
  if (f[0]) {
    static const char* what[3] = {"row done flag", "z ready flag", "chunk count"};
    throw std::runtime_error(
        "solve: a triangular sweep gave up waiting (item " + std::to_string(f[1]) + ", " +
        what[f[2] >= 0 && f[2] < 3 ? f[2] : 0] + " " + std::to_string(f[3]) + ", awaited " +
        std::to_string(f[4]) + ", read " + std::to_string(f[5]) + ")");
  }
}

template void SweepSolver::solve<double>(Matrix&, const Store<double>&, double*, int, int, hipStream_t);
template void SweepSolver::solve<float>(Matrix&, const Store<float>&, float*, int, int, hipStream_t);

}  // namespace hmx
