// Matrix construction, HBM storage layout, GPU assembly and the H-gemv driver.
#include "engine.hpp"

#include <algorithm>
#include <utility>

#include "factor_kernels.hpp"
#include <chrono>
#include <cstring>
#include <functional>
#include <stdexcept>

namespace hmx {

void DevBuf::release() {
  if (p) (void)hipFree(p);
  p = nullptr;
  bytes = 0;
}

void DevBuf::alloc(size_t nbytes) {
  release();
  if (nbytes == 0) nbytes = 256;
  HMX_CHECK(hipMalloc(&p, nbytes));
  bytes = nbytes;
}

void DevBuf::ensure(size_t nbytes) {
  if (nbytes > bytes) alloc(nbytes);
}

void upload(DevBuf& b, const void* src, size_t bytes, hipStream_t s) {
  b.ensure(bytes);
  if (!bytes) return;
  // the source is ordinary (pageable) host memory that callers reuse or free
  // right after: wait until the copy has read it
  HMX_CHECK(hipMemcpyAsync(b.p, src, bytes, hipMemcpyHostToDevice, s));
  HMX_CHECK(hipStreamSynchronize(s));
}

template <>
Store<double>& Matrix::store<double>() {
  return s64;
}
template <>
Store<float>& Matrix::store<float>() {
  return s32;
}

Matrix::Matrix(const ClusterTree* r, const ClusterTree* c, bool sym,
               const AdmissibilityParams& adm, int vt, int rank_cap)
    : adm(adm), value_type(vt), rows(r), cols(c), lower_sym(sym), rk_cap_max(rank_cap) {
  if (sym && r != c) throw std::runtime_error("lower-symmetric storage needs rows tree == cols tree");
  bt = build_block_tree(*r, *c, adm, sym);
  build_storage();
}

Matrix::Matrix(const ClusterTree* r, const ClusterTree* c, bool sym, std::unique_ptr<BlockTree> tree,
               int vt, int rank_cap, const std::vector<int>& cap_hint)
    : value_type(vt), rows(r), cols(c), lower_sym(sym), rk_cap_max(rank_cap) {
  if (sym && r != c) throw std::runtime_error("lower-symmetric storage needs rows tree == cols tree");
  bt = std::move(tree);
  build_storage(&cap_hint);
}

Matrix* Matrix::create_complex(const ClusterTree* r, const ClusterTree* c,
                               const AdmissibilityParams& adm, int rank_cap,
                               std::shared_ptr<ClusterTree> er, std::shared_ptr<ClusterTree> ec,
                               int vt) {
  if (!er) er = std::shared_ptr<ClusterTree>(embed_cluster_tree(*r));
  if (!ec) ec = c == r ? er : std::shared_ptr<ClusterTree>(embed_cluster_tree(*c));
  auto zt = build_block_tree(*r, *c, adm, false);
  auto et = embed_block_tree(*zt, *r, *er, *c, *ec);
  // a complex rank k is a real rank 2k in the embedding
  auto* m = new Matrix(er.get(), ec.get(), false, std::move(et), vt, 2 * rank_cap, {});
  m->adm = adm;
  m->embedded = true;
  m->zrows = r;
  m->zcols = c;
  m->erows = er;
  m->ecols = ec;
  return m;
}

void Matrix::device_init() {
  if (own_stream) return;
  HMX_CHECK(hipStreamCreateWithFlags(&own_stream, hipStreamNonBlocking));
  if (!stream) stream = own_stream;
  overflow_.alloc(sizeof(int));
}

Matrix::~Matrix() {
  fstate_.reset();
  if (own_stream) {
    (void)hipStreamSynchronize(own_stream);
    (void)hipStreamDestroy(own_stream);
  }
}

// One arena layout shared by both precisions: full leaves m*n, Rk leaves
// (m + n) * cap with cap = min(m, n, rk_cap_max); every leaf 64-element aligned.
void Matrix::build_storage(const std::vector<int>* cap_hint) {
  std::vector<LeafDesc> d(bt->leaves.size());
  int64_t off = 0;
  auto align = [](int64_t x) { return (x + 63) / 64 * 64; };
  for (size_t i = 0; i < bt->leaves.size(); ++i) {
    const Block* b = bt->leaves[i];
    LeafDesc& L = d[i];
    L.m = b->m();
    L.n = b->n();
    L.r0 = b->rows->offset;
    L.c0 = b->cols->offset;
    if (b->kind == Kind::Full) {
      L.kind = 0;
      L.cap = 0;
      L.a = off;
      L.b = -1;
      off = align(off + (int64_t)L.m * L.n);
    } else {
      L.kind = 1;
      L.cap = std::min(std::min(L.m, L.n),
                       std::max(rk_cap_max, cap_hint && i < cap_hint->size() ? (*cap_hint)[i] : 0));
      L.a = off;
      off = align(off + (int64_t)L.m * L.cap);
      L.b = off;
      off = align(off + (int64_t)L.n * L.cap);
    }
  }
  s64.hdesc = d;
  s32.hdesc = d;
  s64.elems = s32.elems = (size_t)off;
}

template <typename T>
void Matrix::init_store(Store<T>& st) {
  if (!st.arena.p) {
    st.arena.alloc(st.elems * sizeof(T));
    st.ranks.alloc(st.hdesc.size() * sizeof(int));
    upload(st.desc, st.hdesc.data(), st.hdesc.size() * sizeof(LeafDesc), stream);
    std::vector<int> rk(st.hdesc.size());
    for (size_t i = 0; i < rk.size(); ++i) rk[i] = st.hdesc[i].kind == 0 ? -1 : 0;
    HMX_CHECK(hipMemcpyAsync(st.ranks.p, rk.data(), rk.size() * sizeof(int),
                             hipMemcpyHostToDevice, stream));
    HMX_CHECK(hipStreamSynchronize(stream));
  }
}
template void Matrix::init_store<double>(Store<double>&);
template void Matrix::init_store<float>(Store<float>&);

namespace {
double now() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

void upload_points(const KernelSpec& k, const ClusterTree& t, DevBuf& pts, DevBuf& user,
                   hipStream_t s) {
  if (k.n != t.n) throw std::runtime_error("kernel point count differs from the cluster tree size");
  std::vector<double> soa(3 * (size_t)t.n, 0.0);
  for (int i = 0; i < t.n; ++i) {
    const int u = t.indices[i];
    for (int d = 0; d < k.dim && d < 3; ++d) soa[(size_t)d * t.n + i] = k.points[(size_t)u * k.dim + d];
  }
  upload(pts, soa.data(), soa.size() * sizeof(double), s);
  upload(user, t.indices.data(), t.n * sizeof(int), s);
  HMX_CHECK(hipStreamSynchronize(s));
}
}  // namespace

// truncatedSvd (compression.cpp:224-239) of leaves of 100 rows / cols or more:
// each dense block (column-major, ld m) at dense + off[i] becomes the leaf's
// Rk panels by a dense recompression task (a one-sided Jacobi SVD of the
// block, cut at eps) in batches bounded by the tasks' workspace.
void Matrix::svd_compress_dense(const std::vector<std::pair<int, int64_t>>& blocks,
                                const double* dense, double eps) {
  const int64_t budget = (int64_t)1 << 28;  // doubles of workspace per launch (2 GiB)
  DevBuf dt, dws, dfl;
  dfl.alloc(2 * sizeof(int));
  size_t i = 0;
  while (i < blocks.size()) {
    std::vector<FTask> tasks;
    int64_t ws = 0;
    while (i < blocks.size()) {
      const LeafDesc& L = s64.hdesc[blocks[i].first];
      const int64_t need = fadd_dense_workspace(L.m, L.n);
      if (!tasks.empty() && ws + need > budget) break;
      FTask t{};
      t.type = T_FADD_DENSE;
      t.mode = 1;  // the target's panels are replaced
      t.eps = eps;
      t.alpha = 1.0;
      t.A = const_cast<double*>(dense) + blocks[i].second;
      t.lda = L.m;
      t.m = Dim{L.m, nullptr};
      t.n = Dim{L.n, nullptr};
      t.tgt.a = s64.data() + L.a;
      t.tgt.b = s64.data() + L.b;
      t.tgt.la = L.m;
      t.tgt.lb = L.n;
      t.tgt.m = L.m;
      t.tgt.n = L.n;
      t.tgt.cap = L.cap;
      t.tgt.rank = s64.ranks.as<int>() + blocks[i].first;
      t.ws = ws;
      ws += need;
      tasks.push_back(t);
      ++i;
    }
    upload(dt, tasks.data(), tasks.size() * sizeof(FTask), stream);
    dws.ensure(sizeof(double) * (size_t)std::max<int64_t>(ws, 1));
    HMX_CHECK(hipMemsetAsync(dfl.p, 0, 2 * sizeof(int), stream));
    launch_task_kernel<double>(T_FADD_DENSE, dt.as<FTask>(), (int)tasks.size(), nullptr, 0,
                               dws.as<double>(), dfl.as<int>(), stream);
    int fl[2];
    HMX_CHECK(hipMemcpyAsync(fl, dfl.p, sizeof(fl), hipMemcpyDeviceToHost, stream));
    HMX_CHECK(hipStreamSynchronize(stream));
    if (fl[0]) HMX_CHECK(hipMemcpy(overflow_.p, &fl[0], sizeof(int), hipMemcpyHostToDevice));
  }
}

// An Rk rank above the panel capacity during assembly: every leaf's capacity
// limit doubles (up to its min(m, n)) and the assembly runs again.
bool Matrix::assembly_overflow() {
  HMX_CHECK(hipStreamSynchronize(stream));
  int ov = 0;
  HMX_CHECK(hipMemcpy(&ov, overflow_.p, sizeof(int), hipMemcpyDeviceToHost));
  return ov != 0;
}

void Matrix::grow_rank_capacity() {
  int widest = 0;
  for (const LeafDesc& L : s64.hdesc)
    if (L.kind == 1) widest = std::max(widest, std::min(L.m, L.n));
  if (rk_cap_max >= widest) throw std::runtime_error("assembly: an Rk rank exceeded the panel capacity");
  rk_cap_max = std::min(2 * rk_cap_max, widest);
  s64.arena.release();
  s64.ranks.release();
  s64.desc.release();
  s64.valid = false;
  build_storage(cap_hint_.empty() ? nullptr : &cap_hint_);
}

void Matrix::assemble(const KernelSpec& k, const CompressionSpec& comp) {
  const double t0 = now();
  giant_.clear();
  for (;;) {
    assemble_pass(k, comp);
    if (!assembly_overflow()) break;
    grow_rank_capacity();
  }
  giant_.clear();
  finish_assembly(t0);
}

void Matrix::assemble_host(const HostFunctions& f, const CompressionSpec& comp) {
  const double t0 = now();
  for (;;) {
    assemble_host_pass(f, comp);
    if (!assembly_overflow()) break;
    grow_rank_capacity();
  }
  finish_assembly(t0);
}

namespace {
// ACA panels of one block: room for twice the leaf's Rk capacity (at most
// 1024 columns).  A block whose ACA reaches that many iterations below its
// own limit sets the overflow flag, so the assembly runs again with doubled
// capacities instead of keeping an approximation stopped early.
int aca_kmax(const LeafDesc& L) {
  return std::min(std::min(L.m, L.n), std::min(1024, std::max(2 * L.cap, 64)));
}

// ACA workspace of one launch, in doubles: a quarter of the free device
// memory, between 4 and 64 GiB (a larger launch runs more blocks side by side;
// each launch lasts as long as its slowest block)
int64_t aca_budget_doubles() {
  size_t freeb = 0, total = 0;
  HMX_CHECK(hipMemGetInfo(&freeb, &total));
  const int64_t lo = (int64_t)4 << 30, hi = (int64_t)64 << 30;
  return std::min(hi, std::max(lo, (int64_t)(freeb / 4))) / (int64_t)sizeof(double);
}

// the largest blocks first: blocks of similar cost share a launch
void sort_aca_by_size(std::vector<AcaTask>& tasks, const std::vector<LeafDesc>& desc) {
  std::stable_sort(tasks.begin(), tasks.end(), [&](const AcaTask& a, const AcaTask& b) {
    const LeafDesc &x = desc[a.leaf], &y = desc[b.leaf];
    return (int64_t)(x.m + x.n) * aca_kmax(x) > (int64_t)(y.m + y.n) * aca_kmax(y);
  });
}
}  // namespace

void Matrix::assemble_pass(const KernelSpec& k, const CompressionSpec& comp) {
  device_init();
  upload_points(k, *rows, pts_rows, user_rows, stream);
  if (cols != rows) upload_points(k, *cols, pts_cols, user_cols, stream);
  PointSet rp{pts_rows.as<double>(), user_rows.as<int>(), rows->n};
  PointSet cp = (cols == rows) ? rp : PointSet{pts_cols.as<double>(), user_cols.as<int>(), cols->n};
  KernelParams kp{k.kind, k.length, k.nugget, k.shift, k.wavenumber, k.reg};
  if (embedded != (k.kind == kHelmholtz))
    throw std::runtime_error(embedded ? "assembly: a complex matrix needs the complex (Helmholtz) kernel"
                                      : "assembly: the Helmholtz kernel needs a complex matrix");
  // very large complex blocks first: their ranks may raise their leaves'
  // capacities, which changes the storage layout
  // (a block whose dense image would take more than a third of the free
  // device memory stays with partial ACA)
  std::vector<int> giant;
  if (embedded && comp.method != 0) {
    size_t freeb = 0, total = 0;
    HMX_CHECK(hipMemGetInfo(&freeb, &total));
    for (size_t i = 0; i < bt->leaves.size(); ++i) {
      const LeafDesc& L = s64.hdesc[i];
      if (L.kind == 1 && std::min(L.m, L.n) >= giant_rows &&
          (double)L.m * L.n * sizeof(double) <= freeb / 3.0)
        giant.push_back((int)i);
    }
  }
  if (!giant.empty()) {
    compress_giants(giant, rp, cp, kp, std::max(comp.eps, lowrank_eps));
    bool grow = false;
    for (int l : giant) {
      const int r = giant_.at(l)->rank;
      if (r > s64.hdesc[l].cap) {
        if (cap_hint_.size() < bt->leaves.size()) cap_hint_.resize(bt->leaves.size(), 0);
        cap_hint_[l] = std::max(cap_hint_[l], r);
        grow = true;
      }
    }
    if (grow) {
      s64.arena.release();
      s64.ranks.release();
      s64.desc.release();
      s64.valid = false;
      build_storage(&cap_hint_);
    }
  }
  init_store(s64);
  HMX_CHECK(hipMemsetAsync(overflow_.p, 0, sizeof(int), stream));
  if (!giant.empty()) {
    std::vector<int> gr;
    for (int l : giant) {
      const LeafDesc& L = s64.hdesc[l];
      const GiantPanels& g = *giant_.at(l);
      gr.push_back(g.rank);
      if (g.rank > 0) {
        HMX_CHECK(hipMemcpyAsync(s64.data() + L.a, g.a.p, sizeof(double) * (size_t)L.m * g.rank,
                                 hipMemcpyDeviceToDevice, stream));
        HMX_CHECK(hipMemcpyAsync(s64.data() + L.b, g.b.p, sizeof(double) * (size_t)L.n * g.rank,
                                 hipMemcpyDeviceToDevice, stream));
      }
    }
    for (size_t q = 0; q < giant.size(); ++q)
      HMX_CHECK(hipMemcpyAsync(s64.ranks.as<int>() + giant[q], &gr[q], sizeof(int),
                               hipMemcpyHostToDevice, stream));
    HMX_CHECK(hipStreamSynchronize(stream));
  }

  std::vector<int> full, svd, big_svd;
  std::vector<int> aca;
  for (size_t i = 0; i < bt->leaves.size(); ++i) {
    const LeafDesc& L = s64.hdesc[i];
    if (L.kind == 0) {
      full.push_back((int)i);
    } else if (std::binary_search(giant.begin(), giant.end(), (int)i)) {
      // compressed above
    } else if (embedded) {
      // complex blocks (twice the rows and columns in the embedding): ACA and
      // the truncation at any size, or the dense SVD when asked for
      if (comp.method == 0)
        big_svd.push_back((int)i);
      else
        aca.push_back((int)i);
    } else if (std::max(L.m, L.n) < 100) {
      svd.push_back((int)i);
    } else if (comp.method == 0 /*svd*/) {
      big_svd.push_back((int)i);
    } else {
      aca.push_back((int)i);
    }
  }
  DevBuf dfull, dsvd;
  upload(dfull, full.data(), full.size() * sizeof(int), stream);
  upload(dsvd, svd.data(), svd.size() * sizeof(int), stream);
  launch_assemble_full<double>(s64.desc.as<LeafDesc>(), dfull.as<int>(), (int)full.size(),
                               s64.data(), rp, cp, kp, stream);
  // compress() follows the SVD with rk->truncate(lowRankEpsilon) (compression.cpp:759): the
  // QR + core SVD of an SVD factorisation has the same singular values, so the two cuts
  // together keep sigma_i > max(eps, lowrank_eps) * sigma_0
  launch_svd_compress<double>(s64.desc.as<LeafDesc>(), dsvd.as<int>(), (int)svd.size(), s64.data(),
                              s64.ranks.as<int>(), overflow_.as<int>(), rp, cp, kp,
                              std::max(comp.eps, lowrank_eps), stream);
  // larger blocks with SVD compression: evaluated densely into a staging
  // buffer (the leaf list's descriptors pointed at it), then recompressed
  {
    const int64_t kStageElems = (int64_t)1 << 27;  // 1 GiB of doubles
    size_t i = 0;
    DevBuf stage, sdesc, slist;
    while (i < big_svd.size()) {
      std::vector<LeafDesc> d;
      std::vector<int> lst;
      std::vector<std::pair<int, int64_t>> blocks;
      int64_t off = 0;
      while (i < big_svd.size()) {
        LeafDesc L = s64.hdesc[big_svd[i]];
        const int64_t e = (int64_t)L.m * L.n;
        if (!blocks.empty() && off + e > kStageElems) break;
        blocks.push_back({big_svd[i], off});
        L.a = off;
        lst.push_back((int)d.size());
        d.push_back(L);
        off += (e + 63) / 64 * 64;
        ++i;
      }
      HMX_CHECK(hipStreamSynchronize(stream));
      stage.ensure(sizeof(double) * (size_t)off);
      upload(sdesc, d.data(), d.size() * sizeof(LeafDesc), stream);
      upload(slist, lst.data(), lst.size() * sizeof(int), stream);
      launch_assemble_full<double>(sdesc.as<LeafDesc>(), slist.as<int>(), (int)lst.size(),
                                   stage.as<double>(), rp, cp, kp, stream);
      svd_compress_dense(blocks, stage.as<double>(), std::max(comp.eps, lowrank_eps));
    }
  }
  // ACA blocks, the largest first, in launches bounded by a workspace budget.
  std::vector<AcaTask> acat;
  for (int l : aca) acat.push_back(AcaTask{l, aca_kmax(s64.hdesc[l]), 0});
  sort_aca_by_size(acat, s64.hdesc);
  const int64_t budget = acat.empty() ? 0 : aca_budget_doubles();
  const int partial = comp.method == 2;
  size_t i = 0;
  DevBuf ws, dt;
  const int nleaf = (int)bt->leaves.size();
  while (i < acat.size()) {
    HMX_CHECK(hipStreamSynchronize(stream));
    report_assembly(nleaf - (int)(acat.size() - i), nleaf);
    std::vector<AcaTask> batch;
    int64_t used = 0;
    while (i < acat.size()) {
      const LeafDesc& L = s64.hdesc[acat[i].leaf];
      const int64_t need = aca_workspace(L.m, L.n, acat[i].kmax);
      if (!batch.empty() && used + need > budget) break;
      batch.push_back(AcaTask{acat[i].leaf, acat[i].kmax, used});
      used += need;
      ++i;
    }
    HMX_CHECK(hipStreamSynchronize(stream));
    ws.ensure((size_t)used * sizeof(double));
    upload(dt, batch.data(), batch.size() * sizeof(AcaTask), stream);
    launch_aca<double>(s64.desc.as<LeafDesc>(), dt.as<AcaTask>(), (int)batch.size(), s64.data(),
                       s64.ranks.as<int>(), overflow_.as<int>(), ws.as<double>(), rp, cp, kp,
                       comp.eps, lowrank_eps, partial, stream);
  }
}

// hmat_progress_t of the assembly context: max = the number of leaves,
// current = the leaves whose blocks are assembled (reported between launches)
void Matrix::report_assembly(int done, int total) {
  if (!progress || !progress->update) return;
  progress->max = total;
  progress->current = done;
  progress->update(progress);
}

void Matrix::finish_assembly(double t0) {
  HMX_CHECK(hipStreamSynchronize(stream));
  report_assembly((int)bt->leaves.size(), (int)bt->leaves.size());
  s64.valid = true;
  if (value_type == HMAT_SIMPLE_PRECISION) {
    init_store(s32);
    launch_convert<double, float>(s64.data(), s32.data(), s64.elems, stream);
    HMX_CHECK(hipMemcpyAsync(s32.ranks.p, s64.ranks.p, s64.hdesc.size() * sizeof(int),
                             hipMemcpyDeviceToDevice, stream));
    HMX_CHECK(hipStreamSynchronize(stream));
    s32.valid = true;
    s64.arena.release();
    s64.valid = false;
  }
  assembled = true;
  plan_.built = false;
  plan_.nrhs = 0;
  build_gemv_plan();
  stats.t_assemble = now() - t0;
}

namespace {
// One leaf block (rows R, cols C) from the caller's host functions, column-major
// with ld m (SimpleFunction::assemble / BlockFunction::assemble, assembly.cpp:75-185);
// w doubles per entry (2: complex).
void evaluate_block(const HostFunctions& f, const ClusterTree& rt, const ClusterTree& ct,
                    const Cluster* R, const Cluster* C, double* out, int w = 1) {
  const int m = R->size, n = C->size;
  if (f.simple) {
    const int* ri = rt.indices.data() + R->offset;
    const int* ci = ct.indices.data() + C->offset;
    for (int j = 0; j < n; ++j)
      for (int i = 0; i < m; ++i)
        f.simple(f.user_context, ri[i], ci[j], out + w * (i + (size_t)j * m));
    return;
  }
  hmat_block_info_t info;
  std::memset(&info, 0, sizeof(info));
  info.block_type = hmat_block_full;
  info.number_of_strata = 1;
  auto prepare = [&]() {
    f.prepare(R->offset, m, C->offset, n, const_cast<int*>(rt.indices.data()),
              const_cast<int*>(rt.indices_rev.data()), const_cast<int*>(ct.indices.data()),
              const_cast<int*>(ct.indices_rev.data()), f.user_context, &info);
  };
  prepare();
  if (info.needed_memory != 0) prepare();  // BlockFunction::prepareBlock calls it again
  if (info.block_type == hmat_block_null) {
    std::memset(out, 0, sizeof(double) * w * (size_t)m * n);
  } else if (f.advanced) {
    hmat_block_compute_context_t ac;
    ac.block = out;
    ac.row_start = 0;
    ac.row_count = m;
    ac.col_start = 0;
    ac.col_count = n;
    ac.stratum = -1;
    ac.user_data = info.user_data;
    f.advanced(&ac);
  } else {
    f.block(info.user_data, 0, m, 0, n, out);
  }
  if (info.release_user_data) info.release_user_data(info.user_data);
}

// The block of a complex matrix: the caller's complex block over the complex
// clusters (half the embedded offsets and sizes), expanded to its 2 x 2 real
// representation (2m x 2n, ld 2m).
void evaluate_block_complex(const HostFunctions& f, const ClusterTree& zr, const ClusterTree& zc,
                            const Cluster* R, const Cluster* C, double* out,
                            std::vector<double>& z) {
  Cluster r, c;
  r.offset = R->offset / 2;
  r.size = R->size / 2;
  c.offset = C->offset / 2;
  c.size = C->size / 2;
  const int m = r.size, n = c.size;
  z.resize(2 * (size_t)m * n);
  evaluate_block(f, zr, zc, &r, &c, z.data(), 2);
  const size_t ld = 2 * (size_t)m;
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < m; ++i) {
      const double re = z[2 * (i + (size_t)j * m)], im = z[2 * (i + (size_t)j * m) + 1];
      out[2 * i + 2 * j * ld] = re;
      out[2 * i + 1 + 2 * j * ld] = im;
      out[2 * i + (2 * j + 1) * ld] = -im;
      out[2 * i + 1 + (2 * j + 1) * ld] = re;
    }
}
}  // namespace

void Matrix::assemble_host_pass(const HostFunctions& f, const CompressionSpec& comp) {
  if (!f.simple && !((f.block || f.advanced) && f.prepare))
    throw std::runtime_error("assembly: set simple_compute, or block_compute / advanced_compute with prepare");
  device_init();
  init_store(s64);
  HMX_CHECK(hipMemsetAsync(overflow_.p, 0, sizeof(int), stream));
  PointSet none{nullptr, nullptr, 0};
  KernelParams kp{kExponential, 1.0, 0.0, 0.0};
  // leaves in staging batches of at most kStage doubles (a leaf larger than
  // that is a batch of its own)
  const int64_t kStageElems = (int64_t)1 << 26;  // 512 MB of doubles
  const int64_t aca_budget = aca_budget_doubles();  // doubles of ACA workspace per launch
  const int nleaf = (int)bt->leaves.size();
  std::vector<double> host, zbuf;
  DevBuf table, ws, dl, doff, dt;
  int l0 = 0;
  while (l0 < nleaf) {
    int64_t used = 0;
    int l1 = l0;
    while (l1 < nleaf) {
      const int64_t e = (int64_t)s64.hdesc[l1].m * s64.hdesc[l1].n;
      if (l1 > l0 && used + e > kStageElems) break;
      used += e;
      ++l1;
    }
    host.resize((size_t)std::max<int64_t>(used, 1));
    std::vector<int> full, svd;
    std::vector<int64_t> full_off, svd_off;
    std::vector<std::pair<int, int64_t>> big_svd;
    std::vector<AcaTask> aca;
    int64_t off = 0;
    for (int l = l0; l < l1; ++l) {
      const Block* b = bt->leaves[l];
      const LeafDesc& L = s64.hdesc[l];
      if (embedded)
        evaluate_block_complex(f, *zrows, *zcols, b->rows, b->cols, host.data() + off, zbuf);
      else
        evaluate_block(f, *rows, *cols, b->rows, b->cols, host.data() + off);
      if (L.kind == 0) {
        full.push_back(l);
        full_off.push_back(off);
      } else if (embedded && comp.method == 0) {
        big_svd.push_back({l, off});
      } else if (embedded) {
        AcaTask t{l, aca_kmax(L), 0};
        t.tab = off;
        aca.push_back(t);
      } else if (std::max(L.m, L.n) < 100) {
        svd.push_back(l);
        svd_off.push_back(off);
      } else if (comp.method == 0 /*svd*/) {
        big_svd.push_back({l, off});
      } else {
        AcaTask t{l, aca_kmax(L), 0};
        t.tab = off;
        aca.push_back(t);
      }
      off += (int64_t)L.m * L.n;
    }
    HMX_CHECK(hipStreamSynchronize(stream));
    upload(table, host.data(), (size_t)used * sizeof(double), stream);
    upload(dl, full.data(), full.size() * sizeof(int), stream);
    upload(doff, full_off.data(), full_off.size() * sizeof(int64_t), stream);
    launch_table_full<double>(s64.desc.as<LeafDesc>(), dl.as<int>(), doff.as<int64_t>(),
                              (int)full.size(), table.as<double>(), s64.data(), stream);
    HMX_CHECK(hipStreamSynchronize(stream));
    upload(dl, svd.data(), svd.size() * sizeof(int), stream);
    upload(doff, svd_off.data(), svd_off.size() * sizeof(int64_t), stream);
    launch_svd_compress<double>(s64.desc.as<LeafDesc>(), dl.as<int>(), (int)svd.size(), s64.data(),
                                s64.ranks.as<int>(), overflow_.as<int>(), none, none, kp,
                                std::max(comp.eps, lowrank_eps), stream, doff.as<int64_t>(),
                                table.as<double>());
    if (!big_svd.empty())
      svd_compress_dense(big_svd, table.as<double>(), std::max(comp.eps, lowrank_eps));
    sort_aca_by_size(aca, s64.hdesc);
    size_t i = 0;
    while (i < aca.size()) {
      std::vector<AcaTask> batch;
      int64_t wsu = 0;
      while (i < aca.size()) {
        const LeafDesc& L = s64.hdesc[aca[i].leaf];
        const int64_t need = aca_workspace(L.m, L.n, aca[i].kmax);
        if (!batch.empty() && wsu + need > aca_budget) break;
        AcaTask t = aca[i];
        t.ws = wsu;
        batch.push_back(t);
        wsu += need;
        ++i;
      }
      HMX_CHECK(hipStreamSynchronize(stream));
      ws.ensure((size_t)wsu * sizeof(double));
      upload(dt, batch.data(), batch.size() * sizeof(AcaTask), stream);
      launch_aca<double>(s64.desc.as<LeafDesc>(), dt.as<AcaTask>(), (int)batch.size(), s64.data(),
                         s64.ranks.as<int>(), overflow_.as<int>(), ws.as<double>(), none, none,
                         kp, comp.eps, lowrank_eps, comp.method == 2, stream, table.as<double>());
    }
    if (progress) {
      HMX_CHECK(hipStreamSynchronize(stream));
      report_assembly(l1, nleaf);
    }
    l0 = l1;
  }
}

// ---------------------------------------------------------------------------
// H-gemv plan (host) and driver
// ---------------------------------------------------------------------------
// The item lists hold offsets in units of one right-hand side (the offsets for
// nrhs right-hand sides are nrhs times these; the kernels scale them), so one
// plan serves every nrhs and only the scratch buffers depend on it.
void Matrix::build_gemv_plan() {
  GemvPlan& P = plan_;
  P.max_rows = 0;
  // the y pass takes destination leaf clusters of at most 512 rows: a matrix
  // with wider ones has no H-gemv plan (its gemv reports it, the rest works)
  P.built = false;
  for (const Cluster* c : rows->leaves)
    if (c->size > 512) return;
  const int nrhs = 1;
  const auto& D = s64.hdesc;
  const int nleaf = (int)D.size();
  const int chunk = kZChunk;
  std::vector<ZItem> zi;
  std::vector<ZReduce> zr;
  std::vector<int64_t> zoff_side0(nleaf, -1), zoff_side1(nleaf, -1);
  int64_t zpart = 0, zbuf = 0;
  for (int l = 0; l < nleaf; ++l) {
    const LeafDesc& L = D[l];
    if (L.kind != 1) continue;
    const bool mir = bt->leaves[l]->mirrored() && lower_sym;
    for (int side = 0; side < (mir ? 2 : 1); ++side) {
      const int len = side == 0 ? L.n : L.m;
      ZReduce red{l, (int)zi.size(), 0, 0, zbuf};
      for (int r = 0; r < len; r += chunk) {
        zi.push_back(ZItem{l, side, r, std::min(chunk, len - r), zpart});
        zpart += (int64_t)L.cap * nrhs;
      }
      red.end = (int)zi.size();
      zr.push_back(red);
      (side == 0 ? zoff_side0 : zoff_side1)[l] = zbuf;
      zbuf += (int64_t)L.cap * nrhs;
    }
  }
  // y pass: contributions grouped by destination leaf cluster of the rows tree
  const auto& RL = rows->leaves;
  std::vector<std::vector<YContrib>> per(RL.size());
  std::vector<std::vector<int64_t>> slots_for(RL.size());
  int64_t slot = 0;
  for (int l = 0; l < nleaf; ++l) {
    const Block* b = bt->leaves[l];
    const LeafDesc& L = D[l];
    const bool mir = b->mirrored() && lower_sym;
    if (L.kind == 0) {
      if (!b->rows->is_leaf() || !b->cols->is_leaf())
        throw std::runtime_error("gemv: full leaf spanning several leaf clusters");
      YContrib c{l, 0, 0, 0, -1};
      if (mir) {
        c.z = slot;
        slots_for[b->cols->leaf_index].push_back(slot);
        slot += (int64_t)L.n * nrhs;
      }
      per[b->rows->leaf_index].push_back(c);
    } else {
      for (int q = b->rows->first_leaf; q < b->rows->first_leaf + b->rows->nleaves; ++q)
        per[q].push_back(YContrib{l, 2, RL[q]->offset - L.r0, 0, zoff_side0[l]});
      if (mir)
        for (int q = b->cols->first_leaf; q < b->cols->first_leaf + b->cols->nleaves; ++q)
          per[q].push_back(YContrib{l, 3, RL[q]->offset - L.c0, 0, zoff_side1[l]});
    }
  }
  std::vector<YItem> yi;
  std::vector<YContrib> contrib;
  std::vector<int64_t> slist;
  for (size_t q = 0; q < RL.size(); ++q) {
    P.max_rows = std::max(P.max_rows, RL[q]->size);
    YItem it{RL[q]->offset, RL[q]->size, (int)contrib.size(), 0, (int)slist.size(), 0};
    // full leaves first, then the Rk slices, which the y pass stages in batches
    std::stable_partition(per[q].begin(), per[q].end(), [](const YContrib& c) { return c.kind == 0; });
    contrib.insert(contrib.end(), per[q].begin(), per[q].end());
    slist.insert(slist.end(), slots_for[q].begin(), slots_for[q].end());
    it.end = (int)contrib.size();
    it.send = (int)slist.size();
    yi.push_back(it);
  }
  upload(P.zitems, zi.data(), zi.size() * sizeof(ZItem), stream);
  upload(P.zred, zr.data(), zr.size() * sizeof(ZReduce), stream);
  upload(P.yitems, yi.data(), yi.size() * sizeof(YItem), stream);
  upload(P.contrib, contrib.data(), contrib.size() * sizeof(YContrib), stream);
  upload(P.slotlist, slist.data(), slist.size() * sizeof(int64_t), stream);
  P.zpart1 = zpart;
  P.z1 = zbuf;
  P.slot1 = slot;
  P.nz = (int)zi.size();
  P.nzred = (int)zr.size();
  P.ny = (int)yi.size();
  P.built = true;
  P.nrhs = 0;
  HMX_CHECK(hipStreamSynchronize(stream));
}

size_t Matrix::gemv_scratch_elems(int nrhs) const {
  const GemvPlan& P = plan_;
  return (size_t)(P.zpart1 + P.z1 + P.slot1 + rows->n) * (size_t)nrhs;
}

template <typename T>
void Matrix::gemv_launch(double alpha, const T* x, double beta, T* y, int nrhs, int ldx, int ldy,
                         const Store<T>& st, double* zpart, double* z, double* slots,
                         double* acc) {
  const GemvPlan& P = plan_;
  const LeafDesc* desc = st.desc.template as<LeafDesc>();
  const int* ranks = st.ranks.template as<int>();
  launch_gemv_zpass<T>(desc, ranks, st.data(), P.zitems.as<ZItem>(), P.nz, x, ldx, nrhs, zpart,
                       stream);
  launch_gemv_zreduce(P.zred.as<ZReduce>(), P.nzred, ranks, zpart, P.zitems.as<ZItem>(), nrhs,
                      desc, z, stream);
  launch_gemv_ypass<T>(desc, ranks, st.data(), P.yitems.as<YItem>(), P.ny,
                       P.contrib.as<YContrib>(), z, slots, x, ldx, nrhs, acc, rows->n, P.max_rows,
                       stream);
  launch_gemv_yfinal<T>(P.yitems.as<YItem>(), P.ny, P.slotlist.as<int64_t>(), slots, acc,
                        rows->n, y, ldy, nrhs, alpha, beta, stream);
}

template <typename T>
void Matrix::gemv_device(double alpha, const T* x, double beta, T* y, int nrhs, int ldx, int ldy,
                         const Store<T>& st) {
  if (!st.valid) throw std::runtime_error("gemv: matrix data not available in this precision");
  if (nrhs < 1) throw std::runtime_error("gemv: nrhs must be positive");
  if (!plan_.built) build_gemv_plan();
  if (!plan_.built) throw std::runtime_error("gemv: leaf clusters above 512 rows are not provided");
  GemvPlan& P = plan_;
  if (P.nrhs < nrhs) {
    HMX_CHECK(hipStreamSynchronize(stream));
    P.zpart.ensure(std::max<int64_t>(P.zpart1 * nrhs, 1) * sizeof(double));
    P.z.ensure(std::max<int64_t>(P.z1 * nrhs, 1) * sizeof(double));
    P.slots.ensure(std::max<int64_t>(P.slot1 * nrhs, 1) * sizeof(double));
    P.acc.ensure((size_t)rows->n * nrhs * sizeof(double));
    P.nrhs = nrhs;
  }
  gemv_launch<T>(alpha, x, beta, y, nrhs, ldx, ldy, st, P.zpart.as<double>(), P.z.as<double>(),
                 P.slots.as<double>(), P.acc.as<double>());
}

template <typename T>
void Matrix::gemv_device_scratch(double alpha, const T* x, double beta, T* y, int nrhs,
                                 const Store<T>& st, double* scratch, size_t scratch_elems) {
  if (!st.valid) throw std::runtime_error("gemv: matrix data not available in this precision");
  if (!plan_.built)
    throw std::runtime_error("gemv: no H-gemv plan (matrix not assembled, or leaf clusters above 512 rows)");
  if (nrhs < 1) throw std::runtime_error("gemv: nrhs must be positive");
  const size_t need = gemv_scratch_elems(nrhs);
  if (need > scratch_elems)
    throw std::runtime_error("gemv: needs " + std::to_string(need * sizeof(double)) +
                             " bytes of scratch, the allocation-free entry point has " +
                             std::to_string(scratch_elems * sizeof(double)));
  const GemvPlan& P = plan_;
  double* zpart = scratch;
  double* z = zpart + P.zpart1 * nrhs;
  double* slots = z + P.z1 * nrhs;
  double* acc = slots + P.slot1 * nrhs;
  gemv_launch<T>(alpha, x, beta, y, nrhs, n_cols(), n_rows(), st, zpart, z, slots, acc);
}
template void Matrix::gemv_device_scratch<double>(double, const double*, double, double*, int,
                                                  const Store<double>&, double*, size_t);
template void Matrix::gemv_device_scratch<float>(double, const float*, double, float*, int,
                                                 const Store<float>&, double*, size_t);
template void Matrix::gemv_device<double>(double, const double*, double, double*, int, int, int,
                                          const Store<double>&);
template void Matrix::gemv_device<float>(double, const float*, double, float*, int, int, int,
                                         const Store<float>&);

namespace {
template <typename T>
void permute_in(const T* user, T* hm, const std::vector<int>& idx, int n, int nrhs) {
  for (int r = 0; r < nrhs; ++r)
    for (int i = 0; i < n; ++i) hm[i + (size_t)r * n] = user[idx[i] + (size_t)r * n];
}
template <typename T>
void permute_out(const T* hm, T* user, const std::vector<int>& idx, int n, int nrhs) {
  for (int r = 0; r < nrhs; ++r)
    for (int i = 0; i < n; ++i) user[idx[i] + (size_t)r * n] = hm[i + (size_t)r * n];
}

template <typename T>
void gemv_host_t(Matrix& M, const void* alpha, void* x, const void* beta, void* y, int nrhs,
                 bool permute) {
  const int n = M.n_rows(), nc = M.n_cols();
  std::vector<T> hx((size_t)nc * nrhs), hy((size_t)n * nrhs);
  const T a = *(const T*)alpha, b = *(const T*)beta;
  if (permute) {
    permute_in((const T*)x, hx.data(), M.cols->indices, nc, nrhs);
    permute_in((const T*)y, hy.data(), M.rows->indices, n, nrhs);
  } else {
    std::memcpy(hx.data(), x, hx.size() * sizeof(T));
    std::memcpy(hy.data(), y, hy.size() * sizeof(T));
  }
  DevBuf dx, dy;
  upload(dx, hx.data(), hx.size() * sizeof(T), M.stream);
  upload(dy, hy.data(), hy.size() * sizeof(T), M.stream);
  M.gemv_device<T>((double)a, dx.as<T>(), (double)b, dy.as<T>(), nrhs, nc, n, M.store<T>());
  HMX_CHECK(hipMemcpyAsync(hy.data(), dy.p, hy.size() * sizeof(T), hipMemcpyDeviceToHost, M.stream));
  HMX_CHECK(hipStreamSynchronize(M.stream));
  if (permute)
    permute_out(hy.data(), (T*)y, M.rows->indices, n, nrhs);
  else
    std::memcpy(y, hy.data(), hy.size() * sizeof(T));
}

// y = alpha op(F) x + beta y for a factorised matrix (F: see factor_gemv_device)
template <typename T>
void factor_gemv_host_t(Matrix& M, char trans, const void* alpha, void* x, const void* beta,
                        void* y, int nrhs, bool permute) {
  const int n = M.n_rows();
  std::vector<T> hx((size_t)n * nrhs), ht((size_t)n * nrhs);
  if (permute)
    permute_in((const T*)x, hx.data(), M.cols->indices, n, nrhs);
  else
    std::memcpy(hx.data(), x, hx.size() * sizeof(T));
  DevBuf dx, dt;
  upload(dx, hx.data(), hx.size() * sizeof(T), M.stream);
  dt.alloc(hx.size() * sizeof(T));
  M.factor_gemv_device<T>(trans, dx.as<T>(), dt.as<T>(), nrhs, M.store<T>());
  HMX_CHECK(hipMemcpy(hx.data(), dt.p, hx.size() * sizeof(T), hipMemcpyDeviceToHost));
  if (permute)
    permute_out(hx.data(), ht.data(), M.rows->indices, n, nrhs);
  else
    ht.swap(hx);
  const T a = *(const T*)alpha, b = *(const T*)beta;
  T* yy = (T*)y;
  for (size_t i = 0; i < ht.size(); ++i) yy[i] = a * ht[i] + (b == (T)0 ? (T)0 : b * yy[i]);
}
}  // namespace

void Matrix::gemv_host(char trans, const void* alpha, void* x, const void* beta, void* y, int nrhs,
                       bool permute) {
  if (!assembled) throw std::runtime_error("gemv: matrix not assembled");
  if (factorization != -1) {
    // HMatrix::gemv of the factor (DefaultEngine::gemv default_engine.cpp:133-140)
    if (factor32 && value_type == HMAT_DOUBLE_PRECISION)
      throw std::runtime_error("gemv: the matrix holds a mixed-precision factor");
    if (value_type == HMAT_DOUBLE_PRECISION)
      factor_gemv_host_t<double>(*this, trans, alpha, x, beta, y, nrhs, permute);
    else if (value_type == HMAT_SIMPLE_PRECISION)
      factor_gemv_host_t<float>(*this, trans, alpha, x, beta, y, nrhs, permute);
    else
      throw std::runtime_error("gemv: unsupported value type");
    return;
  }
  if (trans != 'N' && !lower_sym) throw std::runtime_error("gemv: only 'N' for non-symmetric matrices");
  if (value_type == HMAT_DOUBLE_PRECISION)
    gemv_host_t<double>(*this, alpha, x, beta, y, nrhs, permute);
  else if (value_type == HMAT_SIMPLE_PRECISION)
    gemv_host_t<float>(*this, alpha, x, beta, y, nrhs, permute);
  else
    throw std::runtime_error("gemv: unsupported value type");
}

void Matrix::info(hmat_info_t& out) const {
  // HMatrix::info h_matrix.cpp:464-502
  std::memset(&out, 0, sizeof(out));
  std::vector<int> rk(bt->leaves.size(), 0);
  const DevBuf& rb = s64.ranks.p ? s64.ranks : s32.ranks;
  if (rb.p) HMX_CHECK(hipMemcpy(rk.data(), rb.p, rk.size() * sizeof(int), hipMemcpyDeviceToHost));
  std::function<void(const Block*)> rec;
  int count_nodes = 0;
  rec = [&](const Block* b) {
    ++count_nodes;
    const size_t r = b->m(), c = b->n();
    if (r == 0 || c == 0) return;
    if (b->is_leaf()) {
      const size_t s = r * c;
      out.uncompressed_size += s;
      if (b->kind == Kind::Rk) {
        const size_t k = std::max(rk[b->leaf], 0);
        const size_t mem = k * (r + c);
        out.compressed_size += mem;
        if ((int)(r + c) > out.largest_rk_dim_cols + out.largest_rk_dim_rows) {
          out.largest_rk_dim_cols = (int)c;
          out.largest_rk_dim_rows = (int)r;
        }
        const size_t old = ((size_t)out.largest_rk_mem_cols + out.largest_rk_mem_rows) * out.largest_rk_mem_rank;
        if (mem > old) {
          out.largest_rk_mem_cols = (int)c;
          out.largest_rk_mem_rows = (int)r;
          out.largest_rk_mem_rank = (int)k;
        }
        out.rk_count++;
        out.rk_size += s;
      } else {
        out.compressed_size += s;
        out.full_count++;
        out.full_size += s;
      }
    } else {
      for (int i = 0; i < b->nrow * b->ncol; ++i)
        if (b->child[i]) rec(b->child[i]);
    }
  };
  rec(bt->root);
  out.nr_block_clusters = count_nodes;
  if (embedded) {
    // counts of the complex matrix: a complex entry is 2 x 2 real entries, a
    // complex rank k a real rank 2k
    out.full_size /= 4;
    out.rk_size /= 4;
    out.uncompressed_size /= 4;
    out.compressed_size /= 4;
    out.largest_rk_dim_rows /= 2;
    out.largest_rk_dim_cols /= 2;
    out.largest_rk_mem_rows /= 2;
    out.largest_rk_mem_cols /= 2;
    out.largest_rk_mem_rank = (out.largest_rk_mem_rank + 1) / 2;
  }
}

}  // namespace hmx
