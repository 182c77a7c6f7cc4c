// Host-visible declarations of the batched HIP kernels (launch wrappers).
// Every wrapper enqueues on the given stream and never synchronises.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "dev_common.hpp"

namespace hmx {

// Per-leaf descriptor in HBM.  Full leaves: column-major m x n at 'a' (ld = m).
// Rk leaves: A panel m x cap at 'a' (ld = m), B panel n x cap at 'b' (ld = n),
// current rank in the separate rank array.
struct LeafDesc {
  int m, n;
  int r0, c0;   // offsets in hmat numbering
  int kind;     // 0 full, 1 rk
  int cap;
  int64_t a, b;
};

// Points of one cluster tree, hmat order, structure of arrays (x | y | z).
struct PointSet {
  const double* xyz;   // 3 * n
  const int* user;     // hmat index -> user index (for the nugget test)
  int n;
};

// ---- assembly ---------------------------------------------------------------
template <typename T>
void launch_assemble_full(const LeafDesc* desc, const int* list, int count, T* arena,
                          PointSet rp, PointSet cp, KernelParams kp, hipStream_t s);

// Admissible leaves with max(m, n) < 100: dense evaluation + Jacobi SVD +
// truncation (compression.cpp:224-251).  ranks[leaf] set; overflow[0] |= 1 if
// a rank exceeds the leaf capacity.
// tab_off (one offset per listed leaf, device) and table: take the entries
// from blocks staged by the host (column-major, ld m) instead of the kernel.
template <typename T>
void launch_svd_compress(const LeafDesc* desc, const int* list, int count, T* arena, int* ranks,
                         int* overflow, PointSet rp, PointSet cp, KernelParams kp, double eps,
                         hipStream_t s, const int64_t* tab_off = nullptr,
                         const double* table = nullptr);

// Full leaves from blocks staged by the host (one offset per listed leaf).
template <typename T>
void launch_table_full(const LeafDesc* desc, const int* list, const int64_t* tab_off, int count,
                       const double* table, T* arena, hipStream_t s);

struct AcaTask {
  int leaf;
  int kmax;       // capacity of the ACA panels
  int64_t ws;     // offset (doubles) of this task's workspace
  int64_t tab = -1;  // offset of the block staged by the host in the table, or -1
};

// Workspace (doubles) needed by one ACA task followed by its truncation.
int64_t aca_workspace(int m, int n, int kmax);

// ACA+ (compression.cpp:555-724; delegates to ACA partial when n*100 < m,
// or always when partial != 0), then RkMatrix::truncate (rk_matrix.cpp:490-585)
// into the leaf storage.
template <typename T>
void launch_aca(const LeafDesc* desc, const AcaTask* tasks, int count, T* arena, int* ranks,
                int* overflow, double* ws, PointSet rp, PointSet cp, KernelParams kp,
                double aca_eps, double trunc_eps, int partial, hipStream_t s,
                const double* table = nullptr);

// ---- H-gemv -----------------------------------------------------------------
// Work items of the z pass: z_part = P^T x over a row range of a panel of at
// most kZChunk rows.
constexpr int kZChunk = 1024;
struct ZItem {
  int leaf;
  int side;       // 0: B panel against x[cols]  (z),  1: A panel against x[rows] (z')
  int row0, rows; // local row range inside the panel
  int64_t zslot;  // output slot offset per right-hand side (the slot holds cap * nrhs doubles
                  // at nrhs * zslot)
};
// y pass: one work item per destination leaf cluster.  Contributions:
//   kind 0: full leaf with rows == destination: acc += M x[cols]; if the leaf is
//           mirrored (symmetric storage) also slot[fslot] = M^T x[rows]
//   kind 2: Rk A panel rows [row0, row0+len): acc += A z
//   kind 3: Rk B panel rows [row0, row0+len) (mirrored Rk): acc += B z'
// and the final pass adds the mirrored-full slots that target the cluster:
//   y[r0:r0+len] = beta y + alpha (acc + sum slots)
struct YItem {
  int r0, len;
  int beg, end;     // contribution range
  int sbeg, send;   // slot-list range (indices into slot_list)
};
struct YContrib {
  int leaf;
  int kind;
  int row0;
  int pad;
  int64_t z;       // z / z' offset for rk, slot offset for mirrored full (doubles per
                   // right-hand side: the data sits at nrhs * z), else -1
};

// ---- whole-matrix leaf operations (k_ops.hip); one workgroup per listed leaf
template <typename T>
void launch_leaf_norm2(const LeafDesc* desc, const int* ranks, const T* arena, const int* list,
                       int count, double* out, hipStream_t s);
template <typename T>
void launch_leaf_scale(const LeafDesc* desc, const int* list, int count, T* arena, double alpha,
                       hipStream_t s);
template <typename T>
void launch_leaf_add_diag(const LeafDesc* desc, const int* list, int count, T* arena,
                          double alpha, hipStream_t s);
template <typename T>
void launch_leaf_get_diag(const LeafDesc* desc, const int* list, int count, const T* arena,
                          double* out, hipStream_t s);

// Transposition of the leaves (Matrix::transpose): kind 0, a full leaf
// dst (m x n) = the partner's block at src (n x m) transposed, read from a
// copy of the arena (with conj, the real representation of a complex
// transpose: entries of differing row / column parity negated); kind 1, the
// odd rows of the m x n panel at dst negated.
struct TransposeItem {
  int64_t dst, src;
  int m, n;
  int kind, conj;
};
template <typename T>
void launch_leaf_transpose(const TransposeItem* items, int count, T* arena, const T* copy,
                           hipStream_t s);

// Whole-device compression of very large blocks (giant.cpp): seeded standard
// normal numbers; the last `count` columns of V (ld ldv, eigenvalues lam
// ascending), largest first, column j scaled by lam[j]^e, into out (ld rows).
void launch_gauss(double* out, size_t n, uint64_t seed, hipStream_t s);
void launch_pick_scale(const double* V, int ldv, int rows, int first, int count,
                       const double* lam, double e, double* out, hipStream_t s);

// Fixed device scratch of the allocation-free H-gemv (64 MiB per GPU); returns
// its device address on the current device (looked up once per device).
constexpr size_t kNoAllocScratchBytes = size_t(64) << 20;
double* noalloc_gemv_scratch(size_t* elems);

template <typename T>
void launch_gemv_zpass(const LeafDesc* desc, const int* ranks, const T* arena, const ZItem* items,
                       int count, const T* x, int ldx, int nrhs, double* zpart, hipStream_t s);
// Sums z partial slots per (leaf, side): z[z] = sum of slots of items [beg, end).
struct ZReduce {
  int leaf;
  int beg, end;
  int pad;
  int64_t z;
};
void launch_gemv_zreduce(const ZReduce* red, int count, const int* ranks, const double* zpart,
                         const ZItem* items, int nrhs, const LeafDesc* desc, double* z,
                         hipStream_t s);
template <typename T>
// (max_rows: the largest destination leaf cluster, at most 512 rows)
void launch_gemv_ypass(const LeafDesc* desc, const int* ranks, const T* arena, const YItem* items,
                       int count, const YContrib* contrib, const double* z, double* slots,
                       const T* x, int ldx, int nrhs, double* acc, int ldacc, int max_rows,
                       hipStream_t s);
template <typename T>
void launch_gemv_yfinal(const YItem* items, int count, const int64_t* slot_list,
                        const double* slots, const double* acc, int ldacc, T* y, int ldy,
                        int nrhs, double alpha, double beta, hipStream_t s);

// Refinement helpers on column-major N x nrhs vectors (ld = n).
void launch_vec_add_f2d(double* x, const float* d, size_t n, hipStream_t s);   // x += d
void launch_col_sumsq(const double* v, int n, int nrhs, double* out, hipStream_t s);  // out[r] = sum v^2
void launch_vec_add(double* x, const double* d, size_t n, hipStream_t s);  // x += d
// conjugate gradients on n x nrhs blocks, one system per column
void launch_col_dot(const double* a, const double* b, int n, int nrhs, double* out, hipStream_t s);
void launch_cg_xr(double* x, double* r, const double* p, const double* q, const double* rz, const double* pq,
                  int n, int nrhs, hipStream_t s);  // alpha = rz / pq: x += alpha p, r -= alpha q
void launch_cg_p(double* p, const double* z, const double* rz_new, const double* rz_old, int n, int nrhs,
                 hipStream_t s);  // p = z + (rz_new / rz_old) p

// Elementwise conversion of a whole arena (same layout in both precisions).
template <typename S, typename D>
void launch_convert(const S* src, D* dst, size_t n, hipStream_t s);

}  // namespace hmx
