// Persistent triangular sweeps of the H-LLt solve (HMatrix::solveLowerTriangularLeft
// h_matrix.cpp:2338-2372 forward, solveUpperTriangularLeft h_matrix.cpp:2504-2540
// backward, for a factor in lower-symmetric storage).
//
// One launch per sweep.  Its work items, in an order that respects every
// dependency, are handed out by a device ticket counter:
//   ROW q   : x_q <- D_q (x_q - sum of the off-diagonal leaves of block row q
//             (forward) / block column q (backward) applied to finished rows),
//             D_q the precomputed inverse of the diagonal leaf (forward) or its
//             transpose (backward); then raises the row's done flag
//   ZPART c : partial z = P[chunk]^T x[chunk] of one Rk panel chunk (at most
//             kSweepChunk rows) once the rows of the chunk are done; P = B
//             (forward) or A (backward)
//   ZRED l  : z_l = fixed-order sum of the chunk partials of Rk leaf l
// A ROW item waits for the sources of its contributions (finished rows, or
// the z of an Rk leaf), which always precede it in ticket order, so the wait
// ends; every wait is bounded in time and gives up on an error flag.
// Rows finish in sweep order (a row waits for its predecessor's flag before
// raising its own), so one raised flag implies all earlier rows are done.
#pragma once
#include <cstdint>

#include "kernels.hpp"

namespace hmx {

constexpr int kSweepChunk = 256;  // rows of an Rk panel per ZPART item
constexpr int kSweepNR = 8;       // right-hand sides per sweep launch

enum SweepItemType : int { SW_ROW = 0, SW_ZPART = 1, SW_ZRED = 2 };

struct SwItem {
  int type;
  int idx;   // row (sweep position), chunk or reduction index
};

// One leaf cluster of rows, in sweep order (position p).
struct SwRow {
  int r0, m;          // rows [r0, r0 + m) of x
  int cbeg, cend;     // contributions
  int64_t inv;        // inverse of the diagonal leaf (m x m, ld m) in the inverse buffer
};

enum SweepContribKind : int {
  SC_FULL = 0,   // x_row -= L (m x n, ld m) x[src rows]
  SC_FULL_T = 1, // x_row -= L^T x[src rows], L stored n x m (ld n)
  SC_RK = 2      // x_row -= P[row0 .. row0 + m) z_slot (P = A forward, B backward)
};

struct SwContrib {
  int kind;
  int leaf;
  int src;       // SC_FULL*: first row of the source in x; SC_RK: Rk slot
  int dep;       // SC_FULL*: sweep position of the source row; SC_RK: unused
  int row0;      // SC_RK: first panel row of this destination
  int pad;
  int64_t z;     // SC_RK: offset of the slot's z (doubles)
};

struct SwChunk {
  int leaf, slot;
  int row0, rows;   // panel rows of the chunk
  int dep;          // sweep position after which every row of the chunk is done
  int single;       // the leaf has one chunk: write z directly and mark it ready
  int64_t out;      // z offset (single) or partial offset (doubles)
};

struct SwRed {
  int leaf, slot;
  int cbeg, cend;   // chunks of the leaf
  int64_t z;        // offset of the slot's z (doubles)
};

// Device layout of the synchronisation ints of one sweep launch: the ticket
// counter, the error flag and every row's done flag on lines of their own
// (each written once per launch, by one workgroup), then zcount[nslots] and
// zready[nslots].
constexpr int kSwLine = 32;  // ints per 128-byte line
constexpr int kSwTicket = 0, kSwErr = kSwLine, kSwRows = 2 * kSwLine;
inline int64_t sweep_sync_ints(int rows, int nslots) {
  return kSwRows + (int64_t)kSwLine * rows + 2 * (int64_t)nslots;
}
// Diagnostics of the first wait that gave up: fail[0] = 1, then the ticket of
// the item, what it waited for (0 row flag, 1 z ready, 2 chunk count), which
// row / slot, the value awaited and the value last read.
constexpr int kSwFailInts = 6;

// Arguments of one sweep launch (device pointers).
struct SweepLaunch {
  const SwItem* items;
  const SwRow* rows;
  const SwContrib* contrib;
  const SwChunk* chunks;
  const SwRed* reds;
  const LeafDesc* desc;
  const int* ranks;
  double* z;
  double* zpart;
  int* sync;          // sweep_sync_ints(rows, nslots) ints, cleared by the launch
  int* fail;          // kSwFailInts ints, sticky: set when a wait gave up
  int64_t wait_ticks; // wall-clock ticks before a wait gives up
  int nitems, nslots, nrows;
  int fwd;            // 1 forward (L y = b), 0 backward (L^T x = y)
  int nrhs, ldx;      // nrhs <= kSweepNR columns of x, leading dimension ldx
  int max_groups;     // workgroups of the persistent launch
};

// Clears the synchronisation ints and runs one sweep over x in place.
template <typename T>
void launch_sweep(const SweepLaunch& L, const T* arena, const T* inv, T* x, hipStream_t s);

// Unit m x m matrices at rows[i].inv of inv, i < count.
template <typename T>
void launch_sweep_identity(T* inv, const SwRow* rows, int count, hipStream_t s);

}  // namespace hmx
