// Persistent forward / backward sweeps of the H-LLt solve (the item kinds and
// their order are described in sweep.hpp).  One workgroup of 256 threads per
// item at a time; a workgroup takes the next item from a ticket counter until
// the list is exhausted.  Arithmetic in FP64 whatever the storage type.
//
// Synchronisation: a producer's threads store their results, the workgroup
// meets at a barrier, then thread 0 publishes with an agent-scope release
// read-modify-write (a row's done flag, the chunk counter or the z-ready flag
// of an Rk leaf).  A consumer's thread 0 polls with agent-scope
// read-modify-writes that add a run-time zero (performed where device-wide
// atomics are, so a copy of the line held in another cache cannot answer
// them), then takes an agent-scope acquire fence before the workgroup reads
// the data.  Each wait is bounded by the wall clock: past the limit, or when
// another workgroup has failed, the item gives up, records what it waited for
// and every workgroup leaves.
#include "factor_kernels.hpp"
#include "kernels.hpp"
#include "sweep.hpp"

namespace hmx {

namespace {

struct SweepArgs {
  const SwItem* items;
  const SwRow* rows;
  const SwContrib* contrib;
  const SwChunk* chunks;
  const SwRed* reds;
  const LeafDesc* desc;
  const int* ranks;
  double* z;
  double* zpart;
  int* sync;      // ticket | err | row flags | zcount[nslots] | zready[nslots]
  int* fail;      // kSwFailInts, sticky (not cleared between launches)
  int64_t wait_ticks;
  int nitems, nslots, nrows;
  int fwd, nrhs, ldx;
  int zero;       // 0, unknown to the compiler: polls stay read-modify-writes
};

__device__ inline int* row_flag(const SweepArgs& a, int pos) {
  return a.sync + kSwRows + kSwLine * pos;
}
__device__ inline int* zcount(const SweepArgs& a, int slot) {
  return a.sync + kSwRows + kSwLine * a.nrows + slot;
}
__device__ inline int* zready(const SweepArgs& a, int slot) {
  return zcount(a, slot) + a.nslots;
}

__device__ inline int poll(const SweepArgs& a, int* p) {
  return __hip_atomic_fetch_add(p, a.zero, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ inline void raise(int* p) {
  __hip_atomic_fetch_add(p, 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}

// Thread 0: wait until *p >= v (then acquire).  False if the wait gave up.
__device__ bool wait_geq(const SweepArgs& a, int* p, int v, int ticket, int what, int which) {
  int cur = poll(a, p);
  if (cur < v) {
    const uint64_t t0 = wall_clock64();
    for (int it = 1;; ++it) {
      __builtin_amdgcn_s_sleep(2);
      cur = poll(a, p);
      if (cur >= v) break;
      if ((it & 31) == 0) {
        if (poll(a, a.sync + kSwErr)) return false;
        if (wall_clock64() - t0 > (uint64_t)a.wait_ticks) {
          __hip_atomic_fetch_or(a.sync + kSwErr, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          if (atomicCAS(a.fail, 0, 1) == 0) {
            atomicExch(a.fail + 1, ticket);
            atomicExch(a.fail + 2, what);
            atomicExch(a.fail + 3, which);
            atomicExch(a.fail + 4, v);
            atomicExch(a.fail + 5, cur);
          }
          return false;
        }
      }
    }
  }
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
  return true;
}

// Workgroup-wide: thread 0 waits, the others meet it at the barrier.
__device__ bool wg_wait(const SweepArgs& a, int* p, int v, int ticket, int what, int which,
                        int* s_ok) {
  if (threadIdx.x == 0) *s_ok = wait_geq(a, p, v, ticket, what, which) ? 1 : 0;
  __syncthreads();
  const bool ok = *s_ok != 0;
  __syncthreads();
  return ok;
}

// Copies rows [0, n) x NR of a column-major source into LDS (sv[j * NR + r]),
// zeros for r >= nr.
template <typename S, int NR>
__device__ inline void stage_rows(double* sv, const S* src, long ld, int n, int nr) {
  for (int e = threadIdx.x; e < n * NR; e += kBlock) {
    const int j = e / NR, r = e - j * NR;
    sv[e] = r < nr ? (double)src[j + (long)r * ld] : 0.0;
  }
}

template <typename T, int NR>
__device__ bool do_row(const SweepArgs& a, int ticket, int pos, const T* arena, const T* inv, T* x,
                       double* sv, double* red, int* s_ok) {
  const SwRow R = a.rows[pos];
  const int m = R.m;
  const int RW = (m + kWave - 1) / kWave * kWave;
  const int G = kBlock / RW;
  const int tid = threadIdx.x, g = tid / RW, i = tid - g * RW;
  const bool act = g < G && i < m;
  double acc[NR];
#pragma unroll
  for (int r = 0; r < NR; ++r) acc[r] = 0.0;
  int seen = 0;  // rows [0, seen) known done (uniform: every thread takes the same branches)
  for (int c = R.cbeg; c < R.cend; ++c) {
    const SwContrib C = a.contrib[c];
    const LeafDesc L = a.desc[C.leaf];
    int n_in;
    if (C.kind == SC_RK) {
      if (!wg_wait(a, zready(a, C.src), 1, ticket, 1, C.src, s_ok)) return false;
      n_in = a.ranks[C.leaf];
      for (int e = tid; e < n_in * NR; e += kBlock) sv[e] = a.z[C.z + e];
    } else {
      if (C.dep >= seen) {
        if (!wg_wait(a, row_flag(a, C.dep), 1, ticket, 0, C.dep, s_ok)) return false;
        seen = C.dep + 1;
      }
      n_in = C.kind == SC_FULL ? L.n : L.m;
      stage_rows<T, NR>(sv, x + C.src, a.ldx, n_in, a.nrhs);
    }
    __syncthreads();
    if (act) {
      if (C.kind == SC_FULL) {
        const T* M = arena + L.a + i;
        const long ld = L.m;
        for (int j = g; j < n_in; j += G) {
          const double v = (double)M[j * ld];
#pragma unroll
          for (int r = 0; r < NR; ++r) acc[r] -= v * sv[j * NR + r];
        }
      } else if (C.kind == SC_FULL_T) {
        const T* M = arena + L.a + (long)i * L.m;
        for (int j = g; j < n_in; j += G) {
          const double v = (double)M[j];
#pragma unroll
          for (int r = 0; r < NR; ++r) acc[r] -= v * sv[j * NR + r];
        }
      } else {
        const T* P = arena + (a.fwd ? L.a : L.b) + C.row0 + i;
        const long ld = a.fwd ? L.m : L.n;
        for (int l = g; l < n_in; l += G) {
          const double v = (double)P[l * ld];
#pragma unroll
          for (int r = 0; r < NR; ++r) acc[r] -= v * sv[l * NR + r];
        }
      }
    }
    __syncthreads();
  }
  // right-hand side of the rows, then the sum of the groups' partials
  if (act && g == 0)
    for (int r = 0; r < a.nrhs && r < NR; ++r) acc[r] += (double)x[R.r0 + i + (long)r * a.ldx];
  if (g < G)
#pragma unroll
    for (int r = 0; r < NR; ++r) red[tid * NR + r] = acc[r];
  __syncthreads();
  for (int e = tid; e < m * NR; e += kBlock) {
    const int ii = e / NR, r = e - ii * NR;
    double s = 0.0;
    for (int gg = 0; gg < G; ++gg) s += red[(gg * RW + ii) * NR + r];
    sv[e] = s;
  }
  __syncthreads();
  // diagonal block: x_q = D s (forward, D lower triangular) or D^T s (backward)
#pragma unroll
  for (int r = 0; r < NR; ++r) acc[r] = 0.0;
  if (act) {
    const T* D = inv + R.inv;
    if (a.fwd) {
      for (int j = g; j <= i; j += G) {
        const double v = (double)D[i + (long)j * m];
#pragma unroll
        for (int r = 0; r < NR; ++r) acc[r] += v * sv[j * NR + r];
      }
    } else {
      const int j0 = i + ((g - i) % G + G) % G;  // first j >= i with j = g (mod G)
      for (int j = j0; j < m; j += G) {
        const double v = (double)D[j + (long)i * m];
#pragma unroll
        for (int r = 0; r < NR; ++r) acc[r] += v * sv[j * NR + r];
      }
    }
  }
  if (g < G)
#pragma unroll
    for (int r = 0; r < NR; ++r) red[tid * NR + r] = acc[r];
  __syncthreads();
  for (int e = tid; e < m * NR; e += kBlock) {
    const int ii = e / NR, r = e - ii * NR;
    if (r >= a.nrhs) continue;
    double s = 0.0;
    for (int gg = 0; gg < G; ++gg) s += red[(gg * RW + ii) * NR + r];
    x[R.r0 + ii + (long)r * a.ldx] = (T)s;
  }
  // rows finish in sweep order: the previous row is done already (it is a
  // source of this one); the wait makes "flag raised" mean "all rows so far"
  if (pos > 0 && pos - 1 >= seen)
    if (!wg_wait(a, row_flag(a, pos - 1), 1, ticket, 0, pos - 1, s_ok)) return false;
  __syncthreads();
  if (tid == 0) raise(row_flag(a, pos));
  return true;
}

template <typename T, int NR>
__device__ bool do_zpart(const SweepArgs& a, int ticket, int idx, const T* arena, const T* x,
                         double* sv, int* s_ok) {
  const SwChunk K = a.chunks[idx];
  const LeafDesc L = a.desc[K.leaf];
  const int k = a.ranks[K.leaf];
  if (!wg_wait(a, row_flag(a, K.dep), 1, ticket, 0, K.dep, s_ok)) return false;
  const T* P = arena + (a.fwd ? L.b : L.a) + K.row0;
  const long ld = a.fwd ? L.n : L.m;
  const T* xs = x + (a.fwd ? L.c0 : L.r0) + K.row0;
  stage_rows<T, NR>(sv, xs, a.ldx, K.rows, a.nrhs);
  __syncthreads();
  double* out = (K.single ? a.z : a.zpart) + K.out;
  const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
  for (int l = w; l < k; l += kWaves) {
    const T* col = P + l * ld;
    double s[NR];
#pragma unroll
    for (int r = 0; r < NR; ++r) s[r] = 0.0;
    for (int ii = lane; ii < K.rows; ii += kWave) {
      const double v = (double)col[ii];
#pragma unroll
      for (int r = 0; r < NR; ++r) s[r] += v * sv[ii * NR + r];
    }
#pragma unroll
    for (int r = 0; r < NR; ++r) {
      const double t = wave_sum(s[r]);
      if (lane == 0) out[l * NR + r] = t;
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) raise(K.single ? zready(a, K.slot) : zcount(a, K.slot));
  return true;
}

template <int NR>
__device__ bool do_zred(const SweepArgs& a, int ticket, int idx, int* s_ok) {
  const SwRed D = a.reds[idx];
  const int k = a.ranks[D.leaf];
  if (!wg_wait(a, zcount(a, D.slot), D.cend - D.cbeg, ticket, 2, D.slot, s_ok)) return false;
  for (int e = threadIdx.x; e < k * NR; e += kBlock) {
    double s = 0.0;
    for (int c = D.cbeg; c < D.cend; ++c) s += a.zpart[a.chunks[c].out + e];
    a.z[D.z + e] = s;
  }
  __syncthreads();
  if (threadIdx.x == 0) raise(zready(a, D.slot));
  return true;
}

template <typename T, int NR>
__global__ void __launch_bounds__(kBlock) k_sweep(SweepArgs a, const T* arena, const T* inv, T* x) {
  __shared__ double sv[kSweepChunk * NR];
  __shared__ double red[kBlock * NR];
  __shared__ int s_t, s_ok;
  for (;;) {
    if (threadIdx.x == 0)
      s_t = poll(a, a.sync + kSwErr) ? a.nitems
                                     : __hip_atomic_fetch_add(a.sync + kSwTicket, 1, __ATOMIC_RELAXED,
                                                              __HIP_MEMORY_SCOPE_AGENT);
    __syncthreads();
    const int t = s_t;
    __syncthreads();
    if (t >= a.nitems) return;
    const SwItem it = a.items[t];
    bool ok;
    if (it.type == SW_ROW)
      ok = do_row<T, NR>(a, t, it.idx, arena, inv, x, sv, red, &s_ok);
    else if (it.type == SW_ZPART)
      ok = do_zpart<T, NR>(a, t, it.idx, arena, x, sv, &s_ok);
    else
      ok = do_zred<NR>(a, t, it.idx, &s_ok);
    if (!ok) return;
  }
}

// Unit matrices in the inverse buffer (one workgroup per diagonal block).
template <typename T>
__global__ void k_identity(T* inv, const SwRow* rows) {
  const SwRow R = rows[blockIdx.x];
  T* D = inv + R.inv;
  for (int e = threadIdx.x; e < R.m * R.m; e += blockDim.x) D[e] = (e % (R.m + 1)) == 0 ? T(1) : T(0);
}

}  // namespace

template <typename T>
void launch_sweep(const SweepLaunch& L, const T* arena, const T* inv, T* x, hipStream_t s) {
  SweepArgs a;
  a.items = L.items;
  a.rows = L.rows;
  a.contrib = L.contrib;
  a.chunks = L.chunks;
  a.reds = L.reds;
  a.desc = L.desc;
  a.ranks = L.ranks;
  a.z = L.z;
  a.zpart = L.zpart;
  a.sync = L.sync;
  a.fail = L.fail;
  a.wait_ticks = L.wait_ticks;
  a.nitems = L.nitems;
  a.nslots = L.nslots;
  a.nrows = L.nrows;
  a.zero = 0;
  a.fwd = L.fwd;
  a.nrhs = L.nrhs;
  a.ldx = L.ldx;
  if (L.nitems == 0) return;
  launch_fill_int(L.sync, (size_t)sweep_sync_ints(L.nrows, L.nslots), 0, s);
  const int grid = L.nitems < L.max_groups ? L.nitems : L.max_groups;
  if (L.nrhs == 1)
    hipLaunchKernelGGL((k_sweep<T, 1>), dim3(grid), dim3(kBlock), 0, s, a, arena, inv, x);
  else
    hipLaunchKernelGGL((k_sweep<T, kSweepNR>), dim3(grid), dim3(kBlock), 0, s, a, arena, inv, x);
  HMX_CHECK(hipGetLastError());
}

template <typename T>
void launch_sweep_identity(T* inv, const SwRow* rows, int count, hipStream_t s) {
  if (count == 0) return;
  hipLaunchKernelGGL(k_identity<T>, dim3(count), dim3(kBlock), 0, s, inv, rows);
  HMX_CHECK(hipGetLastError());
}

template void launch_sweep<double>(const SweepLaunch&, const double*, const double*, double*,
                                   hipStream_t);
template void launch_sweep<float>(const SweepLaunch&, const float*, const float*, float*,
                                  hipStream_t);
template void launch_sweep_identity<double>(double*, const SwRow*, int, hipStream_t);
template void launch_sweep_identity<float>(float*, const SwRow*, int, hipStream_t);

}  // namespace hmx
