// Leaf-level H-gemv (h_matrix.cpp:792-857, rk_matrix.cpp:156-206) as four
// deterministic passes over HBM-resident leaves:
//   zpass   : z = B^T x[cols] (and z' = A^T x[rows] for mirrored Rk) per panel chunk
//   zreduce : sum the chunk partials in a fixed order
//   ypass   : one workgroup per destination leaf cluster: full leaves of that
//             block row (M x, plus M^T x into a slot when mirrored), Rk A z and
//             mirrored Rk B z' restricted to the cluster's rows
//   yfinal  : y = beta y + alpha (acc + mirrored full slots), fixed order
// Every stored leaf is streamed once per pass that needs it; accumulation is
// in FP64 whatever the storage type.
#include <algorithm>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "kernels.hpp"

namespace hmx {

namespace {

constexpr int kRhsChunk = 8;

template <typename T>
__global__ void __launch_bounds__(kBlock) k_zpass(const LeafDesc* desc, const int* ranks,
                                                  const T* arena, const ZItem* items, const T* x,
                                                  int ldx, int nrhs, double* zpart) {
  const ZItem it = items[blockIdx.x];
  const LeafDesc d = desc[it.leaf];
  const int k = ranks[it.leaf];
  const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
  const int ld = it.side == 0 ? d.n : d.m;
  const T* P = arena + (it.side == 0 ? d.b : d.a) + it.row0;
  const T* xs = x + (it.side == 0 ? d.c0 : d.r0) + it.row0;
  double* z = zpart + it.zslot * nrhs;
  for (int l = w; l < k; l += kWaves) {
    const T* col = P + (size_t)l * ld;
    for (int r0 = 0; r0 < nrhs; r0 += kRhsChunk) {
      const int nr = nrhs - r0 < kRhsChunk ? nrhs - r0 : kRhsChunk;
      double s[kRhsChunk];
#pragma unroll
      for (int r = 0; r < kRhsChunk; ++r) s[r] = 0.0;
      for (int i = lane; i < it.rows; i += kWave) {
        const double p = (double)col[i];
#pragma unroll
        for (int r = 0; r < kRhsChunk; ++r)
          if (r < nr) s[r] += p * (double)xs[i + (size_t)(r0 + r) * ldx];
      }
#pragma unroll
      for (int r = 0; r < kRhsChunk; ++r) {
        if (r < nr) {
          const double v = wave_sum(s[r]);
          if (lane == 0) z[l + (size_t)(r0 + r) * d.cap] = v;
        }
      }
    }
  }
}

// One right-hand side.  A chunk of at most 64 kR64 rows: a lane keeps its
// entries of x (rows lane + 64 v) in registers, and a wavefront issues the
// loads of kZChunk / (64 kR64) panel columns together (indices clamped to the
// chunk and the rank) before it adds them up, column by column in row order.
template <typename T, int kR64>
__device__ void zpass1_chunk(const T* P, int ld, const T* xs, int rows, int k, double* z) {
  constexpr int kC = kZChunk / kWave / kR64;
  const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
  const int rm = rows - 1;
  double xr[kR64];
#pragma unroll
  for (int v = 0; v < kR64; ++v) {
    const int i = lane + kWave * v;
    const double t = (double)xs[i < rm ? i : rm];
    xr[v] = i < rows ? t : 0.0;
  }
  for (int l0 = w * kC; l0 < k; l0 += kWaves * kC) {
    double p[kC][kR64];
#pragma unroll
    for (int c = 0; c < kC; ++c) {
      const T* col = P + (size_t)(l0 + c < k ? l0 + c : k - 1) * ld;
#pragma unroll
      for (int v = 0; v < kR64; ++v) {
        const int i = lane + kWave * v;
        p[c][v] = (double)col[i < rm ? i : rm];
      }
    }
#pragma unroll
    for (int c = 0; c < kC; ++c) {
      double sum = 0.0;
#pragma unroll
      for (int v = 0; v < kR64; ++v) sum += p[c][v] * xr[v];
      sum = wave_sum(sum);
      if (lane == 0 && l0 + c < k) z[l0 + c] = sum;
    }
  }
}

template <typename T>
__global__ void __launch_bounds__(kBlock) k_zpass1(const LeafDesc* desc, const int* ranks,
                                                   const T* arena, const ZItem* items, const T* x,
                                                   double* zpart) {
  const ZItem it = items[blockIdx.x];
  const LeafDesc d = desc[it.leaf];
  const int k = ranks[it.leaf];
  const int ld = it.side == 0 ? d.n : d.m;
  const T* P = arena + (it.side == 0 ? d.b : d.a) + it.row0;
  const T* xs = x + (it.side == 0 ? d.c0 : d.r0) + it.row0;
  double* z = zpart + it.zslot;
  if (it.rows <= 64) zpass1_chunk<T, 1>(P, ld, xs, it.rows, k, z);
  else if (it.rows <= 128) zpass1_chunk<T, 2>(P, ld, xs, it.rows, k, z);
  else if (it.rows <= 256) zpass1_chunk<T, 4>(P, ld, xs, it.rows, k, z);
  else if (it.rows <= 512) zpass1_chunk<T, 8>(P, ld, xs, it.rows, k, z);
  else zpass1_chunk<T, 16>(P, ld, xs, it.rows, k, z);
}

__global__ void k_zreduce(const ZReduce* red, const int* ranks, const double* zpart,
                          const ZItem* items, int nrhs, const LeafDesc* desc, double* z) {
  const ZReduce rd = red[blockIdx.x];
  const int k = ranks[rd.leaf];
  const int cap = desc[rd.leaf].cap;
  for (int e = threadIdx.x; e < k * nrhs; e += blockDim.x) {
    const int l = e % k, r = e / k;
    double s = 0.0;
    for (int q = rd.beg; q < rd.end; ++q) s += zpart[items[q].zslot * nrhs + l + (size_t)r * cap];
    z[rd.z * nrhs + l + (size_t)r * cap] = s;
  }
}

// y pass.  One workgroup per destination leaf cluster I (at most 128 kQ rows).
// The cluster's contribution list is first described in LDS (one parallel
// load of its records), so the loops below issue data loads only.
//  * Full leaves (I x J) are staged in LDS in column chunks with x_J; threads
//    < 128 form M x_J (rows t + 128 q), and for a mirrored leaf threads >= 128
//    form M^T x_I column by column into the leaf's slot.  No cross-lane
//    reductions.
//  * Rk slices (A z, or B z' of a mirrored Rk leaf) are taken four at a time,
//    one per wavefront; a lane covers rows lane + 64 q of I, so one round trip
//    to HBM serves four slices.
// The partial sums of the row owners are added in a fixed order at the end.
// kR right-hand sides per pass.
// doubles of the leaf tile: 30 KiB with one right-hand side (the whole
// workgroup then takes 39 KiB of LDS, four per CU), 40 KiB with more
template <int kR>
constexpr int y_tile() { return kR == 1 ? 3840 : 5120; }
constexpr int kYMeta = 128;    // contribution records described in LDS at a time

template <typename T, int kR, int kQ>
__global__ void __launch_bounds__(kBlock) k_ypass(const LeafDesc* desc, const int* ranks,
                                                  const T* arena, const YItem* items,
                                                  const YContrib* contrib, const double* z,
                                                  double* slots, const T* x, int ldx, int nrhs,
                                                  double* acc_out, int ldacc) {
  constexpr int kYTile = y_tile<kR>();
  constexpr int kRows = 128 * kQ;  // rows of a destination leaf cluster (at most)
  static_assert(5 * kRows <= kYTile, "the partial sums are gathered in the leaf tile");
  __shared__ double tile[kYTile];
  __shared__ double xI[kRows * kR];  // x_I: kRows rows, x_J: 256 columns per right-hand side
  __shared__ double xJ[256 * kR];
  __shared__ const T* mp[kYMeta];      // full: leaf data; rk: first row of the slice
  __shared__ const double* mz[kYMeta]; // rk: z; full mirrored: slot (as double*)
  __shared__ int mkind[kYMeta], mk[kYMeta], mld[kYMeta], mcap[kYMeta], mc0[kYMeta], mn[kYMeta];
  const YItem it = items[blockIdx.x];
  const int tid = threadIdx.x, lane = tid & (kWave - 1), wv = tid / kWave;
  const bool lo = tid < 128;
  const int h = lo ? tid : tid - 128;
  for (int rc = 0; rc < nrhs; rc += kR) {
    const int nr = nrhs - rc < kR ? nrhs - rc : kR;
    __syncthreads();
    for (int e = tid; e < it.len * kR; e += kBlock) {
      const int i = e % it.len, r = e / it.len;
      xI[i + r * kRows] = r < nr ? (double)x[it.r0 + i + (size_t)(rc + r) * ldx] : 0.0;
    }
    double accf[kQ][kR];      // full leaves: rows h + 128 q (threads < 128)
    double accr[2 * kQ][kR];  // Rk slices: rows lane + 64 q
#pragma unroll
    for (int r = 0; r < kR; ++r) {
#pragma unroll
      for (int q = 0; q < kQ; ++q) accf[q][r] = 0.0;
#pragma unroll
      for (int q = 0; q < 2 * kQ; ++q) accr[q][r] = 0.0;
    }
    for (int cb = it.beg; cb < it.end; cb += kYMeta) {
      const int nc = it.end - cb < kYMeta ? it.end - cb : kYMeta;
      __syncthreads();
      if (tid < nc) {
        const YContrib c = contrib[cb + tid];
        const LeafDesc d = desc[c.leaf];
        mkind[tid] = c.kind;
        if (c.kind == 0) {
          mp[tid] = arena + d.a;
          mz[tid] = c.z >= 0 ? slots + c.z * nrhs : nullptr;
          mk[tid] = d.m;
          mn[tid] = d.n;
          mc0[tid] = d.c0;
        } else {
          const bool isA = c.kind == 2;
          mp[tid] = arena + (isA ? d.a : d.b) + c.row0;
          mz[tid] = z + c.z * nrhs;
          mk[tid] = ranks[c.leaf];
          mld[tid] = isA ? d.m : d.n;
          mcap[tid] = d.cap;
        }
      }
      __syncthreads();
      int ci = 0;
      // full leaves (the plan lists them first)
      for (; ci < nc && mkind[ci] == 0; ++ci) {
        const T* M = mp[ci];
        const int m = mk[ci], n = mn[ci], c0 = mc0[ci];
        double* slot = const_cast<double*>(mz[ci]);
        const int ldt = m | 1;
        const int cw_max = kYTile / ldt < 256 ? kYTile / ldt : 256;
        const int ti0 = tid % m, tj0 = tid / m, tdr = kBlock % m, tdq = kBlock / m;
        // m cw <= kYTile: every load of a column chunk is issued in one round
        // (indices clamped) into registers; with one right-hand side the next
        // chunk's loads are in flight while the current one is used from LDS
        // (with more, the registers go to the accumulators)
        constexpr int kTileLoads = (kYTile + kBlock - 1) / kBlock;
        constexpr bool kPrefetch = kR == 1;
        double v[kTileLoads];
        auto load_chunk = [&](int j0) {
          const int tot0 = m * (n - j0 < cw_max ? n - j0 : cw_max);
          const T* src = M + (size_t)j0 * m;
#pragma unroll
          for (int u = 0; u < kTileLoads; ++u) {
            const int e = tid + u * kBlock;
            v[u] = (double)src[e < tot0 ? e : tot0 - 1];
          }
        };
        if (kPrefetch) load_chunk(0);
        for (int j0 = 0; j0 < n; j0 += cw_max) {
          const int cw = n - j0 < cw_max ? n - j0 : cw_max;
          __syncthreads();
          if (!kPrefetch) load_chunk(j0);
          // element e = i + m j of the chunk goes to i + ldt j, with (i, j)
          // stepped by kBlock without a division
          const int tot = m * cw;
          int ti = ti0, tj = tj0;
#pragma unroll
          for (int u = 0; u < kTileLoads; ++u) {
            if (tid + u * kBlock < tot) tile[ti + tj * ldt] = v[u];
            ti += tdr;
            tj += tdq;
            if (ti >= m) {
              ti -= m;
              ++tj;
            }
          }
          for (int e = tid; e < cw * kR; e += kBlock) {
            const int j = e % cw, r = e / cw;
            xJ[j + r * 256] = r < nr ? (double)x[c0 + j0 + j + (size_t)(rc + r) * ldx] : 0.0;
          }
          __syncthreads();
          if (kPrefetch && j0 + cw_max < n) load_chunk(j0 + cw_max);
          if (lo) {
#pragma unroll
            for (int q = 0; q < kQ; ++q) {
              const int i = h + 128 * q;
              if (i >= m) continue;
              const double* row = tile + i;
#pragma unroll 4
              for (int jj = 0; jj < cw; ++jj) {
                const double v = row[jj * ldt];
#pragma unroll
                for (int r = 0; r < kR; ++r) accf[q][r] += v * xJ[jj + r * 256];
              }
            }
          } else if (slot) {
            for (int jj = h; jj < cw; jj += 128) {
              double t[kR];
#pragma unroll
              for (int r = 0; r < kR; ++r) t[r] = 0.0;
              const double* col = tile + jj * ldt;
#pragma unroll 4
              for (int i = 0; i < m; ++i) {
                const double v = col[i];
#pragma unroll
                for (int r = 0; r < kR; ++r) t[r] += v * xI[i + r * kRows];
              }
#pragma unroll
              for (int r = 0; r < kR; ++r)
                if (r < nr) slot[j0 + jj + (size_t)(rc + r) * n] = t[r];
            }
          }
        }
      }
      // Rk slices: wavefront wv takes slices ci + wv, ci + wv + 4, ...; with
      // one right-hand side it streams two slices together (twice the loads
      // in flight), with more the accumulators use the registers instead
      if constexpr (kR > 1) {
        for (int cj = ci + wv; cj < nc; cj += kWaves) {
          const T* P = mp[cj];
          const double* zz = mz[cj] + (size_t)rc * mcap[cj];
          const int k = mk[cj], ld = mld[cj], cap = mcap[cj];
#pragma unroll 4
          for (int l = 0; l < k; ++l) {
            double zl[kR];
#pragma unroll
            for (int r = 0; r < kR; ++r) zl[r] = r < nr ? zz[l + (size_t)r * cap] : 0.0;
            const T* pl = P + (size_t)l * ld;
#pragma unroll
            for (int q = 0; q < 2 * kQ; ++q) {
              const int i = lane + 64 * q;
              if (i < it.len) {
                const double v = (double)pl[i];
#pragma unroll
                for (int r = 0; r < kR; ++r) accr[q][r] += v * zl[r];
              }
            }
          }
        }
      } else
      for (int cj = ci + wv; cj < nc; cj += 2 * kWaves) {
        const int cj2 = cj + kWaves < nc ? cj + kWaves : cj;
        const T* P0 = mp[cj];
        const T* P1 = mp[cj2];
        const double* z0 = mz[cj] + (size_t)rc * mcap[cj];
        const double* z1 = mz[cj2] + (size_t)rc * mcap[cj2];
        const int k0 = mk[cj], ld0 = mld[cj], cap0 = mcap[cj];
        const int k1 = cj2 != cj ? mk[cj2] : 0, ld1 = mld[cj2], cap1 = mcap[cj2];
        const int kmax = k0 > k1 ? k0 : k1;
#pragma unroll 2
        for (int l = 0; l < kmax; ++l) {
          double za[kR], zb[kR];
#pragma unroll
          for (int r = 0; r < kR; ++r) {
            za[r] = (l < k0 && r < nr) ? z0[l + (size_t)r * cap0] : 0.0;
            zb[r] = (l < k1 && r < nr) ? z1[l + (size_t)r * cap1] : 0.0;
          }
          const T* pa = P0 + (size_t)(l < k0 ? l : 0) * ld0;
          const T* pb = P1 + (size_t)(l < k1 ? l : 0) * ld1;
#pragma unroll
          for (int q = 0; q < 2 * kQ; ++q) {
            const int i = lane + 64 * q;
            if (i < it.len) {
              const double va = l < k0 ? (double)pa[i] : 0.0;
              const double vb = l < k1 ? (double)pb[i] : 0.0;
#pragma unroll
              for (int r = 0; r < kR; ++r) accr[q][r] += va * za[r] + vb * zb[r];
            }
          }
        }
      }
    }
    // fixed-order sum of the partials: full part (threads < 128) + the four waves
#pragma unroll
    for (int r = 0; r < kR; ++r) {
      __syncthreads();
      if (lo) {
#pragma unroll
        for (int q = 0; q < kQ; ++q) tile[4 * kRows + h + 128 * q] = accf[q][r];
      }
#pragma unroll
      for (int q = 0; q < 2 * kQ; ++q) tile[wv * kRows + lane + 64 * q] = accr[q][r];
      __syncthreads();
      if (r < nr)
        for (int i = tid; i < it.len; i += kBlock)
          acc_out[it.r0 + i + (size_t)(rc + r) * ldacc] =
              tile[4 * kRows + i] +
              ((tile[i] + tile[kRows + i]) + (tile[2 * kRows + i] + tile[3 * kRows + i]));
    }
  }
}

template <typename T>
__global__ void k_yfinal(const YItem* items, const int64_t* slot_list, const double* slots,
                         const double* acc, T* y, int ldy, int ldacc, int nrhs, double alpha,
                         double beta) {
  const YItem it = items[blockIdx.x];
  for (int e = threadIdx.x; e < it.len * nrhs; e += blockDim.x) {
    const int i = e % it.len, r = e / it.len;
    double s = acc[it.r0 + i + (size_t)r * ldacc];
    for (int q = it.sbeg; q < it.send; ++q) s += slots[slot_list[q] * nrhs + i + (size_t)r * it.len];
    T* py = y + it.r0 + i + (size_t)r * ldy;
    const double old = (beta == 0.0) ? 0.0 : beta * (double)(*py);
    *py = (T)(old + alpha * s);
  }
}

template <typename S, typename D>
__global__ void k_convert(const S* src, D* dst, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    dst[i] = (D)src[i];
}

}  // namespace

template <typename S, typename D>
void launch_convert(const S* src, D* dst, size_t n, hipStream_t s) {
  if (n == 0) return;
  size_t blocks = (n + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  hipLaunchKernelGGL((k_convert<S, D>), dim3((unsigned)blocks), dim3(256), 0, s, src, dst, n);
  HMX_CHECK(hipGetLastError());
}
template void launch_convert<double, float>(const double*, float*, size_t, hipStream_t);

namespace {
__global__ void k_add_f2d(double* x, const float* d, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    x[i] += (double)d[i];
}
__global__ void k_col_sumsq(const double* v, int n, double* out) {
  __shared__ double red[kWaves];
  const double* c = v + (size_t)blockIdx.x * n;
  double s = 0.0;
  for (int i = threadIdx.x; i < n; i += kBlock) s += c[i] * c[i];
  s = block_sum(s, red);
  if (threadIdx.x == 0) out[blockIdx.x] = s;
}
}  // namespace

void launch_vec_add_f2d(double* x, const float* d, size_t n, hipStream_t s) {
  size_t blocks = std::min<size_t>((n + 255) / 256, 65536);
  hipLaunchKernelGGL(k_add_f2d, dim3((unsigned)blocks), dim3(256), 0, s, x, d, n);
  HMX_CHECK(hipGetLastError());
}
void launch_col_sumsq(const double* v, int n, int nrhs, double* out, hipStream_t s) {
  hipLaunchKernelGGL(k_col_sumsq, dim3(nrhs), dim3(kBlock), 0, s, v, n, out);
  HMX_CHECK(hipGetLastError());
}

// ---- conjugate-gradient steps on n x nrhs column blocks (one system per column)
namespace {
__global__ void k_col_dot(const double* a, const double* b, int n, double* out) {
  __shared__ double red[kWaves];
  const double* ca = a + (size_t)blockIdx.x * n;
  const double* cb = b + (size_t)blockIdx.x * n;
  double s = 0.0;
  for (int i = threadIdx.x; i < n; i += kBlock) s += ca[i] * cb[i];
  s = block_sum(s, red);
  if (threadIdx.x == 0) out[blockIdx.x] = s;
}
// alpha_j = rz_j / pq_j (0 when pq_j is 0): x += alpha p, r -= alpha q
__global__ void k_cg_xr(double* x, double* r, const double* p, const double* q, const double* rz,
                        const double* pq, int n, int nrhs) {
  const size_t cnt = (size_t)n * nrhs;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < cnt; e += (size_t)gridDim.x * blockDim.x) {
    const int j = (int)(e / n);
    const double a = pq[j] != 0.0 ? rz[j] / pq[j] : 0.0;
    x[e] += a * p[e];
    r[e] -= a * q[e];
  }
}
// beta_j = rz_new_j / rz_old_j (0 when rz_old_j is 0): p = z + beta p
__global__ void k_cg_p(double* p, const double* z, const double* rz_new, const double* rz_old, int n, int nrhs) {
  const size_t cnt = (size_t)n * nrhs;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < cnt; e += (size_t)gridDim.x * blockDim.x) {
    const int j = (int)(e / n);
    const double b = rz_old[j] != 0.0 ? rz_new[j] / rz_old[j] : 0.0;
    p[e] = z[e] + b * p[e];
  }
}
unsigned grid_for(size_t cnt) { return (unsigned)std::min<size_t>((cnt + 255) / 256, 65536); }
}  // namespace

namespace {
__global__ void k_vec_add(double* x, const double* d, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    x[i] += d[i];
}
}  // namespace
void launch_vec_add(double* x, const double* d, size_t n, hipStream_t s) {
  hipLaunchKernelGGL(k_vec_add, dim3(grid_for(n)), dim3(256), 0, s, x, d, n);
  HMX_CHECK(hipGetLastError());
}
void launch_col_dot(const double* a, const double* b, int n, int nrhs, double* out, hipStream_t s) {
  hipLaunchKernelGGL(k_col_dot, dim3(nrhs), dim3(kBlock), 0, s, a, b, n, out);
  HMX_CHECK(hipGetLastError());
}
void launch_cg_xr(double* x, double* r, const double* p, const double* q, const double* rz, const double* pq,
                  int n, int nrhs, hipStream_t s) {
  hipLaunchKernelGGL(k_cg_xr, dim3(grid_for((size_t)n * nrhs)), dim3(256), 0, s, x, r, p, q, rz, pq, n, nrhs);
  HMX_CHECK(hipGetLastError());
}
void launch_cg_p(double* p, const double* z, const double* rz_new, const double* rz_old, int n, int nrhs,
                 hipStream_t s) {
  hipLaunchKernelGGL(k_cg_p, dim3(grid_for((size_t)n * nrhs)), dim3(256), 0, s, p, z, rz_new, rz_old, n, nrhs);
  HMX_CHECK(hipGetLastError());
}
template void launch_convert<float, double>(const float*, double*, size_t, hipStream_t);

template <typename T>
void launch_gemv_zpass(const LeafDesc* desc, const int* ranks, const T* arena, const ZItem* items,
                       int count, const T* x, int ldx, int nrhs, double* zpart, hipStream_t s) {
  if (count == 0) return;
  if (nrhs == 1)
    hipLaunchKernelGGL(k_zpass1<T>, dim3(count), dim3(kBlock), 0, s, desc, ranks, arena, items, x,
                       zpart);
  else
    hipLaunchKernelGGL(k_zpass<T>, dim3(count), dim3(kBlock), 0, s, desc, ranks, arena, items, x,
                       ldx, nrhs, zpart);
  HMX_CHECK(hipGetLastError());
}

void launch_gemv_zreduce(const ZReduce* red, int count, const int* ranks, const double* zpart,
                         const ZItem* items, int nrhs, const LeafDesc* desc, double* z,
                         hipStream_t s) {
  if (count == 0) return;
  hipLaunchKernelGGL(k_zreduce, dim3(count), dim3(128), 0, s, red, ranks, zpart, items, nrhs,
                     desc, z);
  HMX_CHECK(hipGetLastError());
}

template <typename T>
void launch_gemv_ypass(const LeafDesc* desc, const int* ranks, const T* arena, const YItem* items,
                       int count, const YContrib* contrib, const double* z, double* slots,
                       const T* x, int ldx, int nrhs, double* acc, int ldacc, int max_rows,
                       hipStream_t s) {
  if (count == 0) return;
  if (max_rows > 512) throw std::runtime_error("gemv: leaf clusters above 512 rows");
  // destination leaf clusters of at most 256 rows (the usual case) or 512
  // (twice the registers and the x_I staging, fewer workgroups per CU)
  if (max_rows <= 256) {
    if (nrhs == 1)
      hipLaunchKernelGGL((k_ypass<T, 1, 2>), dim3(count), dim3(kBlock), 0, s, desc, ranks, arena,
                         items, contrib, z, slots, x, ldx, nrhs, acc, ldacc);
    else
      hipLaunchKernelGGL((k_ypass<T, kRhsChunk, 2>), dim3(count), dim3(kBlock), 0, s, desc, ranks,
                         arena, items, contrib, z, slots, x, ldx, nrhs, acc, ldacc);
  } else {
    if (nrhs == 1)
      hipLaunchKernelGGL((k_ypass<T, 1, 4>), dim3(count), dim3(kBlock), 0, s, desc, ranks, arena,
                         items, contrib, z, slots, x, ldx, nrhs, acc, ldacc);
    else
      hipLaunchKernelGGL((k_ypass<T, kRhsChunk, 4>), dim3(count), dim3(kBlock), 0, s, desc, ranks,
                         arena, items, contrib, z, slots, x, ldx, nrhs, acc, ldacc);
  }
  HMX_CHECK(hipGetLastError());
}

template <typename T>
void launch_gemv_yfinal(const YItem* items, int count, const int64_t* slot_list,
                        const double* slots, const double* acc, int ldacc, T* y, int ldy,
                        int nrhs, double alpha, double beta, hipStream_t s) {
  if (count == 0) return;
  hipLaunchKernelGGL(k_yfinal<T>, dim3(count), dim3(128), 0, s, items, slot_list, slots, acc, y,
                     ldy, ldacc, nrhs, alpha, beta);
  HMX_CHECK(hipGetLastError());
}

#define HMX_INST(T)                                                                            \
  template void launch_gemv_zpass<T>(const LeafDesc*, const int*, const T*, const ZItem*, int, \
                                     const T*, int, int, double*, hipStream_t);                 \
  template void launch_gemv_ypass<T>(const LeafDesc*, const int*, const T*, const YItem*, int, \
                                     const YContrib*, const double*, double*, const T*, int,   \
                                     int, double*, int, int, hipStream_t);                      \
  template void launch_gemv_yfinal<T>(const YItem*, int, const int64_t*, const double*,        \
                                      const double*, int, T*, int, int, double, double,         \
                                      hipStream_t);
HMX_INST(double)
HMX_INST(float)

// Device scratch of the allocation-free H-gemv entry point: a fixed array in
// the code object, so each GPU that loads the library holds its own copy (at
// its own address) until the process ends.  Every call on one GPU shares it
// and leaves its contents behind; the H-gemv passes write every scratch entry
// they read, so nothing needs clearing between calls.  Host code must not use
// the array's name as a device address: noalloc_gemv_scratch() returns the
// device address looked up with hipGetSymbolAddress, once per device.
static __device__ __align__(256) double scratch[kNoAllocScratchBytes / sizeof(double)];

double* noalloc_gemv_scratch(size_t* elems) {
  static std::mutex mu;
  static std::vector<double*> ptr;
  int dev = 0;
  HMX_CHECK(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lock(mu);
  if ((int)ptr.size() <= dev) ptr.resize(dev + 1, nullptr);
  if (!ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      throw std::runtime_error(std::string("hipGetSymbolAddress(scratch) failed: ") +
                               hipGetErrorString(e));
    ptr[dev] = static_cast<double*>(p);
  }
  *elems = kNoAllocScratchBytes / sizeof(double);
  return ptr[dev];
}

}  // namespace hmx
