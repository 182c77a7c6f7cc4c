// Device-side helpers shared by the HIP kernels: wavefront/workgroup
// reductions for 64-wide wavefronts and the built-in covariance kernels.
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>

#define HMX_CHECK(expr)                                                                    \
  do {                                                                                     \
    hipError_t e_ = (expr);                                                                \
    if (e_ != hipSuccess)                                                                  \
      throw std::runtime_error(std::string("HIP error ") + hipGetErrorString(e_) + " at " + \
                               __FILE__ + ":" + std::to_string(__LINE__));                 \
  } while (0)

namespace hmx {

constexpr int kWave = 64;
constexpr int kBlock = 256;  // threads per workgroup for the one-workgroup-per-task kernels
constexpr int kWaves = kBlock / kWave;

// Built-in interaction kernels (user numbering for the nugget).
enum KernelKind : int {
  kExponential = 0,
  kGaussian = 1,
  kShiftedExponential = 2,
  kExponentialDLength = 3,  // d/dl of exp(-r / l): (r / l^2) exp(-r / l)
  kGaussianDLength = 4,     // d/dl of exp(-r^2 / l^2): (2 r^2 / l^3) exp(-r^2 / l^2)
  kHelmholtz = 5            // complex, see helmholtz_embedded
};

struct KernelParams {
  int kind;
  double length;
  double nugget;
  double shift;             // kShiftedExponential: z offset of the source (non-symmetric kernel)
  double wavenumber = 0.0;  // kHelmholtz: k
  double reg = 0.0;         // kHelmholtz: h, the distance regularisation
};

// Regularised Helmholtz kernel z = exp(i k r) / (4 pi r), r = sqrt(d^2 + h^2),
// and 1 / (4 pi h) + nugget on the diagonal (same point).
__host__ __device__ inline void helmholtz_value(const KernelParams& k, double dx, double dy,
                                                double dz, bool same, double* re, double* im) {
  constexpr double kFourPi = 12.566370614359172;
  if (same) {
    *re = 1.0 / (kFourPi * k.reg) + k.nugget;
    *im = 0.0;
    return;
  }
  const double r = sqrt(dx * dx + dy * dy + dz * dz + k.reg * k.reg);
  const double s = 1.0 / (kFourPi * r);
  *re = cos(k.wavenumber * r) * s;
  *im = sin(k.wavenumber * r) * s;
}

// Entry (2u + a, 2v + b) of the 2 x 2 real representation [[Re z, -Im z],
// [Im z, Re z]] of the complex entry z = K(u, v): ur = 2u + a, uc = 2v + b.
__host__ __device__ inline double helmholtz_embedded(const KernelParams& k, double dx, double dy,
                                                     double dz, int ur, int uc) {
  double re, im;
  helmholtz_value(k, dx, dy, dz, (ur >> 1) == (uc >> 1), &re, &im);
  const int a = ur & 1, b = uc & 1;
  return a == b ? re : (a ? im : -im);
}

__host__ __device__ inline double kernel_value(const KernelParams& k, double dx, double dy,
                                               double dz, bool same) {
  if (k.kind == kShiftedExponential) dz -= k.shift;
  const double r2 = dx * dx + dy * dy + dz * dz;
  double v;
  if (k.kind == kGaussian) {
    v = exp(-r2 / (k.length * k.length));
  } else if (k.kind == kExponentialDLength) {
    const double r = sqrt(r2);
    v = r / (k.length * k.length) * exp(-r / k.length);
  } else if (k.kind == kGaussianDLength) {
    const double l2 = k.length * k.length;
    v = 2.0 * r2 / (l2 * k.length) * exp(-r2 / l2);
  } else {
    v = exp(-sqrt(r2) / k.length);
  }
  if (same) v += k.nugget;
  return v;
}

// ---- staging global -> LDS ----------------------------------------------------
// Copies an m x n block into LDS, dst[i + j * ldd] = src[i * si + j * sj] for
// i < mv, j < nv and 0 elsewhere.  Each thread issues kStage independent loads
// before storing them: a plain "for (e...) lds[..] = global[..]" loop waits
// for every load before its store, one memory round trip per iteration.
// kS (default kStage) loads per thread per round; one-shot staging of a
// whole leaf takes a larger kS so that it is one or two round trips.
constexpr int kStage = 8;

template <typename D, typename S, int kS = kStage>
__device__ inline void stage_in(D* dst, int ldd, const S* src, long si, long sj, int m, int n,
                                int mv, int nv, double scale = 1.0) {
  const int total = m * n;
  if (total <= 0) return;
  // element e = i + j m of thread t is t + kBlock u: the row and column step by
  // kBlock mod m and kBlock / m (one division per thread, not per element);
  // the loads read a clamped entry and select, so they issue back to back
  const int dq = kBlock / m, dr = kBlock % m;
  const bool any = mv > 0 && nv > 0;
  const int im = (mv < m ? mv : m) - 1, jm = (nv < n ? nv : n) - 1;
  int i = threadIdx.x % m, j = threadIdx.x / m;
  for (int base = threadIdx.x; base < total; base += kS * kBlock) {
    D v[kS];
    int pi[kS], pj[kS];
#pragma unroll
    for (int u = 0; u < kS; ++u) {
      const int e = base + u * kBlock;
      pi[u] = i;
      pj[u] = j;
      const bool ok = any && e < total && i < mv && j < nv;
      const int ic = any ? (i < im ? i : im) : 0, jc = any ? (j < jm ? j : jm) : 0;
      const S x = any ? src[ic * si + jc * sj] : (S)0;
      v[u] = ok ? (D)(scale * (double)x) : (D)0;
      i += dr;
      j += dq;
      if (i >= m) {
        i -= m;
        ++j;
      }
    }
#pragma unroll
    for (int u = 0; u < kS; ++u) {
      const int e = base + u * kBlock;
      if (e < total) dst[pi[u] + pj[u] * ldd] = v[u];
    }
  }
}
template <typename D, typename S, int kS = kStage>
__device__ inline void stage_in(D* dst, int ldd, const S* src, long si, long sj, int m, int n) {
  stage_in<D, S, kS>(dst, ldd, src, si, sj, m, n, m, n);
}

// ---- reductions -------------------------------------------------------------
// Lane exchange through DPP (a VALU operand modifier, no LDS round trip).
// kCtrl: 0xB1 swaps neighbours (quad_perm 1,0,3,2), 0x4E swaps pairs
// (quad_perm 2,3,0,1), 0x141 mirrors each half-row of 8, 0x140 each row of 16.
template <int kCtrl>
__device__ inline double dpp_swap(double v) {
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), kCtrl, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), kCtrl, 0xF, 0xF, false);
  return __hiloint2double(hi, lo);
}
template <int kCtrl>
__device__ inline float dpp_swap(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), kCtrl, 0xF, 0xF, false));
}

// Value of a lane, read into every lane (readlane; ignores the exec mask).
__device__ inline double lane_read(double v, int lane) {
  const int lo = __builtin_amdgcn_readlane(__double2loint(v), lane);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(v), lane);
  return __hiloint2double(hi, lo);
}
__device__ inline float lane_read(float v, int lane) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}

// Sum over aligned groups of g lanes (g a power of two <= 64, uniform across
// the wavefront; every lane of a group must be active).  Each step adds the
// value of a partner lane, so every lane of a group ends with the same,
// bitwise identical sum.
template <typename T>
__device__ inline T group_sum(T v, int g) {
  if (g >= 2) v += dpp_swap<0xB1>(v);
  if (g >= 4) v += dpp_swap<0x4E>(v);
  if (g >= 8) v += dpp_swap<0x141>(v);
  if (g >= 16) v += dpp_swap<0x140>(v);
  if (g == 64) return (lane_read(v, 0) + lane_read(v, 16)) + (lane_read(v, 32) + lane_read(v, 48));
  if (g >= 32) v += __shfl_xor(v, 16, kWave);
  return v;
}

// Sum over the whole wavefront (all 64 lanes active): DPP sums within each
// row of 16 lanes, then the four row sums are read out with readlane and
// added in a fixed order, so every lane returns the same value.
template <typename T>
__device__ inline T wave_sum(T v) {
  v += dpp_swap<0xB1>(v);
  v += dpp_swap<0x4E>(v);
  v += dpp_swap<0x141>(v);
  v += dpp_swap<0x140>(v);
  return (lane_read(v, 0) + lane_read(v, 16)) + (lane_read(v, 32) + lane_read(v, 48));
}

// Sum over the whole workgroup (blockDim.x == kBlock).  All threads get the result.
template <typename T>
__device__ inline T block_sum(T v, T* red /* >= kWaves */) {
  v = wave_sum(v);
  const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
  __syncthreads();
  if (lane == 0) red[w] = v;
  __syncthreads();
  T s = 0;
#pragma unroll
  for (int i = 0; i < kWaves; ++i) s += red[i];
  return s;
}

// Arg-max of |value| with the lowest index winning ties (BLAS i_amax rule).
struct ArgMax {
  double v;
  int i;
};

__device__ inline ArgMax better_max(ArgMax a, ArgMax b) {
  if (b.v > a.v || (b.v == a.v && b.i >= 0 && (a.i < 0 || b.i < a.i))) return b;
  return a;
}

__device__ inline ArgMax wave_argmax(ArgMax a) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    ArgMax b;
    b.v = __shfl_xor(a.v, o, kWave);
    b.i = __shfl_xor(a.i, o, kWave);
    a = better_max(a, b);
  }
  return a;
}

__device__ inline ArgMax block_argmax(ArgMax a, double* redv, int* redi) {
  a = wave_argmax(a);
  const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
  __syncthreads();
  if (lane == 0) {
    redv[w] = a.v;
    redi[w] = a.i;
  }
  __syncthreads();
  ArgMax r{-1.0, -1};
#pragma unroll
  for (int i = 0; i < kWaves; ++i) r = better_max(r, ArgMax{redv[i], redi[i]});
  return r;
}

// Arg-min with the lowest index winning ties (strict '<' scan order).
__device__ inline ArgMax better_min(ArgMax a, ArgMax b) {
  if (b.i < 0) return a;
  if (a.i < 0) return b;
  if (b.v < a.v || (b.v == a.v && b.i < a.i)) return b;
  return a;
}

__device__ inline ArgMax block_argmin(ArgMax a, double* redv, int* redi) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    ArgMax b;
    b.v = __shfl_xor(a.v, o, kWave);
    b.i = __shfl_xor(a.i, o, kWave);
    a = better_min(a, b);
  }
  const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
  __syncthreads();
  if (lane == 0) {
    redv[w] = a.v;
    redi[w] = a.i;
  }
  __syncthreads();
  ArgMax r{0.0, -1};
#pragma unroll
  for (int i = 0; i < kWaves; ++i) r = better_min(r, ArgMax{redv[i], redi[i]});
  return r;
}

}  // namespace hmx
