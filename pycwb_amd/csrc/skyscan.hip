// Coherent sky-scan likelihood on MI355X (gfx950): DPF regulator pass, sky scan
// and argmax, as a plain (one cluster per launch) family and as a batched
// persistent work-queue family captured in a hipGraph.
//
// Data layout in HBM
//   sky tables   FP, FX : (S, n) float32 row-major (setup FP_t / FX_t)
//                ml     : (n, S) int32 delay taps
//   cluster      rms    : (P, n) float32
//                td     : (P, n, 2K+1) float2 {phase 00, phase 90}; one lane reads
//                         the 8-byte pair at its own delay tap, and the taps that
//                         neighbouring sky directions need sit in the same lines.
//   batch        all clusters' rms / td concatenated, addressed by 64-bit offsets.
//
// One lane = one sky direction; the pixel sweep is sequential per lane, so all
// per-direction sums have the reference's forward pixel order.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include <mutex>

#include "skyscan_lane.hh"
#include "pycwb_amd.h"

using namespace pycwb;

namespace pycwb {
static thread_local char g_last_error[256] = "";
void record_error(const char* what, const char* why) {
  snprintf(g_last_error, sizeof(g_last_error), "%s: %s", what, why);
}
const char* last_error() { return g_last_error; }
}  // namespace pycwb

namespace {

constexpr int kMaxIfo = 8;

int fail_hip(hipError_t e, const char* what) {
  pycwb::record_error(what, hipGetErrorString(e));
  return PYCWB_ERR_HIP;
}

#define HIP_TRY(expr)                                  \
  do {                                                 \
    hipError_t _e = (expr);                            \
    if (_e != hipSuccess) return fail_hip(_e, #expr);  \
  } while (0)

// Clears an error left by earlier, unrelated runtime calls of the process so
// that the launch checks below only report this call's own failures.
inline void clear_stale_error() { (void)hipGetLastError(); }

__device__ __forceinline__ unsigned long long wave_max_u64(unsigned long long v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    unsigned long long o = __shfl_xor(v, off, 64);
    v = o > v ? o : v;
  }
  return v;
}

__device__ __forceinline__ unsigned wave_sum_u32(unsigned v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Block-wide max of a key, one atomicMax per block.
template <int BLOCK>
__device__ __forceinline__ void block_argmax(unsigned long long key, unsigned long long* dst) {
  __shared__ unsigned long long s_key[BLOCK / 64];
  key = wave_max_u64(key);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) s_key[wave] = key;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long k = s_key[0];
#pragma unroll
    for (int w = 1; w < BLOCK / 64; ++w) k = s_key[w] > k ? s_key[w] : k;
    if (k) atomicMax(dst, k);
  }
  __syncthreads();
}

template <int BLOCK>
__device__ __forceinline__ void block_count(unsigned v, unsigned* dst) {
  __shared__ unsigned s_cnt[BLOCK / 64];
  v = wave_sum_u32(v);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) s_cnt[wave] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned t = 0;
#pragma unroll
    for (int w = 0; w < BLOCK / 64; ++w) t += s_cnt[w];
    if (t) atomicAdd(dst, t);
  }
  __syncthreads();
}

// ---------------------------------------------------------------------------
// Per-tile bodies shared by the plain and batched kernels.
// ---------------------------------------------------------------------------

// Noise-weight statistics of a cluster for the regulator bounds, float64:
// [0] min_j R2_j, [1] max_j R2_j with R2_j = sum_i r_ji^2, [2 + i] min_j
// r_ji^2 / R2_j and [10 + i] max_j r_ji^2 / R2_j (i < n).  kWStats doubles per
// cluster.
constexpr int kWStats = 24;
constexpr int kWMin = 2, kWMax = 10;

template <int BLK>
__device__ void weight_stats_block(const float* __restrict__ rms, int P, int n,
                                   double* __restrict__ out) {
  __shared__ double s_red[BLK / 64][2 + 16];
  double lo = __builtin_inf(), hi = 0.0, wmin[8], wmax[8];
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    wmin[i] = __builtin_inf();
    wmax[i] = 0.0;
  }
  for (int j = threadIdx.x; j < P; j += BLK) {
    double r2[8], R2 = 0.0;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      if (i < n) {
        const double r = (double)rms[(int64_t)j * n + i];
        r2[i] = r * r;
        R2 += r2[i];
      }
    }
    lo = R2 < lo ? R2 : lo;
    hi = R2 > hi ? R2 : hi;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      if (i < n) {
        const double w = r2[i] / R2;
        wmin[i] = (w < wmin[i] || w != w) ? w : wmin[i];  // NaN sticks
        wmax[i] = (w > wmax[i] || w != w) ? w : wmax[i];
      }
    }
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const double a = __shfl_xor(lo, off, 64), b = __shfl_xor(hi, off, 64);
    lo = (a < lo || a != a) ? a : lo;
    hi = (b > hi || b != b) ? b : hi;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      const double c = __shfl_xor(wmin[i], off, 64);
      wmin[i] = (c < wmin[i] || c != c) ? c : wmin[i];
      const double d = __shfl_xor(wmax[i], off, 64);
      wmax[i] = (d > wmax[i] || d != d) ? d : wmax[i];
    }
  }
  if (lane == 0) {
    s_red[wave][0] = lo;
    s_red[wave][1] = hi;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      s_red[wave][kWMin + i] = wmin[i];
      s_red[wave][kWMax + i] = wmax[i];
    }
  }
  __syncthreads();
  if (threadIdx.x < 2 + 16) {
    const int f = threadIdx.x;
    const bool take_max = f == 1 || f >= kWMax;
    double v = s_red[0][f];
    for (int w = 1; w < BLK / 64; ++w) {
      const double x = s_red[w][f];
      if (take_max) v = (x > v || x != x) ? x : v;
      else v = (x < v || x != x) ? x : v;
    }
    out[f] = v;
  }
  for (int f = 2 + 16 + (int)threadIdx.x; f < kWStats; f += BLK) out[f] = 0.0;
  __syncthreads();
}

// Bounds on one direction's regulator terms t_j = fx_j / (ni_j + 1e-9).
// With M_j = sum_i r_ji^2 v_i v_i^T (v_i = (F+_i, Fx_i)) = R2_j M(w_j), w_j the
// pixel's normalised weights (w_ji = r_ji^2 / R2_j), exact arithmetic gives
// fx_j = lambda_min(M_j), fp_j = lambda_max(M_j) and ni_j = sum_i alpha_i^2 with
// alpha_i = w_ji (u_j.v_i)^2 / lambda_max(M(w_j)) (u_j the dominant eigenvector,
// sum_i alpha_i = 1), so 1/n <= ni_j <= 1; the float32 pipeline keeps fx within
// 1e-4 R2max max|v|^2 of lambda_min (its rounding is ~1e-6 of fp <= R2 max|v|^2).
// Every w_j lies in the polytope W = {wmin <= w <= wmax, sum w = 1}; for up to
// four detectors its vertices are enumerated (n - 1 coordinates at a bound, the
// last from the sum), for more the larger simplex {w >= wmin, sum w = 1} and
// its n vertices stand in for it.  Over the vertices w_v:
//   lower: lambda_min is concave in w, so lambda_min(M(w_j)) >= C = min_v
//     lambda_min(M(w_v)).  ni_j <= N_hi: the point (x, y) = (g11 - g22, 2 g12)
//     of M(w) (its dominant direction at angle 2 theta) is linear in w, so for
//     w in W it lies in the convex hull of the vertices' points; when that hull
//     sits in an open half-plane (every vertex projects positively on the
//     reference direction, the point of the box centre), 2 theta_j lies between
//     the hull's extreme rays e_lo, e_hi, and (u_j.v_i)^2 = |v_i|^2 (1 +
//     cos(2 theta_j - 2 phi_i)) / 2 with cos at most 1 if 2 phi_i lies between
//     them, else max(e_lo.q_i, e_hi.q_i) (q_i the unit vector at 2 phi_i),
//     widened by delta = 1e-5 (tr_max / rho_min + 1) for the float32 rotation's
//     angle; lambda_max >= (tr_min + rho_min) / 2 (rho_min the smallest
//     projection); hence alpha_i <= cap_i = wmax_i |v_i|^2 (1 + cos_max) / 2 /
//     lambda_max_min, and sum alpha_i^2 is at most the greedy fill of the
//     largest caps up to a total of 1.  t_lb = (R2min C - slack) / (N_hi (1 +
//     1e-3) + 1e-9), or / (1 + 2e-5) without N_hi (ni <= 1).
//   upper: lambda_min(M(w)) <= u^T M(w) u for any unit u, linear in w, so
//     lambda_min(M_j) <= R2max min(B1, B2) with B1 = min over 32 sampled u of
//     max_i (u.v_i)^2 (any weights) and B2 = max_v u^T M(w_v) u for u the minor
//     eigenvector of the box centre's Gram matrix; with ni >= 1/n, t_ub = n (1 +
//     2e-5) (R2max min(B1, B2) + slack).  Used only when fp is bounded away from
//     0 (fp^2 >= 1e-3), which also makes every pixel count in NN.
struct RegBounds {
  double t_lb;  // >= 0
  double t_ub;  // < 0: no upper bound
};

// cos and sin of q pi / 32, q < 32: the sampled unit vectors of the upper bound
__constant__ double kUnitCS[32][2] = {
    {1.0, 0.0},
    {0.9951847266721969, 0.0980171403295606},
    {0.9807852804032304, 0.19509032201612825},
    {0.9569403357322088, 0.29028467725446233},
    {0.9238795325112867, 0.3826834323650898},
    {0.881921264348355, 0.47139673682599764},
    {0.8314696123025452, 0.5555702330196022},
    {0.773010453362737, 0.6343932841636455},
    {0.7071067811865476, 0.7071067811865475},
    {0.6343932841636455, 0.773010453362737},
    {0.5555702330196023, 0.8314696123025452},
    {0.4713967368259978, 0.8819212643483549},
    {0.38268343236508984, 0.9238795325112867},
    {0.29028467725446233, 0.9569403357322089},
    {0.19509032201612833, 0.9807852804032304},
    {0.09801714032956077, 0.9951847266721968},
    {6.123233995736766e-17, 1.0},
    {-0.09801714032956065, 0.9951847266721969},
    {-0.1950903220161282, 0.9807852804032304},
    {-0.29028467725446216, 0.9569403357322089},
    {-0.3826834323650897, 0.9238795325112867},
    {-0.4713967368259977, 0.881921264348355},
    {-0.555570233019602, 0.8314696123025455},
    {-0.6343932841636454, 0.7730104533627371},
    {-0.7071067811865475, 0.7071067811865476},
    {-0.773010453362737, 0.6343932841636455},
    {-0.8314696123025453, 0.5555702330196022},
    {-0.8819212643483549, 0.47139673682599786},
    {-0.9238795325112867, 0.3826834323650899},
    {-0.9569403357322088, 0.2902846772544624},
    {-0.9807852804032304, 0.1950903220161286},
    {-0.9951847266721968, 0.09801714032956083},
};

// ---------------------------------------------------------------------------
// Pixel classes of a cluster for the regulator sweep.  The pixels are split
// by their normalised noise weights w_ji = r_ji^2 / R2_j into 2^L classes (L =
// 0 below 64 pixels, 2 below 256, 3 below 1,024, else 4; 0 with NaN or
// infinite statistics or 65,536 pixels or more): L rounds in which every
// class splits at the mean of its widest weight coordinate (fixed-point sums,
// so the classes do not depend on the order of the atomics).  Per class, the
// weight statistics in the layout of the cluster's (R2 range, per-detector
// weight minima and maxima, each rounded outwards to float32), the pixel
// count in slot 18 and, in class 0's slot 19, the number of classes; and a
// table rem[q][g] (q <= ceil(P / 16), 16 classes per row) of the pixels j >=
// 16 q in class g.  The sweep bounds the terms of its remaining pixels class
// by class: a class's weight box is much narrower than the cluster's.
// ---------------------------------------------------------------------------
constexpr int kMaxClasses = 16;
constexpr int kClassStats = kMaxClasses * kWStats;  // doubles per cluster
constexpr int kClassSlotCount = 18, kClassSlotClasses = 19;

#ifndef PYCWB_CLASS_P0
#define PYCWB_CLASS_P0 128
#endif
#ifndef PYCWB_CLASS_MAX_LEVEL
#define PYCWB_CLASS_MAX_LEVEL 4
#endif
#ifndef PYCWB_CLASS_MAX_LEVEL_WIDE
#define PYCWB_CLASS_MAX_LEVEL_WIDE 4  // four or more detectors
#endif
__host__ __device__ inline int class_levels(int P, int n) {
  const int p0 = PYCWB_CLASS_P0;
  const int L = P < p0 ? 0 : P < 4 * p0 ? 2 : P < 16 * p0 ? 3 : P < 65536 ? 4 : 0;
  const int cap = n >= 4 ? PYCWB_CLASS_MAX_LEVEL_WIDE : PYCWB_CLASS_MAX_LEVEL;
  return L < cap ? L : cap;
}

__host__ __device__ inline int64_t class_rem_entries(int P) {
  return ((int64_t)(P + 15) / 16 + 1) * kMaxClasses;
}

__device__ __forceinline__ double pixel_weight(const float* __restrict__ rms, int64_t j, int n,
                                               int i) {
  double R2 = 0.0, ri = 0.0;
  for (int k = 0; k < n; ++k) {
    const double r = (double)rms[j * n + k];
    R2 += r * r;
    if (k == i) ri = r * r;
  }
  return ri / R2;
}

template <int BLK>
__device__ void pixel_classes_block(const float* __restrict__ rms, int P, int n,
                                    const double* __restrict__ ws, uint8_t* __restrict__ pclass,
                                    double* __restrict__ cstats, uint16_t* __restrict__ crem) {
  __shared__ unsigned s_min[kMaxClasses][8], s_max[kMaxClasses][8];
  __shared__ unsigned s_r2min[kMaxClasses], s_r2max[kMaxClasses], s_cnt[kMaxClasses];
  __shared__ unsigned long long s_sum[kMaxClasses];
  __shared__ int s_axis[kMaxClasses];
  __shared__ double s_split[kMaxClasses];
  __shared__ unsigned s_tot[BLK][kMaxClasses];
  bool finite = __builtin_isfinite(ws[0]) && __builtin_isfinite(ws[1]);
  for (int i = 0; i < n; ++i)
    finite = finite && __builtin_isfinite(ws[kWMin + i]) && __builtin_isfinite(ws[kWMax + i]);
  const int L = finite ? class_levels(P, n) : 0;
  const int G = 1 << L;
  const int t = threadIdx.x;
  for (int j = t; j < P; j += BLK) pclass[j] = 0;
  __syncthreads();
  for (int level = 0; level < L; ++level) {
    const int cells = 1 << level;
    for (int e = t; e < cells * 8; e += BLK) {
      s_min[e >> 3][e & 7] = 0x7f800000u;  // +inf
      s_max[e >> 3][e & 7] = 0u;
    }
    __syncthreads();
    for (int j = t; j < P; j += BLK) {
      const int c = pclass[j];
      for (int i = 0; i < n; ++i) {
        const unsigned w = __float_as_uint((float)pixel_weight(rms, j, n, i));
        atomicMin(&s_min[c][i], w);
        atomicMax(&s_max[c][i], w);
      }
    }
    __syncthreads();
    if (t < cells) {
      int ax = 0;
      float wid = -1.f;
      for (int i = 0; i < n; ++i) {
        const float d = __uint_as_float(s_max[t][i]) - __uint_as_float(s_min[t][i]);
        if (d > wid) {
          wid = d;
          ax = i;
        }
      }
      s_axis[t] = ax;
      s_sum[t] = 0ull;
      s_cnt[t] = 0u;
    }
    __syncthreads();
    for (int j = t; j < P; j += BLK) {
      const int c = pclass[j];
      const double w = pixel_weight(rms, j, n, s_axis[c]);
      atomicAdd(&s_sum[c], (unsigned long long)(w * 4294967296.0));
      atomicAdd(&s_cnt[c], 1u);
    }
    __syncthreads();
    if (t < cells)
      s_split[t] = s_cnt[t] ? (double)s_sum[t] / 4294967296.0 / (double)s_cnt[t] : 0.0;
    __syncthreads();
    for (int j = t; j < P; j += BLK) {
      const int c = pclass[j];
      const double w = pixel_weight(rms, j, n, s_axis[c]);
      pclass[j] = (uint8_t)(2 * c + (w > s_split[c] ? 1 : 0));
    }
    __syncthreads();
  }
  // class statistics, rounded outwards
  for (int e = t; e < kMaxClasses * 8; e += BLK) {
    s_min[e >> 3][e & 7] = 0x7f800000u;
    s_max[e >> 3][e & 7] = 0u;
  }
  if (t < kMaxClasses) {
    s_r2min[t] = 0x7f800000u;
    s_r2max[t] = 0u;
    s_cnt[t] = 0u;
  }
  __syncthreads();
  for (int j = t; j < P; j += BLK) {
    const int c = pclass[j];
    double r2[8], R2 = 0.0;
    for (int i = 0; i < n; ++i) {
      const double r = (double)rms[(int64_t)j * n + i];
      r2[i] = r * r;
      R2 += r2[i];
    }
    atomicMin(&s_r2min[c], __float_as_uint(__double2float_rd(R2)));
    atomicMax(&s_r2max[c], __float_as_uint(__double2float_ru(R2)));
    for (int i = 0; i < n; ++i) {
      const double w = r2[i] / R2;
      atomicMin(&s_min[c][i], __float_as_uint(__double2float_rd(w)));
      atomicMax(&s_max[c][i], __float_as_uint(__double2float_ru(w)));
    }
    atomicAdd(&s_cnt[c], 1u);
  }
  __syncthreads();
  for (int e = t; e < kClassStats; e += BLK) {
    const int g = e / kWStats, f = e % kWStats;
    double v = 0.0;
    if (g < G && s_cnt[g]) {
      if (f == 0) v = __uint_as_float(s_r2min[g]);
      else if (f == 1) v = __uint_as_float(s_r2max[g]);
      else if (f >= kWMin && f < kWMin + n) v = __uint_as_float(s_min[g][f - kWMin]);
      else if (f >= kWMax && f < kWMax + n) v = __uint_as_float(s_max[g][f - kWMax]);
    }
    if (f == kClassSlotCount) v = g < G ? (double)s_cnt[g] : 0.0;
    if (f == kClassSlotClasses && g == 0) v = (double)G;
    cstats[e] = v;
  }
  // rem[q][g]: thread t owns the checkpoint rows [q0, q1)
  const int nq = (P + 15) / 16;
  const int per = (nq + BLK - 1) / BLK;
  const int q0 = t * per < nq ? t * per : nq, q1 = (t + 1) * per < nq ? (t + 1) * per : nq;
  for (int g = 0; g < kMaxClasses; ++g) s_tot[t][g] = 0u;
  for (int j = q0 * 16; j < q1 * 16 && j < P; ++j) s_tot[t][pclass[j]] += 1u;
  __syncthreads();
  if (t < kMaxClasses) {  // suffix totals over the threads after each
    unsigned run = 0u;
    for (int u = BLK - 1; u >= 0; --u) {
      const unsigned x = s_tot[u][t];
      s_tot[u][t] = run;
      run += x;
    }
  }
  __syncthreads();
  if (G > 1) {
    if (t == 0)
      for (int g = 0; g < kMaxClasses; ++g) crem[(int64_t)nq * kMaxClasses + g] = 0;
    for (int q = q1 - 1; q >= q0; --q) {
      for (int j = q * 16 + 15; j >= q * 16; --j)
        if (j < P) s_tot[t][pclass[j]] += 1u;
      for (int g = 0; g < kMaxClasses; ++g)
        crem[(int64_t)q * kMaxClasses + g] = (uint16_t)s_tot[t][g];
    }
  }
  __syncthreads();
}

__device__ __forceinline__ float half_bits_value(unsigned short b) {
  return (float)__builtin_bit_cast(_Float16, b);
}

// ---- term bounds in float32 --------------------------------------------
// The bounds described above (RegBounds), for a cluster's statistics or a
// pixel class's, computed in float32 with the hardware's reciprocal and
// square root (about 1 ulp): every margin of the bounds (1e-4 R2max max|v|^2
// on lambda_min, the angle widening 1e-5 (tr / rho + 1), 1e-5 on the caps,
// 1e-3 on ni, 2e-5 on the upper bound) is at least a hundred times the
// float32 error of the quantity it covers, so the bounds stay valid.  The
// extreme rays of the cone are found by cross-product comparisons instead of
// angles, and the result is kept as a pair of half-precision values rounded
// outwards (lower bound down, upper bound up).
__device__ __forceinline__ float fast_rcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float fast_sqrt(float x) { return __builtin_amdgcn_sqrtf(x); }

__device__ __forceinline__ unsigned half_down_f(float f) {  // largest half <= f, f >= 0
  _Float16 h = (_Float16)f;
  unsigned b = __builtin_bit_cast(unsigned short, h);
  if ((float)h > f) b -= 1u;
  return b;
}

__device__ __forceinline__ unsigned half_up_f(float f) {  // smallest half >= f, f >= 0
  _Float16 h = (_Float16)f;
  unsigned b = __builtin_bit_cast(unsigned short, h);
  if ((float)h < f) b += 1u;
  return b;
}

// Per-direction quantities every class's bounds share.
template <int N>
struct DirPre {
  float A2[N], B2[N], AB[N], v2[N], qx[N], qy[N];
  float v2max, B1;
};

template <int N>
__device__ __forceinline__ void dir_pre(const float (&fp0)[N], const float (&fx0)[N],
                                        DirPre<N>& d) {
  d.v2max = 0.f;
#pragma unroll
  for (int i = 0; i < N; ++i) {
    const float a = fp0[i], b = fx0[i];
    d.A2[i] = a * a;
    d.B2[i] = b * b;
    d.AB[i] = a * b;
    d.v2[i] = d.A2[i] + d.B2[i];
    d.v2max = d.v2[i] > d.v2max ? d.v2[i] : d.v2max;
    const float r = d.v2[i] > 0.f ? fast_rcp(d.v2[i]) : 0.f;
    d.qx[i] = (d.A2[i] - d.B2[i]) * r;
    d.qy[i] = 2.f * d.AB[i] * r;
  }
  float B = __builtin_inff();
#pragma unroll 1
  for (int q = 0; q < 32; ++q) {
    const float cu = (float)kUnitCS[q][0], su = (float)kUnitCS[q][1];
    float mx = 0.f;
#pragma unroll
    for (int i = 0; i < N; ++i) {
      const float p = cu * fp0[i] + su * fx0[i];
      mx = p * p > mx ? p * p : mx;
    }
    B = mx < B ? mx : B;
  }
  d.B1 = B;
}

// The (lower, upper) term bounds of one direction over one class (weight
// statistics cg, the layout of weight_stats) as a half-precision pair: lower
// rounded down in the low 16 bits, upper rounded up (or +inf: none) above.
template <int N>
__device__ __forceinline__ unsigned class_bound_pair(const DirPre<N>& d,
                                                     const double* __restrict__ cg) {
  // statistics to float32 rounded outwards (exact for a class's, which are floats)
  const float R2min = __double2float_rd(cg[0]), R2max = __double2float_ru(cg[1]);
  float lo[N], hi[N];
  float sw = 0.f, fpw = 0.f, cs = 0.f;
#pragma unroll
  for (int i = 0; i < N; ++i) {
    lo[i] = __double2float_rd(cg[kWMin + i]);
    hi[i] = __double2float_ru(cg[kWMax + i]);
    sw += lo[i];
    fpw += lo[i] * d.v2[i];
    cs += lo[i] + hi[i];
  }
  const float slack = 1e-4f * R2max * d.v2max;
  const float rcs = fast_rcp(cs);
  float m11 = 0.f, m12 = 0.f, m22 = 0.f;
#pragma unroll
  for (int i = 0; i < N; ++i) {
    const float w = (lo[i] + hi[i]) * rcs;
    m11 += w * d.A2[i];
    m12 += w * d.AB[i];
    m22 += w * d.B2[i];
  }
  float dx = m11 - m22, dy = 2.f * m12;
  const float dn2 = dx * dx + dy * dy;
  if (dn2 > 0.f) {
    const float r = fast_rcp(fast_sqrt(dn2));
    dx *= r;
    dy *= r;
  } else {
    dx = 1.f;
    dy = 0.f;
  }
  float u1 = 1.f, u2 = 0.f;
  {
    const float h = 0.5f * (m11 + m22), disc = h * h - (m11 * m22 - m12 * m12);
    const float lmin = h - fast_sqrt(disc > 0.f ? disc : 0.f);
    const float p1 = m12, p2 = lmin - m11, q1 = lmin - m22, q2 = m12;
    const float np = p1 * p1 + p2 * p2, nq = q1 * q1 + q2 * q2;
    if (np >= nq && np > 0.f) {
      const float r = fast_rcp(fast_sqrt(np));
      u1 = p1 * r;
      u2 = p2 * r;
    } else if (nq > 0.f) {
      const float r = fast_rcp(fast_sqrt(nq));
      u1 = q1 * r;
      u2 = q2 * r;
    }
  }
  const int nbox = N <= 4 ? N << (N - 1) : 0;
  float C = __builtin_inff(), Bq = 0.f, rho = __builtin_inff();
  float trmin = __builtin_inff(), trmax = 0.f;
  float xlo = 0.f, ylo = 0.f, xhi = 0.f, yhi = 0.f;
  int count = 0;
#pragma unroll 1
  for (int v = 0; v < nbox + N; ++v) {
    if (v == nbox && count > 0) break;
    const bool box = v < nbox;
    const int k = box ? v >> (N - 1) : v - nbox;
    const int bits = v & ((1 << (N - 1)) - 1);
    float rest = 0.f, s11 = 0.f, s12 = 0.f, s22 = 0.f, k11 = 0.f, k12 = 0.f, k22 = 0.f;
    float lok = 0.f, hik = 0.f;
#pragma unroll
    for (int i = 0; i < N; ++i) {
      if (i == k) {
        k11 = d.A2[i];
        k12 = d.AB[i];
        k22 = d.B2[i];
        lok = lo[i];
        hik = hi[i];
      } else {
        const int bit = i < k ? i : i - 1;
        const float w = box && ((bits >> bit) & 1) ? hi[i] : lo[i];
        rest += w;
        s11 += w * d.A2[i];
        s12 += w * d.AB[i];
        s22 += w * d.B2[i];
      }
    }
    const float wk = 1.f - rest;
    if (box && !(wk >= lok - 1e-6f && wk <= hik + 1e-6f)) continue;
    const float g11 = s11 + wk * k11, g12 = s12 + wk * k12, g22 = s22 + wk * k22;
    const float h = 0.5f * (g11 + g22), disc = h * h - (g11 * g22 - g12 * g12);
    const float lm = h - fast_sqrt(disc > 0.f ? disc : 0.f);
    C = lm < C ? lm : C;
    const float q = u1 * u1 * g11 + 2.f * u1 * u2 * g12 + u2 * u2 * g22;
    Bq = q > Bq ? q : Bq;
    const float x = g11 - g22, y = 2.f * g12, tr = g11 + g22;
    trmin = tr < trmin ? tr : trmin;
    trmax = tr > trmax ? tr : trmax;
    const float pr = dx * x + dy * y;
    rho = pr < rho ? pr : rho;
    if (count == 0) {
      xlo = xhi = x;
      ylo = yhi = y;
    } else {
      if (xlo * y - ylo * x < 0.f) {  // clockwise of the current lower ray
        xlo = x;
        ylo = y;
      }
      if (xhi * y - yhi * x > 0.f) {  // counter-clockwise of the upper ray
        xhi = x;
        yhi = y;
      }
    }
    ++count;
  }
  float ni_hi = 1.f;
  const bool cone = count > 0 && rho > 1e-3f * trmax;
  const float delta = cone ? 1e-5f * (trmax * fast_rcp(rho) + 1.f) : 1.f;
  if (cone && delta < 0.1f) {
    const float rl = fast_rcp(fast_sqrt(xlo * xlo + ylo * ylo));
    const float rh = fast_rcp(fast_sqrt(xhi * xhi + yhi * yhi));
    const float elx = xlo * rl, ely = ylo * rl, ehx = xhi * rh, ehy = yhi * rh;
    const float rlam = fast_rcp(0.5f * (trmin + rho));
    float cap[N];
#pragma unroll
    for (int i = 0; i < N; ++i) {
      float mc = 1.f;
      if (d.v2[i] > 0.f) {
        const float qx = d.qx[i], qy = d.qy[i];
        const bool inside = elx * qy - ely * qx >= 0.f && qx * ehy - qy * ehx >= 0.f;
        if (!inside) {
          const float c1 = elx * qx + ely * qy, c2 = ehx * qx + ehy * qy;
          mc = (c1 > c2 ? c1 : c2) + delta;
          mc = mc < 1.f ? mc : 1.f;
        }
      }
      cap[i] = hi[i] * d.v2[i] * 0.5f * (1.f + mc) * rlam * (1.f + 1e-5f);
    }
    float tot = 0.f, all = 0.f;
#pragma unroll
    for (int i = 0; i < N; ++i) {
      float ahead = 0.f;
#pragma unroll
      for (int k = 0; k < N; ++k)
        if (k != i && (cap[k] > cap[i] || (cap[k] == cap[i] && k < i))) ahead += cap[k];
      float take = 1.f - ahead;
      take = take < cap[i] ? take : cap[i];
      take = take > 0.f ? take : 0.f;
      tot += take * take;
      all += cap[i];
    }
    if (tot == tot && all >= 1.f) ni_hi = tot < 1.f ? tot : 1.f;
  }
  float den = ni_hi * (1.f + 1e-3f) + 1e-9f;
  den = den < 1.f + 2e-5f ? den : 1.f + 2e-5f;
  float t_lb = (R2min * C - slack) * fast_rcp(den);
  t_lb = t_lb > 0.f ? t_lb : 0.f;  // NaN -> 0
  unsigned hi16 = 0x7c00u;          // +inf: no upper bound
  const float fp_lb = 0.5f * R2min * fpw * (1.f - 1e-5f);
  if (fp_lb * fp_lb >= 1e-3f) {
    const float B = d.B1 < Bq ? d.B1 : Bq;
    const float t_ub = (float)N * (1.f + 2e-5f) * (R2max * B + slack);
    if (t_ub == t_ub) hi16 = half_up_f(t_ub);
  }
  return half_down_f(t_lb) | (hi16 << 16);
}

// DPF quality of the direction s; returns 1 if quality > gamma.
// The reference sums NI = sum_j fx_j / (ni_j + 1e-9) in float32 over j = 0..P-1
// and counts sqrt(NI / (NN + 0.01)) > gamma (dpf.py:25-34, 444-453).  Every
// term is >= 0, so with the partial sum NI_k after k pixels and per-term bounds
// t_lb <= t_j <= t_ub the final sum lies in
// [(NI_k + (P-k) t_lb)(1 - P u'), (NI_k + (P-k) t_ub)(1 + P u')], u' = 1.2e-7
// covering float32 rounding of the remaining additions.  The sweep stops at a
// checkpoint (every 16 pixels, and before the first) as soon as that interval
// lies entirely above gamma^2 (P + 0.01) (NN <= P, so the count is 1) or,
// with an upper bound (which guarantees NN = P), entirely below (count 0);
// the margins 1e-9 cover the float64 division and square root.  Without
// bounds (t_lb = 0, no t_ub) this is the early exit on the partial sum alone.
template <int N>
__device__ __forceinline__ unsigned regulator_lane(const float* __restrict__ FP,
                                                   const float* __restrict__ FX,
                                                   const float* __restrict__ rms, int P,
                                                   int64_t s, double gamma, bool early,
                                                   const double* __restrict__ ws) {
  float fp0[N], fx0[N];
#pragma unroll
  for (int i = 0; i < N; ++i) {
    fp0[i] = FP[s * N + i];
    fx0[i] = FX[s * N + i];
  }
  RegBounds bd;
  bd.t_lb = 0.0;
  bd.t_ub = -1.0;
  if (ws) {  // the float32 bounds of the batched regulator's first pass
    DirPre<N> dp;
    dir_pre<N>(fp0, fx0, dp);
    const unsigned pr = class_bound_pair<N>(dp, ws);
    bd.t_lb = half_bits_value((unsigned short)(pr & 0xffffu));
    const float uh = half_bits_value((unsigned short)(pr >> 16));
    bd.t_ub = __builtin_isinf(uh) ? -1.0 : (double)uh;
  }
  const double den = (double)P + 0.01;
  const double thr_hi = gamma * gamma * den * (1.0 + 1e-9);
  const double thr_lo = gamma * gamma * den * (1.0 - 1e-9);
  const double sum_lo = 1.0 - (double)P * 1.2e-7, sum_hi = 1.0 + (double)P * 1.2e-7;
  float NI = 0.f;
  unsigned NN = 0;
  if (early) {
    const double rest = (double)P;
    if (rest * bd.t_lb * sum_lo > thr_hi) return 1u;
    if (bd.t_ub >= 0.0 && rest * bd.t_ub * sum_hi < thr_lo) return 0u;
  }
  for (int j = 0; j < P; ++j) {
    float r[N];
#pragma unroll
    for (int i = 0; i < N; ++i) r[i] = rms[(int64_t)j * N + i];
    regulator_pixel<N>(r, fp0, fx0, NI, NN);
    if (early && (j & 15) == 15) {
      const double rest = (double)(P - 1 - j);
      if (((double)NI + rest * bd.t_lb) * sum_lo > thr_hi) return 1u;
      if (bd.t_ub >= 0.0 && ((double)NI + rest * bd.t_ub) * sum_hi < thr_lo) return 0u;
    }
  }
  const double q = __builtin_sqrt((double)NI / ((double)NN + 0.01));
  return q > gamma ? 1u : 0u;
}

// Wave-wide min / max of v over all 64 lanes (every lane active): the result
// is the same on every lane, so it is read into a scalar register.
__device__ __forceinline__ int wave_min_int(int v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = min(v, __shfl_xor(v, off, 64));
  return __builtin_amdgcn_readfirstlane(v);
}

__device__ __forceinline__ int wave_max_int(int v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = max(v, __shfl_xor(v, off, 64));
  return __builtin_amdgcn_readfirstlane(v);
}

// Widest summed tap range (over detectors) for which the tile energy bound is
// evaluated; wider ranges scan every pixel.
constexpr int kMaxBoundSpan = 32;

// Tile energy bound of pixel j: U = sum_i max_{t in [lo_i, hi_i]} (v00^2 + v90^2)
// at taps t of detector i.  Every direction of the wavefront reads taps inside
// these ranges, so its pixel energy e = (sum v00^2) + (sum v90^2) + 1e-12 is at
// most U up to float32 rounding (relative 1e-6 at n <= 8, all terms >= 0) plus
// the 1e-12 addend.  Returns false when U * (1 + 1e-5) + 1e-11 <= En: then no
// direction of the wavefront has e > En, and with finite amplitudes the pixel
// contributes nothing (pixel_contributes).  A NaN or inf amplitude in range
// makes U non-finite and the pixel a candidate.
template <int N>
__device__ __forceinline__ bool pixel_candidate(const float2* __restrict__ row, int nd,
                                                const int (&lo)[N], const int (&hi)[N],
                                                double En) {
  float U = 0.f;
#pragma unroll
  for (int i = 0; i < N; ++i) {
    float mx = 0.f;
    for (int t = lo[i]; t <= hi[i]; ++t) {
      const float2 v = row[i * nd + t];
      const float E = v.x * v.x + v.y * v.y;
      mx = (E > mx || E != E) ? E : mx;  // NaN sticks
    }
    U += mx;
  }
  return !((double)U * (1.0 + 1e-5) + 1e-11 <= En);
}

// Wavefront-level ordering point for LDS data shared between lanes.
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

__device__ __forceinline__ unsigned long long lanemask_lt(int lane) {
  return lane ? (~0ull >> (64 - lane)) : 0ull;
}

// Compaction of the full-path evaluations of one wavefront.  Lane d owns
// direction d.  For each pixel, the lanes whose energy test passes enqueue a
// (pixel, owner) pair with the delayed amplitudes they loaded; once 64 pairs
// wait, every lane evaluates one of them (pixel_rest) and stores its
// contributions, and each owner adds its own contributions to its sums in
// queue order, which is its pixel order (a direction enqueues at most one pair
// per pixel, and pixels are enqueued in increasing order).  Pairs beyond the
// first 64 move to the front and wait for the next flush.  The sums therefore
// see exactly the terms of the per-lane sweep in the same order (bit-identical
// results) while the expensive chain runs with every lane busy.
constexpr int kQueue = 128;  // pair slots: at most 63 waiting + 64 of one pixel

template <int N>
struct WaveQueue {
  float2 qv[kQueue][N];           // (v00, v90) of each detector of each pair
  int qjd[kQueue];                // (pixel << 6) | owner lane
  unsigned long long own[64];     // per owner lane: its slots among those evaluated
  float pat[64][2 * N];           // F+ (0..N) and Fx (N..2N) of each lane's direction
  float c_et[64], c_ec[64], c_gn[64], c_rn[64], c_fpet[64], c_fxet[64];
  double c_llcc[64], c_ll[64];
  int c_flags[64];                // bit 0: m, bit 1: ec > 1e-9, bit 2: mk > 0
};

// Evaluate slots [0, cnt), add them to their owners' sums, drop them from the
// queue.  own[] is zero on entry and on exit.  Without MAPS only the sums the
// sky statistic needs are carried (the map-only terms are neither formed nor
// stored).
template <int N, bool MAPS>
__device__ __forceinline__ void queue_flush(WaveQueue<N>& q, int cnt, int& T, int lane,
                                            const float* __restrict__ rms, int P, float rr,
                                            float RR, double En, SkySums& acc) {
  if (lane < cnt) {
    const int jd = q.qjd[lane];
    int j = jd >> 6;
    j = j < 0 ? 0 : (j >= P ? P - 1 : j);
    const int d = jd & 63;
    float v00[N], v90[N], r[N], fp0[N], fx0[N];
#pragma unroll
    for (int i = 0; i < N; ++i) {
      const float2 v = q.qv[lane][i];
      v00[i] = v.x;
      v90[i] = v.y;
      fp0[i] = q.pat[d][i];
      fx0[i] = q.pat[d][N + i];
      r[i] = rms[(int64_t)j * N + i];
    }
    float e, et;
    int m;
    pixel_energy<N>(v00, v90, En, e, m, et);
    PixelEval<N> o;
    pixel_rest<N>(v00, v90, r, fp0, fx0, rr, RR, e, m, et, o);
    q.c_et[lane] = o.et;
    q.c_ec[lane] = o.ec;
    q.c_gn[lane] = o.gn;
    q.c_rn[lane] = o.rn;
    if constexpr (MAPS) {
      q.c_fpet[lane] = o.d.fp * o.et;
      q.c_fxet[lane] = o.d.fx * o.et;
    }
    q.c_llcc[lane] = o.ll * o.ccd;
    q.c_ll[lane] = o.ll;
    q.c_flags[lane] = MAPS ? (o.m | (o.ec > 1e-9f ? 2 : 0) | (o.mk > 0.f ? 4 : 0)) : o.m;
    __hip_atomic_fetch_or(&q.own[d], 1ull << lane, __ATOMIC_RELAXED,
                          __HIP_MEMORY_SCOPE_WORKGROUP);
  }
  wave_sync();
  // each owner walks its own slots in increasing order
  unsigned long long mine = q.own[lane];
  q.own[lane] = 0ull;
  while (mine) {
    const int slot = __ffsll((long long)mine) - 1;
    mine &= mine - 1ull;
    const int f = q.c_flags[slot];
    const float et = q.c_et[slot];
    acc.M += f & 1;
    acc.EE += et;
    acc.Lr += q.c_llcc[slot];
    acc.LL += q.c_ll[slot];
    acc.EC += q.c_ec[slot];
    acc.GN += q.c_gn[slot];
    acc.RN += q.c_rn[slot];
    if constexpr (!MAPS) continue;
    acc.Mp += (f & 2) ? 1.0 : 0.0;
    if (f & 4) {
      acc.sde += et;
      acc.pwe += q.c_fpet[slot];
      acc.cwe += q.c_fxet[slot];
    }
  }
  // move the waiting pairs to the front
  const bool mv = lane + cnt < T;
  float2 v[N];
  int jd = 0;
  if (mv) {
#pragma unroll
    for (int i = 0; i < N; ++i) v[i] = q.qv[lane + cnt][i];
    jd = q.qjd[lane + cnt];
  }
  wave_sync();
  if (mv) {
#pragma unroll
    for (int i = 0; i < N; ++i) q.qv[lane][i] = v[i];
    q.qjd[lane] = jd;
  }
  T -= cnt;
  wave_sync();
}

// Sky scan of direction s; returns the argmax key for position k.
// Pixels are taken in chunks of 64.  With tile_bound (which needs skip_masked)
// the wavefront first marks the chunk's candidate pixels with the tile energy
// bound and every lane then tests only those, in increasing order.  With
// skip_masked the pairs passing the energy test go through the wavefront's
// queue (queue_flush); otherwise every lane evaluates every pixel itself.
// COUNT_TESTED (a measurement pass) adds to *tested the (direction, pixel)
// pairs whose energy the wavefront evaluates: candidate pixels times its
// directions.
template <int N, int WAVES, bool COUNT_TESTED = false, bool MAPS = true>
__device__ __forceinline__ unsigned long long scan_lane(
    const float* __restrict__ FP, const float* __restrict__ FX, const int32_t* __restrict__ ml,
    int64_t S, const float* __restrict__ rms, const float2* __restrict__ td, int P, int K,
    float rr, float RR, double En, double netCC, int64_t s, int64_t k, bool active,
    bool skip_masked, bool tile_bound, float* __restrict__ maps,
    unsigned* __restrict__ tested = nullptr) {
  __shared__ WaveQueue<N> queues[WAVES];
  const int nd = 2 * K + 1;
  const int lane = (int)(threadIdx.x & 63);
  WaveQueue<N>& q = queues[threadIdx.x >> 6];
  float fp0[N], fx0[N];
  int tap[N], lo[N], hi[N];
  int span = 0;
#pragma unroll
  for (int i = 0; i < N; ++i) {
    fp0[i] = FP[s * N + i];
    fx0[i] = FX[s * N + i];
    int t = ml[(int64_t)i * S + s] + K;
    t = t < 0 ? 0 : (t >= nd ? nd - 1 : t);  // host validates |ml| <= K
    tap[i] = i * nd + t;
    lo[i] = wave_min_int(t);
    hi[i] = wave_max_int(t);
    span += hi[i] - lo[i] + 1;
    q.pat[lane][i] = fp0[i];
    q.pat[lane][N + i] = fx0[i];
  }
  q.own[lane] = 0ull;
  const bool use_bound = skip_masked && tile_bound && span <= kMaxBoundSpan;  // wave-uniform
  int T = 0;  // queued pairs (wave-uniform)
  wave_sync();
  SkySums acc;
  sums_init(acc);
  unsigned n_tested = 0;  // candidate pixels (COUNT_TESTED)
  const int64_t pstride = (int64_t)N * nd;
  // the cluster's amplitudes are the same for every lane of the wavefront:
  // held in scalar registers, each candidate's row address is scalar and the
  // lanes add only their 32-bit tap offsets
  const float2* tdw;
  {
    const uint64_t u = (uint64_t)td;
    const uint32_t lo32 = __builtin_amdgcn_readfirstlane((uint32_t)u);
    const uint32_t hi32 = __builtin_amdgcn_readfirstlane((uint32_t)(u >> 32));
    tdw = (const float2*)(((uint64_t)hi32 << 32) | lo32);
  }
  uint32_t toff[N];  // byte offset of each detector's tap within a pixel row
#pragma unroll
  for (int i = 0; i < N; ++i) toff[i] = (uint32_t)tap[i] * (uint32_t)sizeof(float2);
  // the empty asm keeps each offset a 32-bit register read at the load, so the
  // load takes the scalar row base plus the offset (no 64-bit copies of the
  // offsets held through the loop, no address arithmetic per candidate)
  auto tap_load = [&](const float2* row, int i) -> float2 {
    asm volatile("" : "+v"(toff[i]));
    return *(const float2*)((const char*)row + toff[i]);
  };
  if (!skip_masked) {
    // every lane evaluates every pixel itself, in increasing order
    if constexpr (COUNT_TESTED) n_tested = (unsigned)P;
    for (int j = 0; j < P; ++j) {
      float v00[N], v90[N], r[N];
#pragma unroll
      for (int i = 0; i < N; ++i) {
        const float2 v = td[(int64_t)j * pstride + tap[i]];
        v00[i] = v.x;
        v90[i] = v.y;
        r[i] = rms[(int64_t)j * N + i];
      }
      scan_pixel<N>(v00, v90, r, fp0, fx0, rr, RR, En, false, acc);
    }
  } else {
    for (int j0 = 0; j0 < P; j0 += 64) {
      const int left = P - j0;
      unsigned long long cand = left >= 64 ? ~0ull : ((1ull << left) - 1ull);
      if (use_bound) {
        const int j = j0 + lane;
        const bool c = j < P && pixel_candidate<N>(td + (int64_t)j * pstride, nd, lo, hi, En);
        cand = __ballot(c);
      }
      if constexpr (COUNT_TESTED) n_tested += (unsigned)__popcll(cand);
      if (!cand) continue;
      // software pipeline: the next candidate's amplitudes load while this one
      // is tested, into the other of two buffers (no register copies).  The
      // inner loops only test and enqueue; the flush runs between their passes,
      // at the same points of the pixel order (right after the enqueue that
      // fills 64 slots), so the sums keep no copies live through the loop.
      int b = __ffsll((long long)cand) - 1;
      cand &= cand - 1ull;
      float2 bufA[N], bufB[N];
      {
        const float2* row = tdw + (int64_t)(j0 + b) * pstride;
#pragma unroll
        for (int i = 0; i < N; ++i) bufA[i] = tap_load(row, i);
      }
      bool more = true;
      // tests pixel j0 + b from c, loads the next candidate into n; false when
      // the chunk is done or the queue holds 64 pairs
      auto step = [&](const float2 (&c)[N], float2 (&n)[N]) -> bool {
        const int j = j0 + b;
        more = cand != 0ull;
        const int bn = more ? __ffsll((long long)cand) - 1 : b;
        cand &= cand - 1ull;
        const float2* row = tdw + (int64_t)(j0 + bn) * pstride;
#pragma unroll
        for (int i = 0; i < N; ++i) n[i] = tap_load(row, i);
        float v00[N], v90[N];
#pragma unroll
        for (int i = 0; i < N; ++i) {
          v00[i] = c[i].x;
          v90[i] = c[i].y;
        }
        float e, et;
        int m;
        pixel_energy<N>(v00, v90, En, e, m, et);
        const bool contrib = pixel_contributes(e, m);
        const unsigned long long bm = __ballot(contrib);
        if (contrib) {
          const int slot = T + __popcll(bm & lanemask_lt(lane));
#pragma unroll
          for (int i = 0; i < N; ++i) q.qv[slot][i] = c[i];
          q.qjd[slot] = (j << 6) | lane;
        }
        T += __popcll(bm);
        if (!more) return false;
        b = bn;
        return T < 64;
      };
      bool curA = true;  // the next candidate's amplitudes are in bufA (wave-uniform)
      while (more) {
        if (curA) {
          for (;;) {
            if (!step(bufA, bufB)) { curA = false; break; }
            if (!step(bufB, bufA)) { curA = true; break; }
          }
        } else {
          for (;;) {
            if (!step(bufB, bufA)) { curA = true; break; }
            if (!step(bufA, bufB)) { curA = false; break; }
          }
        }
        if (T >= 64) {
          wave_sync();
          queue_flush<N, MAPS>(q, 64, T, lane, rms, P, rr, RR, En, acc);
        }
      }
    }
    if (T > 0) {
      wave_sync();
      queue_flush<N, MAPS>(q, T, T, lane, rms, P, rr, RR, En, acc);
    }
  }
  if constexpr (COUNT_TESTED) {
    const unsigned lanes = (unsigned)__popcll(__ballot(active));
    if (lane == 0 && tested) atomicAdd(tested, n_tested * lanes);
  }
  SkyStat st;
  sky_epilogue(acc, N, netCC, st);
  if (MAPS && maps && active) {
#pragma unroll
    for (int m = 0; m < kNumMaps; ++m) maps[(int64_t)m * S + s] = st.gated ? 0.f : st.maps[m];
  }
  if (!active || st.gated) return 0ull;
  return argmax_key(st.AA, (uint32_t)k);
}

// Regulator formula (dpf.py:31-34), float64.
__device__ __forceinline__ double regulator_value(int64_t n_valid, unsigned count, double En) {
  const double FF = (double)n_valid, ff = (double)count;
  return (FF * FF / (ff * ff + 1.e-9) - 1) * En;
}

// ---------------------------------------------------------------------------
// Plain family: one cluster per launch.
// ---------------------------------------------------------------------------

// Re-lay td00/td90 from the reference's (2K+1, n, P) planes into (P, n, 2K+1)
// float2 pairs, and clear the per-call counters.
__global__ void k_relayout_td(const float* __restrict__ td00, const float* __restrict__ td90,
                              int n, int P, int nd, float2* __restrict__ out,
                              unsigned* __restrict__ counters, int n_counters) {
  const int64_t total = (int64_t)n * P * nd;
  const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (gid < n_counters) counters[gid] = 0u;
  // gid enumerates the source (tap, ifo, pixel) so the loads are coalesced
  for (int64_t e = gid; e < total; e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t j = e % P;
    const int64_t rest = e / P;
    const int64_t i = rest % n;
    const int64_t t = rest / n;
    out[(j * n + i) * nd + t] = make_float2(td00[e], td90[e]);
  }
}

__global__ void k_fill(float* __restrict__ p, int64_t count, float v) {
  const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  for (int64_t e = gid; e < count; e += (int64_t)gridDim.x * blockDim.x) p[e] = v;
}

template <int N, int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_regulator_plain(
    const float* __restrict__ FP, const float* __restrict__ FX, const float* __restrict__ rms,
    int P, const int64_t* __restrict__ valid, int64_t n_valid, double gamma,
    unsigned* __restrict__ count, int flags, const double* __restrict__ ws) {
  const int64_t k = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
  unsigned v = 0;
  if (k < n_valid)
    v = regulator_lane<N>(FP, FX, rms, P, valid[k], gamma,
                          (flags & PYCWB_REGULATOR_EARLY_EXIT) != 0,
                          (flags & PYCWB_REGULATOR_BOUNDS) ? ws : nullptr);
  block_count<BLOCK>(v, count);
}

template <int N>
__global__ __launch_bounds__(256) void k_term_bounds(const float* __restrict__ FP,
                                                     const float* __restrict__ FX, int64_t n_dirs,
                                                     const double* __restrict__ ws,
                                                     double* __restrict__ t_lb,
                                                     double* __restrict__ t_ub) {
  const int64_t s = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (s >= n_dirs) return;
  float fp0[N], fx0[N];
#pragma unroll
  for (int i = 0; i < N; ++i) {
    fp0[i] = FP[s * N + i];
    fx0[i] = FX[s * N + i];
  }
  DirPre<N> dp;
  dir_pre<N>(fp0, fx0, dp);
  const unsigned pr = class_bound_pair<N>(dp, ws);
  t_lb[s] = half_bits_value((unsigned short)(pr & 0xffffu));
  const float uh = half_bits_value((unsigned short)(pr >> 16));
  t_ub[s] = __builtin_isinf(uh) ? -1.0 : (double)uh;
}

__global__ __launch_bounds__(256) void k_classes_plain(const float* __restrict__ rms, int P,
                                                       int n, const double* __restrict__ ws,
                                                       uint8_t* __restrict__ pclass,
                                                       double* __restrict__ cstats,
                                                       uint16_t* __restrict__ crem) {
  pixel_classes_block<256>(rms, P, n, ws, pclass, cstats, crem);
}

template <int N>
__global__ __launch_bounds__(256) void k_class_term_bounds(const float* __restrict__ FP,
                                                           const float* __restrict__ FX,
                                                           int64_t n_dirs,
                                                           const double* __restrict__ cstats,
                                                           double* __restrict__ t_lb,
                                                           double* __restrict__ t_ub) {
  const int64_t s = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (s >= n_dirs) return;
  float fp0[N], fx0[N];
#pragma unroll
  for (int i = 0; i < N; ++i) {
    fp0[i] = FP[s * N + i];
    fx0[i] = FX[s * N + i];
  }
  const int G = (int)cstats[kClassSlotClasses];
  DirPre<N> dp;
  dir_pre<N>(fp0, fx0, dp);
  for (int g = 0; g < kMaxClasses; ++g) {
    double lo = 0.0, hi = -1.0;
    if (g < G && cstats[g * kWStats + kClassSlotCount] > 0.0) {
      const unsigned v = class_bound_pair<N>(dp, cstats + g * kWStats);
      lo = half_bits_value((unsigned short)(v & 0xffffu));
      const float h = half_bits_value((unsigned short)(v >> 16));
      hi = __builtin_isinf(h) ? -1.0 : (double)h;
    }
    t_lb[s * kMaxClasses + g] = lo;
    t_ub[s * kMaxClasses + g] = hi;
  }
}

__global__ __launch_bounds__(256) void k_weight_stats_plain(const float* __restrict__ rms, int P,
                                                            int n, double* __restrict__ out) {
  weight_stats_block<256>(rms, P, n, out);
}

__global__ void k_regulator_finish(const unsigned* __restrict__ count, int64_t n_valid,
                                   double En, float rr, float* __restrict__ REG,
                                   double* __restrict__ reg_f64) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    const double v = regulator_value(n_valid, *count, En);
    if (reg_f64) *reg_f64 = v;
    REG[0] = rr;
    REG[1] = (float)v;
    REG[2] = 0.f;
  }
}

template <int N, int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_scan_plain(
    const float* __restrict__ FP, const float* __restrict__ FX, const int32_t* __restrict__ ml,
    int64_t S, const float* __restrict__ rms, const float2* __restrict__ td, int P, int K,
    const float* __restrict__ REG, double En, double netCC, const int64_t* __restrict__ valid,
    int64_t n_valid, const int32_t* __restrict__ perm, int flags, float* __restrict__ maps,
    unsigned long long* __restrict__ best) {
  const int64_t t = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
  const bool active = t < n_valid;
  // lane t evaluates valid-list position k (delay-tuple order when perm is given)
  int64_t k = active ? (perm ? (int64_t)perm[t] : t) : 0;
  k = k < 0 ? 0 : (k >= n_valid ? n_valid - 1 : k);
  const int64_t s = valid[k];
  const unsigned long long key =
      scan_lane<N, BLOCK / 64>(FP, FX, ml, S, rms, td, P, K, REG[0], REG[1], En, netCC, s, k,
                               active,
                   (flags & PYCWB_SKIP_MASKED_PIXELS) != 0,
                   (flags & PYCWB_TILE_ENERGY_BOUND) != 0, maps);
  block_argmax<BLOCK>(key, best);
}

// l_max / sky_stat_max from the packed key (sky_scan.py:183-189).
__global__ void k_argmax_finish(const unsigned long long* __restrict__ best,
                                const int64_t* __restrict__ valid, int64_t* __restrict__ l_max,
                                double* __restrict__ stat_max) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    const unsigned long long k = *best;
    const float v = k ? key_value(k) : -1.f;
    if (k && v >= 0.f) {
      *l_max = valid[(uint32_t)(k & 0xffffffffu)];
      *stat_max = (double)v;
    } else {
      *l_max = valid[0];
      *stat_max = 0.0;
    }
  }
}

// ---------------------------------------------------------------------------
// Batched family.  Clusters are grouped by detector count n (one persistent
// kernel instantiation per n, so each keeps the register budget of its own n).
// Inside a group, persistent blocks pull (cluster, sky tile) items from
// ticket counters, one queue per XCD; each queue is cluster-major in the
// group's order (largest work first) over its eighth of the tiles, so an
// XCD's concurrent blocks share one cluster's amplitudes in its L2.
// ---------------------------------------------------------------------------
struct BatchDev {
  int n_clusters;
  int tiles_per_cluster;
  int reg_pool;             // tiles per regulator work item (1, 2, 4 or kRegPool)
  int64_t S;
  const int64_t* valid;
  int64_t n_valid;
  const int32_t* order;     // processing order, grouped by n
  const int32_t* npix;      // pixels per cluster
  const int32_t* net;       // network id per cluster
  const int32_t* K;         // delay half-width per cluster
  const int64_t* rms_off;   // float offset into rms
  const int64_t* td_off;    // float2 offset into td
  const float* rms;
  const float2* td;
  const int32_t* net_nifo;  // ifos per network
  const int64_t* net_off;   // element offset of the network's FP/FX (S, n) and ml (n, S)
  const double* net_En;     // network energy threshold per network
  const float* FP;
  const float* FX;
  const int32_t* ml;
  const int32_t* perm;      // per network: valid-list positions in delay-tuple order
  int flags;
  double gamma;
  float rr;
  double netCC;
  unsigned* tickets;        // 2 per group: regulator, scan
  unsigned* counts;         // per cluster regulator count
  unsigned long long* best; // per cluster argmax key
  float* REG1;              // per cluster REG[1] (float32)
  int64_t* l_max;
  double* stat_max;
  double* reg_f64;
  double* wstats;           // per cluster kWStats doubles, or NULL (no regulator bounds)
  uint8_t* pclass;          // class bytes (cluster c's P at rms_off[c]), or NULL (no classes)
  double* cstats;           // per cluster kClassStats doubles
  uint16_t* crem;           // per cluster class_rem_entries(P) counts
  const int64_t* crem_off;  // per cluster offset into crem
  unsigned* cbound;         // class bounds scratch: per wavefront of the grid kList x 16 words
};

__global__ void k_batch_init(BatchDev b, int n_tickets) {
  const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (gid < n_tickets) b.tickets[gid] = 0u;
  for (int64_t c = gid; c < b.n_clusters; c += (int64_t)gridDim.x * blockDim.x) {
    b.counts[c] = 0u;
    b.best[c] = 0ull;
  }
}

__global__ __launch_bounds__(256) void k_batch_weights(BatchDev b) {
  const int c = blockIdx.x;
  weight_stats_block<256>(b.rms + b.rms_off[c], b.npix[c], b.net_nifo[b.net[c]],
                          b.wstats + (int64_t)c * kWStats);
}

__global__ __launch_bounds__(256) void k_batch_classes(BatchDev b) {
  const int c = blockIdx.x;
  const int n = b.net_nifo[b.net[c]];
  // the class bytes of cluster c start at its first noise weight's index
  // (rms_off[c]): a buffer of one byte per noise weight holds every cluster's
  // P class bytes apart from the others whatever their detector counts
  pixel_classes_block<256>(b.rms + b.rms_off[c], b.npix[c], n, b.wstats + (int64_t)c * kWStats,
                           b.pclass + b.rms_off[c], b.cstats + (int64_t)c * kClassStats,
                           b.crem + b.crem_off[c]);
}

// Waves per SIMD the compiler must leave room for (caps VGPRs), per phase and
// detector count (small: up to three detectors).  Measured on MI355X, the
// regulator runs fastest at 5 and the scan at 4 for both three and four
// detectors (a 3-wave scan build, which spilled nothing when the 4-wave one
// still spilled, was 3 % slower: occupancy hides gather latency).
// Work queues per group and phase: 8, one per XCD (1.2 % faster than one
// shared queue on 4,096 clusters of 16-4,096 pixels at HEALPix order 7: 1839
// against 1862 ms per step), or 1 with PYCWB_XCD_QUEUES=0.
#ifndef PYCWB_XCD_QUEUES
#define PYCWB_XCD_QUEUES 1
#endif
constexpr int kTicketsPerGroup = PYCWB_XCD_QUEUES ? 16 : 2;  // per phase: 8 queues or 1

#ifndef PYCWB_OCC_REG_SMALL
#define PYCWB_OCC_REG_SMALL 4
#endif
#ifndef PYCWB_OCC_REG_LARGE
#define PYCWB_OCC_REG_LARGE 4
#endif
#ifndef PYCWB_OCC_SCAN_SMALL
#define PYCWB_OCC_SCAN_SMALL 4
#endif
#ifndef PYCWB_OCC_SCAN_LARGE
#define PYCWB_OCC_SCAN_LARGE 4
#endif
// five or more detectors: at 4 waves the scan spills 18 VGPRs, at 3 none;
// measured on MI355X (five-detector order-8 batch of 2,607 clusters) 0.778
// against 0.799 s for the regulator + scan + argmax
#ifndef PYCWB_OCC_SCAN_WIDE
#define PYCWB_OCC_SCAN_WIDE 3
#endif
template <int PHASE, int N>
struct PhaseOccupancy {
  static constexpr int waves =
      PHASE == 0 ? (N <= 3 ? PYCWB_OCC_REG_SMALL : PYCWB_OCC_REG_LARGE)
                 : ((PHASE == 1 || PHASE == 3)
                        ? (N <= 3 ? PYCWB_OCC_SCAN_SMALL
                                  : (N == 4 ? PYCWB_OCC_SCAN_LARGE : PYCWB_OCC_SCAN_WIDE))
                        : 8);
};

// Regulator work item: reg_pool (at most kRegPool) tiles of BLOCK directions,
// each wavefront taking reg_pool * 64 consecutive delay-order positions.  A direction's sweep
// stops as soon as its term bounds decide its count (regulator_lane), so sweep
// lengths differ widely between neighbouring directions and a wavefront that
// swept one tile ran until its slowest lane ended (at order 7 that doubles the
// sweep's wavefront time).  Here a first pass computes the bounds of all the
// wavefront's directions with every lane busy and counts those decided before
// the first pixel; the others go to a list in LDS, and the sweep hands the
// next listed direction to each lane whose direction is decided, checking
// every 16 pixels.  A direction leaves the sweep only on a bound that decides
// its count, so the count equals that of the full sweep.  Bounds are kept in
// float32 rounded outwards (lower bound down, upper bound up).
#ifndef PYCWB_REG_POOL
#define PYCWB_REG_POOL 8
#endif
constexpr int kRegPool = PYCWB_REG_POOL;
// Class-bound scratch of the batched regulator: per wavefront of the grid, a
// half-precision bound pair per class for each of the kRegPool * 64 listed
// directions.
inline size_t class_bound_scratch_bytes(int block, int grid) {
  return (size_t)grid * (size_t)(block / 64) * (size_t)(kRegPool * 64) * kMaxClasses *
         sizeof(unsigned);
}

__device__ __forceinline__ int lane_rank(unsigned long long mask) {
  return (int)__builtin_amdgcn_mbcnt_hi((unsigned)(mask >> 32),
                                        __builtin_amdgcn_mbcnt_lo((unsigned)mask, 0u));
}

template <int N, int BLOCK>
__device__ __forceinline__ void regulator_item(const BatchDev& b, int c, int item) {
  constexpr int kList = kRegPool * 64;
  static_assert(2 * kList >= kMaxClasses * 64, "class bounds fit the bound area");
  __shared__ int s_dir[BLOCK / 64][kList];
  // per listed direction: float bits of its lower and upper term bound
  // (without pixel classes)
  __shared__ unsigned s_aux[BLOCK / 64][2 * kList];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int net = b.net[c];
  const float* FP = b.FP + b.net_off[net];
  const float* FX = b.FX + b.net_off[net];
  const float* rms = b.rms + b.rms_off[c];
  const int P = b.npix[c];
  const bool early = (b.flags & PYCWB_REGULATOR_EARLY_EXIT) != 0;
  const double* ws = (b.wstats && (b.flags & PYCWB_REGULATOR_BOUNDS))
                         ? b.wstats + (int64_t)c * kWStats : nullptr;
  const double* cls = (ws && early && b.cstats) ? b.cstats + (int64_t)c * kClassStats : nullptr;
  const int G = cls ? (int)cls[kClassSlotClasses] : 1;
  const bool classes = G > 1;
  const uint16_t* crem = classes ? b.crem + b.crem_off[c] : nullptr;
  // this wavefront's class bounds: kMaxClasses half pairs per listed direction
  unsigned* cb = classes ? b.cbound + ((int64_t)blockIdx.x * (BLOCK / 64) + wave) * kList *
                               kMaxClasses
                         : nullptr;
  const double den = (double)P + 0.01;
  const double thr_hi = b.gamma * b.gamma * den * (1.0 + 1e-9);
  const double thr_lo = b.gamma * b.gamma * den * (1.0 - 1e-9);
  const double sum_lo = 1.0 - (double)P * 1.2e-7, sum_hi = 1.0 + (double)P * 1.2e-7;
  const int pool = b.reg_pool;
  const int64_t t0 = ((int64_t)item * (BLOCK / 64) + wave) * (pool * 64);
  unsigned cnt = 0;

  // pass 1: bounds and the check before the first pixel
  int n_open = 0;
  for (int g = 0; g < pool; ++g) {
    const int64_t t = t0 + g * 64 + lane;
    const bool active = t < b.n_valid;
    int64_t k = active ? (b.perm ? (int64_t)b.perm[(int64_t)net * b.n_valid + t] : t) : 0;
    k = k < 0 ? 0 : (k >= b.n_valid ? b.n_valid - 1 : k);
    const int64_t s = b.valid[k];
    bool open = active;
    float lb = 0.f, ub = -1.f;
    if (active && ws) {
      float fp0[N], fx0[N];
#pragma unroll
      for (int i = 0; i < N; ++i) {
        fp0[i] = FP[s * N + i];
        fx0[i] = FX[s * N + i];
      }
      // the cluster-wide bounds, float32 as a class's (the cluster's statistics
      // in place of a class's)
      DirPre<N> dp;
      dir_pre<N>(fp0, fx0, dp);
      const unsigned pr = class_bound_pair<N>(dp, ws);
      lb = half_bits_value((unsigned short)(pr & 0xffffu));
      const float uh = half_bits_value((unsigned short)(pr >> 16));
      ub = __builtin_isinf(uh) ? -1.f : uh;
      if (early) {
        if ((double)P * (double)lb * sum_lo > thr_hi) {
          cnt += 1u;
          open = false;
        } else if (ub >= 0.f && (double)P * (double)ub * sum_hi < thr_lo) {
          open = false;
        }
      }
    }
    const unsigned long long m = __ballot(open);
    if (open) {
      const int slot = n_open + lane_rank(m);
      s_dir[wave][slot] = (int)s;
      if (!classes) {
        s_aux[wave][slot] = __float_as_uint(lb);
        s_aux[wave][kList + slot] = __float_as_uint(ub);
      }
    }
    n_open += __popcll(m);
  }
  __syncthreads();

  // pass 1b (pixel classes): each listed direction's bounds per class, one
  // direction per lane, into the wavefront's area of the class-bound scratch
  if (classes) {
    for (int slot = lane; slot < n_open; slot += 64) {
      const int64_t s = s_dir[wave][slot];
      float fp1[N], fx1[N];
#pragma unroll
      for (int i = 0; i < N; ++i) {
        fp1[i] = FP[s * N + i];
        fx1[i] = FX[s * N + i];
      }
      DirPre<N> dp;
      dir_pre<N>(fp1, fx1, dp);
#pragma unroll 1
      for (int g = 0; g < kMaxClasses; ++g) {
        unsigned v = 0u;
        if (g < G && cls[g * kWStats + kClassSlotCount] > 0.0)
          v = class_bound_pair<N>(dp, cls + g * kWStats);
        cb[(int64_t)slot * kMaxClasses + g] = v;
      }
    }
    __syncthreads();  // the other lanes of the wavefront read these
  }

  // pass 2: the sweep, a listed direction per idle lane
  int next = 0;
  bool busy = false;
  int j = 0;
  float NI = 0.f, lb = 0.f, ub = -1.f;
  unsigned NN = 0;
  int cur = 0;  // the lane's list slot
  float fp0[N], fx0[N];
#pragma unroll
  for (int i = 0; i < N; ++i) fp0[i] = fx0[i] = 0.f;
  for (;;) {
    const unsigned long long idle = __ballot(!busy);
    if (!busy) {
      const int slot = next + lane_rank(idle);
      if (slot < n_open) {
        const int64_t s = s_dir[wave][slot];
#pragma unroll
        for (int i = 0; i < N; ++i) {
          fp0[i] = FP[s * N + i];
          fx0[i] = FX[s * N + i];
        }
        cur = slot;
        if (!classes) {
          lb = __uint_as_float(s_aux[wave][slot]);
          ub = __uint_as_float(s_aux[wave][kList + slot]);
        }
        j = 0;
        NI = 0.f;
        NN = 0u;
        busy = true;
      }
    }
    next += __popcll(idle);
    if (__ballot(busy) == 0ull) break;  // every listed direction decided
    if (busy && early && classes && j < P) {
      // remaining pixels of each class times its bounds (j is a multiple of
      // 16; at j = 0 the class bounds may decide what the cluster-wide bounds
      // could not); a count of 0 also needs every pixel so far in NN (fp > 0)
      const uint4* rq = reinterpret_cast<const uint4*>(crem + (int64_t)(j >> 4) * kMaxClasses);
      const uint4 r0 = rq[0], r1 = rq[1];
      const unsigned rw[8] = {r0.x, r0.y, r0.z, r0.w, r1.x, r1.y, r1.z, r1.w};
      const uint4* bq = reinterpret_cast<const uint4*>(cb + (int64_t)cur * kMaxClasses);
      const uint4 b0 = bq[0], b1 = bq[1], b2 = bq[2], b3 = bq[3];
      const unsigned bw[16] = {b0.x, b0.y, b0.z, b0.w, b1.x, b1.y, b1.z, b1.w,
                               b2.x, b2.y, b2.z, b2.w, b3.x, b3.y, b3.z, b3.w};
      double ls = 0.0, us = 0.0;
      bool hub = NN == (unsigned)j;
#pragma unroll
      for (int g = 0; g < kMaxClasses; ++g) {
        const unsigned rem = (rw[g >> 1] >> ((g & 1) * 16)) & 0xffffu;
        if (g < G && rem) {
          const unsigned v = bw[g];
          const float lo = half_bits_value((unsigned short)(v & 0xffffu));
          const float hi = half_bits_value((unsigned short)(v >> 16));
          ls += (double)rem * (double)lo;
          if (__builtin_isinf(hi)) hub = false;
          else us += (double)rem * (double)hi;
        }
      }
      if (((double)NI + ls) * sum_lo > thr_hi) {
        cnt += 1u;
        busy = false;
      } else if (hub && ((double)NI + us) * sum_hi < thr_lo) {
        busy = false;
      }
    }
    if (busy) {
      const int stop = j + 16 < P ? j + 16 : P;
      for (; j < stop; ++j) {
        float r[N];
#pragma unroll
        for (int i = 0; i < N; ++i) r[i] = rms[(int64_t)j * N + i];
        regulator_pixel<N>(r, fp0, fx0, NI, NN);
      }
      if (j >= P) {
        const double q = __builtin_sqrt((double)NI / ((double)NN + 0.01));
        cnt += q > b.gamma ? 1u : 0u;
        busy = false;
      } else if (early && !classes) {
        const double rest = (double)(P - j);
        if (((double)NI + rest * (double)lb) * sum_lo > thr_hi) {
          cnt += 1u;
          busy = false;
        } else if (ub >= 0.f && ((double)NI + rest * (double)ub) * sum_hi < thr_lo) {
          busy = false;
        }
      }
    }
  }
  block_count<BLOCK>(cnt, &b.counts[c]);
}

template <int N, int BLOCK, int PHASE>
__device__ __forceinline__ void batch_item(const BatchDev& b, int c, int tile) {
  const int64_t t = (int64_t)tile * BLOCK + threadIdx.x;
  const bool active = t < b.n_valid;
  const int net = b.net[c];
  int64_t k = active ? (b.perm ? (int64_t)b.perm[(int64_t)net * b.n_valid + t] : t) : 0;
  k = k < 0 ? 0 : (k >= b.n_valid ? b.n_valid - 1 : k);
  const int64_t s = b.valid[k];
  const float* FP = b.FP + b.net_off[net];
  const float* FX = b.FX + b.net_off[net];
  const float* rms = b.rms + b.rms_off[c];
  const int P = b.npix[c];
  if (PHASE == 2) {
    // energy test only: how many (direction, pixel) pairs take the full path
    const int K = b.K[c], nd = 2 * K + 1;
    const int32_t* ml = b.ml + b.net_off[net];
    const float2* td = b.td + b.td_off[c];
    const double En = b.net_En[net];
    int tap[N];
#pragma unroll
    for (int i = 0; i < N; ++i) {
      int t = ml[(int64_t)i * b.S + s] + K;
      tap[i] = i * nd + (t < 0 ? 0 : (t >= nd ? nd - 1 : t));
    }
    unsigned cnt = 0;
    if (active) {
      for (int j = 0; j < P; ++j) {
        float v00[N], v90[N];
#pragma unroll
        for (int i = 0; i < N; ++i) {
          const float2 v = td[(int64_t)j * N * nd + tap[i]];
          v00[i] = v.x;
          v90[i] = v.y;
        }
        float e, et;
        int m;
        pixel_energy<N>(v00, v90, En, e, m, et);
        cnt += pixel_contributes(e, m) ? 1u : 0u;
      }
    }
    block_count<BLOCK>(cnt, &b.counts[c]);
  } else if (PHASE == 3) {
    // the scan with its regulator from the last finished launch, counting the
    // pairs whose energy it evaluates into counts[] (no argmax)
    const double En = b.net_En[net];
    const int32_t* ml = b.ml + b.net_off[net];
    const float2* td = b.td + b.td_off[c];
    (void)scan_lane<N, BLOCK / 64, true, false>(FP, FX, ml, b.S, rms, td, P, b.K[c], b.rr,
                                                b.REG1[c], En, b.netCC, s, k, active,
                                                (b.flags & PYCWB_SKIP_MASKED_PIXELS) != 0,
                                                (b.flags & PYCWB_TILE_ENERGY_BOUND) != 0,
                                                nullptr, &b.counts[c]);
  } else {
    const double En = b.net_En[net];
    const float RR = (float)regulator_value(b.n_valid, b.counts[c], En);
    const int32_t* ml = b.ml + b.net_off[net];
    const float2* td = b.td + b.td_off[c];
    const unsigned long long key =
        scan_lane<N, BLOCK / 64, false, false>(FP, FX, ml, b.S, rms, td, P, b.K[c], b.rr, RR,
                                               En, b.netCC, s, k, active,
                                               (b.flags & PYCWB_SKIP_MASKED_PIXELS) != 0,
                                               (b.flags & PYCWB_TILE_ENERGY_BOUND) != 0,
                                               nullptr);
    block_argmax<BLOCK>(key, &b.best[c]);
  }
}

// Work items per cluster: tiles of BLOCK directions (scan, energy count) or
// groups of reg_pool tiles (regulator).
template <int PHASE>
__device__ __forceinline__ int item_count(const BatchDev& b) {
  return PHASE == 0 ? (b.tiles_per_cluster + b.reg_pool - 1) / b.reg_pool
                    : b.tiles_per_cluster;
}

template <int N, int BLOCK, int PHASE>
__device__ __forceinline__ void run_item(const BatchDev& b, int c, int item) {
  if constexpr (PHASE == 0) {
    regulator_item<N, BLOCK>(b, c, item);
  } else {
    batch_item<N, BLOCK, PHASE>(b, c, item);
  }
}

template <int N, int BLOCK, int PHASE>
__global__ __launch_bounds__(BLOCK, (PhaseOccupancy<PHASE, N>::waves)) void k_batch_persistent(
    BatchDev b, int first, int count, unsigned* ticket) {
  __shared__ unsigned s_item;
#if PYCWB_XCD_QUEUES
  // Eight queues per group, one per XCD (workgroup i runs on XCD i mod 8):
  // queue q holds the tiles [q * per, (q + 1) * per) of every cluster, cluster
  // by cluster, so an XCD's workgroups take neighbouring tiles (neighbouring
  // delay tuples, the same amplitude lines) and share its L2.  A queue that
  // runs dry moves on to the next XCD's queue; every queue's counter only
  // grows, so each item is taken once and every workgroup ends.
  const unsigned tiles = (unsigned)item_count<PHASE>(b);
  const unsigned per = (tiles + 7u) / 8u;
  const unsigned n_items = (unsigned)count * per;
  const int xcd = (int)(blockIdx.x & 7u);
  for (int qi = 0; qi < 8; ++qi) {
    const int q = (xcd + qi) & 7;
    for (;;) {
      if (threadIdx.x == 0) s_item = atomicAdd(ticket + q, 1u);
      __syncthreads();
      // the same on every lane: scalar from here on
      const unsigned item = __builtin_amdgcn_readfirstlane(s_item);
      __syncthreads();
      if (item >= n_items) break;
      const unsigned tile = (unsigned)q * per + item % per;
      if (tile >= tiles) continue;
      run_item<N, BLOCK, PHASE>(b, b.order[first + (int)(item / per)], (int)tile);
    }
  }
#else
  const unsigned tiles = (unsigned)item_count<PHASE>(b);
  const unsigned n_items = (unsigned)count * tiles;
  for (;;) {
    if (threadIdx.x == 0) s_item = atomicAdd(ticket, 1u);
    __syncthreads();
    const unsigned item = __builtin_amdgcn_readfirstlane(s_item);
    __syncthreads();
    if (item >= n_items) break;  // every block reaches this once the tickets run out
    run_item<N, BLOCK, PHASE>(b, b.order[first + (int)(item / tiles)], (int)(item % tiles));
  }
#endif
}

__global__ void k_batch_finish(BatchDev b) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= b.n_clusters) return;
  const int net = b.net[c];
  const double reg = regulator_value(b.n_valid, b.counts[c], b.net_En[net]);
  b.REG1[c] = (float)reg;
  if (b.reg_f64) b.reg_f64[c] = reg;
  const unsigned long long k = b.best[c];
  const float v = k ? key_value(k) : -1.f;
  if (k && v >= 0.f) {
    b.l_max[c] = b.valid[(uint32_t)(k & 0xffffffffu)];
    b.stat_max[c] = (double)v;
  } else {
    b.l_max[c] = b.valid[0];
    b.stat_max[c] = 0.0;
  }
}

// ---------------------------------------------------------------------------
// launch helpers
// ---------------------------------------------------------------------------
template <int BLOCK>
int launch_regulator_plain(int n, const float* FP, const float* FX, const float* rms, int P,
                           const int64_t* valid, int64_t n_valid, double gamma,
                           unsigned* count, int flags, const double* ws, hipStream_t st) {
  const unsigned grid = (unsigned)((n_valid + BLOCK - 1) / BLOCK);
  switch (n) {
#define CASE(NN)                                                                          \
  case NN:                                                                                \
    hipLaunchKernelGGL((k_regulator_plain<NN, BLOCK>), dim3(grid), dim3(BLOCK), 0, st, FP, \
                       FX, rms, P, valid, n_valid, gamma, count, flags, ws);              \
    break;
    CASE(1) CASE(2) CASE(3) CASE(4) CASE(5) CASE(6) CASE(7) CASE(8)
#undef CASE
    default: return PYCWB_ERR_ARG;
  }
  HIP_TRY(hipGetLastError());
  return PYCWB_OK;
}

template <int BLOCK>
int launch_scan_plain(int n, const float* FP, const float* FX, const int32_t* ml, int64_t S,
                      const float* rms, const float2* td, int P, int K, const float* REG,
                      double En, double netCC, const int64_t* valid, int64_t n_valid,
                      const int32_t* perm, int flags, float* maps, unsigned long long* best,
                      hipStream_t st) {
  const unsigned grid = (unsigned)((n_valid + BLOCK - 1) / BLOCK);
  switch (n) {
#define CASE(NN)                                                                          \
  case NN:                                                                                \
    hipLaunchKernelGGL((k_scan_plain<NN, BLOCK>), dim3(grid), dim3(BLOCK), 0, st, FP, FX, \
                       ml, S, rms, td, P, K, REG, En, netCC, valid, n_valid, perm, flags,  \
                       maps, best);                                                        \
    break;
    CASE(1) CASE(2) CASE(3) CASE(4) CASE(5) CASE(6) CASE(7) CASE(8)
#undef CASE
    default: return PYCWB_ERR_ARG;
  }
  HIP_TRY(hipGetLastError());
  return PYCWB_OK;
}

template <int BLOCK, int PHASE>
int launch_group(int n, const BatchDev& d, int grid, int first, int count, unsigned* ticket,
                 hipStream_t st) {
  switch (n) {
#define CASE(NN)                                                                           \
  case NN:                                                                                 \
    hipLaunchKernelGGL((k_batch_persistent<NN, BLOCK, PHASE>), dim3(grid), dim3(BLOCK), 0, \
                       st, d, first, count, ticket);                                       \
    break;
    CASE(1) CASE(2) CASE(3) CASE(4) CASE(5) CASE(6) CASE(7) CASE(8)
#undef CASE
    default: return PYCWB_ERR_ARG;
  }
  HIP_TRY(hipGetLastError());
  return PYCWB_OK;
}

// Workspace of the plain entry points: 256 bytes of counters, then packed td.
constexpr size_t kCounterBytes = 256;

size_t plain_ws_bytes(int n, int P, int nd) {
  return kCounterBytes + (size_t)n * P * nd * sizeof(float2);
}

struct PlainWs {
  unsigned long long* best;  // argmax key
  unsigned* count;           // regulator count
  double* wstats;            // noise-weight statistics (kWStats doubles at byte 64)
  float2* td;
};

PlainWs plain_ws(void* ws) {
  PlainWs w;
  w.best = (unsigned long long*)ws;
  w.count = (unsigned*)((char*)ws + 8);
  w.wstats = (double*)((char*)ws + 64);
  w.td = (float2*)((char*)ws + kCounterBytes);
  return w;
}
static_assert(64 + kWStats * sizeof(double) <= kCounterBytes, "plain workspace layout");

bool valid_args(int n, int P, int64_t S, int64_t n_valid) {
  return n >= 1 && n <= kMaxIfo && P >= 1 && S >= 1 && n_valid >= 1 && n_valid <= S &&
         n_valid < (int64_t)0xffffffff;
}

unsigned grid_for(int64_t count, int64_t cap) {
  int64_t g = (count + 255) / 256;
  if (g > cap) g = cap;
  if (g < 1) g = 1;
  return (unsigned)g;
}

}  // namespace

// ===========================================================================
// C ABI (declared in include/pycwb_amd.h)
// ===========================================================================
struct pycwb_batch {
  BatchDev dev;
  int block;
  int grid;
  int n_groups;
  int group_first[kMaxIfo];
  int group_count[kMaxIfo];
  int group_nifo[kMaxIfo];
  hipGraph_t graph;
  hipGraphExec_t exec;
};

extern "C" {

size_t pycwb_sky_scan_workspace_bytes(int n_ifo, int n_pix, int n_delay) {
  return plain_ws_bytes(n_ifo, n_pix, n_delay);
}

int pycwb_td_relayout(int n_ifo, int n_pix, int n_delay, const float* td00, const float* td90,
                      void* workspace, size_t workspace_bytes, void* stream) {
  clear_stale_error();
  if (n_ifo < 1 || n_ifo > kMaxIfo || n_pix < 1 || n_delay < 1 || !(n_delay & 1))
    return PYCWB_ERR_ARG;
  if (!td00 || !td90 || !workspace) return PYCWB_ERR_ARG;
  if (workspace_bytes < plain_ws_bytes(n_ifo, n_pix, n_delay)) return PYCWB_ERR_ARG;
  hipStream_t st = (hipStream_t)stream;
  PlainWs w = plain_ws(workspace);
  const int64_t total = (int64_t)n_ifo * n_pix * n_delay;
  hipLaunchKernelGGL(k_relayout_td, dim3(grid_for(total, 65536)), dim3(256), 0, st, td00, td90,
                     n_ifo, n_pix, n_delay, w.td, (unsigned*)w.best, 4);
  HIP_TRY(hipGetLastError());
  return PYCWB_OK;
}

int pycwb_regulator_term_bounds(int n_ifo, int n_pix, const float* rms, const float* FP,
                                const float* FX, int64_t n_dirs, double* weight_stats,
                                double* t_lb, double* t_ub, void* stream) {
  clear_stale_error();
  if (n_ifo < 1 || n_ifo > kMaxIfo || n_pix < 1 || n_dirs < 1) return PYCWB_ERR_ARG;
  if (!rms || !FP || !FX || !weight_stats || !t_lb || !t_ub) return PYCWB_ERR_ARG;
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(k_weight_stats_plain, dim3(1), dim3(256), 0, st, rms, n_pix, n_ifo,
                     weight_stats);
  HIP_TRY(hipGetLastError());
  const unsigned grid = (unsigned)((n_dirs + 255) / 256);
  switch (n_ifo) {
#define CASE(NN)                                                                         \
  case NN:                                                                               \
    hipLaunchKernelGGL((k_term_bounds<NN>), dim3(grid), dim3(256), 0, st, FP, FX, n_dirs, \
                       weight_stats, t_lb, t_ub);                                        \
    break;
    CASE(1) CASE(2) CASE(3) CASE(4) CASE(5) CASE(6) CASE(7) CASE(8)
#undef CASE
    default: return PYCWB_ERR_ARG;
  }
  HIP_TRY(hipGetLastError());
  return PYCWB_OK;
}

int pycwb_regulator_class_bounds(int n_ifo, int n_pix, const float* rms, const float* FP,
                                 const float* FX, int64_t n_dirs, double* weight_stats,
                                 uint8_t* pixel_class, double* class_stats, uint16_t* class_rem,
                                 double* t_lb, double* t_ub, void* stream) {
  clear_stale_error();
  if (n_ifo < 1 || n_ifo > kMaxIfo || n_pix < 1 || n_dirs < 1) return PYCWB_ERR_ARG;
  if (!rms || !FP || !FX || !weight_stats || !pixel_class || !class_stats || !class_rem ||
      !t_lb || !t_ub)
    return PYCWB_ERR_ARG;
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(k_weight_stats_plain, dim3(1), dim3(256), 0, st, rms, n_pix, n_ifo,
                     weight_stats);
  HIP_TRY(hipGetLastError());
  hipLaunchKernelGGL(k_classes_plain, dim3(1), dim3(256), 0, st, rms, n_pix, n_ifo,
                     (const double*)weight_stats, pixel_class, class_stats, class_rem);
  HIP_TRY(hipGetLastError());
  const unsigned grid = (unsigned)((n_dirs + 255) / 256);
  switch (n_ifo) {
#define CASE(NN)                                                                             \
  case NN:                                                                                   \
    hipLaunchKernelGGL((k_class_term_bounds<NN>), dim3(grid), dim3(256), 0, st, FP, FX, n_dirs, \
                       (const double*)class_stats, t_lb, t_ub);                              \
    break;
    CASE(1) CASE(2) CASE(3) CASE(4) CASE(5) CASE(6) CASE(7) CASE(8)
#undef CASE
    default: return PYCWB_ERR_ARG;
  }
  HIP_TRY(hipGetLastError());
  return PYCWB_OK;
}

int pycwb_dpf_regulator(int n_ifo, int n_pix, int64_t n_sky, const float* FP, const float* FX,
                        const float* rms, double gamma_regulator,
                        double network_energy_threshold, float reg0,
                        const int64_t* sky_valid_indices, int64_t n_valid, int flags,
                        float* REG, double* reg_f64, void* workspace, size_t workspace_bytes,
                        void* stream) {
  clear_stale_error();
  if (!valid_args(n_ifo, n_pix, n_sky, n_valid)) return PYCWB_ERR_ARG;
  if (!FP || !FX || !rms || !sky_valid_indices || !REG || !workspace) return PYCWB_ERR_ARG;
  if (workspace_bytes < kCounterBytes) return PYCWB_ERR_ARG;
  hipStream_t st = (hipStream_t)stream;
  PlainWs w = plain_ws(workspace);
  hipLaunchKernelGGL(k_fill, dim3(1), dim3(64), 0, st, (float*)w.count, (int64_t)1, 0.f);
  HIP_TRY(hipGetLastError());
  if (flags & PYCWB_REGULATOR_BOUNDS) {
    hipLaunchKernelGGL(k_weight_stats_plain, dim3(1), dim3(256), 0, st, rms, n_pix, n_ifo,
                       w.wstats);
    HIP_TRY(hipGetLastError());
  }
  int rc = launch_regulator_plain<256>(n_ifo, FP, FX, rms, n_pix, sky_valid_indices, n_valid,
                                       gamma_regulator, w.count, flags, w.wstats, st);
  if (rc) return rc;
  hipLaunchKernelGGL(k_regulator_finish, dim3(1), dim3(64), 0, st, w.count, n_valid,
                     network_energy_threshold, reg0, REG, reg_f64);
  HIP_TRY(hipGetLastError());
  return PYCWB_OK;
}

int pycwb_sky_scan_plain(int n_ifo, int n_pix, int64_t n_sky, int n_delay, const float* FP,
                         const float* FX, const int32_t* ml, const float* rms,
                         const float* td_packed, const float* REG, double netCC,
                         double network_energy_threshold, const int64_t* sky_valid_indices,
                         int64_t n_valid, const int32_t* sky_order, int flags, float* maps,
                         int64_t* l_max, double* sky_stat_max, void* workspace,
                         size_t workspace_bytes, void* stream) {
  clear_stale_error();
  if (!valid_args(n_ifo, n_pix, n_sky, n_valid) || n_delay < 1 || !(n_delay & 1))
    return PYCWB_ERR_ARG;
  if (!FP || !FX || !ml || !rms || !REG || !sky_valid_indices || !l_max || !sky_stat_max ||
      !workspace)
    return PYCWB_ERR_ARG;
  if (workspace_bytes < (td_packed ? kCounterBytes : plain_ws_bytes(n_ifo, n_pix, n_delay)))
    return PYCWB_ERR_ARG;
  hipStream_t st = (hipStream_t)stream;
  PlainWs w = plain_ws(workspace);
  const float2* td = td_packed ? (const float2*)td_packed : w.td;
  hipLaunchKernelGGL(k_fill, dim3(1), dim3(64), 0, st, (float*)w.best, (int64_t)2, 0.f);
  HIP_TRY(hipGetLastError());
  if (maps && n_valid < n_sky) {
    const int64_t cnt = (int64_t)kNumMaps * n_sky;
    hipLaunchKernelGGL(k_fill, dim3(grid_for(cnt, 65536)), dim3(256), 0, st, maps, cnt, 0.f);
    HIP_TRY(hipGetLastError());
  }
  const int K = n_delay / 2;
  int rc = launch_scan_plain<256>(n_ifo, FP, FX, ml, n_sky, rms, td, n_pix, K, REG,
                                  network_energy_threshold, netCC, sky_valid_indices, n_valid,
                                  sky_order, flags, maps, w.best, st);
  if (rc) return rc;
  hipLaunchKernelGGL(k_argmax_finish, dim3(1), dim3(64), 0, st, w.best, sky_valid_indices,
                     l_max, sky_stat_max);
  HIP_TRY(hipGetLastError());
  return PYCWB_OK;
}

int pycwb_batch_launch_stage(pycwb_batch* b, int stage, void* stream);

static int batch_enqueue(const pycwb_batch* b, hipStream_t st) {
  for (int stage = 0; stage < 4; ++stage) {
    int rc = pycwb_batch_launch_stage(const_cast<pycwb_batch*>(b), stage, (void*)st);
    if (rc) return rc;
  }
  return PYCWB_OK;
}

}  // extern "C"

namespace {

// Checks a batch description and fills `b` from it (graph left empty).  With
// own_scratch the description's tickets/counts/best/weight_stats must be set;
// without, the caller supplies them.
int batch_from_desc(const pycwb_batch_desc* d, bool own_scratch, pycwb_batch* b) {
  if (!d) return PYCWB_ERR_ARG;
  if (d->n_clusters < 1 || d->n_sky < 1 || d->n_valid < 1 || d->n_valid > d->n_sky ||
      d->n_sky >= (int64_t)0x7fffffff)
    return PYCWB_ERR_ARG;
  if (d->block != 64 && d->block != 128 && d->block != 256) return PYCWB_ERR_ARG;
  if (d->grid < 1) return PYCWB_ERR_ARG;
  if (d->n_groups < 1 || d->n_groups > kMaxIfo) return PYCWB_ERR_ARG;
  const int64_t tiles = (d->n_valid + d->block - 1) / d->block;
  if ((int64_t)d->n_clusters * tiles >= (int64_t)0xffffffff) return PYCWB_ERR_ARG;
  int covered = 0;
  for (int g = 0; g < d->n_groups; ++g) {
    if (d->group_nifo[g] < 1 || d->group_nifo[g] > kMaxIfo) return PYCWB_ERR_ARG;
    if (d->group_first[g] != covered || d->group_count[g] < 0) return PYCWB_ERR_ARG;
    covered += d->group_count[g];
  }
  if (covered != d->n_clusters) return PYCWB_ERR_ARG;
  if (!d->sky_valid_indices || !d->order || !d->n_pix || !d->net || !d->K || !d->rms_off ||
      !d->td_off || !d->rms || !d->td || !d->net_nifo || !d->net_off || !d->net_En || !d->FP ||
      !d->FX || !d->ml || !d->REG1 || !d->l_max || !d->sky_stat_max)
    return PYCWB_ERR_ARG;
  if (own_scratch && (!d->tickets || !d->counts || !d->best)) return PYCWB_ERR_ARG;
  BatchDev& v = b->dev;
  v.n_clusters = d->n_clusters;
  v.tiles_per_cluster = (int)tiles;
  // The largest regulator pool that still leaves at least 16 work items per
  // persistent workgroup: with fewer, the largest clusters' last items decide
  // the stage time (16 clusters of 16-4,096 pixels at HEALPix order 7:
  // regulator 7.1 ms with one direction per lane, 9.1 ms with pools of 2).
  v.reg_pool = kRegPool;
  while (v.reg_pool > 1 &&
         (int64_t)d->n_clusters * ((tiles + v.reg_pool - 1) / v.reg_pool) < 16 * (int64_t)d->grid)
    v.reg_pool /= 2;
  v.S = d->n_sky;
  v.valid = d->sky_valid_indices;
  v.n_valid = d->n_valid;
  v.order = d->order;
  v.npix = d->n_pix;
  v.net = d->net;
  v.K = d->K;
  v.rms_off = d->rms_off;
  v.td_off = d->td_off;
  v.rms = d->rms;
  v.td = (const float2*)d->td;
  v.net_nifo = d->net_nifo;
  v.net_off = d->net_off;
  v.net_En = d->net_En;
  v.FP = d->FP;
  v.FX = d->FX;
  v.ml = d->ml;
  v.perm = d->sky_order;
  v.flags = d->flags;
  v.gamma = d->gamma_regulator;
  v.rr = d->reg0;
  v.netCC = d->netCC;
  v.tickets = d->tickets;
  v.counts = d->counts;
  v.best = d->best;
  v.REG1 = d->REG1;
  v.l_max = d->l_max;
  v.stat_max = d->sky_stat_max;
  v.reg_f64 = d->reg_f64;
  v.wstats = d->weight_stats;
  // pixel classes need all five buffers, the weight statistics and a
  // class-bound scratch for every wavefront of the grid
  const bool cls = d->class_stats && d->pixel_class && d->class_rem && d->class_rem_off &&
                   d->class_bounds && d->weight_stats &&
                   d->class_bounds_bytes >= class_bound_scratch_bytes(d->block, d->grid);
  v.pclass = cls ? d->pixel_class : nullptr;
  v.cstats = cls ? d->class_stats : nullptr;
  v.crem = cls ? d->class_rem : nullptr;
  v.crem_off = cls ? d->class_rem_off : nullptr;
  v.cbound = cls ? d->class_bounds : nullptr;
  b->block = d->block;
  b->grid = d->grid;
  b->n_groups = d->n_groups;
  for (int g = 0; g < d->n_groups; ++g) {
    b->group_first[g] = d->group_first[g];
    b->group_count[g] = d->group_count[g];
    b->group_nifo[g] = d->group_nifo[g];
  }
  b->graph = nullptr;
  b->exec = nullptr;
  return PYCWB_OK;
}

// ---------------------------------------------------------------------------
// Scratch of pycwb_batch_run_noalloc: one array per GPU, held until the
// process ends.  Layout: tickets (kTicketsPerGroup x 8 groups uint32, padded
// to 1 KiB), counts (uint32), best (uint64), weight statistics (kWStats
// doubles), each for kNoAllocMaxClusters clusters.
// ---------------------------------------------------------------------------
constexpr int kNoAllocMaxClusters = 16384;
constexpr size_t kNoAllocTicketBytes = 1024;
constexpr size_t kNoAllocCountsOff = kNoAllocTicketBytes;
constexpr size_t kNoAllocBestOff = kNoAllocCountsOff + sizeof(unsigned) * kNoAllocMaxClusters;
constexpr size_t kNoAllocWStatsOff =
    kNoAllocBestOff + sizeof(unsigned long long) * kNoAllocMaxClusters;
constexpr size_t kNoAllocScratchBytes =
    kNoAllocWStatsOff + sizeof(double) * kWStats * kNoAllocMaxClusters;
static_assert(sizeof(unsigned) * kTicketsPerGroup * kMaxIfo <= kNoAllocTicketBytes,
              "ticket area of the no-alloc scratch");
static_assert(kNoAllocBestOff % 8 == 0 && kNoAllocWStatsOff % 8 == 0, "scratch alignment");

static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_lock;
unsigned char* g_scratch_ptr[kMaxDevices] = {};

// Device address of this GPU's scratch array (looked up once per device, with
// that device current).  The array's name is a host-side stand-in and must
// never be passed to a kernel.
int scratch_address(unsigned char** out) {
  int dev = 0;
  HIP_TRY(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) {
    pycwb::record_error("pycwb_batch_run_noalloc", "device index beyond 63");
    return PYCWB_ERR_ARG;
  }
  std::lock_guard<std::mutex> guard(g_scratch_lock);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) {
      pycwb::record_error("pycwb_batch_run_noalloc", "no address for the scratch array");
      return PYCWB_ERR_HIP;
    }
    g_scratch_ptr[dev] = (unsigned char*)p;
  }
  *out = g_scratch_ptr[dev];
  return PYCWB_OK;
}

}  // namespace

extern "C" {

int pycwb_batch_create(const pycwb_batch_desc* d, pycwb_batch** out) {
  if (!d || !out) return PYCWB_ERR_ARG;
  *out = nullptr;
  pycwb_batch* b = new pycwb_batch();
  int frc = batch_from_desc(d, true, b);
  if (frc) {
    delete b;
    return frc;
  }
  if (d->use_graph) {
    hipStream_t cs;
    if (hipStreamCreateWithFlags(&cs, hipStreamNonBlocking) != hipSuccess) {
      delete b;
      return PYCWB_ERR_HIP;
    }
    int rc = PYCWB_OK;
    if (hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal) != hipSuccess)
      rc = PYCWB_ERR_HIP;
    if (!rc) {
      int erc = batch_enqueue(b, cs);
      hipError_t e = hipStreamEndCapture(cs, &b->graph);
      if (erc) rc = erc;
      else if (e != hipSuccess) rc = PYCWB_ERR_HIP;
    }
    if (!rc && hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0) != hipSuccess)
      rc = PYCWB_ERR_HIP;
    hipStreamDestroy(cs);
    if (rc) {
      if (b->graph) hipGraphDestroy(b->graph);
      delete b;
      return rc;
    }
  }
  *out = b;
  return PYCWB_OK;
}

int pycwb_batch_run_noalloc(const pycwb_batch_desc* d, void* stream) {
  clear_stale_error();
  pycwb::g_last_error[0] = '\0';
  if (!d) return PYCWB_ERR_ARG;
  if (d->n_clusters > kNoAllocMaxClusters) {
    char why[128];
    snprintf(why, sizeof(why), "%d clusters exceed the maximum of %d", d->n_clusters,
             kNoAllocMaxClusters);
    pycwb::record_error("pycwb_batch_run_noalloc", why);
    return PYCWB_ERR_ARG;
  }
  pycwb_batch b{};
  int rc = batch_from_desc(d, false, &b);
  if (rc) return rc;
  unsigned char* base = nullptr;
  rc = scratch_address(&base);
  if (rc) return rc;
  b.dev.tickets = (unsigned*)base;
  b.dev.counts = (unsigned*)(base + kNoAllocCountsOff);
  b.dev.best = (unsigned long long*)(base + kNoAllocBestOff);
  b.dev.wstats = (double*)(base + kNoAllocWStatsOff);
  b.dev.pclass = nullptr;  // no pixel classes: their buffers scale with the pixels
  b.dev.cstats = nullptr;
  b.dev.crem = nullptr;
  b.dev.crem_off = nullptr;
  b.dev.cbound = nullptr;
  // stage 0 clears the tickets, counts and argmax keys with a fill kernel and
  // writes every weight statistic it later reads
  return batch_enqueue(&b, (hipStream_t)stream);
}

size_t pycwb_batch_noalloc_max_clusters(void) { return kNoAllocMaxClusters; }

int64_t pycwb_class_rem_entries(int n_pix) { return n_pix < 0 ? 0 : class_rem_entries(n_pix); }

size_t pycwb_class_bound_scratch_bytes(int block, int grid) {
  return (block == 64 || block == 128 || block == 256) && grid > 0
             ? class_bound_scratch_bytes(block, grid)
             : 0;
}

int pycwb_batch_launch_stage(pycwb_batch* b, int stage, void* stream) {
  if (!b || stage < 0 || stage > 5) return PYCWB_ERR_ARG;
  clear_stale_error();
  hipStream_t st = (hipStream_t)stream;
  const BatchDev& d = b->dev;
  if (stage == 0) {
    hipLaunchKernelGGL(k_batch_init, dim3(grid_for(d.n_clusters, 4096)), dim3(256), 0, st, d,
                       kTicketsPerGroup * b->n_groups);
    HIP_TRY(hipGetLastError());
    if (d.wstats && (d.flags & PYCWB_REGULATOR_BOUNDS)) {
      hipLaunchKernelGGL(k_batch_weights, dim3((unsigned)d.n_clusters), dim3(256), 0, st, d);
      HIP_TRY(hipGetLastError());
      if (d.cstats) {
        hipLaunchKernelGGL(k_batch_classes, dim3((unsigned)d.n_clusters), dim3(256), 0, st, d);
        HIP_TRY(hipGetLastError());
      }
    }
    return PYCWB_OK;
  }
  if (stage == 3) {
    hipLaunchKernelGGL(k_batch_finish, dim3(grid_for(d.n_clusters, 1 << 20)), dim3(256), 0, st,
                       d);
    HIP_TRY(hipGetLastError());
    return PYCWB_OK;
  }
  // stage 4: count pass (energy test only) into counts[]; uses the regulator
  // tickets, so it runs after stage 0 and before anything that reads counts.
  // stage 5: the scan counting the pairs whose energy it evaluates into
  // counts[] (measurement pass; runs after stage 0, reads REG1 of the last
  // finished launch)
  const int phase = stage == 4 ? 2 : (stage == 5 ? 3 : stage - 1);
  for (int g = 0; g < b->n_groups; ++g) {
    unsigned* ticket =
        d.tickets + kTicketsPerGroup * g + ((phase == 1 || phase == 3) ? kTicketsPerGroup / 2 : 0);
    int rc = PYCWB_ERR_ARG;
    const int n = b->group_nifo[g], first = b->group_first[g], cnt = b->group_count[g];
#define PHASES(BLK)                                                                \
  rc = phase == 3   ? launch_group<BLK, 3>(n, d, b->grid, first, cnt, ticket, st) \
       : phase == 2 ? launch_group<BLK, 2>(n, d, b->grid, first, cnt, ticket, st) \
       : phase      ? launch_group<BLK, 1>(n, d, b->grid, first, cnt, ticket, st) \
                    : launch_group<BLK, 0>(n, d, b->grid, first, cnt, ticket, st);
    if (b->block == 64) {
      PHASES(64)
    } else if (b->block == 128) {
      PHASES(128)
    } else if (b->block == 256) {
      PHASES(256)
    }
#undef PHASES
    if (rc) return rc;
  }
  return PYCWB_OK;
}

int pycwb_batch_launch(pycwb_batch* b, void* stream) {
  if (!b) return PYCWB_ERR_ARG;
  hipStream_t st = (hipStream_t)stream;
  if (b->exec) {
    HIP_TRY(hipGraphLaunch(b->exec, st));
    return PYCWB_OK;
  }
  return batch_enqueue(b, st);
}

int pycwb_batch_destroy(pycwb_batch* b) {
  if (!b) return PYCWB_OK;
  if (b->exec) hipGraphExecDestroy(b->exec);
  if (b->graph) hipGraphDestroy(b->graph);
  delete b;
  return PYCWB_OK;
}

const char* pycwb_version(void) { return "pycwb_amd 0.1 gfx950"; }

const char* pycwb_last_error(void) { return pycwb::last_error(); }

}  // extern "C"
