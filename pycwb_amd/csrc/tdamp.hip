// Time-delay amplitude extraction on MI355X: the td00/td90 tables of the
// likelihood from the WDM maps.  Restates batch_get_td_vecs and
// _get_pixel_amplitude_nb (pycwb/utils/td_vector_kernels.py:16-199, cWB's
// WDM<T>::getTDamp): for pixel (n, m) and delay dT in [-K, K] (steps of
// tau/L, J = M L per time bin) the amplitude is the same-band filter sum of
// layer m plus the two cross-band sums of layers m -+ 1, over the 2 n_c + 1
// time bins around n, with the whole-bin shift dT = s J + r (truncation
// toward zero) and, for odd s, the quadrature swap with the (n + m) sign.
// Products float32 x float64 and sums in float64 as the Numba kernel; the
// result is rounded to float32.
// One thread per (pixel, delay): threads of one pixel read the same plane
// values (broadcast) and neighbouring filter rows (cached).
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "skyscan_lane.hh"
#include "pycwb_amd.h"

namespace {

struct TdArgs {
  const int32_t* pix;
  int64_t n_pixels;
  const float* p00;
  const float* p90;
  int64_t n_rows;  // rows of the zero-padded planes (n_time + 2 n_c)
  int M;
  const double* T0;
  const double* Tx;
  const double2* trig;  // (sin, cos) of q pi / (2J), q in [0, 4J)
  int nc;
  int J;
  int K;
  float* out;
  int64_t row_stride;  // floats between consecutive pixels' outputs
  int interleave;      // 0: [a00(-K..K), a90(-K..K)]; 1: (a00, a90) pairs
};

// even / odd filter sums of one layer; rows outside the planes read as zero.
// T points at the delay's entry of tap 0 in a tap-major table (stride D = 2J+1
// between taps), so threads of consecutive delays read consecutive addresses.
__device__ __forceinline__ void band_sums(const float* __restrict__ plane, int64_t n_rows, int M1,
                                          int n, int m, const double* __restrict__ T, int64_t D,
                                          int W, double& se, double& so) {
  double e = 0.0, o = 0.0;
  const float* col = plane + m;
  // pairs of taps (even, odd) so the sums keep their sequential order; loads
  // of several pairs are in flight at once
  int k = 0;
#pragma unroll 4
  for (; k + 1 < W; k += 2) {
    const int64_t r0 = (int64_t)n + k, r1 = r0 + 1;
    const double v0 = (r0 >= 0 && r0 < n_rows) ? (double)col[r0 * M1] : 0.0;
    const double v1 = (r1 >= 0 && r1 < n_rows) ? (double)col[r1 * M1] : 0.0;
    e += v0 * T[k * D];
    o += v1 * T[(k + 1) * D];
  }
  if (k < W) {
    const int64_t r0 = (int64_t)n + k;
    const double v0 = (r0 >= 0 && r0 < n_rows) ? (double)col[r0 * M1] : 0.0;
    e += v0 * T[k * D];
  }
  se = e;
  so = o;
}

// band_sums for a fixed number of taps W: every tap's plane value and filter
// coefficient is loaded before any is summed (rows outside the planes at a
// safe index, then taken as zero), and the even and odd sums add them in the
// same order as band_sums, so the results are identical.  Out of line, so
// that the six sums of a thread run one after another, each with its loads in
// flight together, instead of all at once beyond the register file.
template <int W>
__device__ __attribute__((noinline)) void band_sums_w(const float* __restrict__ plane,
                                                      int64_t n_rows, int M1, int n, int m,
                                                      const double* __restrict__ T, int64_t D,
                                                      double& se, double& so) {
  const float* col = plane + m;
  float pv[W];
  double tv[W];
#pragma unroll
  for (int k = 0; k < W; ++k) {
    const int64_t r = (int64_t)n + k;
    const bool in = r >= 0 && r < n_rows;
    pv[k] = col[(in ? r : 0) * M1];
    pv[k] = in ? pv[k] : 0.f;
    tv[k] = T[k * D];
  }
  double e = 0.0, o = 0.0;
#pragma unroll
  for (int k = 0; k < W; ++k) {
    if (k & 1) o += (double)pv[k] * tv[k];
    else e += (double)pv[k] * tv[k];
  }
  se = e;
  so = o;
}

template <int WT>
__device__ __forceinline__ void layer_sums(const float* __restrict__ plane, int64_t n_rows,
                                           int M1, int n, int m, const double* __restrict__ T,
                                           int64_t D, int W, double& se, double& so) {
  if constexpr (WT > 0) {
    band_sums_w<WT>(plane, n_rows, M1, n, m, T, D, se, so);
  } else {
    band_sums(plane, n_rows, M1, n, m, T, D, W, se, so);
  }
}

// sin and cos of q pi / (2J) for q in [0, 4J): every phase of the filter sums
// is such a multiple (m pi dt / M = 2 m r pi / 2J, (2m -+ 1) dt pi / 2M =
// (2m -+ 1) r pi / 2J with dt = r M / J), as WDM<T>::setTDFilter tabulates
// sinTD / cosTD (cwb-core/WDM.cc:688-699).
__device__ __forceinline__ double2 phase(const double2* __restrict__ trig, int J4, int q) {
  int r = q % J4;
  r += r < 0 ? J4 : 0;
  return trig[r];  // (sin, cos)
}

template <int WT>
__device__ __forceinline__ double pixel_amp(const TdArgs& a, int n, int m, int dT,
                                            const float* __restrict__ plane, bool quad) {
  const int M = a.M, M1 = M + 1, W = 2 * a.nc + 1;
  const bool odd_n = n & 1, odd_mn = (m + n) & 1;
  const bool cancel_even = odd_n ^ quad;
  const bool edge = m == 0 || m == M;
  const int64_t D = 2 * (int64_t)a.J + 1;
  const double* t0 = a.T0 + (dT + a.J);
  const double* tx = a.Tx + (dT + a.J);
  double se, so;
  layer_sums<WT>(plane, a.n_rows, M1, n, m, t0, D, W, se, so);
  if (edge && cancel_even) se = 0.0;
  if (edge && !cancel_even) so = 0.0;
  const int J4 = 4 * a.J;
  const double2 ph = phase(a.trig, J4, 2 * m * dT);
  double r = odd_n ? se * ph.y - so * ph.x : se * ph.y + so * ph.x;
  if (m > 0) {
    double le, lo;
    layer_sums<WT>(plane, a.n_rows, M1, n, m - 1, tx, D, W, le, lo);
    if (m == 1 && cancel_even) le = 0.0;
    if (m == 1 && !cancel_even) lo = 0.0;
    const double slp = phase(a.trig, J4, (2 * m - 1) * dT).x;
    double t = odd_n ? (le - lo) * slp : (le + lo) * slp;
    if (m == 1 || m == M) t *= 1.4142135623730951;
    r = odd_mn ? r - t : r + t;
  }
  if (m < M) {
    double he, ho;
    layer_sums<WT>(plane, a.n_rows, M1, n, m + 1, tx, D, W, he, ho);
    if (m == M - 1 && cancel_even) he = 0.0;
    if (m == M - 1 && !cancel_even) ho = 0.0;
    const double shp = phase(a.trig, J4, (2 * m + 1) * dT).x;
    double t = odd_n ? (he + ho) * shp : (he - ho) * shp;
    if (m == 0 || m == M - 1) t *= 1.4142135623730951;
    r = odd_mn ? r - t : r + t;
  }
  if (edge && cancel_even) return 0.0;
  return r;
}

// WT: the filters' tap count 2 n_c + 1 when it is a compile-time case (25),
// else 0 (runtime loop).
template <int WT>
__global__ __launch_bounds__(256) void k_td_vectors(TdArgs a) {
  const int half = 2 * a.K + 1;
  const int64_t total = a.n_pixels * half;
  for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * 256) {
    const int64_t p = e / half;
    const int ki = (int)(e - p * half);
    const int idx = a.pix[p];
    const int n = idx / (a.M + 1), m = idx - n * (a.M + 1);
    const int dT = ki - a.K;
    const int sh = dT >= 0 ? dT / a.J : -((-dT) / a.J);
    const int sub = dT - sh * a.J, ne = n - sh;
    // odd whole-bin shifts swap the quadratures (getTDamp): a00 from the 90
    // plane, a90 from the 00 plane, one of them negated by the (n + m) parity
    const bool swap = (sh & 1) != 0;
    double a00 = pixel_amp<WT>(a, ne, m, sub, swap ? a.p90 : a.p00, swap);
    double a90 = pixel_amp<WT>(a, ne, m, sub, swap ? a.p00 : a.p90, !swap);
    if (swap) {
      if ((n + m) & 1) a00 = -a00;
      else a90 = -a90;
    }
    float* o = a.out + p * a.row_stride;
    if (a.interleave) {
      o[2 * ki] = (float)a00;
      o[2 * ki + 1] = (float)a90;
    } else {
      o[ki] = (float)a00;
      o[half + ki] = (float)a90;
    }
  }
}

// Pixel-major form for 25-tap filters: the 2K + 1 delays split into chunks of
// 192, and block b serves chunk b mod n_chunks over its own pixels, thread t
// owning delay 192 c + t of every pixel.  Its taps (T0 and Tx at its sub-step,
// 25 each) stay in registers; the block stages each pixel's plane values once,
// as doubles in LDS: layers m - 1, m, m + 1 of both planes over the rows
// n - s_hi .. n - s_lo + 24 that the chunk's whole-bin shifts reach (zero
// outside the planes and the layers), and loads the next pixel's values into
// registers while the current one is computed.  Every amplitude is the same sequence of operations as pixel_amp's
// (even and odd sums in ascending tap order, the same phases, edges and
// quadrature swap), so the tables are identical.
constexpr int kPixThreads = 192;
constexpr int kPixTaps = 25;
constexpr int kPixStage = 8;   // staged values per thread and pixel (6 NR <= 1536)

struct PixTaps {
  double t0[kPixTaps], tx[kPixTaps];
};

// v is 16-byte aligned: taps 2j and 2j + 1 are read as one pair (the even
// one to e, the odd one to o, in ascending order as before)
__device__ __forceinline__ void pix_sums(const double* __restrict__ v, const double* tap,
                                         double& se, double& so) {
  // one sum's LDS reads at a time: the scheduler may not hoist the next sum's
  // loads into this one (all six at once need more than the register file)
  __builtin_amdgcn_sched_barrier(0);
  const double2* v2 = reinterpret_cast<const double2*>(v);
  double e = 0.0, o = 0.0;
#pragma unroll
  for (int j = 0; j < kPixTaps / 2; ++j) {
    const double2 x = v2[j];
    e += x.x * tap[2 * j];
    o += x.y * tap[2 * j + 1];
  }
  e += v[kPixTaps - 1] * tap[kPixTaps - 1];
  se = e;
  so = o;
}

// pixel_amp with the layer values at v(L)[k] = plane row ne + k of layer m - 1 + L
__device__ __forceinline__ double pix_amp(const TdArgs& a, int n, int m, int dT,
                                          const double* __restrict__ vm1,
                                          const double* __restrict__ v0,
                                          const double* __restrict__ vp1, const PixTaps& T,
                                          bool quad) {
  const int M = a.M;
  const bool odd_n = n & 1, odd_mn = (m + n) & 1;
  const bool cancel_even = odd_n ^ quad;
  const bool edge = m == 0 || m == M;
  double se, so;
  pix_sums(v0, T.t0, se, so);
  if (edge && cancel_even) se = 0.0;
  if (edge && !cancel_even) so = 0.0;
  const int J4 = 4 * a.J;
  const double2 ph = phase(a.trig, J4, 2 * m * dT);
  double r = odd_n ? se * ph.y - so * ph.x : se * ph.y + so * ph.x;
  if (m > 0) {
    double le, lo;
    pix_sums(vm1, T.tx, le, lo);
    if (m == 1 && cancel_even) le = 0.0;
    if (m == 1 && !cancel_even) lo = 0.0;
    const double slp = phase(a.trig, J4, (2 * m - 1) * dT).x;
    double t = odd_n ? (le - lo) * slp : (le + lo) * slp;
    if (m == 1 || m == M) t *= 1.4142135623730951;
    r = odd_mn ? r - t : r + t;
  }
  if (m < M) {
    double he, ho;
    pix_sums(vp1, T.tx, he, ho);
    if (m == M - 1 && cancel_even) he = 0.0;
    if (m == M - 1 && !cancel_even) ho = 0.0;
    const double shp = phase(a.trig, J4, (2 * m + 1) * dT).x;
    double t = odd_n ? (he + ho) * shp : (he - ho) * shp;
    if (m == 0 || m == M - 1) t *= 1.4142135623730951;
    r = odd_mn ? r - t : r + t;
  }
  if (edge && cancel_even) return 0.0;
  return r;
}

__device__ __forceinline__ int trunc_div(int x, int J) { return x >= 0 ? x / J : -((-x) / J); }

// whole-bin shifts of the delays of chunk c: [s_lo, s_hi]
__host__ __device__ inline int td_chunk_rows(int c, int K, int J, int& s_lo, int& s_hi) {
  const int half = 2 * K + 1;
  const int lo = c * kPixThreads, hi = (lo + kPixThreads < half ? lo + kPixThreads : half) - 1;
  auto tdiv = [](int x, int j) { return x >= 0 ? x / j : -((-x) / j); };
  s_lo = tdiv(lo - K, J);
  s_hi = tdiv(hi - K, J);
  return kPixTaps + s_hi - s_lo;
}

__global__ __launch_bounds__(kPixThreads) void k_td_pixels(TdArgs a, int n_chunks, int NS2) {
  // two buffers of NS2 doubles: (copy, plane, layer, row), rows of each slice
  // NRp = NR rounded up to even; copy 1 starts one row later than copy 0, so
  // every lane's 25 rows begin at an even (16-byte aligned) slice position
  extern __shared__ double2 stage2[];
  double* stage = reinterpret_cast<double*>(stage2);
  const int c = (int)(blockIdx.x % (unsigned)n_chunks);
  const int64_t g = blockIdx.x / (unsigned)n_chunks;
  const int64_t G = gridDim.x / (unsigned)n_chunks;
  int s_min, s_max;
  const int NR = td_chunk_rows(c, a.K, a.J, s_min, s_max);
  const int NRp = (NR + 1) & ~1;
  const int NS = 12 * NRp;
  const int half = 2 * a.K + 1;
  const int M1 = a.M + 1;
  const int ki = c * kPixThreads + threadIdx.x;
  const bool active = ki < half;
  const int dT = (active ? ki : 0) - a.K;
  const int sh = trunc_div(dT, a.J);
  const int sub = dT - sh * a.J;
  const int off = s_max - sh;   // row ne + k of the lane is staged row off + k
  // copy and aligned slice position of the lane's first row
  const int lane_base = (off & 1) ? 6 * NRp + off - 1 : off;
  PixTaps T;
  {
    const int64_t D = 2 * (int64_t)a.J + 1;
#pragma unroll
    for (int k = 0; k < kPixTaps; ++k) {
      T.t0[k] = a.T0[k * D + sub + a.J];
      T.tx[k] = a.Tx[k * D + sub + a.J];
    }
  }
  // staged value i: copy i / (6 NRp), then plane, layer m - 1 + L, and row
  // n - s_max + r + copy
  auto fetch = [&](int64_t p, float (&v)[kPixStage]) {
    const int idx = a.pix[p];
    const int n = idx / M1, m = idx - n * M1;
#pragma unroll
    for (int j = 0; j < kPixStage; ++j) {
      const int i = threadIdx.x + j * kPixThreads;
      float x = 0.f;
      if (i < NS) {
        const int cp = i / (6 * NRp), rem0 = i - cp * 6 * NRp;
        const int pl = rem0 / (3 * NRp), rem = rem0 - pl * 3 * NRp;
        const int layer = m - 1 + rem / NRp;
        const int64_t row = (int64_t)n - s_max + rem % NRp + cp;
        if (layer >= 0 && layer <= a.M && row >= 0 && row < a.n_rows)
          x = (pl ? a.p90 : a.p00)[row * M1 + layer];
      }
      v[j] = x;
    }
  };
  float nx[kPixStage];
  int64_t p = g;
  if (p < a.n_pixels) fetch(p, nx);
  int b = 0;
  for (; p < a.n_pixels; p += G, b ^= 1) {
    double* sb = stage + b * NS2;
#pragma unroll
    for (int j = 0; j < kPixStage; ++j) {
      const int i = threadIdx.x + j * kPixThreads;
      if (i < NS) sb[i] = (double)nx[j];
    }
    __syncthreads();   // this pixel staged; the other buffer is free again
    const int idx = a.pix[p];
    const int n = idx / M1, m = idx - n * M1;
    if (p + G < a.n_pixels) fetch(p + G, nx);
    if (active) {
      const int ne = n - sh;
      // odd whole-bin shifts swap the quadratures (getTDamp): a00 from the 90
      // plane, a90 from the 00 plane, one of them negated by the (n + m) parity
      const bool swap = (sh & 1) != 0;
      const double* q00 = sb + lane_base;             // plane 00, layer m - 1
      const double* q90 = sb + lane_base + 3 * NRp;   // plane 90, layer m - 1
      const double* A = swap ? q90 : q00;
      const double* Bq = swap ? q00 : q90;
      double a00 = pix_amp(a, ne, m, sub, A, A + NRp, A + 2 * NRp, T, swap);
      double a90 = pix_amp(a, ne, m, sub, Bq, Bq + NRp, Bq + 2 * NRp, T, !swap);
      if (swap) {
        if ((n + m) & 1) a00 = -a00;
        else a90 = -a90;
      }
      float* o = a.out + p * a.row_stride;
      if (a.interleave) {
        o[2 * ki] = (float)a00;
        o[2 * ki + 1] = (float)a90;
      } else {
        o[ki] = (float)a00;
        o[half + ki] = (float)a90;
      }
    }
  }
}

}  // namespace

extern "C" {

int pycwb_td_vectors(const int32_t* pixel_index, int64_t n_pixels, const float* padded00,
                     const float* padded90, int64_t n_rows, int M, const double* T0,
                     const double* Tx, const double* trig, int n_coeffs, int J, int K, float* out,
                     int64_t row_stride, int interleave, void* stream) {
  if (n_pixels < 0 || M < 1 || n_coeffs < 0 || J < 1 || K < 0 || n_rows < 1) return PYCWB_ERR_ARG;
  if (n_pixels == 0) return PYCWB_OK;
  if (!pixel_index || !padded00 || !padded90 || !T0 || !Tx || !trig || !out) return PYCWB_ERR_ARG;
  if (K > 64 * J || row_stride < (int64_t)(4 * K + 2)) return PYCWB_ERR_ARG;
  (void)hipGetLastError();
  TdArgs a;
  a.pix = pixel_index; a.n_pixels = n_pixels; a.p00 = padded00; a.p90 = padded90;
  a.n_rows = n_rows; a.M = M; a.T0 = T0; a.Tx = Tx; a.trig = (const double2*)trig;
  a.nc = n_coeffs; a.J = J; a.K = K;
  a.out = out; a.row_stride = row_stride; a.interleave = interleave ? 1 : 0;
  const int n_chunks = (2 * K + 1 + kPixThreads - 1) / kPixThreads;
  int NRmax = 0;
  for (int c = 0; c < n_chunks; ++c) {
    int lo, hi;
    const int nr = td_chunk_rows(c, K, J, lo, hi);
    NRmax = nr > NRmax ? nr : NRmax;
  }
  const int NRpmax = (NRmax + 1) & ~1;
  if (2 * n_coeffs + 1 == kPixTaps && 12 * NRpmax <= kPixStage * kPixThreads) {
    int64_t G = 8192 / n_chunks;
    G = G < 1 ? 1 : (G > n_pixels ? n_pixels : G);
    const int64_t grid = G * n_chunks;
    if (grid > 0x7fffffff) return PYCWB_ERR_ARG;
    hipLaunchKernelGGL(k_td_pixels, dim3((unsigned)grid), dim3(kPixThreads),
                       (size_t)2 * 12 * NRpmax * sizeof(double), (hipStream_t)stream, a,
                       n_chunks, 12 * NRpmax);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
      pycwb::record_error("k_td_pixels", hipGetErrorString(e));
      return PYCWB_ERR_HIP;
    }
    return PYCWB_OK;
  }
  const int64_t total = n_pixels * (2 * K + 1);
  int64_t grid = (total + 255) / 256;
  if (grid > 65536) grid = 65536;
  if (2 * n_coeffs + 1 == 25) {
    hipLaunchKernelGGL((k_td_vectors<25>), dim3((unsigned)grid), dim3(256), 0,
                       (hipStream_t)stream, a);
  } else {
    hipLaunchKernelGGL((k_td_vectors<0>), dim3((unsigned)grid), dim3(256), 0,
                       (hipStream_t)stream, a);
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    pycwb::record_error("k_td_vectors", hipGetErrorString(e));
    return PYCWB_ERR_HIP;
  }
  return PYCWB_OK;
}

}  // extern "C"
