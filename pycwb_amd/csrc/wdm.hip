// WDM forward transform t2w on MI355X, float64.
//
// Restates WDM<T>::t2w (cwb-core/WDM.cc:1385-1534) for a batch of series:
//   mirror padding (left edge reflected without repeating sample 0, right edge
//   reflected with repetition; WDM.cc:1407-1411, tf_batch_generation.py:37-61),
//   per time bin n (stride MM) the symmetric window fold of 2*m_H - 1 samples
//   into 2M bins, a length-2M FFT (radix-2, in LDS), the 1/sqrt(2) edge layers
//   and the parity-alternating 00/90 assignment times sqrt(2).
// One workgroup transforms B consecutive time bins of one series; the folds
// read the series straight from HBM (neighbouring bins overlap, so the
// re-reads hit L2) and the FFTs run on LDS.
// Output layout: (series, N, M+1) complex (map00, map90) pairs, bin-major so
// every store is coalesced; with quadrature off, the power map in map00.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "skyscan_lane.hh"
#include "pycwb_amd.h"

namespace {

constexpr int kThreads = 256;
constexpr int kLdsDoubles = 4096;  // 32 KiB of complex scratch (2048 complex)

__device__ __forceinline__ double ext_sample(const double* __restrict__ x, int64_t L, int64_t k) {
  // k: index relative to the first input sample (may be negative / >= L)
  if (k < 0) k = -k;                  // ts[m_H - n] = x[n]
  else if (k >= L) k = 2 * L - 1 - k; // ts[m_H + L + n] = x[L - 1 - n]
  return (k >= 0 && k < L) ? x[k] : 0.0;
}

__device__ __forceinline__ unsigned bitrev(unsigned v, int bits) {
  return __brev(v) >> (32 - bits);
}

__global__ __launch_bounds__(kThreads) void k_t2w(
    const double* __restrict__ X, int64_t L, int64_t ldx, int n_series, int M, int stride,
    int64_t N, const double* __restrict__ filt, int m_H, const double2* __restrict__ tw,
    int log2_2M, int B, int quadrature, double2* __restrict__ out) {
  __shared__ double2 buf[kLdsDoubles / 2];
  const int M2 = 2 * M;
  const int64_t tiles = (N + B - 1) / B;
  const int64_t blk = blockIdx.x;
  const int s = (int)(blk / tiles);
  const int64_t n0 = (blk % tiles) * B;
  if (s >= n_series) return;
  const double* x = X + (int64_t)s * ldx;
  const int nb = (int)((N - n0) < B ? (N - n0) : B);
  // fold: thread (b, k) sums the window taps t = k + 2M*j over |t| <= m_H - 1
  for (int e = threadIdx.x; e < nb * M2; e += kThreads) {
    const int b = e / M2, k = e % M2;
    const int64_t base = (n0 + b) * stride;
    double acc = 0.0;
    // t = k - 2M*q (q >= 1) on the negative side, t = k + 2M*q (q >= 0) on the positive
    for (int t = k; t <= m_H - 1; t += M2) acc += ext_sample(x, L, base + t) * filt[t];
    for (int t = k - M2; t >= -(m_H - 1); t -= M2) acc += ext_sample(x, L, base + t) * filt[-t];
    buf[b * M2 + bitrev((unsigned)k, log2_2M)] = make_double2(acc, 0.0);
  }
  __syncthreads();
  // radix-2 DIT FFT of length 2M on each bin, exp(-2 pi i f k / 2M)
  for (int h = 1; h < M2; h <<= 1) {
    const int tstride = M2 / (2 * h);
    for (int e = threadIdx.x; e < nb * M; e += kThreads) {
      const int b = e / M, r = e % M;
      const int g = r / h, j = r % h;
      const int i0 = b * M2 + g * 2 * h + j, i1 = i0 + h;
      const double2 w = tw[j * tstride];
      const double2 u = buf[i0], v = buf[i1];
      const double vr = v.x * w.x - v.y * w.y, vi = v.x * w.y + v.y * w.x;
      buf[i0] = make_double2(u.x + vr, u.y + vi);
      buf[i1] = make_double2(u.x - vr, u.y - vi);
    }
    __syncthreads();
  }
  // edges and 00/90 assignment (WDM.cc:1494-1510)
  const double sqrt2 = 1.4142135623730951;
  for (int e = threadIdx.x; e < nb * (M + 1); e += kThreads) {
    const int b = e / (M + 1), f = e % (M + 1);
    const int64_t n = n0 + b;
    double R = buf[b * M2 + f].x, I = buf[b * M2 + f].y;
    if (f == 0 || f == M) {
      R = R / sqrt2;
      I = R;
    }
    double2 o;
    if (!quadrature) {
      o = make_double2(R * R + I * I, 0.0);
    } else if ((f + (int)(n & 1)) & 1) {
      o = make_double2(sqrt2 * I, sqrt2 * R);
    } else {
      o = make_double2(sqrt2 * R, -sqrt2 * I);
    }
    out[((int64_t)s * N + n) * (M + 1) + f] = o;
  }
}

}  // namespace

extern "C" {

int pycwb_wdm_t2w(const double* x, int64_t n_samples, int64_t ldx, int n_series, int M,
                  int stride, const double* filter, int m_H, const double* twiddle,
                  int quadrature, double* out, void* stream) {
  if (!x || !filter || !twiddle || !out) return PYCWB_ERR_ARG;
  if (n_series < 1 || M < 1 || (M & (M - 1)) || 2 * M > kLdsDoubles / 2 || stride < 1 ||
      m_H < 1 || n_samples < 1 || ldx < n_samples || n_samples <= m_H)
    return PYCWB_ERR_ARG;
  (void)hipGetLastError();
  const int64_t padded = n_samples + (n_samples % stride ? stride - n_samples % stride : 0);
  const int64_t N = padded / stride;
  int log2_2M = 0;
  while ((1 << log2_2M) < 2 * M) ++log2_2M;
  int B = (kLdsDoubles / 2) / (2 * M);
  if (B > 64) B = 64;
  if (B < 1) B = 1;
  const int64_t tiles = (N + B - 1) / B;
  const int64_t grid = tiles * n_series;
  if (grid > 0x7fffffff) return PYCWB_ERR_ARG;
  hipLaunchKernelGGL(k_t2w, dim3((unsigned)grid), dim3(kThreads), 0, (hipStream_t)stream, x,
                     n_samples, ldx, n_series, M, stride, N, filter, m_H,
                     (const double2*)twiddle, log2_2M, B, quadrature, (double2*)out);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    pycwb::record_error("k_t2w", hipGetErrorString(e));
    return PYCWB_ERR_HIP;
  }
  return PYCWB_OK;
}

}  // extern "C"
