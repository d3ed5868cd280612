// WDM forward transform t2w on MI355X, float64.
//
// Restates WDM<T>::t2w (cwb-core/WDM.cc:1385-1534) for a batch of series:
//   mirror padding (left edge reflected without repeating sample 0, right edge
//   reflected with repetition; WDM.cc:1407-1411, tf_batch_generation.py:37-61),
//   per time bin n (stride MM) the symmetric window fold of 2*m_H - 1 samples
//   into 2M bins, a length-2M FFT (radix-2, in LDS), the 1/sqrt(2) edge layers
//   and the parity-alternating 00/90 assignment times sqrt(2).
// The stride == M levels run polyphase kernels: k_t2w_stream (M <= 128 and
// M = 512) keeps each block's input window in an LDS ring and streams
// consecutive tiles of one series through it; k_t2w_poly_h (the other M) reads
// each thread's window straight from HBM (neighbouring bins overlap, so the
// re-reads hit L2).  The FFTs run on LDS.
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

// Sample k of the mirror-padded input of a shifted copy of x: the padding
// index is resolved first, then the copy's shift sh (sh > 0: xs[i] = x[i + sh]
// for i < L - sh, sh < 0: xs[i] = x[i + sh] for i >= -sh, the other end
// keeping the original samples), so the transform reads the shifted series
// without it being written out.
__device__ __forceinline__ double ext_sample_s(const double* __restrict__ x, int64_t L, int64_t k,
                                               int64_t sh) {
  if (k < 0) k = -k;
  else if (k >= L) k = 2 * L - 1 - k;
  if (!(k >= 0 && k < L)) return 0.0;
  if (sh > 0 && k < L - sh) k += sh;
  else if (sh < 0 && k >= -sh) k += sh;
  return x[k];
}

// Sample of a shifted copy (SH) or of the series itself: without shifts the
// kernels compile the plain mirror-padding read (no per-sample shift tests).
template <bool SH>
__device__ __forceinline__ double ext_sample_t(const double* __restrict__ x, int64_t L, int64_t k,
                                               int64_t sh) {
  if constexpr (SH) return ext_sample_s(x, L, k, sh);
  else return ext_sample(x, L, k);
}

__device__ __forceinline__ unsigned bitrev(unsigned v, int bits) {
  return __brev(v) >> (32 - bits);
}

// LDS position of complex element i of an FFT buffer in the swizzled layout:
// bits 0-3 of i are XORed with bits 4-5 (once into bits 0-1, once into bits
// 2-3).  Within any aligned block of 64 elements this is a bijection, so a
// buffer keeps its size.  It spreads the radix-4 passes' strided accesses
// (elements 4l + m, and groups of four 16 apart) over distinct 16-byte bank
// groups, and leaves runs of consecutive elements consecutive within each
// aligned group of 16.
__device__ __forceinline__ int swz(int i) { return i ^ (((i >> 4) & 3) * 5); }

// In-place radix-2 DIT FFT, exp(-2 pi i f k / 2M), of `count` sequences of
// length 2M = 2^lg stored back to back in LDS (input in bit-reversed order).
// Stages run in pairs as radix-4 butterflies (one barrier per pair), the last
// stage alone when lg is odd; power-of-two index arithmetic throughout.
// tw[k] = exp(-2 pi i k / 2M), k < M.
__device__ __forceinline__ void lds_fft(double2* __restrict__ buf, int count, int lg,
                                        const double2* __restrict__ tw) {
  const int M2 = 1 << lg;
  const int qpl = lg - 2;  // log2 of the radix-4 butterflies per sequence
  int lh = 0;
  for (; lh + 1 < lg; lh += 2) {
    const int h = 1 << lh;
    const int n4 = count << qpl;
    for (int e = threadIdx.x; e < n4; e += blockDim.x) {
      const int p = e >> qpl, r = e & ((1 << qpl) - 1);
      const int g = r >> lh, j = r & (h - 1);
      const int i0 = (p << lg) + (g << (lh + 2)) + j;
      const double2 w1 = tw[j << (lg - 1 - lh)];
      const double2 w2a = tw[j << (lg - 2 - lh)], w2b = tw[(j + h) << (lg - 2 - lh)];
      const int p0 = i0, p1 = i0 + h, p2 = i0 + 2 * h, p3 = i0 + 3 * h;
      const double2 x0 = buf[p0], x1 = buf[p1], x2 = buf[p2], x3 = buf[p3];
      // stage h: (x0, x1) and (x2, x3) with w1
      double vr = x1.x * w1.x - x1.y * w1.y, vi = x1.x * w1.y + x1.y * w1.x;
      const double2 a0 = make_double2(x0.x + vr, x0.y + vi), a1 = make_double2(x0.x - vr, x0.y - vi);
      vr = x3.x * w1.x - x3.y * w1.y;
      vi = x3.x * w1.y + x3.y * w1.x;
      const double2 a2 = make_double2(x2.x + vr, x2.y + vi), a3 = make_double2(x2.x - vr, x2.y - vi);
      // stage 2h: (a0, a2) with w2a, (a1, a3) with w2b
      vr = a2.x * w2a.x - a2.y * w2a.y;
      vi = a2.x * w2a.y + a2.y * w2a.x;
      buf[p0] = make_double2(a0.x + vr, a0.y + vi);
      buf[p2] = make_double2(a0.x - vr, a0.y - vi);
      vr = a3.x * w2b.x - a3.y * w2b.y;
      vi = a3.x * w2b.y + a3.y * w2b.x;
      buf[p1] = make_double2(a1.x + vr, a1.y + vi);
      buf[p3] = make_double2(a1.x - vr, a1.y - vi);
    }
    __syncthreads();
  }
  if (lh < lg) {  // the last stage, h = M2 / 2
    const int h = 1 << lh;
    const int n2 = count << (lg - 1);
    for (int e = threadIdx.x; e < n2; e += blockDim.x) {
      const int p = e >> (lg - 1), j = e & (h - 1);
      const int i0 = (p << lg) + j, i1 = i0 + h;
      const double2 w = tw[j];
      const double2 u = buf[i0], v = buf[i1];
      const double vr = v.x * w.x - v.y * w.y, vi = v.x * w.y + v.y * w.x;
      buf[i0] = make_double2(u.x + vr, u.y + vi);
      buf[i1] = make_double2(u.x - vr, u.y - vi);
    }
    __syncthreads();
  }
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
    const int b = e >> log2_2M, k = e & (M2 - 1);
    const int64_t base = (n0 + b) * stride;
    double acc = 0.0;
    // t = k - 2M*q (q >= 1) on the negative side, t = k + 2M*q (q >= 0) on the positive
    for (int t = k; t <= m_H - 1; t += M2) acc += ext_sample(x, L, base + t) * filt[t];
    for (int t = k - M2; t >= -(m_H - 1); t -= M2) acc += ext_sample(x, L, base + t) * filt[-t];
    buf[b * M2 + bitrev((unsigned)k, log2_2M)] = make_double2(acc, 0.0);
  }
  __syncthreads();
  // DIT FFT of length 2M on each bin, exp(-2 pi i f k / 2M)
  lds_fft(buf, nb, log2_2M, tw);
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

// Polyphase form for stride == M (the quadrature and power maps of t2w):
// bin n, residue k in [0, 2M) folds x[n M + k + 2M q] h[|k + 2M q|] over
// |k + 2M q| <= m_H - 1, i.e. the FIR sum_q h_k[q] y_k[n + 2q] of the
// decimated sequence y_k[j] = x[j M + k].  A thread owns one residue and R
// consecutive bins and slides an R-wide register window along y_k, so each
// sample is loaded about once per thread (consecutive threads read
// consecutive k: coalesced).  Two real bins share one complex FFT (bin 2p in
// the real part, 2p+1 in the imaginary part), split afterwards with
// X[f] = (Z[f] + conj Z[2M-f]) / 2, Y[f] = (Z[f] - conj Z[2M-f]) / 2i.
constexpr int kPolyLdsMax = 2048;  // complex doubles of LDS per block (32 KiB): B * M

// Bins per block: B * M complex doubles of LDS; measured on MI355X (5 x 2.46 M
// samples), 16 KiB blocks are fastest up to M = 128 and 32 KiB blocks above.
// Levels with m_H = 12 M + 1 and M <= 128 or M = 512 take the streaming
// kernel (k_t2w_stream) instead.

__host__ int poly_bins(int M) {
  const int budget = M <= 128 ? 1024 : kPolyLdsMax;
  const int B = budget / M;
  return B > 64 ? 64 : B;
}

template <int R>
__global__ __launch_bounds__(kThreads) void k_t2w_poly(
    const double* __restrict__ X, int64_t L, int64_t ldx, int n_series, int M, int64_t N,
    const double* __restrict__ filt, int m_H, const double2* __restrict__ tw, int log2_2M, int B,
    int quadrature, double2* __restrict__ out) {
  extern __shared__ double2 buf[];  // B * M complex (B / 2 bin pairs of length 2M)
  const int M2 = 2 * M;
  const int KK = (m_H - 1) / M2;
  const int64_t tiles = (N + B - 1) / B;
  const int64_t blk = blockIdx.x;
  const int s = (int)(blk / tiles);
  const int64_t n0 = (blk % tiles) * B;
  if (s >= n_series) return;
  const double* x = X + (int64_t)s * ldx;
  double* lds = reinterpret_cast<double*>(buf);
  // fold
  const int items = (B / R) * M2;
  for (int it = threadIdx.x; it < items; it += kThreads) {
    const int k = it & (M2 - 1), r = it >> log2_2M;
    const int64_t nb0 = n0 + (int64_t)r * R;
    double acc[R], win[R];
#pragma unroll
    for (int b = 0; b < R; ++b) {
      acc[b] = 0.0;
      win[b] = ext_sample(x, L, (nb0 + b - 2 * KK) * M + k);
    }
    for (int q = -KK; q <= KK; ++q) {
      const int t = k + M2 * q;
      const int at = t < 0 ? -t : t;
      const double h = at <= m_H - 1 ? filt[at] : 0.0;
#pragma unroll
      for (int b = 0; b < R; ++b) acc[b] += h * win[b];
#pragma unroll
      for (int b = 0; b + 2 < R; ++b) win[b] = win[b + 2];
      const int64_t jn = nb0 + R - 2 + 2 * (q + 1);
      win[R - 2] = ext_sample(x, L, jn * M + k);
      win[R - 1] = ext_sample(x, L, (jn + 1) * M + k);
    }
    const int kr = (int)bitrev((unsigned)k, log2_2M);
#pragma unroll
    for (int b = 0; b < R; ++b) {
      const int bb = r * R + b;
      lds[2 * ((bb >> 1) * M2 + kr) + (bb & 1)] = acc[b];
    }
  }
  __syncthreads();
  // DIT FFT of length 2M on each bin pair, exp(-2 pi i f k / 2M)
  lds_fft(buf, B / 2, log2_2M, tw);
  // split the pairs; edges and 00/90 assignment (WDM.cc:1494-1510)
  const double sqrt2 = 1.4142135623730951;
  for (int e = threadIdx.x; e < B * (M + 1); e += kThreads) {
    const int b = e / (M + 1), f = e % (M + 1);
    const int64_t n = n0 + b;
    if (n >= N) continue;
    const int p = b >> 1;
    const double2 z = buf[p * M2 + f];
    const double2 zc = buf[p * M2 + ((M2 - f) & (M2 - 1))];
    double R_, I_;
    if ((b & 1) == 0) {
      R_ = 0.5 * (z.x + zc.x);
      I_ = 0.5 * (z.y - zc.y);
    } else {
      R_ = 0.5 * (z.y + zc.y);
      I_ = -0.5 * (z.x - zc.x);
    }
    if (f == 0 || f == M) {
      R_ = R_ / sqrt2;
      I_ = R_;
    }
    double2 o;
    if (!quadrature) {
      o = make_double2(R_ * R_ + I_ * I_, 0.0);
    } else if ((f + (int)(n & 1)) & 1) {
      o = make_double2(sqrt2 * I_, sqrt2 * R_);
    } else {
      o = make_double2(sqrt2 * R_, -sqrt2 * I_);
    }
    out[((int64_t)s * N + n) * (M + 1) + f] = o;
  }
}

// The polyphase kernel with the fold's loads issued up front: thread (k, r)
// reads the R + 4 KK samples x[(nb0 - 2 KK + u) M + k] (u < R + 4 KK) and
// the 2 KK + 1 filter taps h[|k + 2 M q|] into registers first, then forms
// acc[b] = sum over q = -KK..KK (ascending, the order of k_t2w_poly) of
// h_q x[(nb0 + b + 2 q) M + k]: the same terms in the same order, so the
// maps are identical, with one memory latency per thread instead of one per
// step of the sliding window; the FFT reads its twiddles from LDS.  KK =
// (m_H - 1) / 2M is a template parameter (6 for the WDM filters of every level
// here).
template <int R, int KK, bool SH>
__global__ __launch_bounds__(kThreads) void k_t2w_poly_h(
    const double* __restrict__ X, int64_t L, int64_t ldx, int n_series, int M, int64_t N,
    const double* __restrict__ filt, int m_H, const double2* __restrict__ tw, int log2_2M, int B,
    int quadrature, double2* __restrict__ out, const int64_t* __restrict__ shifts) {
  extern __shared__ double2 buf[];  // B * M complex (B / 2 bin pairs of length 2M), M twiddles
  const int M2 = 2 * M;
  const int64_t tiles = (N + B - 1) / B;
  const int64_t blk = blockIdx.x;
  const int s = (int)(blk / tiles);
  const int64_t n0 = (blk % tiles) * B;
  if (s >= n_series) return;
  const double* x = X + (int64_t)s * ldx;
  const int64_t sh = SH ? shifts[s] : 0;
  double* lds = reinterpret_cast<double*>(buf);
  const int items = (B / R) * M2;
  for (int it = threadIdx.x; it < items; it += kThreads) {
    const int k = it & (M2 - 1), r = it >> log2_2M;
    const int64_t nb0 = n0 + (int64_t)r * R;
    double h[2 * KK + 1];
#pragma unroll
    for (int q = -KK; q <= KK; ++q) {
      const int t = k + M2 * q;
      const int at = t < 0 ? -t : t;
      h[q + KK] = at <= m_H - 1 ? filt[at] : 0.0;
    }
    double acc[R];
#pragma unroll
    for (int b = 0; b < R; ++b) acc[b] = 0.0;
    // sample u = b + 2(q + KK) of the window goes to every bin b it reaches as
    // it arrives (each bin sums in ascending q); at most 8 loads ahead
#pragma unroll
    for (int u = 0; u < R + 4 * KK; ++u) {
      if (u % 8 == 0) __builtin_amdgcn_sched_barrier(0);
      const double v = ext_sample_t<SH>(x, L, (nb0 - 2 * KK + u) * M + k, sh);
#pragma unroll
      for (int b = 0; b < R; ++b) {
        const int d = u - b;
        if (d >= 0 && !(d & 1) && d <= 4 * KK) acc[b] += h[d >> 1] * v;
      }
    }
    const int kr = (int)bitrev((unsigned)k, log2_2M);
#pragma unroll
    for (int b = 0; b < R; ++b) {
      const int bb = r * R + b;
      lds[2 * ((bb >> 1) * M2 + kr) + (bb & 1)] = acc[b];
    }
  }
  // the M twiddles into LDS after the fold's buffer: every FFT stage reads
  // them (0.157 against 0.167 ms at M = 64 with them in global memory)
  double2* tws = buf + (size_t)B * M;
  for (int i = threadIdx.x; i < M; i += kThreads) tws[i] = tw[i];
  __syncthreads();
  lds_fft(buf, B / 2, log2_2M, tws);
  // split the pairs; edges and 00/90 assignment (WDM.cc:1494-1510); e / (M + 1)
  // by a multiply-high (exact for e < 2^16, M + 1 <= 2^16)
  const double sqrt2 = 1.4142135623730951;
  const unsigned M1 = (unsigned)M + 1u;
  const unsigned magic = 0xffffffffu / M1 + 1u;
  for (int e = threadIdx.x; e < B * (M + 1); e += kThreads) {
    const int b = (int)__umulhi((unsigned)e, magic);
    const int f = e - b * (int)M1;
    const int64_t n = n0 + b;
    if (n >= N) continue;
    const int p = b >> 1;
    const double2 z = buf[p * M2 + f];
    const double2 zc = buf[p * M2 + ((M2 - f) & (M2 - 1))];
    double R_, I_;
    if ((b & 1) == 0) {
      R_ = 0.5 * (z.x + zc.x);
      I_ = 0.5 * (z.y - zc.y);
    } else {
      R_ = 0.5 * (z.y + zc.y);
      I_ = -0.5 * (z.x - zc.x);
    }
    if (f == 0 || f == M) {
      R_ = R_ / sqrt2;
      I_ = R_;
    }
    double2 o;
    if (!quadrature) {
      o = make_double2(R_ * R_ + I_ * I_, 0.0);
    } else if ((f + (int)(n & 1)) & 1) {
      o = make_double2(sqrt2 * I_, sqrt2 * R_);
    } else {
      o = make_double2(sqrt2 * R_, -sqrt2 * I_);
    }
    out[((int64_t)s * N + n) * (M + 1) + f] = o;
  }
}

// Streaming form of the polyphase kernel (M = 2^LG, KK = 6): a block of THR
// threads owns C consecutive tiles of B = THR R / 2M bins of one series.
// Thread (k, r) folds residue k of bins r R .. r R + R - 1 of every tile, so
// the block covers the tile's (B / R) x 2M items once.  The input window of the
// current tile, rows n0 - 2 KK .. n0 + B + 2 KK of M samples, sits in an LDS
// ring of B + 4 KK + 1 rows: consecutive tiles share all but B rows, and the B
// rows the next tile adds (B M / THR samples per thread) are loaded into
// registers while the current tile is folded, transformed and stored, then
// written over the ring rows this tile no longer needs.  A thread's 2 KK + 1
// filter taps stay in registers for the whole chunk.  Two shapes:
//   M <= 128:  THR = 256, R = 8 (tiles of 1,024 samples); the FFT twiddles of a
//              thread's butterflies are kept in registers too (stream_fft) and
//              the FFT buffer is in the swizzled layout (swz);
//   M = 256, 512: THR = 1024, R = 8 at M = 256 (tiles of 16 bins, 152 KiB of
//              LDS) and R = 4 at M = 512 (tiles of 4 bins, 156 KiB), one block
//              per CU; the FFT is lds_fft with its twiddles in LDS, the buffer
//              plain (at M = 128 this shape measured slower than the 256-thread
//              one).
// With SH, series s is a shifted copy of one series (shifts, ldx 0), read
// through ext_sample_s; without, the plain mirror-padded read.
// The fold forms the same terms in the same order as k_t2w_poly_h (each window
// sample is added to every bin that uses it as soon as it is read, so each bin
// sums in ascending q), the FFT runs the same stages as lds_fft and the split
// computes the same expressions, so the maps are identical.  The split handles
// both bins of a pair from one read of Z[f] and Z[2M - f], layers 1 .. M - 1
// first and the two edge layers (the only ones with a division) in a loop of
// their own.
constexpr int kStreamMaxM = 128;      // the 256-thread shape
constexpr int kStreamWideM = 512;     // the 1024-thread shape
#ifndef PYCWB_T2W_STREAM_BLOCKS
#define PYCWB_T2W_STREAM_BLOCKS 1024   // 256-thread blocks to aim for: 4 per CU of an MI355X
                                       // (a quarter as many 1024-thread blocks)
#endif

__host__ size_t stream_lds_bytes(int M, int KK, int thr, int R) {
  const int B = thr * R / (2 * M);
  return (size_t)B * M * sizeof(double2) + (size_t)(B + 4 * KK + 1) * M * sizeof(double) +
         (thr == kThreads ? 0 : (size_t)M * sizeof(double2));
}

// The FFT of the streaming kernel: NP = 2^(9 - LG) sequences of 2M = 2^LG2
// points (LG2 = LG + 1), the same radix-2 DIT stages in the same pairing and
// order as lds_fft, with the index arithmetic fixed at compile time.  With 256
// threads every pass is exactly one radix-4 butterfly per thread (the last
// radix-2 stage, for odd LG2, two), and a thread's butterfly sits at the same
// place in every tile, so its twiddles are read once per block into registers
// (LDS then carries only the data: a quarter less traffic per pass).  The
// buffer is in the swizzled layout (swz).
template <int LG2>
struct StreamTw {
  static constexpr int NR4 = LG2 / 2;   // radix-4 passes
  double2 w[NR4 > 0 ? NR4 : 1][3];
  double2 wl;                            // the radix-2 stage's twiddle (odd LG2)
};

template <int LG2>
__device__ __forceinline__ void stream_tw_load(StreamTw<LG2>& T, const double2* __restrict__ tw) {
  const int e = threadIdx.x;
#pragma unroll
  for (int ps = 0; ps < StreamTw<LG2>::NR4; ++ps) {
    const int lh = 2 * ps;
    const int h = 1 << lh;
    const int j = (e & ((1 << (LG2 - 2)) - 1)) & (h - 1);
    T.w[ps][0] = tw[j << (LG2 - 1 - lh)];
    T.w[ps][1] = tw[j << (LG2 - 2 - lh)];
    T.w[ps][2] = tw[(j + h) << (LG2 - 2 - lh)];
  }
  if (LG2 & 1) T.wl = tw[e & ((1 << (LG2 - 1)) - 1)];   // 256 is a multiple of M
}

template <int LG2>
__device__ __forceinline__ void stream_fft(double2* __restrict__ buf, const StreamTw<LG2>& T) {
  const int e = threadIdx.x;
#pragma unroll
  for (int ps = 0; ps < StreamTw<LG2>::NR4; ++ps) {
    constexpr int qpl = LG2 - 2;
    const int lh = 2 * ps;
    const int h = 1 << lh;
    const int p = e >> qpl, r = e & ((1 << qpl) - 1);
    const int g = r >> lh, j = r & (h - 1);
    const int i0 = (p << LG2) + (g << (lh + 2)) + j;
    const double2 w1 = T.w[ps][0], w2a = T.w[ps][1], w2b = T.w[ps][2];
    const int p0 = swz(i0), p1 = swz(i0 + h), p2 = swz(i0 + 2 * h), p3 = swz(i0 + 3 * h);
    const double2 x0 = buf[p0], x1 = buf[p1], x2 = buf[p2], x3 = buf[p3];
    double vr = x1.x * w1.x - x1.y * w1.y, vi = x1.x * w1.y + x1.y * w1.x;
    const double2 a0 = make_double2(x0.x + vr, x0.y + vi), a1 = make_double2(x0.x - vr, x0.y - vi);
    vr = x3.x * w1.x - x3.y * w1.y;
    vi = x3.x * w1.y + x3.y * w1.x;
    const double2 a2 = make_double2(x2.x + vr, x2.y + vi), a3 = make_double2(x2.x - vr, x2.y - vi);
    vr = a2.x * w2a.x - a2.y * w2a.y;
    vi = a2.x * w2a.y + a2.y * w2a.x;
    buf[p0] = make_double2(a0.x + vr, a0.y + vi);
    buf[p2] = make_double2(a0.x - vr, a0.y - vi);
    vr = a3.x * w2b.x - a3.y * w2b.y;
    vi = a3.x * w2b.y + a3.y * w2b.x;
    buf[p1] = make_double2(a1.x + vr, a1.y + vi);
    buf[p3] = make_double2(a1.x - vr, a1.y - vi);
    __syncthreads();
  }
  if constexpr (LG2 & 1) {
    constexpr int h = 1 << (LG2 - 1);
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      const int ee = e + q * kThreads;
      const int p = ee >> (LG2 - 1), j = ee & (h - 1);
      const int i0 = swz((p << LG2) + j), i1 = swz((p << LG2) + j + h);
      const double2 w = T.wl;
      const double2 u = buf[i0], v = buf[i1];
      const double vr = v.x * w.x - v.y * w.y, vi = v.x * w.y + v.y * w.x;
      buf[i0] = make_double2(u.x + vr, u.y + vi);
      buf[i1] = make_double2(u.x - vr, u.y - vi);
    }
    __syncthreads();
  }
}

template <int KK, int LG, int R, int THR, bool SH>
__global__ __launch_bounds__(THR) void k_t2w_stream(
    const double* __restrict__ X, int64_t L, int64_t ldx, int n_series, int64_t N,
    const double* __restrict__ filt, int m_H, const double2* __restrict__ tw, int C,
    int quadrature, double2* __restrict__ out, const int64_t* __restrict__ shifts) {
  constexpr int M = 1 << LG, M2 = 2 * M, LG2 = LG + 1;
  constexpr int B = THR * R / M2;  // bins per tile
  constexpr int P = B * M / THR;   // new samples per thread and tile
  constexpr int WR = B + 4 * KK + 1;
  constexpr int NP = B / 2;        // bin pairs per tile
  constexpr bool kRegTw = THR == kThreads;
  static_assert(B >= R && B % 2 == 0 && P * THR == B * M, "tile shape");
  // FFT buffer position of element i: swizzled in the 256-thread shape (measured
  // faster there), plain in the 1024-thread one (measured slower swizzled)
  auto at = [](int i) { return kRegTw ? swz(i) : i; };
  extern __shared__ double2 buf[];  // B M complex (FFT), the window ring, M twiddles
  const int64_t tiles = (N + B - 1) / B;
  const int64_t chunks = (tiles + C - 1) / C;
  const int s = (int)(blockIdx.x / chunks);
  if (s >= n_series) return;
  const int64_t t0 = (blockIdx.x % chunks) * C;
  const int nt = (int)((tiles - t0) < C ? (tiles - t0) : C);
  const double* x = X + (int64_t)s * ldx;
  const int64_t sh = SH ? shifts[s] : 0;   // shifted copies of one series (ldx 0)
  double* lds = reinterpret_cast<double*>(buf);
  double* win = lds + 2 * B * M;
  double2* tws = reinterpret_cast<double2*>(win + WR * M);   // kRegTw false only
  const int64_t N0 = t0 * B;
  const int64_t base = (N0 - 2 * KK) * M;   // sample of ring row 0, column 0 at the start
  for (int i = threadIdx.x; i < WR * M; i += THR) win[i] = ext_sample_t<SH>(x, L, base + i, sh);
  if constexpr (!kRegTw) {
    for (int i = threadIdx.x; i < M; i += THR) tws[i] = tw[i];
  }
  const int k = threadIdx.x & (M2 - 1), r = threadIdx.x >> LG2;
  const int krow = k >> LG, kcol = k & (M - 1);
  const int kr = (int)bitrev((unsigned)k, LG2);
  double h[2 * KK + 1];
#pragma unroll
  for (int q = -KK; q <= KK; ++q) {
    const int t = k + M2 * q;
    const int at = t < 0 ? -t : t;
    h[q + KK] = at <= m_H - 1 ? filt[at] : 0.0;
  }
  const double sqrt2 = 1.4142135623730951;
  StreamTw<LG2> T;
  if constexpr (kRegTw) stream_tw_load<LG2>(T, tw);
  __syncthreads();
  for (int t = 0; t < nt; ++t) {
    const int64_t n0 = N0 + (int64_t)t * B;
    const bool more = t + 1 < nt;
    double pre[P];
    if (more) {
      // ring rows (t + 1) B + 4 KK + 1 .. + B - 1, relative to the chunk's first row
      const int64_t nb = base + ((int64_t)(t + 1) * B + 4 * KK + 1) * M;
#pragma unroll
      for (int p = 0; p < P; ++p)
        pre[p] = ext_sample_t<SH>(x, L, nb + threadIdx.x + p * THR, sh);
    }
    // fold: sample u of the thread's column (ring row t B + r R + u + krow,
    // wrapping at most once) adds h[(u - b) / 2] x to every bin b of the same
    // parity within reach
    double acc[R];
#pragma unroll
    for (int b = 0; b < R; ++b) acc[b] = 0.0;
    const int sl = (t * B + r * R + krow) % WR;
    const int lim = WR - sl;
    const double* wa = win + sl * M + kcol;
    const double* wb = wa - WR * M;
#pragma unroll
    for (int u = 0; u < R + 4 * KK; ++u) {
      const double v = (u < lim ? wa : wb)[u * M];
#pragma unroll
      for (int b = 0; b < R; ++b) {
        const int d = u - b;
        if (d >= 0 && !(d & 1) && d <= 4 * KK) acc[b] += h[d >> 1] * v;
      }
    }
#pragma unroll
    for (int b = 0; b < R; ++b) {
      const int bb = r * R + b;
      lds[2 * at((bb >> 1) * M2 + kr) + (bb & 1)] = acc[b];
    }
    __syncthreads();
    if constexpr (kRegTw) stream_fft<LG2>(buf, T);
    else lds_fft(buf, NP, LG2, tws);
    // split the pairs (bin 2p in the real part, 2p + 1 in the imaginary part);
    // edges and 00/90 assignment (WDM.cc:1494-1510).  n0 is even, so bin 2p
    // has even n and bin 2p + 1 odd n.
    double2* ot = out + ((int64_t)s * N + n0) * (M + 1);
    const int nb_valid = (int)((N - n0) < B ? (N - n0) : B);
    for (int e = threadIdx.x; e < NP * (M - 1); e += THR) {
      const int p = e / (M - 1);
      const int f = e - p * (M - 1) + 1;
      const double2 z = buf[at(p * M2 + f)];
      const double2 zc = buf[at(p * M2 + M2 - f)];
      const double R0 = 0.5 * (z.x + zc.x), I0 = 0.5 * (z.y - zc.y);
      const double R1 = 0.5 * (z.y + zc.y), I1 = -0.5 * (z.x - zc.x);
      double2 oa, ob;
      if (!quadrature) {
        oa = make_double2(R0 * R0 + I0 * I0, 0.0);
        ob = make_double2(R1 * R1 + I1 * I1, 0.0);
      } else if (f & 1) {
        oa = make_double2(sqrt2 * I0, sqrt2 * R0);
        ob = make_double2(sqrt2 * R1, -sqrt2 * I1);
      } else {
        oa = make_double2(sqrt2 * R0, -sqrt2 * I0);
        ob = make_double2(sqrt2 * I1, sqrt2 * R1);
      }
      if (2 * p < nb_valid) ot[(2 * p) * (M + 1) + f] = oa;
      if (2 * p + 1 < nb_valid) ot[(2 * p + 1) * (M + 1) + f] = ob;
    }
    for (int e = threadIdx.x; e < NP * 2; e += THR) {
      const int p = e >> 1, f = (e & 1) ? M : 0;
      const double2 z = buf[at(p * M2 + f)];   // Z[2M - f] is Z[f] for f = 0, M
      double R0 = 0.5 * (z.x + z.x), R1 = 0.5 * (z.y + z.y);
      R0 = R0 / sqrt2;
      R1 = R1 / sqrt2;
      const double I0 = R0, I1 = R1;
      double2 oa, ob;
      if (!quadrature) {
        oa = make_double2(R0 * R0 + I0 * I0, 0.0);
        ob = make_double2(R1 * R1 + I1 * I1, 0.0);
      } else {
        oa = make_double2(sqrt2 * R0, -sqrt2 * I0);
        ob = make_double2(sqrt2 * I1, sqrt2 * R1);
      }
      if (2 * p < nb_valid) ot[(2 * p) * (M + 1) + f] = oa;
      if (2 * p + 1 < nb_valid) ot[(2 * p + 1) * (M + 1) + f] = ob;
    }
    if (more) {
      // the new rows replace ring rows t B .. t B + B - 1, which only this
      // tile's fold read (finished before the FFT's first barrier)
      const int rel0 = ((t + 1) * B + 4 * KK + 1) % WR;
#pragma unroll
      for (int p = 0; p < P; ++p) {
        const int i = threadIdx.x + p * THR;
        int row = rel0 + (i >> LG);
        row = row >= WR ? row - WR : row;
        win[row * M + (i & (M - 1))] = pre[p];
      }
    }
    __syncthreads();
  }
}

template <int LG, int R, int THR>
void launch_stream(unsigned grid, hipStream_t st, const double* x, int64_t L, int64_t ldx,
                   int n_series, int64_t N, const double* filt, int m_H, const double2* tw,
                   int C, int quadrature, double2* out, const int64_t* shifts) {
  if (shifts)
    hipLaunchKernelGGL((k_t2w_stream<6, LG, R, THR, true>), dim3(grid), dim3(THR),
                       stream_lds_bytes(1 << LG, 6, THR, R), st, x, L, ldx, n_series, N, filt,
                       m_H, tw, C, quadrature, out, shifts);
  else
    hipLaunchKernelGGL((k_t2w_stream<6, LG, R, THR, false>), dim3(grid), dim3(THR),
                       stream_lds_bytes(1 << LG, 6, THR, R), st, x, L, ldx, n_series, N, filt,
                       m_H, tw, C, quadrature, out, shifts);
}

// The 1024-thread shapes need more than 64 KiB of dynamic LDS: allowed once per
// process; false if the device refuses (M = 256 and 512 then take k_t2w_poly_h).
bool stream_wide_ready() {
  static const bool ok = [] {
    const int need = (int)stream_lds_bytes(kStreamWideM, 6, 1024, 4);
    const int need256 = (int)stream_lds_bytes(256, 6, 1024, 8);
    const hipFuncAttribute a = hipFuncAttributeMaxDynamicSharedMemorySize;
    const bool set =
        hipFuncSetAttribute((const void*)k_t2w_stream<6, 9, 4, 1024, false>, a, need) ==
            hipSuccess &&
        hipFuncSetAttribute((const void*)k_t2w_stream<6, 9, 4, 1024, true>, a, need) ==
            hipSuccess &&
        hipFuncSetAttribute((const void*)k_t2w_stream<6, 8, 8, 1024, false>, a, need256) ==
            hipSuccess &&
        hipFuncSetAttribute((const void*)k_t2w_stream<6, 8, 8, 1024, true>, a, need256) ==
            hipSuccess;
    (void)hipGetLastError();
    return set;
  }();
  return ok;
}

template <int R>
void launch_poly(dim3 grid, hipStream_t st, const double* x, int64_t L, int64_t ldx,
                 int n_series, int M, int64_t N, const double* filt, int m_H, const double2* tw,
                 int log2_2M, int B, int quadrature, double2* out, const int64_t* shifts) {
  const size_t lds = (size_t)B * M * sizeof(double2);
  if ((m_H - 1) / (2 * M) == 6) {
    if (shifts)
      hipLaunchKernelGGL((k_t2w_poly_h<R, 6, true>), grid, dim3(kThreads),
                         lds + M * sizeof(double2), st, x, L, ldx, n_series, M, N, filt, m_H, tw,
                         log2_2M, B, quadrature, out, shifts);
    else
      hipLaunchKernelGGL((k_t2w_poly_h<R, 6, false>), grid, dim3(kThreads),
                         lds + M * sizeof(double2), st, x, L, ldx, n_series, M, N, filt, m_H, tw,
                         log2_2M, B, quadrature, out, shifts);
  } else {
    hipLaunchKernelGGL((k_t2w_poly<R>), grid, dim3(kThreads), lds, st, x, L, ldx, n_series, M,
                       N, filt, m_H, tw, log2_2M, B, quadrature, out);
  }
}

// Inverse transform w2t / w2tQ (cwb-core/WDM.cc:1536-1710), two kernels.
// k_w2t_bins: per time bin n, the half spectrum from column n (00 or 90
// coefficients, the reference's phase rules) and its complex-to-real inverse
// of length 2M, r[k] = sum_j X_j e^{+i pi j k / M} (Hermitian extension),
// computed as the forward FFT of the conjugate extension; two bins share one
// complex FFT (bin 2p real, 2p+1 imaginary: both results are real).
// k_w2t_gather: output sample t sums, over the bins n in ascending order
// (the reference's overlap-add order), h[|t - nM|] r_n[(M (n & 1) + t - nM)
// mod 2M] for |t - nM| < m_H.
__global__ __launch_bounds__(kThreads) void k_w2t_bins(
    const double2* __restrict__ map, int64_t N, int n_series, int M,
    const double2* __restrict__ tw, int log2_2M, int B, int quad, double* __restrict__ work) {
  extern __shared__ double2 buf[];  // (B / 2) * 2M complex
  const int M2 = 2 * M;
  const int64_t tiles = (N + B - 1) / B;
  const int s = (int)(blockIdx.x / tiles);
  const int64_t n0 = (blockIdx.x % tiles) * B;
  if (s >= n_series) return;
  const double r2 = 1.4142135623730951;
  const int pairs = B / 2;
  for (int e = threadIdx.x; e < pairs * M2; e += kThreads) {
    const int p = e >> log2_2M, k = e & (M2 - 1);
    double wr[2], wi[2];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int64_t n = n0 + 2 * p + h;
      wr[h] = wi[h] = 0.0;
      if (n >= N) continue;
      const int j = k <= M ? k : M2 - k;          // spectrum index of this slot
      const double2 cc = map[((int64_t)s * N + n) * (M + 1) + j];
      const double c = quad ? cc.y : cc.x;
      double re = 0.0, im = 0.0;
      if (j == 0) {
        if (quad ? (n & 1) : !(n & 1)) re = c;
      } else if (j == M) {
        if (quad ? ((n + M) & 1) : !((n + M) & 1)) re = (!quad && (M & 1)) ? -c : c;
      } else if (!quad) {
        if ((n + j) & 1) im = c / r2;
        else if (j & 1) re = -c / r2;
        else re = c / r2;
      } else {
        if ((n + j) & 1) re = c / r2;
        else if (j & 1) im = c / r2;
        else im = -c / r2;
      }
      // conjugate extension: W_j = conj X_j (j <= M), W_{2M-j} = X_j
      wr[h] = re;
      wi[h] = (k <= M) ? -im : im;
      if (k == 0 || k == M) wi[h] = 0.0;
    }
    // pair: Z = W_a + i W_b
    buf[p * M2 + bitrev((unsigned)k, log2_2M)] = make_double2(wr[0] - wi[1], wi[0] + wr[1]);
  }
  __syncthreads();
  lds_fft(buf, pairs, log2_2M, tw);
  for (int e = threadIdx.x; e < B * M2; e += kThreads) {
    const int b = e >> log2_2M, k = e & (M2 - 1);
    const int64_t n = n0 + b;
    if (n >= N) continue;
    const double2 z = buf[(b >> 1) * M2 + k];
    work[((int64_t)s * N + n) * M2 + k] = (b & 1) ? z.y : z.x;
  }
}

__global__ __launch_bounds__(kThreads) void k_w2t_gather(
    const double* __restrict__ work, int64_t N, int M, int log2_M,
    const double* __restrict__ filt, int m_H, int64_t n_out, double* __restrict__ out) {
  const int64_t t = (int64_t)blockIdx.x * kThreads + threadIdx.x;
  if (t >= n_out) return;
  const int s = blockIdx.y;
  const int M2 = 2 * M;
  const int64_t lo = t - (m_H - 1);
  const int64_t n_lo = lo <= 0 ? 0 : (lo + M - 1) >> log2_M;
  int64_t n_hi = (t + m_H - 1) >> log2_M;
  if (n_hi > N - 1) n_hi = N - 1;
  double acc = 0.0;
  const double* w = work + (int64_t)s * N * M2;
  for (int64_t n = n_lo; n <= n_hi; ++n) {
    const int64_t j = t - (n << log2_M);
    const int64_t aj = j < 0 ? -j : j;
    if (aj >= m_H) continue;
    const int idx = (int)((M * (n & 1) + j) & (M2 - 1));   // mod 2M, also for j < 0
    acc += filt[aj] * w[n * M2 + idx];
  }
  out[(int64_t)s * n_out + t] = acc;
}

// Streaming form of k_w2t_bins for M = 2^LG <= 512: a block of 256 threads owns
// C consecutive tiles of B = 1024 / M bins (B / 2 pairs, 1,024 complex points)
// of one series.  Thread e handles the four (pair, slot) items e + 256 i of a
// tile.  The coefficients the next tile needs (one phase of two bins per item)
// are loaded into registers once the current tile's spectrum is in LDS, so they
// arrive while it is transformed and stored; the FFT twiddles stay in registers
// (stream_fft) and the FFT buffer is in the swizzled layout.  The half
// spectrum, the FFT stages and the outputs are those of k_w2t_bins, so the work
// rows are identical.
template <int LG>
__global__ __launch_bounds__(kThreads) void k_w2t_stream(
    const double2* __restrict__ map, int64_t N, int n_series, const double2* __restrict__ tw,
    int C, int quad, double* __restrict__ work) {
  constexpr int M = 1 << LG, M2 = 2 * M, LG2 = LG + 1;
  constexpr int B = 1024 >> LG;
  constexpr int NP = B / 2;
  constexpr int I = NP * M2 / kThreads;   // items per thread: 4
  static_assert(I * kThreads == NP * M2, "tile shape");
  extern __shared__ double2 buf[];
  const int64_t tiles = (N + B - 1) / B;
  const int64_t chunks = (tiles + C - 1) / C;
  const int s = (int)(blockIdx.x / chunks);
  if (s >= n_series) return;
  const int64_t t0 = (blockIdx.x % chunks) * C;
  const int nt = (int)((tiles - t0) < C ? (tiles - t0) : C);
  const double* mp = reinterpret_cast<const double*>(map + (int64_t)s * N * (M + 1)) + (quad ? 1 : 0);
  const double r2 = 1.4142135623730951;
  StreamTw<LG2> T;
  stream_tw_load<LG2>(T, tw);
  // the spectrum slot of item i: k = e & (2M - 1), j = min(k, 2M - k) (the
  // same for every item when 2M <= 256)
  auto slot = [](int i) { return (int)((threadIdx.x + i * kThreads) & (M2 - 1)); };
  auto half = [](int k) { return k <= M ? k : M2 - k; };
  double cv[I][2];
  auto load = [&](int64_t n0, double (&c)[I][2]) {
#pragma unroll
    for (int i = 0; i < I; ++i) {
      const int p = (threadIdx.x + i * kThreads) >> LG2;
      const int j = half(slot(i));
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const int64_t n = n0 + 2 * p + h;
        c[i][h] = n < N ? mp[2 * (n * (M + 1) + j)] : 0.0;
      }
    }
  };
  load(t0 * B, cv);
  for (int t = 0; t < nt; ++t) {
    const int64_t n0 = (t0 + t) * B;
#pragma unroll
    for (int i = 0; i < I; ++i) {
      const int p = (threadIdx.x + i * kThreads) >> LG2;
      const int k = slot(i), j = half(k);
      double wr[2], wi[2];
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const int64_t n = n0 + 2 * p + h;
        wr[h] = wi[h] = 0.0;
        if (n >= N) continue;
        const double c = cv[i][h];
        double re = 0.0, im = 0.0;
        if (j == 0) {
          if (quad ? (n & 1) : !(n & 1)) re = c;
        } else if (j == M) {
          if (quad ? ((n + M) & 1) : !((n + M) & 1)) re = (!quad && (M & 1)) ? -c : c;
        } else if (!quad) {
          if ((n + j) & 1) im = c / r2;
          else if (j & 1) re = -c / r2;
          else re = c / r2;
        } else {
          if ((n + j) & 1) re = c / r2;
          else if (j & 1) im = c / r2;
          else im = -c / r2;
        }
        wr[h] = re;
        wi[h] = (k <= M) ? -im : im;
        if (k == 0 || k == M) wi[h] = 0.0;
      }
      buf[swz(p * M2 + (int)bitrev((unsigned)k, LG2))] =
          make_double2(wr[0] - wi[1], wi[0] + wr[1]);
    }
    // the next tile's coefficients, in flight during this tile's FFT and stores
    const bool more = t + 1 < nt;
    if (more) load(n0 + B, cv);
    __syncthreads();
    stream_fft<LG2>(buf, T);
    double* wk = work + ((int64_t)s * N + n0) * M2;
#pragma unroll
    for (int i = 0; i < I; ++i) {
      const int p = (threadIdx.x + i * kThreads) >> LG2;
      const int k = slot(i);
      const double2 z = buf[swz(p * M2 + k)];
      if (n0 + 2 * p < N) wk[(2 * p) * M2 + k] = z.x;
      if (n0 + 2 * p + 1 < N) wk[(2 * p + 1) * M2 + k] = z.y;
    }
    __syncthreads();
  }
}

template <int LG>
void launch_w2t_stream(unsigned grid, hipStream_t st, const double2* map, int64_t N, int n_series,
                       const double2* tw, int C, int quad, double* work) {
  hipLaunchKernelGGL((k_w2t_stream<LG>), dim3(grid), dim3(kThreads),
                     (size_t)1024 * sizeof(double2), st, map, N, n_series, tw, C, quad, work);
}

// Polyphase form of k_w2t_gather for m_H = 12 M + 1.  Output t = j M + k
// (k < M) sums, over the bins n = j - d with |d M + k| <= 12 M and 0 <= n < N,
// in ascending n, h[|d M + k|] r_n[(M (n & 1) + d M + k) mod 2M], and that
// index is k + M (j & 1) for every term.  So a thread owns residue k and the R
// outputs j = jb + 2i of one parity (i < R): all of them read column k + M (j &
// 1) of the bins jb - 12 .. jb + 2R + 10, which the thread loads once (2R + 23
// values, consecutive threads reading consecutive k), with its 25 taps
// h[|d M + k|] loaded once too.  Each value is added to every output that uses
// it as it arrives, in ascending n per output and skipping the same terms, so
// the samples are identical to k_w2t_gather's.
template <int R>
__global__ __launch_bounds__(kThreads) void k_w2t_poly(
    const double* __restrict__ work, int64_t N, int M, int log2_M,
    const double* __restrict__ filt, int m_H, int64_t n_out, double* __restrict__ out) {
  constexpr int D = 12;   // m_H - 1 = D M
  const int s = blockIdx.y;
  const int64_t gi = (int64_t)blockIdx.x * kThreads + threadIdx.x;
  const int k = (int)(gi & (M - 1));
  const int p = (int)((gi >> log2_M) & 1);
  const int64_t g = gi >> (log2_M + 1);
  const int64_t jb = 2 * R * g + p;
  const int64_t J = (n_out + M - 1) >> log2_M;
  if (jb >= J) return;
  const int M2 = 2 * M;
  // tap e = u - 2i of output i is d = D - e: h[|(D - e) M + k|] when < m_H
  double tap[2 * D + 1];
  unsigned tmask = 0;
#pragma unroll
  for (int e = 0; e <= 2 * D; ++e) {
    const int t = (D - e) * M + k;
    const int at = t < 0 ? -t : t;
    const bool v = at < m_H;
    tap[e] = v ? filt[at] : 0.0;
    tmask |= (unsigned)v << e;
  }
  const double* w = work + (int64_t)s * N * M2 + k + M * p;
  double acc[R];
#pragma unroll
  for (int i = 0; i < R; ++i) acc[i] = 0.0;
#pragma unroll
  for (int u = 0; u < 2 * (R - 1) + 2 * D + 1; ++u) {
    // at most 8 loads ahead of their use: 154 VGPRs (3 waves per SIMD) instead
    // of 188 with all 39 hoisted (2 waves); w2t of 7 levels 1.57 -> 1.45 ms
    if (u % 8 == 0) __builtin_amdgcn_sched_barrier(0);
    const int64_t n = jb - D + u;
    const bool nv = n >= 0 && n < N;
    const double v = w[(nv ? n : 0) * M2];
#pragma unroll
    for (int i = 0; i < R; ++i) {
      const int e = u - 2 * i;
      if (e >= 0 && e <= 2 * D) {
        const bool take = nv && ((tmask >> e) & 1u);
        const double a = acc[i] + tap[e] * v;
        acc[i] = take ? a : acc[i];
      }
    }
  }
#pragma unroll
  for (int i = 0; i < R; ++i) {
    const int64_t t = ((jb + 2 * i) << log2_M) + k;
    if (jb + 2 * i < J && t < n_out) out[(int64_t)s * n_out + t] = acc[i];
  }
}
constexpr int kW2tPolyR = 8;

// pycwb_wdm_t2w, and with shifts (one per series, ldx 0) the transforms of
// shifted copies of one series; shifts are honoured by the polyphase kernels
// for m_H = 12 M + 1 (k_t2w_stream, k_t2w_poly_h) only.
int t2w_impl(const double* x, int64_t n_samples, int64_t ldx, int n_series, int M, int stride,
             const double* filter, int m_H, const double* twiddle, int quadrature, double* out,
             void* stream, const int64_t* shifts) {
  if (!x || !filter || !twiddle || !out) return PYCWB_ERR_ARG;
  if (n_series < 1 || M < 1 || (M & (M - 1)) || 2 * M > kLdsDoubles / 2 || stride < 1 ||
      m_H < 1 || n_samples < 1 || (shifts ? ldx != 0 : ldx < n_samples) || n_samples <= m_H)
    return PYCWB_ERR_ARG;
  if (shifts && !(stride == M && (m_H - 1) == 12 * M && M >= 2)) return PYCWB_ERR_ARG;
  (void)hipGetLastError();
  const int64_t padded = n_samples + (n_samples % stride ? stride - n_samples % stride : 0);
  const int64_t N = padded / stride;
  int log2_2M = 0;
  while ((1 << log2_2M) < 2 * M) ++log2_2M;
  const bool wide = stride == M && (m_H - 1) == 12 * M && (M == kStreamWideM || M == 256) &&
                    stream_wide_ready();
  if ((stride == M && (m_H - 1) == 12 * M && M >= 2 && M <= kStreamMaxM) || wide) {
    // streaming polyphase path: tiles of B bins (1024 / M with 256 threads; with
    // 1024 threads 16 at M = 256 and 4 at M = 512), C consecutive tiles per block
    const int B = wide ? (M == 256 ? 16 : 4) : 1024 / M;
    const int64_t tiles = (N + B - 1) / B;
    const int64_t target = wide ? PYCWB_T2W_STREAM_BLOCKS / 4 : PYCWB_T2W_STREAM_BLOCKS;
    int64_t C = (tiles * n_series + target - 1) / target;
    if (C < 1) C = 1;
    if (C > 1 << 20) C = 1 << 20;
    const int64_t chunks = (tiles + C - 1) / C;
    const int64_t grid = chunks * n_series;
    if (grid > 0x7fffffff) return PYCWB_ERR_ARG;
    const hipStream_t st = (hipStream_t)stream;
    const double2* tw = (const double2*)twiddle;
    double2* o = (double2*)out;
    const unsigned g = (unsigned)grid;
    switch (log2_2M - 1) {
      case 1: launch_stream<1, 8, 256>(g, st, x, n_samples, ldx, n_series, N, filter, m_H, tw, (int)C, quadrature, o, shifts); break;
      case 2: launch_stream<2, 8, 256>(g, st, x, n_samples, ldx, n_series, N, filter, m_H, tw, (int)C, quadrature, o, shifts); break;
      case 3: launch_stream<3, 8, 256>(g, st, x, n_samples, ldx, n_series, N, filter, m_H, tw, (int)C, quadrature, o, shifts); break;
      case 4: launch_stream<4, 8, 256>(g, st, x, n_samples, ldx, n_series, N, filter, m_H, tw, (int)C, quadrature, o, shifts); break;
      case 5: launch_stream<5, 8, 256>(g, st, x, n_samples, ldx, n_series, N, filter, m_H, tw, (int)C, quadrature, o, shifts); break;
      case 6: launch_stream<6, 8, 256>(g, st, x, n_samples, ldx, n_series, N, filter, m_H, tw, (int)C, quadrature, o, shifts); break;
      case 7: launch_stream<7, 8, 256>(g, st, x, n_samples, ldx, n_series, N, filter, m_H, tw, (int)C, quadrature, o, shifts); break;
      case 8: launch_stream<8, 8, 1024>(g, st, x, n_samples, ldx, n_series, N, filter, m_H, tw, (int)C, quadrature, o, shifts); break;
      default: launch_stream<9, 4, 1024>(g, st, x, n_samples, ldx, n_series, N, filter, m_H, tw, (int)C, quadrature, o, shifts); break;
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
      pycwb::record_error("k_t2w_stream", hipGetErrorString(e));
      return PYCWB_ERR_HIP;
    }
    return PYCWB_OK;
  }
  if (stride == M && (m_H - 1) % (2 * M) == 0 && M >= 2 && M <= kPolyLdsMax / 2) {
    // polyphase path: B bins per block (even, >= 2), R bins per thread
    const int B = poly_bins(M);
    int R = (B * 2 * M) / kThreads;
    R = R > 16 ? 16 : (R < 2 ? 2 : R);
    int Rp = 2;
    while (Rp * 2 <= R) Rp *= 2;
    R = Rp > B ? B : Rp;
    const int64_t tiles = (N + B - 1) / B;
    const int64_t grid = tiles * n_series;
    if (grid > 0x7fffffff) return PYCWB_ERR_ARG;
    const dim3 g((unsigned)grid);
    const hipStream_t st = (hipStream_t)stream;
    const double2* tw = (const double2*)twiddle;
    double2* o = (double2*)out;
    switch (R) {
      case 2: launch_poly<2>(g, st, x, n_samples, ldx, n_series, M, N, filter, m_H, tw, log2_2M, B, quadrature, o, shifts); break;
      case 4: launch_poly<4>(g, st, x, n_samples, ldx, n_series, M, N, filter, m_H, tw, log2_2M, B, quadrature, o, shifts); break;
      case 8: launch_poly<8>(g, st, x, n_samples, ldx, n_series, M, N, filter, m_H, tw, log2_2M, B, quadrature, o, shifts); break;
      default: launch_poly<16>(g, st, x, n_samples, ldx, n_series, M, N, filter, m_H, tw, log2_2M, B, quadrature, o, shifts); break;
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
      pycwb::record_error("k_t2w_poly", hipGetErrorString(e));
      return PYCWB_ERR_HIP;
    }
    return PYCWB_OK;
  }
  if (shifts) return PYCWB_ERR_ARG;
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

}  // namespace

extern "C" {

int pycwb_wdm_t2w(const double* x, int64_t n_samples, int64_t ldx, int n_series, int M,
                  int stride, const double* filter, int m_H, const double* twiddle,
                  int quadrature, double* out, void* stream) {
  return t2w_impl(x, n_samples, ldx, n_series, M, stride, filter, m_H, twiddle, quadrature, out,
                  stream, nullptr);
}

int pycwb_wdm_t2w_shifted(const double* ts, int64_t n_samples, const int64_t* shifts,
                          int n_shifts, int M, const double* filter, int m_H,
                          const double* twiddle, int quadrature, double* out, void* stream) {
  if (!shifts) return PYCWB_ERR_ARG;
  return t2w_impl(ts, n_samples, 0, n_shifts, M, M, filter, m_H, twiddle, quadrature, out, stream,
                  shifts);
}

int pycwb_wdm_w2t(const double* map, int64_t N, int n_series, int M, const double* filter,
                  int m_H, const double* twiddle, int quad, double* work, double* out,
                  int64_t n_out, void* stream) {
  if (!map || !filter || !twiddle || !work || !out) return PYCWB_ERR_ARG;
  if (n_series < 1 || M < 2 || (M & (M - 1)) || M > kPolyLdsMax / 2 || N < 1 || m_H < 1 ||
      n_out < 1)
    return PYCWB_ERR_ARG;
  (void)hipGetLastError();
  int log2_2M = 0;
  while ((1 << log2_2M) < 2 * M) ++log2_2M;
  int B = (kPolyLdsMax / M) & ~1;   // bins per block: (B / 2) * 2M complex = 32 KiB
  if (B > 64) B = 64;
  if (B < 2) B = 2;
  const int64_t tiles = (N + B - 1) / B;
  const int64_t grid = tiles * n_series;
  if (grid > 0x7fffffff || n_out / kThreads > 0x7fffffff || n_series > 65535)
    return PYCWB_ERR_ARG;
  const hipStream_t st = (hipStream_t)stream;
  if (M <= kStreamWideM) {
    // streaming form: tiles of 1024 / M bins, C consecutive tiles per block
    const int Bs = 1024 / M;
    const int64_t tiles_s = (N + Bs - 1) / Bs;
    int64_t C = (tiles_s * n_series + PYCWB_T2W_STREAM_BLOCKS - 1) / PYCWB_T2W_STREAM_BLOCKS;
    if (C < 1) C = 1;
    if (C > 1 << 20) C = 1 << 20;
    const int64_t grid_s = ((tiles_s + C - 1) / C) * n_series;
    if (grid_s > 0x7fffffff) return PYCWB_ERR_ARG;
    const unsigned g = (unsigned)grid_s;
    const double2* mp = (const double2*)map;
    const double2* tw = (const double2*)twiddle;
    switch (log2_2M - 1) {
      case 1: launch_w2t_stream<1>(g, st, mp, N, n_series, tw, (int)C, quad, work); break;
      case 2: launch_w2t_stream<2>(g, st, mp, N, n_series, tw, (int)C, quad, work); break;
      case 3: launch_w2t_stream<3>(g, st, mp, N, n_series, tw, (int)C, quad, work); break;
      case 4: launch_w2t_stream<4>(g, st, mp, N, n_series, tw, (int)C, quad, work); break;
      case 5: launch_w2t_stream<5>(g, st, mp, N, n_series, tw, (int)C, quad, work); break;
      case 6: launch_w2t_stream<6>(g, st, mp, N, n_series, tw, (int)C, quad, work); break;
      case 7: launch_w2t_stream<7>(g, st, mp, N, n_series, tw, (int)C, quad, work); break;
      case 8: launch_w2t_stream<8>(g, st, mp, N, n_series, tw, (int)C, quad, work); break;
      default: launch_w2t_stream<9>(g, st, mp, N, n_series, tw, (int)C, quad, work); break;
    }
  } else {
    hipLaunchKernelGGL(k_w2t_bins, dim3((unsigned)grid), dim3(kThreads),
                       (size_t)(B / 2) * 2 * M * sizeof(double2), st, (const double2*)map, N,
                       n_series, M, (const double2*)twiddle, log2_2M, B, quad, work);
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    pycwb::record_error("k_w2t_bins", hipGetErrorString(e));
    return PYCWB_ERR_HIP;
  }
  if (m_H - 1 == 12 * M) {
    // polyphase gather: (ceil(J / 2R) groups) x 2 parities x M residues per series
    const int64_t J = (n_out + M - 1) / M;
    const int64_t threads = ((J + 2 * kW2tPolyR - 1) / (2 * kW2tPolyR)) * 2 * M;
    const int64_t blocks = (threads + kThreads - 1) / kThreads;
    if (blocks > 0x7fffffff) return PYCWB_ERR_ARG;
    hipLaunchKernelGGL((k_w2t_poly<kW2tPolyR>), dim3((unsigned)blocks, (unsigned)n_series),
                       dim3(kThreads), 0, st, work, N, M, log2_2M - 1, filter, m_H, n_out, out);
  } else {
    hipLaunchKernelGGL(k_w2t_gather,
                       dim3((unsigned)((n_out + kThreads - 1) / kThreads), (unsigned)n_series),
                       dim3(kThreads), 0, st, work, N, M, log2_2M - 1, filter, m_H, n_out, out);
  }
  e = hipGetLastError();
  if (e != hipSuccess) {
    pycwb::record_error("k_w2t_gather", hipGetErrorString(e));
    return PYCWB_ERR_HIP;
  }
  return PYCWB_OK;
}

}  // extern "C"
