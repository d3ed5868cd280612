// Array statistics of the coherence stage's pixel-energy threshold.
//
// Restates the array work of compute_threshold
// (pycwb/modules/coherence_native/veto_threshold.py:98-222); the scalar Gamma
// fits that follow stay on the host, as in the reference.  The detectors'
// energy maps (n_ifo, n_freq, n_time) float64 are summed in detector order and
// clipped to [0, n_ifo * 100] in one of the reference's two orders:
//   time_major = 0  rows of frequency, time bins [begin, begin + count/n_freq)
//                   (the edge-cropped maps summed and raveled, :173-176)
//   time_major = 1  the time-major flat index j in [begin, begin + count)
//                   (flat[t * n_freq + f], the w[nL:nR] region, :136-157)
// Then, per request: the counts of values above 1e-4 and 1e-3; the values
// sorted ascending (the order statistics of np.sort, :186-201); the values
// above 1e-3 compacted in order and summed, with their logs, in numpy's
// add.reduce order per 8192-value chunk (np.mean of them, :162-163; the host
// adds the chunk sums in order).
//   k_th_zero     the counters
//   k_th_values   value   sum over detectors, clip, per-block counts
//                         (frequency-row order)
//   k_th_values_tm tile   the same in time-major order through an LDS
//                         transpose
//   k_th_counts           the counts summed over the blocks
//   (sort)                hipCUB radix sort of the float64 values
//   (select)              hipCUB stable compaction of the values above 1e-3
//   k_th_chunks   chunk   512 threads per 8192 values: 8 lanes per 128-value
//                         leaf (numpy's strided accumulators), shuffle trees
//                         over the accumulators and the 64 leaves (the
//                         partial last chunk walks numpy's general tree)
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>
#include <math.h>
#include <stdint.h>

#include "np_sum.hh"
#include "skyscan_lane.hh"
#include "pycwb_amd.h"

using namespace pycwb;

namespace {

constexpr int kBlock = 256;
constexpr int kWave = 64;

struct ThArgs {
  const double* maps;
  int n_ifo, n_freq;
  int64_t n_time;
  int time_major;
  int64_t begin, count, ncol;
  double hi;
  double* values;
  unsigned long long* partial;  // (2, gridDim.x) per-block counts
};

constexpr int kValueBlocks = 2048;

// np.clip(x, 0, hi): x below 0 becomes 0, above hi becomes hi; -0.0 and NaN
// pass unchanged.
__device__ __forceinline__ double clip(double x, double hi) {
  const double v = x < 0.0 ? 0.0 : x;
  return v > hi ? hi : v;
}

__global__ void k_th_zero(unsigned long long* counts, int64_t* n_sel) {
  if (threadIdx.x < 2) counts[threadIdx.x] = 0ull;
  if (threadIdx.x == 0 && n_sel) *n_sel = 0;
}

// counts[c] = sum of the per-block partial counts (one block).
__global__ void __launch_bounds__(kBlock) k_th_counts(const unsigned long long* partial,
                                                     int n_blocks, unsigned long long* counts) {
  __shared__ unsigned long long s[2][kBlock];
  unsigned long long c4 = 0, c3 = 0;
  for (int b = threadIdx.x; b < n_blocks; b += kBlock) {
    c4 += partial[b];
    c3 += partial[n_blocks + b];
  }
  s[0][threadIdx.x] = c4;
  s[1][threadIdx.x] = c3;
  __syncthreads();
  for (int o = kBlock / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) {
      s[0][threadIdx.x] += s[0][threadIdx.x + o];
      s[1][threadIdx.x] += s[1][threadIdx.x + o];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    counts[0] = s[0][0];
    counts[1] = s[1][0];
  }
}

// Grid-stride over the values; per-thread counts reduced across the wave by
// shuffles and across the block in LDS into the block's partial counts (no
// atomics: same-address atomics from every block serialise).
// Indices fit 32 bits (count and begin + count below 2^31, checked by the
// host), so the row/column split is a 32-bit division.
__global__ void __launch_bounds__(kBlock) k_th_values(ThArgs a) {
  unsigned long long c4 = 0, c3 = 0;
  const int64_t plane = (int64_t)a.n_freq * a.n_time;
  const uint32_t count = (uint32_t)a.count, begin = (uint32_t)a.begin;
  const uint32_t nf = (uint32_t)a.n_freq, ncol = (uint32_t)a.ncol;
  const uint32_t stride = gridDim.x * kBlock;
  for (uint32_t k = blockIdx.x * kBlock + threadIdx.x; k < count; k += stride) {
    uint32_t f, t;
    if (a.time_major) {
      const uint32_t j = begin + k;
      t = j / nf;
      f = j - t * nf;
    } else {
      f = k / ncol;
      t = begin + (k - f * ncol);
    }
    const double* p = a.maps + (int64_t)f * a.n_time + t;
    // np.sum(axis=0) starts from 0.0 (-0.0 sums to +0.0); the time-major
    // path (w = flat0.copy(); w += ...) starts from the first map
    double w = a.time_major ? p[0] : 0.0 + p[0];
    for (int d = 1; d < a.n_ifo; ++d) w = w + p[d * plane];
    const double v = clip(w, a.hi);
    a.values[k] = v;
    c4 += v > 1.0e-4 ? 1ull : 0ull;
    c3 += v > 1.0e-3 ? 1ull : 0ull;
  }
  for (int o = kWave / 2; o > 0; o >>= 1) {
    c4 += __shfl_down(c4, o);
    c3 += __shfl_down(c3, o);
  }
  __shared__ unsigned long long s4[kBlock / kWave], s3[kBlock / kWave];
  const int wave = threadIdx.x / kWave;
  if ((threadIdx.x & (kWave - 1)) == 0) {
    s4[wave] = c4;
    s3[wave] = c3;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long t4 = 0, t3 = 0;
    for (int w = 0; w < kBlock / kWave; ++w) {
      t4 += s4[w];
      t3 += s3[w];
    }
    a.partial[blockIdx.x] = t4;
    a.partial[gridDim.x + blockIdx.x] = t3;
  }
}

// Per-block counts reduced into the block's partial counts.
__device__ __forceinline__ void block_counts(unsigned long long c4, unsigned long long c3,
                                             unsigned long long* partial) {
  for (int o = kWave / 2; o > 0; o >>= 1) {
    c4 += __shfl_down(c4, o);
    c3 += __shfl_down(c3, o);
  }
  __shared__ unsigned long long s4[kBlock / kWave], s3[kBlock / kWave];
  const int wave = threadIdx.x / kWave;
  if ((threadIdx.x & (kWave - 1)) == 0) {
    s4[wave] = c4;
    s3[wave] = c3;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long t4 = 0, t3 = 0;
    for (int w = 0; w < kBlock / kWave; ++w) {
      t4 += s4[w];
      t3 += s3[w];
    }
    partial[blockIdx.x] = t4;
    partial[gridDim.x + blockIdx.x] = t3;
  }
}

// The time-major order through an LDS transpose: a tile of T time bins x
// n_freq rows is read along time (coalesced), summed over detectors, and
// written in time-major order (coalesced).  T is a power of two with
// T * n_freq * 8 bytes <= 32 KB.  Grid-stride over the tiles.
__global__ void __launch_bounds__(kBlock) k_th_values_tm(ThArgs a, int log2_T) {
  extern __shared__ double tile[];  // [T][n_freq]
  const int T = 1 << log2_T;
  const int nf = a.n_freq;
  const uint32_t cells = (uint32_t)T * (uint32_t)nf;
  const int64_t plane = (int64_t)nf * a.n_time;
  const int64_t j0 = a.begin, j1 = a.begin + a.count;
  const int64_t t_first = j0 / nf, t_last = (j1 - 1) / nf;
  const int64_t n_tiles = (t_last - t_first) / T + 1;
  unsigned long long c4 = 0, c3 = 0;
  for (int64_t tile_i = blockIdx.x; tile_i < n_tiles; tile_i += gridDim.x) {
    const int64_t t0 = t_first + tile_i * T;
    for (uint32_t idx = threadIdx.x; idx < cells; idx += kBlock) {
      const int f = (int)(idx >> log2_T), tt = (int)(idx & (uint32_t)(T - 1));
      const int64_t t = t0 + tt;
      double w = 0.0;
      if (t < a.n_time) {
        const double* p = a.maps + (int64_t)f * a.n_time + t;
        w = p[0];
        for (int d = 1; d < a.n_ifo; ++d) w = w + p[d * plane];
      }
      tile[tt * nf + f] = w;
    }
    __syncthreads();
    for (uint32_t idx = threadIdx.x; idx < cells; idx += kBlock) {
      const uint32_t tt = idx / (uint32_t)nf, f = idx - tt * (uint32_t)nf;
      const int64_t j = (t0 + tt) * nf + f;
      if (j >= j0 && j < j1) {
        const double v = clip(tile[idx], a.hi);
        a.values[j - j0] = v;
        c4 += v > 1.0e-4 ? 1ull : 0ull;
        c3 += v > 1.0e-3 ? 1ull : 0ull;
      }
    }
    __syncthreads();
  }
  block_counts(c4, c3, a.partial);
}

struct Above {
  __device__ bool operator()(const double& x) const { return x > 1.0e-3; }
};

struct Inside {  // lo < x < hi
  double lo, hi;
  __device__ bool operator()(const double& x) const { return x > lo && x < hi; }
};

struct Plain {
  const double* x;
  __device__ double operator()(int64_t i) const { return x[i]; }
};

struct Log {
  const double* x;
  __device__ double operator()(int64_t i) const { return log(x[i]); }
};

constexpr int kChunkBlock = 512;  // 64 leaves x 8 accumulators
constexpr int kTreeLevels = 8;    // root + 7 levels of splits (len < 8192)
constexpr int kTreeWidth = 128;   // 2^7 nodes on the deepest level

// out[2c] = np.sum of chunk c of x[0 .. m), out[2c + 1] = np.sum of its logs.
// A full chunk (8192 values) is 64 leaves of 128 added as a balanced binary
// tree.  Leaf l is taken by 8 lanes, lane a adding x[128 l + a + 8 i] for
// i = 0 .. 15 in order (numpy's 8 strided accumulators); shuffles combine
// them as ((r0 + r1) + (r2 + r3)) + ((r4 + r5) + (r6 + r7)), and the first
// wave adds the 64 leaves up the tree by shuffles.  The last, partial chunk
// builds numpy's general tree level by level in LDS, sums its leaves in
// parallel and adds them up level by level.
__global__ void __launch_bounds__(kChunkBlock) k_th_chunks(const double* x, const int64_t* m_ptr,
                                                          double* out) {
  __shared__ int64_t ts[kTreeLevels][kTreeWidth], tn[kTreeLevels][kTreeWidth];
  __shared__ double tv[kTreeLevels][kTreeWidth], tl[kTreeLevels][kTreeWidth];
  __shared__ double lv[kWave], lg[kWave];
  const int64_t m = *m_ptr;
  const int64_t base = (int64_t)blockIdx.x * kNpChunk;
  if (base >= m) return;
  // The chunk's length is taken in each branch, not through a select of
  // min(m - base, 8192): hipcc (ROCm 7.2, gfx950) compiled that select as an
  // s_cselect on a stale SCC, which gave the last chunk 8192 values.
  const int64_t len = m - base;
  const Plain g{x + base};
  const Log gl{x + base};
  if (len >= kNpChunk) {
    const int leaf = threadIdx.x >> 3, acc = threadIdx.x & 7;
    const double* p = x + base + leaf * kNpLeaf + acc;
    double r = p[0], rl = log(p[0]);
    for (int i = 1; i < (int)(kNpLeaf / 8); ++i) {
      const double v = p[8 * i];
      r = r + v;
      rl = rl + log(v);
    }
    for (int o = 1; o < 8; o <<= 1) {
      const double r2 = __shfl_down(r, o), l2 = __shfl_down(rl, o);
      if ((acc & (2 * o - 1)) == 0) {
        r = r + r2;
        rl = rl + l2;
      }
    }
    if (acc == 0) {
      lv[leaf] = r;
      lg[leaf] = rl;
    }
    __syncthreads();
    if (threadIdx.x < kWave) {
      const int lane = threadIdx.x;
      double v = lv[lane], lgv = lg[lane];
      for (int o = 1; o < kWave; o <<= 1) {
        const double v2 = __shfl_down(v, o), l2 = __shfl_down(lgv, o);
        if ((lane & (2 * o - 1)) == 0) {
          v = v + v2;
          lgv = lgv + l2;
        }
      }
      if (lane == 0) {
        out[2 * blockIdx.x] = v;
        out[2 * blockIdx.x + 1] = lgv;
      }
    }
    return;
  }
  // numpy's tree over [0, len), level by level: node i of level d + 1 is
  // half (i & 1) of node i / 2 of level d when that node holds more than 128
  // values (split at len / 2 rounded down to a multiple of 8); nodes of at
  // most 128 values are leaves.  Depth <= 7 below the root for len < 8192.
  if (threadIdx.x == 0) {
    ts[0][0] = 0;
    tn[0][0] = len;
  }
  __syncthreads();
  for (int d = 0; d < kTreeLevels - 1; ++d) {
    const int i = threadIdx.x;
    if (i < (2 << d)) {
      const int64_t pn = tn[d][i >> 1];
      int64_t cs = 0, cn = -1;
      if (pn > kNpLeaf) {
        int64_t n2 = pn / 2;
        n2 -= n2 % 8;
        cs = (i & 1) ? ts[d][i >> 1] + n2 : ts[d][i >> 1];
        cn = (i & 1) ? pn - n2 : n2;
      }
      ts[d + 1][i] = cs;
      tn[d + 1][i] = cn;
    }
    __syncthreads();
  }
  for (int k = threadIdx.x; k < kTreeLevels * kTreeWidth; k += kChunkBlock) {
    const int d = k / kTreeWidth, i = k % kTreeWidth;
    if (i < (1 << d) && tn[d][i] >= 0 && tn[d][i] <= kNpLeaf) {
      tv[d][i] = np_leaf(g, ts[d][i], tn[d][i]);
      tl[d][i] = np_leaf(gl, ts[d][i], tn[d][i]);
    }
  }
  __syncthreads();
  for (int d = kTreeLevels - 2; d >= 0; --d) {
    const int i = threadIdx.x;
    if (i < (1 << d) && tn[d][i] > kNpLeaf) {
      tv[d][i] = tv[d + 1][2 * i] + tv[d + 1][2 * i + 1];
      tl[d][i] = tl[d + 1][2 * i] + tl[d + 1][2 * i + 1];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[2 * blockIdx.x] = tv[0][0];
    out[2 * blockIdx.x + 1] = tl[0][0];
  }
}

// ---- Gamma-to-Gauss of a max-energy map (TimeFrequencyMap.Gamma2Gauss) ----

// partial[block] = number of x[begin .. end) above thr in the block's share.
__global__ void __launch_bounds__(kBlock) k_g2g_count(const double* x, int64_t begin,
                                                     int64_t end, double thr,
                                                     unsigned long long* partial) {
  unsigned long long c = 0;
  const int64_t stride = (int64_t)gridDim.x * kBlock;
  for (int64_t k = begin + (int64_t)blockIdx.x * kBlock + threadIdx.x; k < end; k += stride)
    c += x[k] > thr ? 1ull : 0ull;
  for (int o = kWave / 2; o > 0; o >>= 1) c += __shfl_down(c, o);
  __shared__ unsigned long long sc[kBlock / kWave];
  if ((threadIdx.x & (kWave - 1)) == 0) sc[threadIdx.x / kWave] = c;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long t = 0;
    for (int w = 0; w < kBlock / kWave; ++w) t += sc[w];
    partial[blockIdx.x] = t;
  }
}

__global__ void __launch_bounds__(kBlock) k_g2g_reduce(const unsigned long long* partial,
                                                      int n_blocks, unsigned long long* out) {
  __shared__ unsigned long long s[kBlock];
  unsigned long long c = 0;
  for (int b = threadIdx.x; b < n_blocks; b += kBlock) c += partial[b];
  s[threadIdx.x] = c;
  __syncthreads();
  for (int o = kBlock / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) s[threadIdx.x] += s[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = s[0];
}

// out = amp - ALP (1 + log(amp / ALP)) where amp = x alp / avr >= ALP, else 0
// (time_frequency_map.py:196-204: amp = flat * alp / avr in that order).
__global__ void k_g2g_map(const double* x, int64_t n, double alp, double avr, double ALP,
                          double* out) {
  const int64_t stride = (int64_t)gridDim.x * kBlock;
  for (int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x; k < n; k += stride) {
    const double amp = x[k] * alp / avr;
    out[k] = amp >= ALP ? amp - ALP * (1.0 + log(amp / ALP)) : 0.0;
  }
}

__global__ void k_g2g_scale(double* x, int64_t n, double s) {
  const int64_t stride = (int64_t)gridDim.x * kBlock;
  for (int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x; k < n; k += stride) x[k] *= s;
}

int check_launch(const char* what) {
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    record_error(what, hipGetErrorString(e));
    return PYCWB_ERR_HIP;
  }
  return PYCWB_OK;
}

size_t align_up(size_t x) { return (x + 255) & ~(size_t)255; }

struct Layout {
  size_t compact, n_sel, partial, temp, temp_bytes, total;
};

int layout(const pycwb_threshold_desc* d, Layout* L) {
  const int n = (int)d->count;
  size_t sort_bytes = 0, sel_bytes = 0;
  if (n > 0) {
    if (d->sorted && hipcub::DeviceRadixSort::SortKeys(nullptr, sort_bytes,
                                                       (const double*)nullptr,
                                                       (double*)nullptr, n) != hipSuccess) {
      record_error("pycwb_threshold_work_bytes", "hipCUB sort storage query failed");
      return PYCWB_ERR_HIP;
    }
    if (d->chunk_sums && hipcub::DeviceSelect::If(nullptr, sel_bytes, (const double*)nullptr,
                                                  (double*)nullptr, (int64_t*)nullptr, n,
                                                  Above()) != hipSuccess) {
      record_error("pycwb_threshold_work_bytes", "hipCUB select storage query failed");
      return PYCWB_ERR_HIP;
    }
  }
  size_t o = 0;
  L->compact = o;
  o = align_up(o + (d->chunk_sums ? sizeof(double) * (size_t)d->count : 0));
  L->n_sel = o;
  o = align_up(o + sizeof(int64_t));
  L->partial = o;
  o = align_up(o + 2 * sizeof(unsigned long long) * kValueBlocks);
  L->temp_bytes = sort_bytes > sel_bytes ? sort_bytes : sel_bytes;
  L->temp = o;
  o = align_up(o + (L->temp_bytes ? L->temp_bytes : 1));
  L->total = o;
  return PYCWB_OK;
}

int check_desc(const pycwb_threshold_desc* d) {
  if (!d || !d->maps || !d->values || !d->counts || d->n_ifo < 1 || d->n_freq < 1 ||
      d->n_time < 1 || d->count < 0 || d->count >= INT32_MAX || d->begin < 0 ||
      d->begin + d->count >= INT32_MAX)
    return PYCWB_ERR_ARG;
  const int64_t cells = (int64_t)d->n_freq * d->n_time;
  if (d->time_major) {
    if (d->begin + d->count > cells) return PYCWB_ERR_ARG;
  } else {
    if (d->count % d->n_freq != 0 || d->begin + d->count / d->n_freq > d->n_time)
      return PYCWB_ERR_ARG;
  }
  return PYCWB_OK;
}

}  // namespace

extern "C" {

int pycwb_threshold_work_bytes(const pycwb_threshold_desc* d, size_t* bytes) {
  if (!bytes) return PYCWB_ERR_ARG;
  int rc = check_desc(d);
  if (rc != PYCWB_OK) return rc;
  Layout L;
  if ((rc = layout(d, &L)) != PYCWB_OK) return rc;
  *bytes = L.total;
  return PYCWB_OK;
}

int pycwb_threshold_stats(const pycwb_threshold_desc* d, void* stream) {
  int rc = check_desc(d);
  if (rc != PYCWB_OK) return rc;
  Layout L;
  if ((rc = layout(d, &L)) != PYCWB_OK) return rc;
  if (!d->work || d->work_bytes < L.total) return PYCWB_ERR_ARG;
  const hipStream_t s = (hipStream_t)stream;
  char* w = (char*)d->work;
  int64_t* n_sel = (int64_t*)(w + L.n_sel);
  (void)hipGetLastError();
  hipLaunchKernelGGL(k_th_zero, dim3(1), dim3(kWave), 0, s,
                     (unsigned long long*)d->counts, n_sel);
  if ((rc = check_launch("k_th_zero")) != PYCWB_OK) return rc;
  if (d->count == 0) return PYCWB_OK;
  ThArgs a;
  a.maps = d->maps; a.n_ifo = d->n_ifo; a.n_freq = d->n_freq; a.n_time = d->n_time;
  a.time_major = d->time_major; a.begin = d->begin; a.count = d->count;
  a.ncol = d->time_major ? 0 : d->count / d->n_freq;
  a.hi = d->hi; a.values = d->values;
  a.partial = (unsigned long long*)(w + L.partial);
  int blocks;
  if (d->time_major) {
    int log2_T = 6;
    while (log2_T > 0 && ((int64_t)d->n_freq << log2_T) * 8 > 32768) --log2_T;
    const int64_t t_first = d->begin / d->n_freq;
    const int64_t t_last = (d->begin + d->count - 1) / d->n_freq;
    const int64_t tiles = ((t_last - t_first) >> log2_T) + 1;
    blocks = (int)(tiles < kValueBlocks ? tiles : kValueBlocks);
    const size_t lds = (size_t)d->n_freq * ((size_t)1 << log2_T) * sizeof(double);
    if (lds > 65536) return PYCWB_ERR_ARG;
    hipLaunchKernelGGL(k_th_values_tm, dim3((unsigned)blocks), dim3(kBlock), lds, s, a, log2_T);
    if ((rc = check_launch("k_th_values_tm")) != PYCWB_OK) return rc;
  } else {
    const int64_t want = (d->count + kBlock - 1) / kBlock;
    blocks = (int)(want < kValueBlocks ? want : kValueBlocks);
    hipLaunchKernelGGL(k_th_values, dim3((unsigned)blocks), dim3(kBlock), 0, s, a);
    if ((rc = check_launch("k_th_values")) != PYCWB_OK) return rc;
  }
  hipLaunchKernelGGL(k_th_counts, dim3(1), dim3(kBlock), 0, s,
                     (const unsigned long long*)a.partial, blocks,
                     (unsigned long long*)d->counts);
  if ((rc = check_launch("k_th_counts")) != PYCWB_OK) return rc;
  const int n = (int)d->count;
  if (d->sorted) {
    size_t tb = L.temp_bytes;
    const hipError_t e = hipcub::DeviceRadixSort::SortKeys(w + L.temp, tb, d->values, d->sorted,
                                                           n, 0, 64, s);
    if (e != hipSuccess) {
      record_error("threshold sort", hipGetErrorString(e));
      return PYCWB_ERR_HIP;
    }
  }
  if (d->chunk_sums) {
    size_t tb = L.temp_bytes;
    double* compact = (double*)(w + L.compact);
    const hipError_t e = hipcub::DeviceSelect::If(w + L.temp, tb, d->values, compact, n_sel, n,
                                                  Above(), s);
    if (e != hipSuccess) {
      record_error("threshold select", hipGetErrorString(e));
      return PYCWB_ERR_HIP;
    }
    const int64_t chunks = (d->count + kNpChunk - 1) / kNpChunk;
    hipLaunchKernelGGL(k_th_chunks, dim3((unsigned)chunks), dim3(kChunkBlock), 0, s,
                       (const double*)compact, (const int64_t*)n_sel, d->chunk_sums);
    if ((rc = check_launch("k_th_chunks")) != PYCWB_OK) return rc;
  }
  return PYCWB_OK;
}

}  // extern "C"

namespace {

struct G2gLayout {
  size_t compact, partial, temp, temp_bytes, total;
};

int g2g_layout(const pycwb_g2g_desc* d, G2gLayout* L) {
  const int64_t win = d->nR - d->nL + 1;
  size_t sort_bytes = 0, sel_bytes = 0;
  if (hipcub::DeviceRadixSort::SortKeys(nullptr, sort_bytes, (const double*)nullptr,
                                        (double*)nullptr, (int)win) != hipSuccess ||
      hipcub::DeviceSelect::If(nullptr, sel_bytes, (const double*)nullptr, (double*)nullptr,
                               (int64_t*)nullptr, (int)(win - 1), Inside{0.0, 0.0}) !=
          hipSuccess) {
    record_error("pycwb_g2g_work_bytes", "hipCUB storage query failed");
    return PYCWB_ERR_HIP;
  }
  size_t o = 0;
  L->compact = o;
  o = align_up(o + sizeof(double) * (size_t)win);
  L->partial = o;
  o = align_up(o + sizeof(unsigned long long) * kValueBlocks);
  L->temp_bytes = sort_bytes > sel_bytes ? sort_bytes : sel_bytes;
  L->temp = o;
  o = align_up(o + (L->temp_bytes ? L->temp_bytes : 1));
  L->total = o;
  return PYCWB_OK;
}

int g2g_count(const double* x, int64_t begin, int64_t end, double thr,
              unsigned long long* partial, unsigned long long* out, hipStream_t s) {
  const int64_t want = (end - begin + kBlock - 1) / kBlock;
  const int blocks = (int)(want < 1 ? 1 : (want < kValueBlocks ? want : kValueBlocks));
  hipLaunchKernelGGL(k_g2g_count, dim3((unsigned)blocks), dim3(kBlock), 0, s, x, begin, end,
                     thr, partial);
  int rc = check_launch("k_g2g_count");
  if (rc != PYCWB_OK) return rc;
  hipLaunchKernelGGL(k_g2g_reduce, dim3(1), dim3(kBlock), 0, s,
                     (const unsigned long long*)partial, blocks, out);
  return check_launch("k_g2g_reduce");
}

int g2g_sort(const double* x, int64_t n, double* out, char* temp, size_t temp_bytes,
             hipStream_t s) {
  size_t tb = temp_bytes;
  const hipError_t e = hipcub::DeviceRadixSort::SortKeys(temp, tb, x, out, (int)n, 0, 64, s);
  if (e != hipSuccess) {
    record_error("gamma-to-gauss sort", hipGetErrorString(e));
    return PYCWB_ERR_HIP;
  }
  return PYCWB_OK;
}

int g2g_check(const pycwb_g2g_desc* d) {
  if (!d || !d->flat || d->n < 4 || d->n >= INT32_MAX || d->nL < 0 || d->nR <= d->nL ||
      d->nR > d->n - 1)
    return PYCWB_ERR_ARG;
  return PYCWB_OK;
}

}  // namespace

extern "C" {

int pycwb_g2g_work_bytes(const pycwb_g2g_desc* d, size_t* bytes) {
  if (!bytes) return PYCWB_ERR_ARG;
  int rc = g2g_check(d);
  if (rc != PYCWB_OK) return rc;
  G2gLayout L;
  if ((rc = g2g_layout(d, &L)) != PYCWB_OK) return rc;
  *bytes = L.total;
  return PYCWB_OK;
}

int pycwb_g2g_phase(const pycwb_g2g_desc* d, int phase, void* stream) {
  int rc = g2g_check(d);
  if (rc != PYCWB_OK) return rc;
  G2gLayout L;
  if ((rc = g2g_layout(d, &L)) != PYCWB_OK) return rc;
  if (!d->work || d->work_bytes < L.total) return PYCWB_ERR_ARG;
  const hipStream_t s = (hipStream_t)stream;
  char* w = (char*)d->work;
  unsigned long long* partial = (unsigned long long*)(w + L.partial);
  const int64_t win = d->nR - d->nL + 1;
  (void)hipGetLastError();
  const unsigned grid = (unsigned)((d->n + kBlock - 1) / kBlock < 4096
                                       ? (d->n + kBlock - 1) / kBlock : 4096);
  switch (phase) {
    case 1:  // count(flat > 0.001) over the map; the window flat[nL .. nR] sorted
      if (!d->count || !d->sorted) return PYCWB_ERR_ARG;
      if ((rc = g2g_count(d->flat, 0, d->n, 1.0e-3, partial,
                          (unsigned long long*)d->count, s)) != PYCWB_OK)
        return rc;
      return g2g_sort(d->flat + d->nL, win, d->sorted, w + L.temp, L.temp_bytes, s);
    case 2: {  // region values in (lo, hi), in order; numpy-order sums and log sums
      if (!d->chunk_sums || !d->n_valid) return PYCWB_ERR_ARG;
      size_t tb = L.temp_bytes;
      double* compact = (double*)(w + L.compact);
      const hipError_t e = hipcub::DeviceSelect::If(w + L.temp, tb, d->flat + d->nL, compact,
                                                    d->n_valid, (int)(win - 1),
                                                    Inside{d->lo, d->hi}, s);
      if (e != hipSuccess) {
        record_error("gamma-to-gauss select", hipGetErrorString(e));
        return PYCWB_ERR_HIP;
      }
      const int64_t chunks = (win - 1 + kNpChunk - 1) / kNpChunk;
      hipLaunchKernelGGL(k_th_chunks, dim3((unsigned)(chunks < 1 ? 1 : chunks)),
                         dim3(kChunkBlock), 0, s, (const double*)compact,
                         (const int64_t*)d->n_valid, d->chunk_sums);
      return check_launch("k_th_chunks");
    }
    case 3:  // the transformed map; count(out[nL .. n - nL) > 1e-5); its window sorted
      if (!d->out || !d->count || !d->sorted) return PYCWB_ERR_ARG;
      hipLaunchKernelGGL(k_g2g_map, dim3(grid), dim3(kBlock), 0, s, d->flat, d->n, d->alp,
                         d->avr, d->ALP, d->out);
      if ((rc = check_launch("k_g2g_map")) != PYCWB_OK) return rc;
      if ((rc = g2g_count(d->out, d->nL, d->n - d->nL, 1.0e-5, partial,
                          (unsigned long long*)d->count + 1, s)) != PYCWB_OK)
        return rc;
      return g2g_sort(d->out + d->nL, win, d->sorted, w + L.temp, L.temp_bytes, s);
    case 4:  // out *= scale
      if (!d->out) return PYCWB_ERR_ARG;
      hipLaunchKernelGGL(k_g2g_scale, dim3(grid), dim3(kBlock), 0, s, d->out, d->n, d->scale);
      return check_launch("k_g2g_scale");
    default:
      return PYCWB_ERR_ARG;
  }
}

}  // extern "C"
