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
//   k_th_values   value   sum over detectors, clip, wave-ballot counts
//   (sort)                hipCUB radix sort of the float64 values
//   (select)              hipCUB stable compaction of the values above 1e-3
//   k_th_chunks   chunk   one wave per 8192 values: lane 0 walks the pairwise
//                         tree and lists its leaves, the lanes sum the
//                         leaves, lane 0 adds them up the tree
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
constexpr int kMaxLeaves = 256;  // a chunk's leaves hold > 57 values: at most 143

struct ThArgs {
  const double* maps;
  int n_ifo, n_freq;
  int64_t n_time;
  int time_major;
  int64_t begin, count, ncol;
  double hi;
  double* values;
  unsigned long long* counts;
};

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

__global__ void k_th_values(ThArgs a) {
  const int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  bool above4 = false, above3 = false;
  if (k < a.count) {
    int64_t f, t;
    if (a.time_major) {
      const int64_t j = a.begin + k;
      t = j / a.n_freq;
      f = j - t * a.n_freq;
    } else {
      f = k / a.ncol;
      t = a.begin + (k - f * a.ncol);
    }
    const int64_t plane = (int64_t)a.n_freq * a.n_time;
    const double* p = a.maps + f * a.n_time + t;
    // np.sum(axis=0) starts from 0.0 (-0.0 sums to +0.0); the time-major
    // path (w = flat0.copy(); w += ...) starts from the first map
    double w = a.time_major ? p[0] : 0.0 + p[0];
    for (int d = 1; d < a.n_ifo; ++d) w = w + p[d * plane];
    const double v = clip(w, a.hi);
    a.values[k] = v;
    above4 = v > 1.0e-4;
    above3 = v > 1.0e-3;
  }
  const unsigned long long b4 = __ballot(above4), b3 = __ballot(above3);
  if ((threadIdx.x & (kWave - 1)) == 0 && (b4 | b3)) {
    atomicAdd(a.counts + 0, (unsigned long long)__popcll(b4));
    atomicAdd(a.counts + 1, (unsigned long long)__popcll(b3));
  }
}

struct Above {
  __device__ bool operator()(const double& x) const { return x > 1.0e-3; }
};

struct Plain {
  const double* x;
  __device__ double operator()(int64_t i) const { return x[i]; }
};

struct Log {
  const double* x;
  __device__ double operator()(int64_t i) const { return log(x[i]); }
};

struct Recorder {
  int64_t* s;
  int64_t* n;
  int k;
  __device__ double operator()(int64_t s0, int64_t n0) {
    if (k < kMaxLeaves) {
      s[k] = s0;
      n[k] = n0;
    }
    ++k;
    return 0.0;
  }
};

struct Replay {
  const double* v;
  int k;
  __device__ double operator()(int64_t, int64_t) { return v[k++]; }
};

// out[2c] = np.sum of chunk c of x[0 .. m), out[2c + 1] = np.sum of its logs.
__global__ void __launch_bounds__(kWave) k_th_chunks(const double* x, const int64_t* m_ptr,
                                                    double* out) {
  __shared__ int64_t ls[kMaxLeaves], ln[kMaxLeaves];
  __shared__ double lv[kMaxLeaves], lg[kMaxLeaves];
  __shared__ int n_leaf;
  const int64_t m = *m_ptr;
  const int64_t base = (int64_t)blockIdx.x * kNpChunk;
  if (base >= m) return;
  const int64_t len = m - base < kNpChunk ? m - base : kNpChunk;
  if (threadIdx.x == 0) {
    Recorder rec{ls, ln, 0};
    np_tree(len, rec);
    n_leaf = rec.k < kMaxLeaves ? rec.k : kMaxLeaves;
  }
  __syncthreads();
  const Plain g{x + base};
  const Log gl{x + base};
  for (int i = threadIdx.x; i < n_leaf; i += kWave) {
    lv[i] = np_leaf(g, ls[i], ln[i]);
    lg[i] = np_leaf(gl, ls[i], ln[i]);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    Replay r1{lv, 0};
    out[2 * blockIdx.x] = np_tree(len, r1);
    Replay r2{lg, 0};
    out[2 * blockIdx.x + 1] = np_tree(len, r2);
  }
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
  size_t compact, n_sel, temp, temp_bytes, total;
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
  L->temp_bytes = sort_bytes > sel_bytes ? sort_bytes : sel_bytes;
  L->temp = o;
  o = align_up(o + (L->temp_bytes ? L->temp_bytes : 1));
  L->total = o;
  return PYCWB_OK;
}

int check_desc(const pycwb_threshold_desc* d) {
  if (!d || !d->maps || !d->values || !d->counts || d->n_ifo < 1 || d->n_freq < 1 ||
      d->n_time < 1 || d->count < 0 || d->count >= INT32_MAX || d->begin < 0)
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
  a.hi = d->hi; a.values = d->values; a.counts = (unsigned long long*)d->counts;
  hipLaunchKernelGGL(k_th_values, dim3((unsigned)((d->count + kBlock - 1) / kBlock)),
                     dim3(kBlock), 0, s, a);
  if ((rc = check_launch("k_th_values")) != PYCWB_OK) return rc;
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
    hipLaunchKernelGGL(k_th_chunks, dim3((unsigned)chunks), dim3(kWave), 0, s,
                       (const double*)compact, (const int64_t*)n_sel, d->chunk_sums);
    if ((rc = check_launch("k_th_chunks")) != PYCWB_OK) return rc;
  }
  return PYCWB_OK;
}

}  // extern "C"
