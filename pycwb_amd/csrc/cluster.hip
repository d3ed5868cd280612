// Connected-component clustering of the selected network pixels, many lags
// at once.
//
// Restates cluster_pixels (pycwb/modules/coherence_native/clustering.py:
// 18-157) with _label_components_grid (kernels.py:24-105) and
// _subnet_subrho_batch_numba (kernels.py:141-229).  Two pixels of one lag are
// joined when |df| <= kf and |dt| <= kt, and pixels of one TF cell are always
// joined; pixels outside the (n_freq, n_time) grid stay alone.  A component's
// label is the rank of its smallest pixel index among the lag's components
// (the reference's union-find keeps the smaller root, so its roots are the
// components' smallest indices and its labels follow them in pixel order).
//
// Kernels (every lag of a call in one grid each):
//   k_cl_fill     cell    (pycwb_cluster_map_init) the index map set to -1
//   k_cl_init     pixel   lag of each pixel (binary search of lag_off), parent
//                         and sort values set to the identity
//   k_cl_scatter  pixel   atomicMax of the pixel index into its cell: the cell
//                         holds its last pixel, as the reference's map does
//                         (the caller's index map holds -1 everywhere on
//                         entry: pycwb_cluster_map_init once, then every call
//                         restores it)
//   k_cl_union    pixel   lock-free union-find over the (2kf+1) x (2kt+1)
//                         window; the larger root is linked under the smaller
//                         by compare-and-swap, so a root is always the
//                         component's smallest index
//   k_cl_root     pixel   root of each pixel; root flags
//   (scan)                exclusive sum of the root flags: the global
//                         cluster id of every root
//   k_cl_label    pixel   cluster id and 1-based label of every pixel
//   (sort)                stable radix sort of (cluster id, pixel index): the
//                         members of each cluster in pixel order
//   k_cl_members  slot    members and cluster offsets
//   k_cl_unscatter pixel  the pixels' cells set back to -1
//   k_cl_terms    slot    per-pixel subnet/subrho terms (kernels.py:150-189)
//   k_cl_stats    cluster energy, c_time, c_freq (numpy's add.reduce order,
//                         np_sum.hh), subnet and subrho: the terms summed in
//                         member order (kernels.py:141-191)
// The scan and the sort are hipCUB's device-wide primitives.
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

struct ClArgs {
  int n_lag;
  const int64_t* lag_off;
  int64_t n;
  const int64_t* freq;
  const int64_t* time;
  const double* energy;
  const double* det_energy;
  const double* noise_rms;
  int n_ifo, n_freq, n_time, kf, kt;
  double dt, half_rate, n_sub;
  // work
  int32_t* map;
  int32_t* lag_of;
  int32_t* parent;
  int32_t* root;
  int32_t* flag;
  int32_t* rank;
  uint32_t* key_in;
  uint32_t* key_out;
  int32_t* val_in;
  int32_t* val_out;
  double* terms;  // (n, 4) per-pixel subnet/subrho terms in member order
  // outputs
  int64_t* label;
  int64_t* lag_cluster_off;
  int64_t* cluster_off;
  int64_t* members;
  double *c_energy, *c_time, *c_freq, *c_subnet, *c_subrho;
};

__device__ __forceinline__ int32_t ld(const int32_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ void st(int32_t* p, int32_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Root of x with path halving.  parent[x] <= x always holds and a stored value
// is always an ancestor of x, so concurrent halving and linking keep every
// path acyclic and ending at the component's root.
__device__ int32_t find(int32_t* parent, int32_t x) {
  int32_t p = ld(parent + x);
  while (p != x) {
    const int32_t g = ld(parent + p);
    if (g != p) st(parent + x, g);
    x = p;
    p = ld(parent + x);
  }
  return x;
}

// Links the larger root under the smaller.  A failed compare-and-swap means
// another thread linked that root first; every retry follows such a link, so
// the loop ends after at most n - 1 of them.
__device__ void unite(int32_t* parent, int32_t a, int32_t b) {
  for (;;) {
    a = find(parent, a);
    b = find(parent, b);
    if (a == b) return;
    if (a > b) {
      const int32_t t = a;
      a = b;
      b = t;
    }
    int32_t expect = b;
    if (__hip_atomic_compare_exchange_strong(parent + b, &expect, a, __ATOMIC_RELAXED,
                                             __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
      return;
  }
}

__device__ __forceinline__ bool on_grid(const ClArgs& a, int64_t f, int64_t t) {
  return f >= 0 && f < a.n_freq && t >= 0 && t < a.n_time;
}

__device__ __forceinline__ int64_t cell(const ClArgs& a, int lag, int64_t f, int64_t t) {
  return ((int64_t)lag * a.n_freq + f) * a.n_time + t;
}

__global__ void k_cl_init(ClArgs a) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= a.n) return;
  int lo = 0, hi = a.n_lag;  // last lag with lag_off[lag] <= i
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (a.lag_off[mid] <= i) lo = mid;
    else hi = mid;
  }
  a.lag_of[i] = lo;
  a.parent[i] = (int32_t)i;
  a.val_in[i] = (int32_t)i;
}

__global__ void k_cl_fill(int32_t* map, int64_t size) {
  const int64_t stride = (int64_t)gridDim.x * kBlock;
  for (int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x; k < size; k += stride) map[k] = -1;
}

__global__ void k_cl_scatter(ClArgs a) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= a.n) return;
  const int64_t f = a.freq[i], t = a.time[i];
  if (!on_grid(a, f, t)) return;
  atomicMax(a.map + cell(a, a.lag_of[i], f, t), (int32_t)i);
}

__global__ void k_cl_unscatter(ClArgs a) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= a.n) return;
  const int64_t f = a.freq[i], t = a.time[i];
  if (!on_grid(a, f, t)) return;
  a.map[cell(a, a.lag_of[i], f, t)] = -1;
}

__global__ void k_cl_union(ClArgs a) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= a.n) return;
  const int64_t f = a.freq[i], t = a.time[i];
  if (!on_grid(a, f, t)) return;
  const int lag = a.lag_of[i];
  const int64_t f0 = f - a.kf < 0 ? 0 : f - a.kf;
  const int64_t f1 = f + a.kf + 1 > a.n_freq ? a.n_freq : f + a.kf + 1;
  const int64_t t0 = t - a.kt < 0 ? 0 : t - a.kt;
  const int64_t t1 = t + a.kt + 1 > a.n_time ? a.n_time : t + a.kt + 1;
  for (int64_t ff = f0; ff < f1; ++ff) {
    const int32_t* row = a.map + cell(a, lag, ff, 0);
    for (int64_t tt = t0; tt < t1; ++tt) {
      const int32_t j = row[tt];
      // neighbours below i (the pair is seen from its larger end), and the
      // cell's last pixel for the other pixels of i's own cell
      if (j >= 0 && (j < i || (ff == f && tt == t && j != i))) unite(a.parent, (int32_t)i, j);
    }
  }
}

__global__ void k_cl_root(ClArgs a) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i > a.n) return;
  if (i == a.n) {
    a.flag[i] = 0;
    return;
  }
  const int32_t r = find(a.parent, (int32_t)i);
  a.root[i] = r;
  a.flag[i] = r == i ? 1 : 0;
}

__global__ void k_cl_label(ClArgs a) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i <= a.n_lag) a.lag_cluster_off[i] = a.rank[a.lag_off[i]];
  if (i >= a.n) return;
  const int32_t cid = a.rank[a.root[i]];
  a.key_in[i] = (uint32_t)cid;
  a.label[i] = (int64_t)cid - a.rank[a.lag_off[a.lag_of[i]]] + 1;
}

__global__ void k_cl_members(ClArgs a) {
  const int64_t p = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (p >= a.n) return;
  a.members[p] = a.val_out[p];
  const uint32_t k = a.key_out[p];
  if (p == 0 || a.key_out[p - 1] != k) a.cluster_off[k] = p;
  if (p == a.n - 1) a.cluster_off[k + 1] = a.n;
}

__global__ void k_cl_empty(ClArgs a) {
  const int i = threadIdx.x + blockIdx.x * kBlock;
  if (i <= a.n_lag) a.lag_cluster_off[i] = 0;
  if (i == 0) a.cluster_off[0] = 0;
}

struct EnergyOf {
  const double* e;
  const int64_t* m;
  __device__ double operator()(int64_t k) const { return e[m[k]]; }
};

struct TimeOf {  // t * dt - dt / 2
  const int64_t* t;
  const int64_t* m;
  double dt, hdt;
  __device__ double operator()(int64_t k) const { return (double)t[m[k]] * dt - hdt; }
};

struct FreqOf {  // f * (rate / 2)
  const int64_t* f;
  const int64_t* m;
  double s;
  __device__ double operator()(int64_t k) const { return (double)f[m[k]] * s; }
};

// Per-pixel terms of _subnet_subrho_numba (kernels.py:150-189), one thread per
// member slot p (pixel members[p], asnr = sqrt of its detector energies):
// terms[4p .. 4p+3] = (rho term, e_tot - e_max, e_max, subnet term).  The
// cluster sums add them in member order, as the reference's loop does.
__global__ void k_cl_terms(ClArgs a) {
  const int64_t q = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (q >= a.n) return;
  const int64_t p = a.members[q];
  const int n_ifo = a.n_ifo;
  const double n_sub = a.n_sub;
  double amp_sum = 0.0, e_max = 0.0, e_tot = 0.0, nsd = 0.0, msd = 0.0;
  for (int d = 0; d < n_ifo; ++d) {
    const double amp = sqrt(a.det_energy[p * n_ifo + d]);
    const double x = amp * amp;
    const double r = a.noise_rms ? a.noise_rms[p * n_ifo + d] : 1.0;
    const double v = r * r;
    amp_sum += fabs(amp);
    if (x > e_max) {
      e_max = x;
      msd = v;
    }
    e_tot += x;
    if (v > 0.0) nsd += 1.0 / v;
  }
  const double a_fac = amp_sum > 0.0 ? e_tot / (amp_sum * amp_sum) : 1.0;
  const double y = e_tot - e_max;
  const double x_corr = y * (1.0 + y / (e_max + 1.0e-5));
  if (msd > 0.0) nsd -= 1.0 / msd;
  const double v_corr = (2.0 * e_max - e_tot) * msd * nsd / 10.0;
  const double a_cut = (x_corr + n_sub) != 0.0 ? x_corr / (x_corr + n_sub) : 0.0;
  const double denom = x_corr + (v_corr > 0.0 ? v_corr : 1.0e-5);
  double* t = a.terms + 4 * q;
  t[0] = (1.0 - a_fac) * (e_tot - n_sub * 2.0);
  t[1] = e_tot - e_max;
  t[2] = e_max;
  t[3] = (e_tot * x_corr / denom) * (a_cut > 0.5 ? a_cut : 0.0);
}

__global__ void k_cl_stats(ClArgs a) {
  const int64_t c = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  const int64_t total = a.rank[a.n];
  if (c >= total) return;
  const int64_t b = a.cluster_off[c], m = a.cluster_off[c + 1] - b;
  const int64_t* mem = a.members + b;
  a.c_energy[c] = np_reduce(EnergyOf{a.energy, mem}, m);
  a.c_time[c] = np_reduce(TimeOf{a.time, mem, a.dt, a.dt / 2.0}, m) / (double)m;
  a.c_freq[c] = np_reduce(FreqOf{a.freq, mem, a.half_rate}, m) / (double)m;
  if (a.n_ifo < 2) {
    a.c_subnet[c] = 0.0;
    a.c_subrho[c] = 0.0;
    return;
  }
  // kernels.py:141-191: the loop's four running sums over the pixel terms
  double rho = 0.0, e_sub_total = 0.0, e_max_sum = 0.0, acc = 0.0;
  const double* t = a.terms + 4 * b;
  for (int64_t k = 0; k < m; ++k) {
    rho += t[4 * k];
    e_sub_total += t[4 * k + 1];
    e_max_sum += t[4 * k + 2];
    acc += t[4 * k + 3];
  }
  a.c_subnet[c] = acc / (e_max_sum + e_sub_total + 0.01);
  a.c_subrho[c] = rho >= 0.0 ? sqrt(rho) : __builtin_nan("");
}

int check_launch(const char* what) {
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    record_error(what, hipGetErrorString(e));
    return PYCWB_ERR_HIP;
  }
  return PYCWB_OK;
}

int check_hip(hipError_t e, const char* what) {
  if (e != hipSuccess) {
    record_error(what, hipGetErrorString(e));
    return PYCWB_ERR_HIP;
  }
  return PYCWB_OK;
}

size_t align_up(size_t x) { return (x + 255) & ~(size_t)255; }

int key_bits(int64_t n) {
  int b = 1;
  while (b < 32 && ((int64_t)1 << b) < n) ++b;
  return b;
}

struct Layout {
  size_t lag_of, parent, root, flag, rank, key_in, key_out, val_in, val_out, terms, temp;
  size_t temp_bytes, total;
};

int layout(const pycwb_cluster_desc* d, Layout* L) {
  const int64_t n = d->n_pix;
  const int64_t cells = (int64_t)d->n_lag * d->n_freq * d->n_time;
  size_t o = 0;
  auto take = [&](size_t bytes) {
    const size_t at = o;
    o = align_up(o + bytes);
    return at;
  };
  L->lag_of = take(sizeof(int32_t) * (size_t)n);
  L->parent = take(sizeof(int32_t) * (size_t)n);
  L->root = take(sizeof(int32_t) * (size_t)n);
  L->flag = take(sizeof(int32_t) * (size_t)(n + 1));
  L->rank = take(sizeof(int32_t) * (size_t)(n + 1));
  L->key_in = take(sizeof(uint32_t) * (size_t)n);
  L->key_out = take(sizeof(uint32_t) * (size_t)n);
  L->val_in = take(sizeof(int32_t) * (size_t)n);
  L->val_out = take(sizeof(int32_t) * (size_t)n);
  L->terms = take(4 * sizeof(double) * (size_t)n);
  size_t scan_bytes = 0, sort_bytes = 0;
  if (n > 0) {
    if (hipcub::DeviceScan::ExclusiveSum(nullptr, scan_bytes, (const int32_t*)nullptr,
                                         (int32_t*)nullptr, (int)(n + 1)) != hipSuccess ||
        hipcub::DeviceRadixSort::SortPairs(nullptr, sort_bytes, (const uint32_t*)nullptr,
                                           (uint32_t*)nullptr, (const int32_t*)nullptr,
                                           (int32_t*)nullptr, (int)n, 0, key_bits(n)) !=
            hipSuccess) {
      record_error("pycwb_cluster_work_bytes", "hipCUB temporary-storage query failed");
      return PYCWB_ERR_HIP;
    }
  }
  L->temp_bytes = scan_bytes > sort_bytes ? scan_bytes : sort_bytes;
  L->temp = take(L->temp_bytes > 0 ? L->temp_bytes : 1);
  L->total = o;
  return PYCWB_OK;
}

int check_desc(const pycwb_cluster_desc* d) {
  if (!d || !d->lag_off || d->n_lag < 1 || d->n_lag > 65535 || d->n_pix < 0 ||
      d->n_pix >= INT32_MAX || d->n_freq < 1 || d->n_time < 1 || d->kf < 1 || d->kt < 1 ||
      d->n_ifo < 0)
    return PYCWB_ERR_ARG;
  if (!d->lag_cluster_off || !d->cluster_off) return PYCWB_ERR_ARG;
  if (d->n_pix > 0 && (!d->freq || !d->time || !d->energy || !d->label || !d->members ||
                       !d->c_energy || !d->c_time || !d->c_freq || !d->c_subnet ||
                       !d->c_subrho || (d->n_ifo > 1 && !d->det_energy)))
    return PYCWB_ERR_ARG;
  return PYCWB_OK;
}

}  // namespace

extern "C" {

int pycwb_cluster_map_init(int32_t* index_map, int64_t cells, void* stream) {
  if (!index_map || cells < 1) return PYCWB_ERR_ARG;
  (void)hipGetLastError();
  const int64_t blocks = (cells + kBlock - 1) / kBlock;
  hipLaunchKernelGGL(k_cl_fill, dim3((unsigned)(blocks < 16384 ? blocks : 16384)), dim3(kBlock),
                     0, (hipStream_t)stream, index_map, cells);
  return check_launch("k_cl_fill");
}

int pycwb_cluster_work_bytes(const pycwb_cluster_desc* d, size_t* bytes) {
  if (!bytes) return PYCWB_ERR_ARG;
  int rc = check_desc(d);
  if (rc != PYCWB_OK) return rc;
  Layout L;
  if ((rc = layout(d, &L)) != PYCWB_OK) return rc;
  *bytes = L.total;
  return PYCWB_OK;
}

int pycwb_cluster_pixels(const pycwb_cluster_desc* d, void* stream) {
  int rc = check_desc(d);
  if (rc != PYCWB_OK) return rc;
  Layout L;
  if ((rc = layout(d, &L)) != PYCWB_OK) return rc;
  if (d->n_pix > 0 && (!d->work || d->work_bytes < L.total || !d->index_map))
    return PYCWB_ERR_ARG;
  const hipStream_t s = (hipStream_t)stream;
  (void)hipGetLastError();
  ClArgs a;
  a.n_lag = d->n_lag; a.lag_off = d->lag_off; a.n = d->n_pix;
  a.freq = d->freq; a.time = d->time; a.energy = d->energy; a.det_energy = d->det_energy;
  a.noise_rms = d->noise_rms;
  a.n_ifo = d->n_ifo; a.n_freq = d->n_freq; a.n_time = d->n_time; a.kf = d->kf; a.kt = d->kt;
  a.dt = d->dt; a.half_rate = d->half_rate; a.n_sub = d->n_sub;
  char* w = (char*)d->work;
  a.map = d->index_map; a.lag_of = (int32_t*)(w + L.lag_of);
  a.parent = (int32_t*)(w + L.parent); a.root = (int32_t*)(w + L.root);
  a.flag = (int32_t*)(w + L.flag); a.rank = (int32_t*)(w + L.rank);
  a.key_in = (uint32_t*)(w + L.key_in); a.key_out = (uint32_t*)(w + L.key_out);
  a.val_in = (int32_t*)(w + L.val_in); a.val_out = (int32_t*)(w + L.val_out);
  a.terms = (double*)(w + L.terms);
  a.label = d->label; a.lag_cluster_off = d->lag_cluster_off; a.cluster_off = d->cluster_off;
  a.members = d->members;
  a.c_energy = d->c_energy; a.c_time = d->c_time; a.c_freq = d->c_freq;
  a.c_subnet = d->c_subnet; a.c_subrho = d->c_subrho;
  if (d->n_pix == 0) {
    hipLaunchKernelGGL(k_cl_empty, dim3((unsigned)((d->n_lag + kBlock) / kBlock)), dim3(kBlock),
                       0, s, a);
    return check_launch("k_cl_empty");
  }
  const int64_t n = d->n_pix;
  const unsigned gp = (unsigned)((n + kBlock - 1) / kBlock);
  const unsigned gp1 = (unsigned)((n + kBlock) / kBlock);
  hipLaunchKernelGGL(k_cl_init, dim3(gp), dim3(kBlock), 0, s, a);
  if ((rc = check_launch("k_cl_init")) != PYCWB_OK) return rc;
  hipLaunchKernelGGL(k_cl_scatter, dim3(gp), dim3(kBlock), 0, s, a);
  if ((rc = check_launch("k_cl_scatter")) != PYCWB_OK) return rc;
  hipLaunchKernelGGL(k_cl_union, dim3(gp), dim3(kBlock), 0, s, a);
  if ((rc = check_launch("k_cl_union")) != PYCWB_OK) return rc;
  hipLaunchKernelGGL(k_cl_unscatter, dim3(gp), dim3(kBlock), 0, s, a);
  if ((rc = check_launch("k_cl_unscatter")) != PYCWB_OK) return rc;
  hipLaunchKernelGGL(k_cl_root, dim3(gp1), dim3(kBlock), 0, s, a);
  if ((rc = check_launch("k_cl_root")) != PYCWB_OK) return rc;
  size_t tb = L.temp_bytes;
  if ((rc = check_hip(hipcub::DeviceScan::ExclusiveSum(w + L.temp, tb, a.flag, a.rank,
                                                       (int)(n + 1), s),
                      "cluster root scan")) != PYCWB_OK)
    return rc;
  const int64_t lab_threads = n > d->n_lag + 1 ? n : d->n_lag + 1;
  hipLaunchKernelGGL(k_cl_label, dim3((unsigned)((lab_threads + kBlock - 1) / kBlock)),
                     dim3(kBlock), 0, s, a);
  if ((rc = check_launch("k_cl_label")) != PYCWB_OK) return rc;
  tb = L.temp_bytes;
  if ((rc = check_hip(hipcub::DeviceRadixSort::SortPairs(w + L.temp, tb, a.key_in, a.key_out,
                                                         a.val_in, a.val_out, (int)n, 0,
                                                         key_bits(n), s),
                      "cluster member sort")) != PYCWB_OK)
    return rc;
  hipLaunchKernelGGL(k_cl_members, dim3(gp), dim3(kBlock), 0, s, a);
  if ((rc = check_launch("k_cl_members")) != PYCWB_OK) return rc;
  if (d->n_ifo >= 2) {
    hipLaunchKernelGGL(k_cl_terms, dim3(gp), dim3(kBlock), 0, s, a);
    if ((rc = check_launch("k_cl_terms")) != PYCWB_OK) return rc;
  }
  hipLaunchKernelGGL(k_cl_stats, dim3(gp), dim3(kBlock), 0, s, a);
  return check_launch("k_cl_stats");
}

}  // extern "C"
