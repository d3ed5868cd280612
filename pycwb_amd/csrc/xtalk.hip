// Cross-talk CSR of many clusters from the WDM cross-talk catalog.
//
// Restates getXTalk_pixels_fast (pycwb/modules/xtalk/monster.py:240-440): for
// every pixel i (at resolution r_i, time bin t, frequency f) and every
// resolution r_j <= r_i, the catalog bucket (r_i, r_j, f, t & 1) lists
// neighbours by index relative to (t - odd) * (L_i / L_j) * (L_j + 1); a
// neighbour that is a pixel of the cluster (matched on (L_j + 1, index)) gets
// an entry in row i and, when r_j < r_i, the swapped entry in its own row.
//
// GPU layout: one thread per pixel, all clusters of a batch at once.
//   1. k_xt_hash: per-cluster open-addressing table (nLayer, time) -> pixel,
//      later pixels overwriting earlier ones (the reference's dictionary).
//   2. k_xt_walk<COUNT>: the catalog walk; counts entries per row (atomics).
//   3. k_xt_scan_*: exclusive scan of the counts -> row starts.
//   4. k_xt_walk<FILL>: the same walk writes entries at per-row cursors with a
//      sort key (source pixel, walk position).
//   5. k_xt_sort_rows: each row sorted by that key, which is the order in
//      which the reference's fill loop writes it.
// Entry values are float32 in the reference's expression order.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "skyscan_lane.hh"
#include "pycwb_amd.h"

using namespace pycwb;

namespace {

constexpr int kBlock = 256;
constexpr int kScanBlock = 1024;

struct XtArgs {
  int n_clusters;
  const int64_t* pix_off;
  const int32_t* nlayer;
  const int64_t* time;
  const int64_t* hash_off;
  int n_res;
  int max_layer;
  const int32_t* layers;
  const int32_t* table;
  const float* coeff;
  int64_t n_coeff;
  int64_t* hkey;
  int32_t* hval;
  int32_t* counts;
  int64_t* row_off;
  int64_t* skey;
  int64_t* block_sums;
  int32_t* lookup;
  int64_t* xt_off;
  float* xt;
  int64_t xt_capacity;
};

__device__ __forceinline__ int64_t pix_key(int64_t nlayer, int64_t t) {
  return nlayer * ((int64_t)1 << 32) + t;
}

__device__ __forceinline__ uint64_t mix(int64_t k) {
  uint64_t x = (uint64_t)k;
  x ^= x >> 33;
  x *= 0xff51afd7ed558ccdull;
  x ^= x >> 33;
  x *= 0xc4ceb9fe1a85ec53ull;
  x ^= x >> 33;
  return x;
}

// Cluster of global pixel g (binary search in pix_off).
__device__ int cluster_of(const XtArgs& a, int64_t g) {
  int lo = 0, hi = a.n_clusters - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (a.pix_off[mid] <= g) lo = mid; else hi = mid - 1;
  }
  return lo;
}

__global__ void k_xt_init(XtArgs a, int64_t n_slots, int64_t n_pix) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g < n_slots) {
    a.hkey[g] = -1;
    a.hval[g] = -1;
  }
  if (g < n_pix) a.counts[g] = 0;
}

__global__ void k_xt_hash(XtArgs a, int64_t n_pix) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= n_pix) return;
  const int c = cluster_of(a, g);
  const int local = (int)(g - a.pix_off[c]);
  const int64_t h0 = a.hash_off[c];
  const uint64_t mask = (uint64_t)(a.hash_off[c + 1] - h0) - 1;
  const int64_t key = pix_key(a.nlayer[g], a.time[g]);
  uint64_t s = mix(key) & mask;
  for (uint64_t probe = 0; probe <= mask; ++probe) {
    unsigned long long* slot = (unsigned long long*)(a.hkey + h0 + s);
    const unsigned long long prev = atomicCAS(slot, (unsigned long long)-1ll,
                                              (unsigned long long)key);
    if (prev == (unsigned long long)-1ll || prev == (unsigned long long)key) {
      atomicMax(a.hval + h0 + s, local);
      return;
    }
    s = (s + 1) & mask;
  }
}

__device__ int find_pixel(const XtArgs& a, int c, int64_t key) {
  const int64_t h0 = a.hash_off[c];
  const uint64_t mask = (uint64_t)(a.hash_off[c + 1] - h0) - 1;
  uint64_t s = mix(key) & mask;
  for (uint64_t probe = 0; probe <= mask; ++probe) {
    const int64_t k = a.hkey[h0 + s];
    if (k == key) return a.hval[h0 + s];
    if (k == -1) return -1;
    s = (s + 1) & mask;
  }
  return -1;
}

__device__ __forceinline__ int res_of(const XtArgs& a, int nlayer) {
  int r = -1;
  for (int i = 0; i < a.n_res; ++i)
    if (a.layers[i] == nlayer - 1) r = i;
  return r;
}

// One pixel's catalog walk (monster.py:335-357 and :377-432).  COUNT adds to
// the per-row counts; otherwise entries go to the rows' cursors (counts
// reused as cursors after the scan) with their sort keys.
template <bool COUNT>
__global__ void k_xt_walk(XtArgs a, int64_t n_pix) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= n_pix) return;
  const int c = cluster_of(a, g);
  const int64_t p0 = a.pix_off[c];
  const int i = (int)(g - p0);
  const int ri = res_of(a, a.nlayer[g]);
  if (ri < 0) return;
  const int64_t layer_i = a.layers[ri];
  const int64_t t = a.time[g];
  const int64_t t1 = t / (layer_i + 1);
  const int64_t f1 = t - t1 * (layer_i + 1);
  const int64_t odd = t1 & 1;
  for (int rj = 0; rj <= ri; ++rj) {
    const int64_t layer_j = a.layers[rj];
    const int64_t base = (t1 - odd) * (layer_i / layer_j) * (layer_j + 1);
    const int64_t tb = ((((int64_t)ri * a.n_res + rj) * (a.max_layer + 1) + f1) * 2 + odd) * 2;
    const int64_t e0 = a.table[tb], e1 = a.table[tb + 1];
    for (int64_t e = e0; e < e1; ++e) {
      const float* ce = a.coeff + 5 * e;
      const int64_t indx2 = (int64_t)ce[0] + base;
      if (indx2 < 0) continue;
      const int j = find_pixel(a, c, pix_key(layer_j + 1, indx2));
      if (j < 0) continue;
      if (COUNT) {
        atomicAdd(a.counts + p0 + i, 1);
        if (rj < ri) atomicAdd(a.counts + p0 + j, 1);
        continue;
      }
      const float c1 = ce[1], c2 = ce[2], c3 = ce[3], c4 = ce[4];
      const int64_t key = (int64_t)i * ((int64_t)a.n_res * a.n_coeff) + (int64_t)rj * a.n_coeff + e;
      {
        const int64_t pos = a.row_off[p0 + i] + atomicAdd(a.counts + p0 + i, 1);
        if (pos < a.xt_capacity) {
          float* r = a.xt + 8 * pos;
          r[0] = (float)j;
          r[1] = c1 * c1 + c2 * c2;
          r[2] = c3 * c3 + c4 * c4;
          r[3] = c1 * c1 + c2 * c2 + c3 * c3 + c4 * c4;
          r[4] = c1;
          r[5] = c3;
          r[6] = c2;
          r[7] = c4;
          a.skey[pos] = key;
        }
      }
      if (rj < ri) {
        const int64_t pos = a.row_off[p0 + j] + atomicAdd(a.counts + p0 + j, 1);
        if (pos < a.xt_capacity) {
          float* r = a.xt + 8 * pos;
          r[0] = (float)i;
          r[1] = c1 * c1 + c3 * c3;
          r[2] = c2 * c2 + c4 * c4;
          r[3] = c1 * c1 + c2 * c2 + c3 * c3 + c4 * c4;
          r[4] = c1;
          r[5] = c2;
          r[6] = c3;
          r[7] = c4;
          a.skey[pos] = key;
        }
      }
    }
  }
}

// Exclusive scan of counts -> row_off (n_pix + 1 entries), three phases.
__global__ __launch_bounds__(kScanBlock) void k_xt_scan_blocks(XtArgs a, int64_t n_pix) {
  __shared__ int64_t s[kScanBlock];
  const int64_t g = (int64_t)blockIdx.x * kScanBlock + threadIdx.x;
  const int64_t v = g < n_pix ? (int64_t)a.counts[g] : 0;
  s[threadIdx.x] = v;
  __syncthreads();
  for (int off = 1; off < kScanBlock; off <<= 1) {
    const int64_t t = threadIdx.x >= off ? s[threadIdx.x - off] : 0;
    __syncthreads();
    s[threadIdx.x] += t;
    __syncthreads();
  }
  if (g < n_pix) a.row_off[g] = s[threadIdx.x] - v;   // exclusive, within the block
  if (threadIdx.x == kScanBlock - 1) a.block_sums[blockIdx.x] = s[threadIdx.x];
}

// Single workgroup: exclusive scan of the block sums (chunked per thread).
__global__ __launch_bounds__(kScanBlock) void k_xt_scan_sums(XtArgs a, int64_t n_blocks,
                                                             int64_t n_pix) {
  __shared__ int64_t s[kScanBlock];
  const int64_t per = (n_blocks + kScanBlock - 1) / kScanBlock;
  const int64_t b0 = (int64_t)threadIdx.x * per;
  int64_t tot = 0;
  for (int64_t b = b0; b < b0 + per && b < n_blocks; ++b) tot += a.block_sums[b];
  s[threadIdx.x] = tot;
  __syncthreads();
  for (int off = 1; off < kScanBlock; off <<= 1) {
    const int64_t t = threadIdx.x >= off ? s[threadIdx.x - off] : 0;
    __syncthreads();
    s[threadIdx.x] += t;
    __syncthreads();
  }
  int64_t run = s[threadIdx.x] - tot;
  for (int64_t b = b0; b < b0 + per && b < n_blocks; ++b) {
    const int64_t x = a.block_sums[b];
    a.block_sums[b] = run;
    run += x;
  }
  if (threadIdx.x == kScanBlock - 1) a.row_off[n_pix] = s[threadIdx.x];
}

__global__ void k_xt_scan_add(XtArgs a, int64_t n_pix) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g < n_pix) {
    a.row_off[g] += a.block_sums[g / kScanBlock];
    a.counts[g] = 0;   // cursors for the fill walk
  }
}

__global__ void k_xt_cluster_offsets(XtArgs a) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c <= a.n_clusters) a.xt_off[c] = a.row_off[a.pix_off[c]];
}

// Row i sorted by key (insertion sort; rows hold tens of entries), then its
// [start, end) relative to the cluster's first entry.
__global__ void k_xt_sort_rows(XtArgs a, int64_t n_pix) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= n_pix) return;
  const int64_t r0 = a.row_off[g], r1 = a.row_off[g + 1];
  for (int64_t p = r0 + 1; p < r1 && r1 <= a.xt_capacity; ++p) {
    const int64_t k = a.skey[p];
    float v[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) v[q] = a.xt[8 * p + q];
    int64_t m = p - 1;
    while (m >= r0 && a.skey[m] > k) {
      a.skey[m + 1] = a.skey[m];
#pragma unroll
      for (int q = 0; q < 8; ++q) a.xt[8 * (m + 1) + q] = a.xt[8 * m + q];
      --m;
    }
    a.skey[m + 1] = k;
#pragma unroll
    for (int q = 0; q < 8; ++q) a.xt[8 * (m + 1) + q] = v[q];
  }
  const int c = cluster_of(a, g);
  const int64_t base = a.row_off[a.pix_off[c]];
  a.lookup[2 * g] = (int32_t)(r0 - base);
  a.lookup[2 * g + 1] = (int32_t)(r1 - base);
}

int check_launch(const char* what) {
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    record_error(what, hipGetErrorString(e));
    return PYCWB_ERR_HIP;
  }
  return PYCWB_OK;
}

XtArgs args_of(const pycwb_xtalk_desc* d) {
  XtArgs a;
  a.n_clusters = d->n_clusters; a.pix_off = d->pix_off; a.nlayer = d->nlayer; a.time = d->time;
  a.hash_off = d->hash_off; a.n_res = d->n_res; a.max_layer = d->max_layer; a.layers = d->layers;
  a.table = d->table; a.coeff = d->coeff; a.n_coeff = d->n_coeff; a.hkey = d->hash_keys;
  a.hval = d->hash_vals; a.counts = d->counts; a.row_off = d->row_off; a.skey = d->sort_keys;
  a.block_sums = d->block_sums; a.lookup = d->lookup; a.xt_off = d->xt_off; a.xt = d->xt;
  a.xt_capacity = d->xt_capacity;
  return a;
}

unsigned blocks_for(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

}  // namespace

extern "C" {

size_t pycwb_xtalk_scan_blocks(int64_t n_pixels) {
  return (size_t)((n_pixels + kScanBlock - 1) / kScanBlock + 1);
}

int pycwb_xtalk_count(const pycwb_xtalk_desc* d, void* stream) {
  if (!d || d->n_clusters < 1 || d->n_pixels < 0 || d->n_slots < 1 || !d->pix_off ||
      !d->nlayer || !d->time || !d->hash_off || d->n_res < 1 || d->max_layer < 0 ||
      !d->layers || !d->table || !d->coeff || d->n_coeff < 0 || !d->hash_keys ||
      !d->hash_vals || !d->counts || !d->row_off || !d->block_sums || !d->xt_off)
    return PYCWB_ERR_ARG;
  (void)hipGetLastError();
  hipStream_t st = (hipStream_t)stream;
  const XtArgs a = args_of(d);
  const int64_t n = d->n_pixels;
  const int64_t init_n = n > d->n_slots ? n : d->n_slots;
  hipLaunchKernelGGL(k_xt_init, dim3(blocks_for(init_n, kBlock)), dim3(kBlock), 0, st, a,
                     d->n_slots, n);
  int rc = check_launch("k_xt_init");
  if (rc != PYCWB_OK) return rc;
  if (n > 0) {
    hipLaunchKernelGGL(k_xt_hash, dim3(blocks_for(n, kBlock)), dim3(kBlock), 0, st, a, n);
    if ((rc = check_launch("k_xt_hash")) != PYCWB_OK) return rc;
    hipLaunchKernelGGL(k_xt_walk<true>, dim3(blocks_for(n, kBlock)), dim3(kBlock), 0, st, a,
                       n);
    if ((rc = check_launch("k_xt_walk<count>")) != PYCWB_OK) return rc;
  }
  const int64_t nb = (n + kScanBlock - 1) / kScanBlock;
  if (nb > 0) {
    hipLaunchKernelGGL(k_xt_scan_blocks, dim3((unsigned)nb), dim3(kScanBlock), 0, st, a, n);
    if ((rc = check_launch("k_xt_scan_blocks")) != PYCWB_OK) return rc;
  }
  hipLaunchKernelGGL(k_xt_scan_sums, dim3(1), dim3(kScanBlock), 0, st, a, nb, n);
  if ((rc = check_launch("k_xt_scan_sums")) != PYCWB_OK) return rc;
  if (n > 0) {
    hipLaunchKernelGGL(k_xt_scan_add, dim3(blocks_for(n, kBlock)), dim3(kBlock), 0, st, a, n);
    if ((rc = check_launch("k_xt_scan_add")) != PYCWB_OK) return rc;
  }
  hipLaunchKernelGGL(k_xt_cluster_offsets, dim3(blocks_for(d->n_clusters + 1, kBlock)),
                     dim3(kBlock), 0, st, a);
  return check_launch("k_xt_cluster_offsets");
}

int pycwb_xtalk_fill(const pycwb_xtalk_desc* d, void* stream) {
  if (!d || d->n_clusters < 1 || d->n_pixels < 0 || !d->counts || !d->row_off ||
      !d->sort_keys || !d->lookup || !d->xt || d->xt_capacity < 0)
    return PYCWB_ERR_ARG;
  (void)hipGetLastError();
  hipStream_t st = (hipStream_t)stream;
  const XtArgs a = args_of(d);
  const int64_t n = d->n_pixels;
  if (n == 0) return PYCWB_OK;
  hipLaunchKernelGGL(k_xt_walk<false>, dim3(blocks_for(n, kBlock)), dim3(kBlock), 0, st, a, n);
  int rc = check_launch("k_xt_walk<fill>");
  if (rc != PYCWB_OK) return rc;
  hipLaunchKernelGGL(k_xt_sort_rows, dim3(blocks_for(n, kBlock)), dim3(kBlock), 0, st, a, n);
  return check_launch("k_xt_sort_rows");
}

}  // extern "C"
