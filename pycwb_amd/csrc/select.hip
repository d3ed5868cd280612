// Network-pixel selection of many lags at once.
//
// Restates select_network_pixels (pycwb/modules/coherence_native/selection.py:
// 144-243) with its kernels _align_threshold_map_numba (kernels.py:231-302),
// _candidate_passes_support_numba (:107-137) and _select_candidates_numba
// (:305-368): per lag, the detectors' energy maps summed at circularly
// shifted time bins (float64, detector order), clipped to the support map
// (0 below Eo, 2 Eo + 0.1 above 2 Eo), the core/halo test on each pixel's
// neighbourhood, and the selected pixels emitted in (frequency, time) order.
//
// Kernels (all lags of a call in one grid each):
//   k_sel_live     (lag, time)       shifted-veto live mask
//   k_sel_combined (lag, freq, time) support map
//   k_sel_rows     one block per (lag, freq) row: flags and the row's count
//   k_sel_scan     exclusive scan of the row counts -> row and lag offsets
//   k_sel_emit     one block per row: block-wide scan of the flags -> output
//                  positions, the per-detector payload
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "skyscan_lane.hh"
#include "pycwb_amd.h"

using namespace pycwb;

namespace {

constexpr int kBlock = 256;
constexpr int kScanBlock = 1024;

struct SelArgs {
  int n_ifo, n_freq;
  int64_t n_time;
  const double* energy;
  int n_lag;
  const int64_t* shift;
  const int16_t* veto;
  int edge_bins, ib, ie;
  double eo;
  double* comb;
  uint8_t* flags;
  uint8_t* live;
  int64_t* row_off;
  int64_t* lag_off;
  int64_t capacity;
  int64_t* freq;
  int64_t* time;
  double* value;
  double* det_energy;
  int64_t* det_index;
};

__device__ __forceinline__ int64_t src_bin(const SelArgs& a, int lag, int d, int64_t t) {
  const int64_t vs = a.edge_bins, nn = a.n_time - 2 * (int64_t)a.edge_bins;
  return vs + (t - vs + a.shift[(int64_t)lag * a.n_ifo + d]) % nn;
}

__global__ void k_sel_live(SelArgs a) {
  const int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  const int lag = blockIdx.y;
  if (t >= a.n_time) return;
  const int64_t vs = a.edge_bins, vstop = a.n_time - a.edge_bins;
  uint8_t ok = 0;
  if (t >= vs && t < vstop) {
    ok = 1;
    if (a.veto)
      for (int d = 0; d < a.n_ifo; ++d)
        if (a.veto[src_bin(a, lag, d, t)] == 0) {
          ok = 0;
          break;
        }
  }
  a.live[(int64_t)lag * a.n_time + t] = ok;
}

__global__ void k_sel_combined(SelArgs a) {
  const int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  const int f = blockIdx.y;
  const int lag = blockIdx.z;
  if (t >= a.n_time) return;
  const int64_t vs = a.edge_bins, vstop = a.n_time - a.edge_bins;
  double c = 0.0;
  if (f >= a.ib && t >= vs && t < vstop && !(a.veto && !a.live[(int64_t)lag * a.n_time + t])) {
    double v = 0.0;
    for (int d = 0; d < a.n_ifo; ++d)
      v += a.energy[((int64_t)d * a.n_freq + f) * a.n_time + src_bin(a, lag, d, t)];
    const double em = 2.0 * a.eo;
    if (!(v < a.eo)) c = v > em ? em + 0.1 : v;
  }
  a.comb[((int64_t)lag * a.n_freq + f) * a.n_time + t] = c;
}

__device__ __forceinline__ bool passes(const double* cm, int64_t n_time, int f, int64_t t,
                                       int ii, double eo, double em, double eh) {
#define CB(ff, tt) cm[(int64_t)(ff) * n_time + (tt)]
  const double e = CB(f, t);
  if (e < eo) return false;
  const double ct = CB(f + 1, t) + CB(f, t + 1) + CB(f + 1, t + 1);
  const double cb = CB(f - 1, t) + CB(f, t - 1) + CB(f - 1, t - 1);
  double ht = CB(f + 1, t + 2);
  if (f < ii) ht += CB(f + 2, t + 2) + CB(f + 2, t + 1);
  double hb = CB(f - 1, t - 2);
  if (f >= 2) hb += CB(f - 2, t - 2) + CB(f - 2, t - 1);
#undef CB
  return !((ct + cb) * e < eh && (ct + ht) * e < eh && (cb + hb) * e < eh && e < em);
}

__device__ __forceinline__ void sel_window(const SelArgs& a, int& fs, int& fe, int64_t& ts,
                                           int64_t& te) {
  const int64_t margin = a.edge_bins > 2 ? a.edge_bins : 2;
  ts = margin;
  te = a.n_time - margin;
  fs = a.ib;
  fe = a.ie > fs ? a.ie : fs;
  if (fe > a.n_freq - 1) fe = a.n_freq - 1;
}

// One block per (lag, frequency row): the row's flags and their count.
__global__ __launch_bounds__(kBlock) void k_sel_rows(SelArgs a) {
  __shared__ int s_cnt[kBlock / 64];
  const int f = blockIdx.x, lag = blockIdx.y;
  int fs, fe;
  int64_t ts, te;
  sel_window(a, fs, fe, ts, te);
  const double* cm = a.comb + (int64_t)lag * a.n_freq * a.n_time;
  uint8_t* fl = a.flags + ((int64_t)lag * a.n_freq + f) * a.n_time;
  const double em = 2.0 * a.eo, eh = em * em;
  int cnt = 0;
  for (int64_t t = threadIdx.x; t < a.n_time; t += kBlock) {
    uint8_t k = 0;
    if (f >= fs && f < fe && t >= ts && t < te) k = passes(cm, a.n_time, f, t, a.n_freq - 2,
                                                           a.eo, em, eh) ? 1 : 0;
    fl[t] = k;
    cnt += k;
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) cnt += __shfl_xor(cnt, off, 64);
  if ((threadIdx.x & 63) == 0) s_cnt[threadIdx.x >> 6] = cnt;
  __syncthreads();
  if (threadIdx.x == 0) {
    int tot = 0;
    for (int w = 0; w < kBlock / 64; ++w) tot += s_cnt[w];
    a.row_off[(int64_t)lag * a.n_freq + f] = tot;
  }
}

// Single workgroup: exclusive scan of the n_lag * n_freq row counts in place,
// total at [rows]; lag_off[l] = first pixel of lag l.
__global__ __launch_bounds__(kScanBlock) void k_sel_scan(SelArgs a) {
  __shared__ int64_t s[kScanBlock];
  const int64_t rows = (int64_t)a.n_lag * a.n_freq;
  const int64_t per = (rows + kScanBlock - 1) / kScanBlock;
  const int64_t b0 = (int64_t)threadIdx.x * per;
  int64_t tot = 0;
  for (int64_t r = b0; r < b0 + per && r < rows; ++r) tot += a.row_off[r];
  s[threadIdx.x] = tot;
  __syncthreads();
  for (int off = 1; off < kScanBlock; off <<= 1) {
    const int64_t v = threadIdx.x >= off ? s[threadIdx.x - off] : 0;
    __syncthreads();
    s[threadIdx.x] += v;
    __syncthreads();
  }
  int64_t run = s[threadIdx.x] - tot;
  for (int64_t r = b0; r < b0 + per && r < rows; ++r) {
    const int64_t x = a.row_off[r];
    a.row_off[r] = run;
    run += x;
  }
  if (threadIdx.x == kScanBlock - 1) a.row_off[rows] = s[threadIdx.x];
  __syncthreads();
  for (int l = threadIdx.x; l <= a.n_lag; l += kScanBlock)
    a.lag_off[l] = a.row_off[(int64_t)l * a.n_freq];
}

// One block per row: positions by a block-wide scan of the flags, in time order.
__global__ __launch_bounds__(kBlock) void k_sel_emit(SelArgs a) {
  __shared__ int s_w[kBlock / 64];
  __shared__ int64_t s_base;
  const int f = blockIdx.x, lag = blockIdx.y;
  const int64_t row = (int64_t)lag * a.n_freq + f;
  const uint8_t* fl = a.flags + row * a.n_time;
  if (threadIdx.x == 0) s_base = a.row_off[row];
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int64_t t0 = 0; t0 < a.n_time; t0 += kBlock) {
    const int64_t t = t0 + threadIdx.x;
    const int k = t < a.n_time ? fl[t] : 0;
    const unsigned long long bal = __ballot(k);
    const int before = __popcll(bal & ((1ull << lane) - 1ull));
    if (lane == 0) s_w[wave] = __popcll(bal);
    __syncthreads();
    int off = 0;
    for (int w = 0; w < wave; ++w) off += s_w[w];
    const int64_t base = s_base;
    if (k) {
      const int64_t pos = base + off + before;
      if (pos < a.capacity) {
        a.freq[pos] = f;
        a.time[pos] = t;
        double raw = 0.0;
        for (int d = 0; d < a.n_ifo; ++d) {
          const int64_t dt = src_bin(a, lag, d, t);   // t is inside the valid range
          const double x = a.energy[((int64_t)d * a.n_freq + f) * a.n_time + dt];
          raw += x;
          a.det_energy[pos * a.n_ifo + d] = x > 0.0 ? x : 0.0;
          a.det_index[pos * a.n_ifo + d] = dt * a.n_freq + f;
        }
        a.value[pos] = raw;
      }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      int tot = 0;
      for (int w = 0; w < kBlock / 64; ++w) tot += s_w[w];
      s_base = base + tot;
    }
    __syncthreads();
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

SelArgs args_of(const pycwb_select_desc* d) {
  SelArgs a;
  a.n_ifo = d->n_ifo; a.n_freq = d->n_freq; a.n_time = d->n_time; a.energy = d->energy;
  a.n_lag = d->n_lag; a.shift = d->shift_bins; a.veto = d->veto; a.edge_bins = d->edge_bins;
  a.ib = d->ib; a.ie = d->ie; a.eo = d->eo; a.comb = d->combined; a.flags = d->flags;
  a.live = d->live; a.row_off = d->row_off; a.lag_off = d->lag_off; a.capacity = d->capacity;
  a.freq = d->freq; a.time = d->time; a.value = d->value; a.det_energy = d->det_energy;
  a.det_index = d->det_index;
  return a;
}

}  // namespace

extern "C" {

int pycwb_select_count(const pycwb_select_desc* d, void* stream) {
  if (!d || d->n_ifo < 1 || d->n_freq < 3 || d->n_time < 5 || d->n_lag < 1 ||
      d->n_lag > 65535 || d->n_freq > 65535 || !d->energy || !d->shift_bins || !d->combined ||
      !d->flags || !d->live || !d->row_off || !d->lag_off || d->edge_bins < 0 ||
      d->n_time - 2 * (int64_t)d->edge_bins < 1 || d->ib < 1)
    return PYCWB_ERR_ARG;
  (void)hipGetLastError();
  const hipStream_t st = (hipStream_t)stream;
  const SelArgs a = args_of(d);
  const unsigned tb = (unsigned)((d->n_time + kBlock - 1) / kBlock);
  hipLaunchKernelGGL(k_sel_live, dim3(tb, (unsigned)d->n_lag), dim3(kBlock), 0, st, a);
  int rc = check_launch("k_sel_live");
  if (rc != PYCWB_OK) return rc;
  hipLaunchKernelGGL(k_sel_combined, dim3(tb, (unsigned)d->n_freq, (unsigned)d->n_lag),
                     dim3(kBlock), 0, st, a);
  if ((rc = check_launch("k_sel_combined")) != PYCWB_OK) return rc;
  hipLaunchKernelGGL(k_sel_rows, dim3((unsigned)d->n_freq, (unsigned)d->n_lag), dim3(kBlock), 0,
                     st, a);
  if ((rc = check_launch("k_sel_rows")) != PYCWB_OK) return rc;
  hipLaunchKernelGGL(k_sel_scan, dim3(1), dim3(kScanBlock), 0, st, a);
  return check_launch("k_sel_scan");
}

int pycwb_select_emit(const pycwb_select_desc* d, void* stream) {
  if (!d || !d->flags || !d->row_off || !d->freq || !d->time || !d->value || !d->det_energy ||
      !d->det_index || d->capacity < 0 || d->n_lag < 1 || d->n_lag > 65535 || d->n_freq < 3 ||
      d->n_freq > 65535)
    return PYCWB_ERR_ARG;
  (void)hipGetLastError();
  const SelArgs a = args_of(d);
  hipLaunchKernelGGL(k_sel_emit, dim3((unsigned)d->n_freq, (unsigned)d->n_lag), dim3(kBlock), 0,
                     (hipStream_t)stream, a);
  return check_launch("k_sel_emit");
}

}  // extern "C"
