// Time-delay max-energy maps (the coherence stage's decorrelated TF maps).
//
// Restates the pattern path of time_delay_max_energy_numba
// (pycwb/modules/coherence_native/time_delay_numba.py:183-288, 578-679): the
// detector's series and its copies shifted left and right by k = d, 2d, ...
// <= max_delay samples (the far end keeps the original samples) each go
// through the quadrature t2w (pycwb_wdm_t2w, in batches), every map's packet
// energy (_wdm_packet_energy_nb, :49-104, parameters from
// _compute_packet_energy_params, time_delay_packet.py:8-64) is taken, and the
// pixel-wise maximum is kept; the lowest layer (two for patterns 5, 6, 9) is
// zeroed at the end.
//   k_me_shift       (copy, sample)   the shifted copies of one batch
//   k_me_packet_max  pixel            packet energy of each map of the batch
//                                     and the running maximum
//   k_me_finish      time bin         zero the lowest layers
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "skyscan_lane.hh"
#include "pycwb_amd.h"

using namespace pycwb;

namespace {

constexpr int kBlock = 256;

// shift > 0: left (x[i] = ts[i + k] for i < L - k), < 0: right
// (x[i] = ts[i - k] for i >= k), 0: the series itself.
__global__ void k_me_shift(const double* __restrict__ ts, int64_t L,
                           const int64_t* __restrict__ shifts, double* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  const int s = blockIdx.y;
  if (i >= L) return;
  const int64_t k = shifts[s];
  double v = ts[i];
  if (k > 0 && i < L - k) v = ts[i + k];
  else if (k < 0 && i >= -k) v = ts[i + k];
  out[(int64_t)s * L + i] = v;
}

struct PacketArgs {
  const double2* maps;   // (n_maps, T, M1) (00, 90) pairs
  int n_maps;
  int64_t T;
  int M1;
  int64_t jb, je;
  int mL, mH;
  double mean;
  int64_t p[9];
  int init;              // 1: the first map initialises the maximum
  double* cur;           // (T, M1)
};

__device__ __forceinline__ double packet(const double2* m, int64_t J, int64_t j, int mm,
                                         const PacketArgs& a) {
  const double re0 = m[j].x;
  if (j < a.jb || j >= a.je) return re0 > 0.0 ? re0 : 0.0;
  if (mm < a.mL || mm > a.mH) return 0.0;
  double ss = 0.0, ee = 0.0, EE = 0.0;
#pragma unroll
  for (int n = 1; n < 9; ++n) {
    int64_t idx = j + a.p[n];
    idx = idx < 0 ? 0 : (idx >= J ? J - 1 : idx);
    const double2 q = m[idx];
    ss += q.x * q.y;
    ee += q.x * q.x;
    EE += q.y * q.y;
  }
  const double im0 = m[j].y;
  ss += re0 * im0 * (a.mean - 8.0);
  ee += re0 * re0 * (a.mean - 8.0);
  EE += im0 * im0 * (a.mean - 8.0);
  const double cc = ee - EE, s2 = 2.0 * ss;
  const double nn2 = cc * cc + s2 * s2;
  double nn = nn2 > 0.0 ? sqrt(nn2) : 0.0;
  const double se = ee + EE;
  if (se < nn) nn = se;
  const double h1 = (se + nn) / 2.0, h2 = (se - nn) / 2.0;
  const double aa = (h1 > 0.0 ? sqrt(h1) : 0.0) + (h2 > 0.0 ? sqrt(h2) : 0.0);
  return a.mean == 1.0 ? se / 2.0 : aa * aa / 4.0;
}

__global__ void k_me_packet_max(PacketArgs a) {
  const int64_t J = a.T * a.M1;
  const int64_t j = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (j >= J) return;
  const int mm = (int)(j % a.M1);
  double best = 0.0;
  int s0 = 0;
  if (a.init) {
    best = packet(a.maps, J, j, mm, a);
    s0 = 1;
  } else {
    best = a.cur[j];
  }
  for (int s = s0; s < a.n_maps; ++s) {
    const double e = packet(a.maps + (int64_t)s * J, J, j, mm, a);
    if (e > best) best = e;
  }
  a.cur[j] = best;
}

// The packet pass with each map's pixels staged in LDS: a block owns kPkTile
// consecutive pixels and, map by map, loads the span they and their 8
// neighbours reach (offsets lo .. hi of the flat index) with coalesced reads,
// each position holding map[clamp(index)] as packet() reads it; packet_tile
// then forms the same sums in the same order, so the maximum is identical.
// (packet() alone reads each map value 9 times through the caches.)
constexpr int kPkTile = 1024;
constexpr int kPkPer = kPkTile / kBlock;

__device__ __forceinline__ double packet_tile(const double2* __restrict__ tile, int64_t pos,
                                              int64_t j, int mm, const PacketArgs& a) {
  const double2 c = tile[pos];
  const double re0 = c.x;
  if (j < a.jb || j >= a.je) return re0 > 0.0 ? re0 : 0.0;
  if (mm < a.mL || mm > a.mH) return 0.0;
  double ss = 0.0, ee = 0.0, EE = 0.0;
#pragma unroll
  for (int n = 1; n < 9; ++n) {
    const double2 q = tile[pos + a.p[n]];
    ss += q.x * q.y;
    ee += q.x * q.x;
    EE += q.y * q.y;
  }
  const double im0 = c.y;
  ss += re0 * im0 * (a.mean - 8.0);
  ee += re0 * re0 * (a.mean - 8.0);
  EE += im0 * im0 * (a.mean - 8.0);
  const double cc = ee - EE, s2 = 2.0 * ss;
  const double nn2 = cc * cc + s2 * s2;
  double nn = nn2 > 0.0 ? sqrt(nn2) : 0.0;
  const double se = ee + EE;
  if (se < nn) nn = se;
  const double h1 = (se + nn) / 2.0, h2 = (se - nn) / 2.0;
  const double aa = (h1 > 0.0 ? sqrt(h1) : 0.0) + (h2 > 0.0 ? sqrt(h2) : 0.0);
  return a.mean == 1.0 ? se / 2.0 : aa * aa / 4.0;
}

__global__ __launch_bounds__(kBlock) void k_me_packet_tiled(PacketArgs a, int64_t lo,
                                                            int64_t span) {
  extern __shared__ double2 tile[];   // span values: flat indices j0 + lo .. j0 + lo + span - 1
  const int64_t J = a.T * a.M1;
  const int64_t j0 = (int64_t)blockIdx.x * kPkTile;
  double best[kPkPer];
  int mmv[kPkPer];
#pragma unroll
  for (int q = 0; q < kPkPer; ++q) {
    const int64_t j = j0 + threadIdx.x + q * kBlock;
    mmv[q] = (int)(j % a.M1);
    best[q] = (!a.init && j < J) ? a.cur[j] : 0.0;
  }
  for (int s = 0; s < a.n_maps; ++s) {
    const double2* m = a.maps + (int64_t)s * J;
    __syncthreads();   // the previous map's reads are done
    for (int64_t i = threadIdx.x; i < span; i += kBlock) {
      int64_t idx = j0 + lo + i;
      idx = idx < 0 ? 0 : (idx >= J ? J - 1 : idx);
      tile[i] = m[idx];
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < kPkPer; ++q) {
      const int64_t j = j0 + threadIdx.x + q * kBlock;
      if (j >= J) continue;
      const double e = packet_tile(tile, j - j0 - lo, j, mmv[q], a);
      if (a.init && s == 0) best[q] = e;
      else if (e > best[q]) best[q] = e;
    }
  }
#pragma unroll
  for (int q = 0; q < kPkPer; ++q) {
    const int64_t j = j0 + threadIdx.x + q * kBlock;
    if (j < J) a.cur[j] = best[q];
  }
}

__global__ void k_me_finish(double* cur, int64_t T, int M1, int two) {
  const int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (t >= T) return;
  cur[t * M1] = 0.0;
  if (two && M1 > 2) cur[t * M1 + 1] = 0.0;
}

int check_launch(const char* what) {
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    record_error(what, hipGetErrorString(e));
    return PYCWB_ERR_HIP;
  }
  return PYCWB_OK;
}

}  // namespace

extern "C" {

int pycwb_max_energy_shift(const double* ts, int64_t L, const int64_t* shifts, int n_copies,
                           double* out, void* stream) {
  if (!ts || !shifts || !out || L < 1 || n_copies < 1 || n_copies > 65535) return PYCWB_ERR_ARG;
  (void)hipGetLastError();
  hipLaunchKernelGGL(k_me_shift, dim3((unsigned)((L + kBlock - 1) / kBlock), (unsigned)n_copies),
                     dim3(kBlock), 0, (hipStream_t)stream, ts, L, shifts, out);
  return check_launch("k_me_shift");
}

int pycwb_max_energy_packet(const pycwb_packet_desc* d, void* stream) {
  if (!d || !d->maps || !d->cur || d->n_maps < 1 || d->T < 1 || d->M1 < 2) return PYCWB_ERR_ARG;
  (void)hipGetLastError();
  PacketArgs a;
  a.maps = (const double2*)d->maps; a.n_maps = d->n_maps; a.T = d->T; a.M1 = d->M1;
  a.jb = d->jb; a.je = d->je; a.mL = d->mL; a.mH = d->mH; a.mean = d->mean;
  for (int k = 0; k < 9; ++k) a.p[k] = d->p[k];
  a.init = d->init; a.cur = d->cur;
  const int64_t J = d->T * d->M1;
  int64_t lo = 0, hi = 0;
  for (int k = 1; k < 9; ++k) {
    lo = a.p[k] < lo ? a.p[k] : lo;
    hi = a.p[k] > hi ? a.p[k] : hi;
  }
  const int64_t span = kPkTile + hi - lo;
  if (span * (int64_t)sizeof(double2) <= 64 * 1024) {
    hipLaunchKernelGGL(k_me_packet_tiled, dim3((unsigned)((J + kPkTile - 1) / kPkTile)),
                       dim3(kBlock), (size_t)span * sizeof(double2), (hipStream_t)stream, a, lo,
                       span);
    return check_launch("k_me_packet_tiled");
  }
  hipLaunchKernelGGL(k_me_packet_max, dim3((unsigned)((J + kBlock - 1) / kBlock)), dim3(kBlock),
                     0, (hipStream_t)stream, a);
  return check_launch("k_me_packet_max");
}

int pycwb_max_energy_finish(double* cur, int64_t T, int M1, int zero_two_layers, void* stream) {
  if (!cur || T < 1 || M1 < 1) return PYCWB_ERR_ARG;
  (void)hipGetLastError();
  hipLaunchKernelGGL(k_me_finish, dim3((unsigned)((T + kBlock - 1) / kBlock)), dim3(kBlock), 0,
                     (hipStream_t)stream, cur, T, M1, zero_two_layers);
  return check_launch("k_me_finish");
}

}  // extern "C"
