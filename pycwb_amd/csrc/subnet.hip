// Sub-network cut of many clusters (the loudest pixels of each supercluster).
//
// Stage 1, the sky loop: restates optimze_sky_loc_from_td
// (pycwb/modules/super_cluster_native/sub_net_cut.py:404-517).  One lane owns
// one sky direction and sweeps the cluster's pixels twice, in pixel order:
// pass 1 accumulates the network, sub-network and thresholded energies (both
// the "> En" and ">= En" sets), pass 2 runs the dominant-polarisation frame
// (_dpf_np_loops_vec_into, :520-565) and the projected likelihood sse_like_ps
// (:728-742) over the pixels at or above the threshold.  The running maximum
// `AA > AA_max` (strict, starting at 0: first direction of the largest
// positive AA) becomes a per-tile reduction plus a per-cluster pick.
//
// Stage 2, the multi-resolution statistics at the chosen direction: restates
// mra_statistics_from_td (:659-724) with the greedy principal-pixel
// extraction sse_MRA_ps (:746-825) over the cluster's cross-talk CSR, one
// wave per cluster; the sums that the reference takes in pixel order are
// taken in pixel order (a lane-ordered walk over each 64-pixel chunk).
//
// Typing follows Numba's: float32 arrays, scalars and sequential sums; an
// int64 or float64 operand promotes to float64 (the subcut test, AA, suball,
// submra, rHo).  Built with -ffp-contract=off and correctly rounded divide
// and sqrt, so every operation rounds like the CPU restatement.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "skyscan_lane.hh"
#include "pycwb_amd.h"

using namespace pycwb;

namespace {

constexpr int kSkyBlock = 256;
constexpr int kRec = 8;          // doubles per tile / per-cluster record
constexpr int kMraLdsPix = 4096; // working energies kept in LDS up to this many pixels

struct SubnetArgs {
  const int32_t* which;
  const int32_t* npix;
  const int32_t* net;
  const int32_t* K;
  const int64_t* rms_off;
  const int64_t* td_off;
  const float* rms;
  const float2* td;
  const int64_t* net_off;
  const float* net_En;
  const float* net_Es;
  const float* FP;
  const float* FX;
  const int32_t* ml;
  int64_t S;
  int n_tiles;
  double subcut;
  double* tiles;
  double* sky;
  const int64_t* lk_off;
  const int32_t* lookup;
  const int64_t* xt_off;
  const float* xt;
  float* scratch;
  const int64_t* scr_off;
  double* mra;
  const int32_t* order;
};

// sse_like_ps (sub_net_cut.py:728-742): each np.dot a forward float32 sum.
template <int N>
__device__ __forceinline__ float sse_like(const float (&f)[N], const float (&F)[N],
                                          const float (&a)[N], const float (&A)[N]) {
  float xp = 0.f, XP = 0.f, xx = 0.f, XX = 0.f, gp = 0.f, gx = 0.f;
#pragma unroll
  for (int i = 0; i < N; ++i) {
    xp += f[i] * a[i];
    XP += f[i] * A[i];
    xx += F[i] * a[i];
    XX += F[i] * A[i];
    gp += f[i] * f[i];
    gx += F[i] * F[i];
  }
  gp = gp + 1e-12f;
  gx = gx + 1e-12f;
  xp = xp * xp + XP * XP;
  xx = xx * xx + XX * XX;
  return xp / gp + xx / gx;
}

// Per-direction record: the reference's (stat, Em, Am, Vm, suball, EE) with
// stat = AA; valid only for AA > 0 (a NaN AA never wins `AA > AA_max`).
struct SkyRec {
  double AA;
  int64_t l;
  float Em;
  double Am;
  int64_t Vm;
  double suball;
  double EE;
};

__device__ __forceinline__ bool rec_better(double a, int64_t la, double b, int64_t lb) {
  return a > b || (a == b && la < lb);
}

// Per-cluster constants and the direction's delays and antenna factors.
template <int N>
struct DirCtx {
  int P, nd;
  float En, Es;
  const float2* td;
  const float* rms;
  int d[N];
  float fp0[N], fx0[N];
};

template <int N>
__device__ __forceinline__ void dir_ctx(const SubnetArgs& a, int c, int64_t l, DirCtx<N>& x) {
  x.P = a.npix[c];
  const int net = a.net[c];
  const int K = a.K[c];
  x.nd = 2 * K + 1;
  x.En = a.net_En[net];
  x.Es = a.net_Es[net];
  x.td = a.td + a.td_off[c];
  x.rms = a.rms + a.rms_off[c];
  const int64_t no = a.net_off[net];
#pragma unroll
  for (int i = 0; i < N; ++i) {
    x.d[i] = a.ml[no + (int64_t)i * a.S + l] + K;
    x.fp0[i] = a.FP[no + l * N + i];
    x.fx0[i] = a.FX[no + l * N + i];
  }
}

// What pass 2 needs from pass 1 (sub_net_cut.py:466-486).
struct Pass1 {
  float aa, Eo, Ls, Ln;
};

// Pass-1 sums of one direction (sub_net_cut.py:432-438).
struct Acc1 {
  float m_gt, Eo_gt, Ls_gt, Ln_gt, Eo_ge, Ls_ge, Ln_ge;
};

// One pixel of pass 1 (sub_net_cut.py:439-464) from its per-detector
// energies de[i] = v00^2 + v90^2 at the direction's delays.
template <int N>
__device__ __forceinline__ void acc1_pixel(Acc1& s, const float (&de)[N], float En, float Es) {
  float rE = 0.f, em = 0.f;
#pragma unroll
  for (int i = 0; i < N; ++i) {
    rE += de[i];
    if (de[i] > em) em = de[i];
  }
  const float sub = rE - em;
  if (rE >= En) {
    s.Eo_ge += rE;
    s.Ls_ge += sub;
    if (sub > Es) s.Ln_ge += rE;
  }
  if (rE > En) {
    s.m_gt += 1.f;
    s.Eo_gt += rE;
    s.Ls_gt += sub;
    if (sub > Es) s.Ln_gt += rE;
  }
}

// The end of pass 1 (sub_net_cut.py:466-490): the subcut pre-cut and the
// Eo <= 0 skip.  True when the direction goes on to pass 2.
__device__ __forceinline__ bool acc1_finish(const SubnetArgs& a, const Acc1& s, Pass1& q) {
  const float Eo_cut = s.Eo_gt + 0.01f;
  const int64_t m_cut = (int64_t)(2.0 * (double)s.m_gt + 0.01);
  const float aa = s.Ls_gt * s.Ln_gt / (Eo_cut - s.Ls_gt);
  if (a.subcut >= 0.0 &&
      ((double)aa - (double)m_cut) / ((double)aa + (double)m_cut + (double)1e-16f) < a.subcut)
    return false;
  if (s.Eo_ge <= 0.f) return false;
  q.aa = aa;
  q.Eo = s.Eo_ge;
  q.Ls = s.Ls_ge;
  q.Ln = s.Ln_ge;
  return true;
}

// Pass 1 with the lane's own gathers from the amplitude table.  (Staging a
// block's pixel rows in LDS instead is slower: the 256 delay-ordered
// directions of a block read a narrow delay range, and staging every delay
// costs more than the gathers it saves.)
template <int N>
__device__ void pass1_global(const DirCtx<N>& x, Acc1& s) {
  for (int j = 0; j < x.P; ++j) {
    const float2* row = x.td + (int64_t)j * N * x.nd;
    float de[N];
#pragma unroll
    for (int i = 0; i < N; ++i) {
      const float2 v = row[i * x.nd + x.d[i]];
      de[i] = v.x * v.x + v.y * v.y;
    }
    acc1_pixel<N>(s, de, x.En, x.Es);
  }
}

// Pass 2 (sub_net_cut.py:470-515): DPF likelihood over the pixels with
// E >= En, then the direction's statistics.  r stays invalid unless AA > 0.
template <int N>
__device__ void sky_pass2(const DirCtx<N>& x, int64_t l, const Pass1& q, SkyRec& r) {
  int64_t m = 0;
  float Lo = 0.f;
  for (int j = 0; j < x.P; ++j) {
    const float2* row = x.td + (int64_t)j * N * x.nd;
    float v00[N], v90[N];
    float rE = 0.f;
#pragma unroll
    for (int i = 0; i < N; ++i) {
      const float2 v = row[i * x.nd + x.d[i]];
      v00[i] = v.x;
      v90[i] = v.y;
      rE += v.x * v.x + v.y * v.y;
    }
    if (rE < x.En) continue;
    m += 1;
    float rr[N];
#pragma unroll
    for (int i = 0; i < N; ++i) rr[i] = x.rms[(int64_t)j * N + i];
    Dpf<N> dp;
    dpf_pixel<N>(rr, x.fp0, x.fx0, dp);
    Lo += sse_like<N>(dp.f, dp.F, v00, v90);
  }
  const float aa = q.aa, Eo = q.Eo, Ls = q.Ls, Ln = q.Ln;
  const double AA = (double)aa / ((double)(fabsf(aa) + fabsf(Eo - Lo)) +
                                  (double)(2 * m) * (double)(Eo - Ln) / (double)Eo);
  if (!(AA > 0.0)) return;
  if (!rec_better(AA, l, r.AA, r.l)) return;
  const float ee = Ls * Eo / (Eo - Ls);
  const double emm = (double)fabsf(Eo - Lo) + (double)(2 * m);
  r.AA = AA;
  r.l = l;
  r.Em = Eo;
  r.Am = ((double)aa - (double)m) / ((double)aa + (double)m);
  r.Vm = m;
  r.suball = (double)ee / ((double)ee + emm);
  r.EE = emm;
}

__device__ __forceinline__ void wave_best(double& AA, int64_t& l) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const double oA = __shfl_xor(AA, off, 64);
    const int64_t ol = __shfl_xor(l, off, 64);
    if (rec_better(oA, ol, AA, l)) {
      AA = oA;
      l = ol;
    }
  }
}

__device__ __forceinline__ void store_rec(double* o, const SkyRec& r) {
  o[0] = r.AA;
  o[1] = (double)r.l;
  o[2] = (double)r.Em;
  o[3] = r.Am;
  o[4] = (double)r.Vm;
  o[5] = r.suball;
  o[6] = r.EE;
  o[7] = 0.0;
}

// Grid (n_tiles, n_which): block b.y evaluates the directions at positions
// [b.x * 256, b.x * 256 + 256) of its network's direction order (identity
// without one) for cluster which[b.y] and writes its best record.
// Pass 1 runs one direction per lane; the directions that survive the pre-cut
// (typically a few per cent) are compacted into an LDS queue so that pass 2,
// the costly DPF sweep, runs on dense lanes instead of a few lanes per wave.
template <int N>
__global__ __launch_bounds__(kSkyBlock) void k_subnet_sky(SubnetArgs a) {
  __shared__ double s_rec[kSkyBlock / 64][kRec];
  __shared__ int s_ql[kSkyBlock];
  __shared__ Pass1 s_qp[kSkyBlock];
  __shared__ int s_count;
  const int slot = blockIdx.y;
  const int c = a.which[slot];
  const int64_t pos = (int64_t)blockIdx.x * kSkyBlock + threadIdx.x;
  const int32_t* order = a.order ? a.order + (int64_t)a.net[c] * a.S : nullptr;
  if (threadIdx.x == 0) s_count = 0;
  __syncthreads();
  if (pos < a.S) {
    const int64_t l = order ? (int64_t)order[pos] : pos;
    DirCtx<N> x;
    dir_ctx<N>(a, c, l, x);
    Acc1 acc = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    pass1_global<N>(x, acc);
    Pass1 q;
    if (acc1_finish(a, acc, q)) {
      const int k = atomicAdd(&s_count, 1);
      s_ql[k] = (int)l;
      s_qp[k] = q;
    }
  }
  __syncthreads();
  const int count = s_count;
  SkyRec r;
  r.AA = -1.0;
  r.l = INT64_MAX;
  r.Em = 0.f; r.Am = 0.0; r.Vm = 0; r.suball = 0.0; r.EE = 0.0;
  for (int k = threadIdx.x; k < count; k += kSkyBlock) {
    const int64_t lk = s_ql[k];
    DirCtx<N> x;
    dir_ctx<N>(a, c, lk, x);
    sky_pass2<N>(x, lk, s_qp[k], r);
  }
  double bA = r.AA;
  int64_t bl = r.l;
  wave_best(bA, bl);
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  if (lane == 0) s_rec[wave][0] = -1.0;
  __syncthreads();
  if (bA > 0.0 && r.l == bl) store_rec(s_rec[wave], r);
  __syncthreads();
  if (threadIdx.x == 0) {
    int best = -1;
    for (int w = 0; w < kSkyBlock / 64; ++w) {
      if (!(s_rec[w][0] > 0.0)) continue;
      if (best < 0 || rec_better(s_rec[w][0], (int64_t)s_rec[w][1], s_rec[best][0],
                                 (int64_t)s_rec[best][1]))
        best = w;
    }
    double* o = a.tiles + ((int64_t)slot * a.n_tiles + blockIdx.x) * kRec;
    if (best < 0) {
      o[0] = -1.0;
    } else {
      for (int k = 0; k < kRec; ++k) o[k] = s_rec[best][k];
    }
  }
}

// One wave per evaluated cluster: the tiles' best records in direction order
// -> the reference's 8-tuple (l_max, stat, Em, Am, lm, Vm, suball, EE), all
// zero when no direction has AA > 0.
__global__ __launch_bounds__(64) void k_subnet_pick(SubnetArgs a) {
  const int slot = blockIdx.x;
  const int c = a.which[slot];
  const double* t = a.tiles + (int64_t)slot * a.n_tiles * kRec;
  double bA = -1.0;
  int64_t bl = INT64_MAX;
  int bt = -1;
  for (int k = threadIdx.x; k < a.n_tiles; k += 64) {
    const double A = t[(int64_t)k * kRec];
    if (A > 0.0) {
      const int64_t l = (int64_t)t[(int64_t)k * kRec + 1];
      if (rec_better(A, l, bA, bl)) {
        bA = A;
        bl = l;
        bt = k;
      }
    }
  }
  double wA = bA;
  int64_t wl = bl;
  wave_best(wA, wl);
  double* o = a.sky + (int64_t)c * kRec;
  if (wA > 0.0) {
    if (bt >= 0 && bl == wl && bA == wA) {
      const double* r = t + (int64_t)bt * kRec;
      o[0] = r[1];
      o[1] = r[0];
      o[2] = r[2];
      o[3] = r[3];
      o[4] = r[1];
      o[5] = r[4];
      o[6] = r[5];
      o[7] = r[6];
    }
  } else if (threadIdx.x == 0) {
    for (int k = 0; k < kRec; ++k) o[k] = 0.0;
  }
}

// Stage 2: mra_statistics_from_td at l_max = sky[c][0], one wave per cluster.
// Scratch per cluster (4 n P floats): working amplitudes a00/a90 and the
// extracted components amp/AMP, each (n, P); plus P working energies (LDS up
// to kMraLdsPix pixels, else the scratch's tail).
template <int N>
__global__ __launch_bounds__(64) void k_subnet_mra(SubnetArgs a) {
  __shared__ float s_rN[kMraLdsPix];
  const int c = a.which[blockIdx.x];
  const int lane = threadIdx.x;
  const int P = a.npix[c];
  const int net = a.net[c];
  const int K = a.K[c];
  const int nd = 2 * K + 1;
  const float En = a.net_En[net], Es = a.net_Es[net];
  const float2* td = a.td + a.td_off[c];
  const float* rms = a.rms + a.rms_off[c];
  const int64_t no = a.net_off[net];
  const int64_t l = (int64_t)a.sky[(int64_t)c * kRec];
  const int64_t nP = (int64_t)N * P;
  float* a00 = a.scratch + a.scr_off[c];
  float* a90 = a00 + nP;
  float* amp = a90 + nP;
  float* AMP = amp + nP;
  float* rN = P <= kMraLdsPix ? s_rN : AMP + nP;
  const int32_t* lookup = a.lookup + 2 * a.lk_off[c];
  const float* xt = a.xt + 8 * a.xt_off[c];
  int d[N];
  float fp0[N], fx0[N];
#pragma unroll
  for (int i = 0; i < N; ++i) {
    d[i] = a.ml[no + (int64_t)i * a.S + l] + K;
    fp0[i] = a.FP[no + l * N + i];
    fx0[i] = a.FX[no + l * N + i];
  }
  // energies and the above-threshold count (sub_net_cut.py:671-686)
  int cnt = 0;
  for (int j = lane; j < P; j += 64) {
    const float2* row = td + (int64_t)j * N * nd;
    float rE = 0.f;
#pragma unroll
    for (int i = 0; i < N; ++i) {
      const float2 v = row[i * nd + d[i]];
      rE += v.x * v.x + v.y * v.y;
      a00[i * P + j] = v.x;
      a90[i * P + j] = v.y;
      amp[i * P + j] = 0.f;
      AMP[i * P + j] = 0.f;
    }
    rN[j] = rE;
    cnt += rE > En ? 1 : 0;
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) cnt += __shfl_xor(cnt, off, 64);
  __syncthreads();
  // sse_MRA_ps (sub_net_cut.py:746-825)
  float EE = 0.f;
  int k = 0;
  while (k < cnt) {
    float bv = -INFINITY;
    int bi = P;
    for (int j = lane; j < P; j += 64) {
      const float v = rN[j];
      if (v > bv) {
        bv = v;
        bi = j;
      }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      const float ov = __shfl_xor(bv, off, 64);
      const int oi = __shfl_xor(bi, off, 64);
      if (ov > bv || (ov == bv && oi < bi)) {
        bv = ov;
        bi = oi;
      }
    }
    const int m = bi;
    if (!(m < P) || rN[m] <= En) break;
    float E = 0.f;
    float mam[N], mAM[N];
#pragma unroll
    for (int i = 0; i < N; ++i) {
      mam[i] = a00[i * P + m];
      mAM[i] = a90[i * P + m];
      E += mam[i] * mam[i] + mAM[i] * mAM[i];
    }
    EE += E;
    if ((double)(E / EE) < 0.01) break;
    __syncthreads();
    if (lane == 0) {
#pragma unroll
      for (int i = 0; i < N; ++i) {
        amp[i * P + m] += mam[i];
        AMP[i * P + m] += mAM[i];
      }
    }
    const int e0 = lookup[2 * m], e1 = lookup[2 * m + 1];
    for (int e = e0 + lane; e < e1; e += 64) {
      const float* cx = xt + 8 * (int64_t)e;
      const int q = (int)cx[0];
      if (q < 0 || q >= P) continue;
      if (rN[q] > En) {
        float r = 0.f;
#pragma unroll
        for (int i = 0; i < N; ++i) {
          const float u = a00[i * P + q] - (mam[i] * cx[4] + mAM[i] * cx[5]);
          const float w = a90[i * P + q] - (mam[i] * cx[6] + mAM[i] * cx[7]);
          a00[i * P + q] = u;
          a90[i * P + q] = w;
          r += u * u + w * w;
        }
        rN[q] = r;
      }
    }
    __syncthreads();
    ++k;
  }
  // statistics of the extracted components (sub_net_cut.py:693-724)
  int64_t m = 0;
  float Ls = 0.f, Ln = 0.f, Eo = 0.f, Lo = 0.f;
  for (int base = 0; base < P; base += 64) {
    const int j = base + lane;
    float ee = 0.f, em = 0.f, val = 0.f;
    bool keep = false;
    if (j < P) {
      float x[N], X[N];
#pragma unroll
      for (int i = 0; i < N; ++i) {
        x[i] = amp[i * P + j];
        X[i] = AMP[i * P + j];
        ee += x[i] * x[i] + X[i] * X[i];
      }
      keep = !(ee < En);
      if (keep) {
        float rr[N];
#pragma unroll
        for (int i = 0; i < N; ++i) {
          const float e = x[i] * x[i] + X[i] * X[i];
          if (e > em) em = e;
          rr[i] = rms[(int64_t)j * N + i];
        }
        Dpf<N> dp;
        dpf_pixel<N>(rr, fp0, fx0, dp);
        val = sse_like<N>(dp.f, dp.F, x, X);
      }
    }
    for (int t = 0; t < 64; ++t) {
      const int kt = __shfl((int)keep, t, 64);
      const float et = __shfl(ee, t, 64);
      const float mt = __shfl(em, t, 64);
      const float vt = __shfl(val, t, 64);
      if (kt) {
        m += 1;
        Ls += et - mt;
        Eo += et;
        if (et - mt > Es) Ln += et;
        Lo += vt;
      }
    }
  }
  if (lane == 0) {
    const float sm = Ls * Eo / (Eo - Ls + 1e-16f);
    double* o = a.mra + (int64_t)c * kRec;
    o[0] = (double)sm / ((double)(fabsf(sm) + fabsf(Eo - Lo)) + (double)(2 * (m + 6)));
    o[1] = sqrt((double)(Lo * Lo) / ((double)Eo + (double)(2 * m) + (double)1e-16f) / 2.0);
    o[2] = Eo;
    o[3] = Lo;
    o[4] = Ls;
    o[5] = (double)m;
    o[6] = (double)k;
    o[7] = (double)cnt;
  }
}

template <template <int> class Launch>
int dispatch(int n, const SubnetArgs& a, int n_which, hipStream_t st) {
  switch (n) {
    case 1: return Launch<1>::run(a, n_which, st);
    case 2: return Launch<2>::run(a, n_which, st);
    case 3: return Launch<3>::run(a, n_which, st);
    case 4: return Launch<4>::run(a, n_which, st);
    case 5: return Launch<5>::run(a, n_which, st);
    case 6: return Launch<6>::run(a, n_which, st);
    case 7: return Launch<7>::run(a, n_which, st);
    case 8: return Launch<8>::run(a, n_which, st);
    default: return PYCWB_ERR_ARG;
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

template <int N>
struct LaunchSky {
  static int run(const SubnetArgs& a, int n_which, hipStream_t st) {
    hipLaunchKernelGGL(k_subnet_sky<N>, dim3((unsigned)a.n_tiles, (unsigned)n_which),
                       dim3(kSkyBlock), 0, st, a);
    int rc = check_launch("k_subnet_sky");
    if (rc != PYCWB_OK) return rc;
    hipLaunchKernelGGL(k_subnet_pick, dim3((unsigned)n_which), dim3(64), 0, st, a);
    return check_launch("k_subnet_pick");
  }
};

template <int N>
struct LaunchMra {
  static int run(const SubnetArgs& a, int n_which, hipStream_t st) {
    hipLaunchKernelGGL(k_subnet_mra<N>, dim3((unsigned)n_which), dim3(64), 0, st, a);
    return check_launch("k_subnet_mra");
  }
};

SubnetArgs args_of(const pycwb_subnet_desc* d) {
  SubnetArgs a;
  a.which = d->which; a.npix = d->n_pix; a.net = d->net; a.K = d->K; a.rms_off = d->rms_off;
  a.td_off = d->td_off; a.rms = d->rms; a.td = (const float2*)d->td; a.net_off = d->net_off;
  a.net_En = d->net_En; a.net_Es = d->net_Es; a.FP = d->FP; a.FX = d->FX; a.ml = d->ml;
  a.S = d->n_sky; a.n_tiles = (int)pycwb_subnet_tiles(d->n_sky); a.subcut = d->subcut;
  a.tiles = d->tiles; a.sky = d->sky; a.lk_off = d->lk_off; a.lookup = d->lookup;
  a.xt_off = d->xt_off; a.xt = d->xt; a.scratch = d->scratch; a.scr_off = d->scr_off;
  a.mra = d->mra;
  a.order = d->sky_order;
  return a;
}

bool common_ok(const pycwb_subnet_desc* d) {
  return d->which && d->n_pix && d->net && d->K && d->rms_off && d->td_off && d->rms && d->td &&
         d->net_off && d->net_En && d->net_Es && d->FP && d->FX && d->ml && d->sky &&
         d->n_sky >= 1 && d->n_ifo >= 1 && d->n_ifo <= 8;
}

}  // namespace

extern "C" {

size_t pycwb_subnet_tiles(int64_t n_sky) {
  return n_sky < 1 ? 0 : (size_t)((n_sky + kSkyBlock - 1) / kSkyBlock);
}

size_t pycwb_subnet_mra_scratch(int n_ifo, int n_pix) {
  return (size_t)(4 * n_ifo + 1) * (size_t)n_pix;
}

int pycwb_subnet_sky(const pycwb_subnet_desc* d, void* stream) {
  if (!d || d->n_which < 0) return PYCWB_ERR_ARG;
  if (d->n_which == 0) return PYCWB_OK;
  if (!common_ok(d) || !d->tiles || d->n_which > 65535 || d->n_sky > 0x7fffffff)
    return PYCWB_ERR_ARG;
  (void)hipGetLastError();
  return dispatch<LaunchSky>(d->n_ifo, args_of(d), d->n_which, (hipStream_t)stream);
}

int pycwb_subnet_mra(const pycwb_subnet_desc* d, void* stream) {
  if (!d || d->n_which < 0) return PYCWB_ERR_ARG;
  if (d->n_which == 0) return PYCWB_OK;
  if (!common_ok(d) || !d->lk_off || !d->lookup || !d->xt_off || !d->xt || !d->scratch ||
      !d->scr_off || !d->mra)
    return PYCWB_ERR_ARG;
  (void)hipGetLastError();
  return dispatch<LaunchMra>(d->n_ifo, args_of(d), d->n_which, (hipStream_t)stream);
}

}  // extern "C"
