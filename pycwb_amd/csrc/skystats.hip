// Best-sky statistics of many clusters at their selected directions: one
// workgroup per cluster.  Restates compute_statistics_at_sky_position
// (pycwb/modules/likelihoodWP/sky_statistics.py:26-290) with the packet
// operations of pycwb/modules/likelihoodWP/packet_ops.py:
//   per-pixel scan quantities (pixel_eval, identical arithmetic to the scan),
//   avx_packet_ps (per-detector sequential sums, one lane per detector),
//   packet_norm_numpy (cross-talk CSR), gw_norm_numpy, avx_noise_ps,
//   avx_setAMP_ps, avx_loadNULL_ps and the two xtalk energy sums,
// then Lm = Em - Np - Gn, the time-domain normalisation, rho and the rest.
// The NumPy stages are float64 with block-tree sums (tolerance-level parity);
// the Numba stages keep their float32/float64 typing and sequential order.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>

#include "skyscan_lane.hh"
#include "pycwb_amd.h"

using namespace pycwb;

namespace {

constexpr int kBlock = 256;

struct StatsWs32 {
  float *v00, *v90, *p, *q, *d00, *d90, *s00, *s90, *n00, *n90, *f, *F;
  float *et, *mk, *ee, *EE, *ec, *gn, *fp, *fx;
  float *pr, *pa, *pR, *pA;   // avx_pol_ps: p00_POL (radius, angle), p90_POL
};
struct StatsWs64 {
  double *qn, *pn, *pd, *pD, *ps, *pS, *rnx;
};

__device__ inline StatsWs32 ws32(float* b, int n, int P) {
  StatsWs32 w;
  const int64_t nP = (int64_t)n * P;
  w.v00 = b; w.v90 = b + nP; w.p = b + 2 * nP; w.q = b + 3 * nP; w.d00 = b + 4 * nP;
  w.d90 = b + 5 * nP; w.s00 = b + 6 * nP; w.s90 = b + 7 * nP; w.n00 = b + 8 * nP;
  w.n90 = b + 9 * nP; w.f = b + 10 * nP; w.F = b + 11 * nP;
  float* t = b + 12 * nP;
  w.et = t; w.mk = t + P; w.ee = t + 2 * P; w.EE = t + 3 * P; w.ec = t + 4 * P;
  w.gn = t + 5 * P; w.fp = t + 6 * P; w.fx = t + 7 * P;
  w.pr = t + 8 * P; w.pa = t + 9 * P; w.pR = t + 10 * P; w.pA = t + 11 * P;
  return w;
}

__device__ inline StatsWs64 ws64(double* b, int n, int P) {
  StatsWs64 w;
  const int64_t nP = (int64_t)n * P;
  w.qn = b; w.pn = b + nP; w.pd = b + 2 * nP; w.pD = b + 3 * nP; w.ps = b + 4 * nP;
  w.pS = b + 5 * nP; w.rnx = b + 6 * nP;
  return w;
}

// Block-wide float64 sums of NV values (every thread passes its partials).
template <int NV>
__device__ void block_sum(double (&v)[NV], double* sh) {
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    double x = v[k];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
    v[k] = x;
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) {
#pragma unroll
    for (int k = 0; k < NV; ++k) sh[wave * NV + k] = v[k];
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    double s = 0.0;
    for (int w = 0; w < kBlock / 64; ++w) s += sh[w * NV + k];
    v[k] = s;
  }
  __syncthreads();
}

struct StatsArgs {
  const int32_t* npix;
  const int32_t* net;
  const int32_t* K;
  const int64_t* rms_off;
  const int64_t* td_off;
  const float* rms;
  const float2* td;
  const int64_t* net_off;
  const double* net_En;
  const float* FP;
  const float* FX;
  const int32_t* ml;
  int64_t S;
  const int64_t* l_max;
  const float* REG1;
  float reg0;
  const int64_t* lk_off;
  const int32_t* lookup;
  const int64_t* xt_off;
  const float* xt;
  int xgb;
  float* s32;
  const int64_t* s32_off;
  double* s64;
  const int64_t* s64_off;
  double* scal;
  double* S_snr;
};

// avx_packet_ps for one detector (packet_ops.py:709-738), sequential over pixels.
__device__ void packet_ifo(const float* v00, const float* v90, const float* mask, int P,
                           float& si_o, float& co_o, float& ai_o, float& Ai_o, float& a_o,
                           float& A_o, float& E_o) {
  const double o = 1.e-9;
  float aa = 0.f, AA = 0.f, aA = 0.f;
  for (int i = 0; i < P; ++i) {
    const float mk = mask[i] > 0.f ? 1.f : 0.f;
    aa += mk * (v00[i] * v00[i]);
    AA += mk * (v90[i] * v90[i]);
    aA += mk * (v00[i] * v90[i]);
  }
  const float _si = 2.f * aA;
  const float _co = aa - AA;
  const double _x = (double)(aa + AA) + o;
  const float _cc = _co * _co;
  const float _ss = _si * _si;
  const float _nn = __builtin_sqrtf(_cc + _ss);
  const float a = (float)__builtin_sqrt((_x + (double)_nn) / 2.0);
  const float A = (float)__builtin_sqrt(fabs((_x - (double)_nn) / 2.0));
  const double cc = (double)_co / ((double)_nn + o);
  const float sgn = _si > 0.f ? 1.f : -1.f;
  si_o = (float)__builtin_sqrt((1.0 - cc) / 2.0);
  co_o = (float)(__builtin_sqrt((1.0 + cc) / 2.0) * (double)sgn);
  const float s = a + A;
  E_o = (s * s) / 2.f;
  a_o = a;
  A_o = A;
  ai_o = (float)(1.0 / ((double)a + o));
  Ai_o = (float)(1.0 / ((double)A + o));
}

// NumPy's float64 sum of the N products of one row (add.reduce over a short
// contiguous axis: from +0.0 in order below eight terms, the eight-accumulator
// pairing at eight).
template <int N>
__device__ __forceinline__ double row_sum(const double (&t)[N]) {
  if constexpr (N < 8) {
    double s = 0.0;
#pragma unroll
    for (int i = 0; i < N; ++i) s += t[i];
    return s;
  } else {
    return ((t[0] + t[1]) + (t[2] + t[3])) + ((t[4] + t[5]) + (t[6] + t[7]));
  }
}

// avx_pol_ps (packet_ops.py:564-633): polar coordinates of each pixel's masked
// data in the DPF frame, float64 with float32 results --
// r = sqrt(xp^2/(fp+1e-9) + xx^2/(fx+1e-9)), a = atan2(xx/(sqrt fx + 1e-9),
// xp/(sqrt fp + 1e-9)) for the 00 phase (p00_POL) and the same for 90 (p90_POL).
template <int N>
__device__ __forceinline__ void pol_pixel(const float (&v00)[N], const float (&v90)[N],
                                          const PixelEval<N>& o, const StatsWs32& w, int j) {
  const double o9 = 1.e-9;
  const double m = o.mk > 0.f ? 1.0 : 0.0;
  double t0[N], t1[N], t2[N], t3[N];
#pragma unroll
  for (int i = 0; i < N; ++i) {
    const double p = (double)v00[i] * m, q = (double)v90[i] * m;
    t0[i] = (double)o.d.f[i] * p;
    t1[i] = (double)o.d.f[i] * q;
    t2[i] = (double)o.d.F[i] * p;
    t3[i] = (double)o.d.F[i] * q;
  }
  const double xp = row_sum<N>(t0), XP = row_sum<N>(t1);
  const double xx = row_sum<N>(t2), XX = row_sum<N>(t3);
  const double fp = (double)o.d.fp, fx = (double)o.d.fx;
  const double sfp = __builtin_sqrt(fp) + o9, sfx = __builtin_sqrt(fx) + o9;
  const double rp = xp * xp / (fp + o9) + xx * xx / (fx + o9);
  const double RP = XP * XP / (fp + o9) + XX * XX / (fx + o9);
  w.pr[j] = (float)__builtin_sqrt(rp);
  w.pa[j] = (float)atan2(xx / sfx, xp / sfp);
  w.pR[j] = (float)__builtin_sqrt(RP);
  w.pA[j] = (float)atan2(XX / sfx, XP / sfp);
}

template <int N>
__device__ void stats_cluster(const StatsArgs& g, int c) {
  __shared__ double sh[kBlock / 64 * 16];
  __shared__ float pk[2][7][8];  // [data/signal][si, co, ai, Ai, a, A, E][ifo]
  __shared__ double nrm[8];
  const int P = g.npix[c];
  const int K = g.K[c];
  const int nd = 2 * K + 1;
  const int net = g.net[c];
  const int64_t S = g.S;
  int64_t l = g.l_max[c];
  l = l < 0 ? 0 : (l >= S ? S - 1 : l);
  const float* FP = g.FP + g.net_off[net];
  const float* FX = g.FX + g.net_off[net];
  const int32_t* ml = g.ml + g.net_off[net];
  const double En = g.net_En[net];
  const float* rms = g.rms + g.rms_off[c];
  const float2* td = g.td + g.td_off[c];
  const int32_t* lookup = g.lookup + 2 * g.lk_off[c];
  const float* xt = g.xt + 8 * g.xt_off[c];
  StatsWs32 w = ws32(g.s32 + g.s32_off[c], N, P);
  StatsWs64 W = ws64(g.s64 + g.s64_off[c], N, P);
  const float rr = g.reg0, RR = g.REG1[c];
  float fp0[N], fx0[N];
  int tap[N];
#pragma unroll
  for (int i = 0; i < N; ++i) {
    fp0[i] = FP[l * N + i];
    fx0[i] = FX[l * N + i];
    int t = ml[(int64_t)i * S + l] + K;
    tap[i] = i * nd + (t < 0 ? 0 : (t >= nd ? nd - 1 : t));
  }
  // ---- per-pixel scan quantities at l ----
  for (int j = threadIdx.x; j < P; j += kBlock) {
    float v00[N], v90[N], r[N];
#pragma unroll
    for (int i = 0; i < N; ++i) {
      const float2 v = td[(int64_t)j * N * nd + tap[i]];
      v00[i] = v.x;
      v90[i] = v.y;
      r[i] = rms[(int64_t)j * N + i];
    }
    PixelEval<N> o;
    pixel_eval<N>(v00, v90, r, fp0, fx0, rr, RR, En, o);
#pragma unroll
    for (int i = 0; i < N; ++i) {
      w.v00[i * P + j] = v00[i];
      w.v90[i * P + j] = v90[i];
      w.p[i * P + j] = o.p[i];
      w.q[i * P + j] = o.q[i];
      w.f[j * N + i] = o.d.f[i];
      w.F[j * N + i] = o.d.F[i];
    }
    w.et[j] = o.et;
    w.mk[j] = o.mk;
    w.ee[j] = o.ee;
    w.EE[j] = o.EE;
    w.ec[j] = o.ec;
    w.gn[j] = o.gn;
    w.fp[j] = o.d.fp;
    w.fx[j] = o.d.fx;
    pol_pixel<N>(v00, v90, o, w, j);
  }
  __syncthreads();
  // ---- avx_packet_ps: data packet on lanes [0, N), signal packet on [64, 64+N) ----
  {
    const int t = threadIdx.x;
    const int which = t >= 64 ? 1 : 0;
    const int i = t - 64 * which;
    if (i < N && t < 128) {
      const float* a = which ? w.p : w.v00;
      const float* b = which ? w.q : w.v90;
      float si, co, ai, Ai, av, Av, E;
      packet_ifo(a + i * P, b + i * P, w.mk, P, si, co, ai, Ai, av, Av, E);
      pk[which][0][i] = si;
      pk[which][1][i] = co;
      pk[which][2][i] = ai;
      pk[which][3][i] = Ai;
      pk[which][4][i] = av;
      pk[which][5][i] = Av;
      pk[which][6][i] = E;
    }
  }
  __syncthreads();
  // packet amplitudes (packet_ops.py:746-751)
  for (int j = threadIdx.x; j < P; j += kBlock) {
    const float mk = w.mk[j] > 0.f ? 1.f : 0.f;
#pragma unroll
    for (int i = 0; i < N; ++i) {
      {
        const float x = w.v00[i * P + j], y = w.v90[i * P + j];
        const float _a = x * pk[0][1][i] + y * pk[0][0][i];
        const float _A = y * pk[0][1][i] - x * pk[0][0][i];
        w.d00[i * P + j] = mk * _a * pk[0][2][i];
        w.d90[i * P + j] = mk * _A * pk[0][3][i];
      }
      {
        const float x = w.p[i * P + j], y = w.q[i * P + j];
        const float _a = x * pk[1][1][i] + y * pk[1][0][i];
        const float _A = y * pk[1][1][i] - x * pk[1][0][i];
        w.s00[i * P + j] = mk * _a * pk[1][2][i];
        w.s90[i * P + j] = mk * _A * pk[1][3][i];
      }
    }
  }
  __syncthreads();
  // ---- packet_norm_numpy on the data packet (packet_ops.py:843-901) ----
  {
    double part[N];
#pragma unroll
    for (int i = 0; i < N; ++i) part[i] = 0.0;
    for (int j = threadIdx.x; j < P; j += kBlock) {
      double rn = 0.0;
#pragma unroll
      for (int i = 0; i < N; ++i) W.qn[i * P + j] = 0.0;
      if (w.mk[j] > 0.f) {
        double x[4][N];
#pragma unroll
        for (int i = 0; i < N; ++i) x[0][i] = x[1][i] = x[2][i] = x[3][i] = 0.0;
        for (int e = lookup[2 * j]; e < lookup[2 * j + 1]; ++e) {
          int nb = (int)xt[8 * e];
          nb = nb < 0 ? 0 : (nb >= P ? P - 1 : nb);
          const double c0 = xt[8 * e + 4], c1 = xt[8 * e + 5], c2 = xt[8 * e + 6],
                       c3 = xt[8 * e + 7];
#pragma unroll
          for (int i = 0; i < N; ++i) {
            const double pv = w.d00[i * P + nb], qv = w.d90[i * P + nb];
            x[0][i] += pv * c0;
            x[1][i] += pv * c1;
            x[2][i] += qv * c2;
            x[3][i] += qv * c3;
          }
        }
#pragma unroll
        for (int i = 0; i < N; ++i) {
          const double pi = w.d00[i * P + j], qi = w.d90[i * P + j];
          const double t = x[0][i] * pi + x[1][i] * qi + x[2][i] * pi + x[3][i] * qi;
          const double tc = t < 0 ? 0.0 : t;
          part[i] += tc;
          const double e = (pi * pi + qi * qi) / (tc + 1.e-12);
          W.qn[i * P + j] = e >= 1 ? 0.0 : e;
          const double u = x[0][i] + x[2][i], v = x[1][i] + x[3][i];
          rn += u * u + v * v;
        }
      }
      W.rnx[j] = rn;
    }
    block_sum<N>(part, sh);
    if (threadIdx.x < N) nrm[threadIdx.x] = part[threadIdx.x] < 2.0 ? 2.0 : part[threadIdx.x];
  }
  __syncthreads();
  // ---- gw_norm_numpy: S_snr and p_norm (packet_ops.py:936-948) ----
  if (threadIdx.x < N) g.S_snr[(int64_t)c * 8 + threadIdx.x] =
      ((double)pk[1][6][threadIdx.x] * 2) / nrm[threadIdx.x];
  for (int j = threadIdx.x; j < P; j += kBlock) {
#pragma unroll
    for (int i = 0; i < N; ++i) W.pn[i * P + j] = w.ec[j] > 0.f ? W.qn[i * P + j] : 0.0;
  }
  __syncthreads();
  // ---- avx_noise_ps (packet_ops.py:120-153) + avx_setAMP_ps (:266-285) ----
  double acc[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
  for (int j = threadIdx.x; j < P; j += kBlock) {
    double ns = 0.0, nx = 0.0;
#pragma unroll
    for (int i = 0; i < N; ++i) {
      ns += W.pn[i * P + j];
      nx += W.qn[i * P + j];
    }
    ns /= N;
    nx /= N;
    const double mk = w.mk[j] > 0.f ? 1.0 : 0.0;
    const double core = mk * (nx > 0 ? 1.0 : 0.0);
    const double ec = (double)w.ec[j];
    acc[0] += core * ec;               // EC
    acc[1] += core;                    // NC
    acc[2] += (mk - core) * W.rnx[j];  // ES
    const double gn = (double)w.gn[j];
    double rc = gn < 2.0 ? 1.0 : 0.0;
    const double nn = gn * (1.0 - rc);
    rc = rc + nn * 0.5;
    rc = ec / (rc + 1e-9);
    const double sig = mk * (ns > 0 ? 1.0 : 0.0);
    const float gnew = (float)(mk * gn * nx);
    w.gn[j] = gnew;
    acc[3] += sig * ec;                // SC
    acc[4] += sig * rc;                // RC
    acc[5] += sig * (double)gnew;      // GN
    acc[6] += sig;                     // NS
    acc[7] += (mk - sig) * (double)w.et[j];  // EH
    // avx_setAMP_ps for the data (q_norm) and signal (p_norm) packets
    const float mk5 = 0.5f * (w.mk[j] > 0.f ? 1.f : 0.f);
    double col = 0.0;
#pragma unroll
    for (int i = 0; i < N; ++i) {
      const double nq = (double)((pk[0][4][i] + pk[0][5][i]) * mk5) * W.qn[i * P + j];
      const float dp = w.d00[i * P + j], dq = w.d90[i * P + j];
      const float rp = dp * pk[0][1][i] - dq * pk[0][0][i];
      const float rq = dq * pk[0][1][i] + dp * pk[0][0][i];
      const double pd = nq * (double)rp, pD = nq * (double)rq;
      W.pd[i * P + j] = pd;
      W.pD[i * P + j] = pD;
      col += W.qn[i * P + j];
      const double np_ = (double)((pk[1][4][i] + pk[1][5][i]) * mk5) * W.pn[i * P + j];
      const float sp = w.s00[i * P + j], sq = w.s90[i * P + j];
      const float tp = sp * pk[1][1][i] - sq * pk[1][0][i];
      const float tq = sq * pk[1][1][i] + sp * pk[1][0][i];
      const double ps = np_ * (double)tp, pS = np_ * (double)tq;
      W.ps[i * P + j] = ps;
      W.pS[i * P + j] = pS;
      // avx_loadNULL_ps: float32(data) - float32(signal)
      w.n00[i * P + j] = (float)pd - (float)ps;
      w.n90[i * P + j] = (float)pD - (float)pS;
    }
    acc[8] += col * (double)mk5;       // _Np
  }
  block_sum<10>(acc, sh);
  // ---- xtalk energy sums Em (data) and Np (null) (packet_ops.py:756-810) ----
  double g2[2 * N];
#pragma unroll
  for (int i = 0; i < 2 * N; ++i) g2[i] = 0.0;
  for (int j = threadIdx.x; j < P; j += kBlock) {
    if (w.mk[j] <= 0.f) continue;
#pragma unroll
    for (int which = 0; which < 2; ++which) {
#pragma unroll
      for (int i = 0; i < N; ++i) {
        double x0 = 0, x1 = 0, x2 = 0, x3 = 0;
        for (int e = lookup[2 * j]; e < lookup[2 * j + 1]; ++e) {
          int nb = (int)xt[8 * e];
          nb = nb < 0 ? 0 : (nb >= P ? P - 1 : nb);
          const double pv = which ? (double)w.n00[i * P + nb] : W.pd[i * P + nb];
          const double qv = which ? (double)w.n90[i * P + nb] : W.pD[i * P + nb];
          x0 += pv * (double)xt[8 * e + 4];
          x1 += pv * (double)xt[8 * e + 5];
          x2 += qv * (double)xt[8 * e + 6];
          x3 += qv * (double)xt[8 * e + 7];
        }
        const double pi = which ? (double)w.n00[i * P + j] : W.pd[i * P + j];
        const double qi = which ? (double)w.n90[i * P + j] : W.pD[i * P + j];
        const double t = x0 * pi + x1 * qi + x2 * pi + x3 * qi;
        if (t > 0.0) g2[which * N + i] += t;
      }
    }
  }
  block_sum<2 * N>(g2, sh);
  if (threadIdx.x == 0) {
    double Em = 0.0, Np = 0.0;
    for (int i = 0; i < N; ++i) {
      Em += g2[i];
      Np += g2[N + i];
    }
    double Ep = 0.0, Lp = 0.0;
    for (int i = 0; i < N; ++i) {
      Ep += (double)pk[0][6][i];
      Lp += (double)pk[1][6][i];
    }
    const double Eo = Ep / 2.0, Lo = Lp / 2.0;
    const double Gn = acc[5], SC = acc[3];
    double Ec = acc[0], Dc = SC - acc[0];
    const double Rc = acc[4] / (SC + 0.01), Eh = acc[7] / 2, Es = acc[2] / 2;
    const double Neff = acc[8] * 4 / N - 1;
    const double Lm = Em - Np - Gn;
    double norm = Em > 0 ? (Eo - Eh) / Em : 1.e9;
    if (norm < 1) norm = 1;
    Ec /= norm;
    Dc /= norm;
    const double chi = (Np + Gn) / (Neff * N);
    const double cc = chi > 1 ? chi : 1;
    double rho, xrho = 0.0;
    if (!g.xgb) {
      rho = Ec > 0 ? __builtin_sqrt(Ec * Rc / 2.) : 0.0;
    } else {
      rho = Ec > 0 ? __builtin_sqrt(Ec / (1 + chi * ((chi > 1 ? chi : 1.0) - 1))) : 0.0;
      xrho = Ec > 0 ? __builtin_sqrt(Ec * Rc / 2.) : 0.0;
    }
    double* out = g.scal + (int64_t)c * 16;
    out[0] = Gn; out[1] = Ec; out[2] = Dc; out[3] = Rc; out[4] = Eh; out[5] = Es;
    out[6] = Np; out[7] = Em; out[8] = Lm; out[9] = norm; out[10] = cc; out[11] = rho;
    out[12] = xrho; out[13] = Lo; out[14] = Eo; out[15] = Neff;
  }
}

__global__ __launch_bounds__(kBlock) void k_sky_stats(StatsArgs g, const int32_t* __restrict__ nifo,
                                                      const int32_t* __restrict__ which) {
  const int c = which ? which[blockIdx.x] : (int)blockIdx.x;
  switch (nifo[c]) {
    case 1: stats_cluster<1>(g, c); break;
    case 2: stats_cluster<2>(g, c); break;
    case 3: stats_cluster<3>(g, c); break;
    case 4: stats_cluster<4>(g, c); break;
    case 5: stats_cluster<5>(g, c); break;
    case 6: stats_cluster<6>(g, c); break;
    case 7: stats_cluster<7>(g, c); break;
    case 8: stats_cluster<8>(g, c); break;
    default: break;
  }
}

}  // namespace

extern "C" {

// Scratch per cluster: (12 n + 12) P floats and (6 n + 1) P doubles.
size_t pycwb_sky_stats_scratch32(int n_ifo, int n_pix) {
  return (size_t)(12 * n_ifo + 12) * (size_t)n_pix;
}
size_t pycwb_sky_stats_scratch64(int n_ifo, int n_pix) {
  return (size_t)(6 * n_ifo + 1) * (size_t)n_pix;
}

int pycwb_sky_statistics(const pycwb_stats_desc* d, void* stream) {
  if (!d || d->n_clusters < 0) return PYCWB_ERR_ARG;
  if (d->n_clusters == 0) return PYCWB_OK;
  if (!d->n_pix || !d->nifo || !d->net || !d->K || !d->rms_off || !d->td_off || !d->rms ||
      !d->td || !d->net_off || !d->net_En || !d->FP || !d->FX || !d->ml || !d->l_max ||
      !d->REG1 || !d->lk_off || !d->lookup || !d->xt_off || !d->xt || !d->s32 || !d->s32_off ||
      !d->s64 || !d->s64_off || !d->scal || !d->S_snr || d->n_sky < 1)
    return PYCWB_ERR_ARG;
  (void)hipGetLastError();
  StatsArgs g;
  g.npix = d->n_pix; g.net = d->net; g.K = d->K; g.rms_off = d->rms_off; g.td_off = d->td_off;
  g.rms = d->rms; g.td = (const float2*)d->td; g.net_off = d->net_off; g.net_En = d->net_En;
  g.FP = d->FP; g.FX = d->FX; g.ml = d->ml; g.S = d->n_sky; g.l_max = d->l_max;
  g.REG1 = d->REG1; g.reg0 = d->reg0; g.lk_off = d->lk_off; g.lookup = d->lookup;
  g.xt_off = d->xt_off; g.xt = d->xt; g.xgb = d->xgb_rho_mode; g.s32 = d->s32;
  g.s32_off = d->s32_off; g.s64 = d->s64; g.s64_off = d->s64_off; g.scal = d->scal;
  g.S_snr = d->S_snr;
  hipLaunchKernelGGL(k_sky_stats, dim3((unsigned)d->n_clusters), dim3(kBlock), 0,
                     (hipStream_t)stream, g, d->nifo, d->which);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    record_error("k_sky_stats", hipGetErrorString(e));
    return PYCWB_ERR_HIP;
  }
  return PYCWB_OK;
}

}  // extern "C"
