// Continuous-direction sky likelihood and its gradients (float64).
//
// Objective (per cluster, direction (theta, phi), theta = colatitude):
//   tau_i = (R_i - R_ref) . n(theta, phi) / c * rate     (continuous delay, taps)
//   v_i   = linear interpolation of the delayed amplitudes at tau_i + K
//   F+_i, Fx_i from the detector response tensor at psi = 0, gha = -phi
//           (pycwb/types/detector.py:272-378, 1292-1308)
//   then the scan chain of sky_stat.py / dpf.py / sky_scan.py:131-141 in float64
//   with REG fixed, and AA = L * Co (the sky statistic; the netCC gate is not
//   applied so the objective is defined everywhere).
// Branches (energy mask, max(., 0), DPF sign choices, interpolation cell) are
// those of the evaluation point; the gradient is that of the active piece.
//
// Gradients are propagated in tangent form inside the kernel:
//   d/d(theta, phi): every quantity carries two tangents (Dual<2>);
//   d/d(amplitudes): AA depends on the pixel sums (EE, EC, GN, RN) only, so the
//   direction-level adjoints dAA/dS are formed once (Dual<4>) and each pixel's
//   contributions are differentiated with respect to its own 2n amplitudes
//   (Dual<2n>), then spread onto the two interpolation taps.
// One workgroup per cluster; masked pixels (energy below threshold) contribute
// exactly zero to value and gradients.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "skyscan_lane.hh"
#include "pycwb_amd.h"

namespace {

constexpr int kBlock = 256;

template <int K>
struct Dual {
  double v;
  double d[K];
  __device__ Dual() {}
  __device__ Dual(double x) : v(x) {
#pragma unroll
    for (int i = 0; i < K; ++i) d[i] = 0.0;
  }
};

#define DUAL_BIN(op, expr_d)                                                       \
  template <int K>                                                                 \
  __device__ __forceinline__ Dual<K> operator op(const Dual<K>& a, const Dual<K>& b) { \
    Dual<K> r;                                                                     \
    r.v = a.v op b.v;                                                              \
    _Pragma("unroll") for (int i = 0; i < K; ++i) r.d[i] = expr_d;                 \
    return r;                                                                      \
  }
DUAL_BIN(+, a.d[i] + b.d[i])
DUAL_BIN(-, a.d[i] - b.d[i])
DUAL_BIN(*, a.d[i] * b.v + a.v * b.d[i])
#undef DUAL_BIN

template <int K>
__device__ __forceinline__ Dual<K> operator/(const Dual<K>& a, const Dual<K>& b) {
  Dual<K> r;
  r.v = a.v / b.v;
  const double ib = 1.0 / b.v;
#pragma unroll
  for (int i = 0; i < K; ++i) r.d[i] = (a.d[i] - r.v * b.d[i]) * ib;
  return r;
}
template <int K>
__device__ __forceinline__ Dual<K> operator-(const Dual<K>& a) {
  Dual<K> r;
  r.v = -a.v;
#pragma unroll
  for (int i = 0; i < K; ++i) r.d[i] = -a.d[i];
  return r;
}
template <int K> __device__ __forceinline__ Dual<K> operator+(const Dual<K>& a, double b) { return a + Dual<K>(b); }
template <int K> __device__ __forceinline__ Dual<K> operator+(double a, const Dual<K>& b) { return Dual<K>(a) + b; }
template <int K> __device__ __forceinline__ Dual<K> operator-(const Dual<K>& a, double b) { return a - Dual<K>(b); }
template <int K> __device__ __forceinline__ Dual<K> operator-(double a, const Dual<K>& b) { return Dual<K>(a) - b; }
template <int K> __device__ __forceinline__ Dual<K> operator*(const Dual<K>& a, double b) {
  Dual<K> r;
  r.v = a.v * b;
#pragma unroll
  for (int i = 0; i < K; ++i) r.d[i] = a.d[i] * b;
  return r;
}
template <int K> __device__ __forceinline__ Dual<K> operator*(double a, const Dual<K>& b) { return b * a; }
template <int K> __device__ __forceinline__ Dual<K> operator/(const Dual<K>& a, double b) { return a * (1.0 / b); }
template <int K> __device__ __forceinline__ Dual<K> operator/(double a, const Dual<K>& b) { return Dual<K>(a) / b; }

template <int K>
__device__ __forceinline__ Dual<K> dsqrt(const Dual<K>& a) {
  Dual<K> r;
  r.v = __builtin_sqrt(a.v);
  const double g = r.v > 0.0 ? 0.5 / r.v : 0.0;
#pragma unroll
  for (int i = 0; i < K; ++i) r.d[i] = a.d[i] * g;
  return r;
}
template <int K>
__device__ __forceinline__ Dual<K> dabs(const Dual<K>& a) { return a.v < 0.0 ? -a : a; }
template <int K>
__device__ __forceinline__ Dual<K> pos(const Dual<K>& a) { return a.v > 0.0 ? a : Dual<K>(0.0); }

// Contributions of one pixel to the sums the objective reads.
template <int K>
struct PixSums {
  Dual<K> EE, EC, GN, RN;
  int M;
};

// The float64 scan chain of one pixel (sky_stat.py:74-405, dpf.py:396-450) on
// already delayed amplitudes v (2N values) with weights r and patterns fp0/fx0.
template <int N, int K>
__device__ void cont_pixel(const Dual<K> (&v00)[N], const Dual<K> (&v90)[N], const double (&r)[N],
                           const Dual<K> (&fp0)[N], const Dual<K> (&fx0)[N], double rr,
                           double RR, double En, PixSums<K>& o) {
  Dual<K> aa(0.0), AA(0.0);
#pragma unroll
  for (int i = 0; i < N; ++i) {
    aa = aa + v00[i] * v00[i];
    AA = AA + v90[i] * v90[i];
  }
  const Dual<K> e = aa + AA + 1e-12;
  o.EE = Dual<K>(0.0);
  o.EC = Dual<K>(0.0);
  o.GN = Dual<K>(0.0);
  o.RN = Dual<K>(0.0);
  o.M = 0;
  if (!(e.v > En)) return;  // masked: every contribution is exactly zero
  o.M = 1;
  o.EE = e;
  const Dual<K>& et = e;
  // dominant polarisation frame
  Dual<K> f[N], F[N];
  Dual<K> ff(0.0), FF(0.0), fF(0.0);
#pragma unroll
  for (int i = 0; i < N; ++i) {
    f[i] = fp0[i] * r[i];
    F[i] = fx0[i] * r[i];
    ff = ff + f[i] * f[i];
    FF = FF + F[i] * F[i];
    fF = fF + F[i] * f[i];
  }
  const Dual<K> s = 2.0 * fF, c = ff - FF, AP = ff + FF;
  const Dual<K> nn = dsqrt(c * c + s * s);
  const Dual<K> cc = c / (nn + 1e-9);
  const Dual<K> fp = (AP + nn) / 2.0;
  const Dual<K> si = dsqrt((1.0 - cc) / 2.0);
  const Dual<K> cq = dsqrt((1.0 + cc) / 2.0);
  const Dual<K> co = s.v > 0.0 ? cq : -cq;
  Dual<K> t(0.0);
#pragma unroll
  for (int i = 0; i < N; ++i) {
    const Dual<K> a = f[i], b = F[i];
    f[i] = a * co + b * si;
    F[i] = b * co - a * si;
    t = t + f[i] * F[i];
  }
  t = t / (fp + 1e-9);
  Dual<K> fx(0.0), ni(0.0);
#pragma unroll
  for (int i = 0; i < N; ++i) {
    F[i] = F[i] - f[i] * t;
    fx = fx + F[i] * F[i];
    const Dual<K> f2 = f[i] * f[i];
    ni = ni + f2 * f2;
  }
  ni = ni / (fp * fp + 1e-9);
  // avx_GW_ps
  Dual<K> xp(0.0), XP(0.0), xx(0.0), XX(0.0);
#pragma unroll
  for (int i = 0; i < N; ++i) {
    xp = xp + v00[i] * f[i];
    XP = XP + v90[i] * f[i];
    xx = xx + v00[i] * F[i];
    XX = XX + v90[i] * F[i];
  }
  Dual<K> tf = dsqrt(ni * (xp * xp + XP * XP) / (et + 1e-9)) * rr - fp;
  tf = pos(tf);
  const Dual<K> fd = 1.0 / (fp + tf + 1e-9);
  Dual<K> h = xp * fd;
  const Dual<K> H = XP * fd;
  h = h * h + H * H;
  const Dual<K> Hx = xx * xx + XX * XX;
  Dual<K> Fd = dsqrt(Hx / (h + 1e-9));
  const Dual<K> R = 0.1 + RR / (et + 1e-9);
  Fd = pos(Fd * R - fx);
  Fd = 1.0 / (fx + Fd + 1e-9);
  const Dual<K> au = xp * fd, AU = XP * fd, av = xx * Fd, AV = XX * Fd;
  const Dual<K> mk = fd * fp + Fd * fx;  // (a + 1) - 1
  Dual<K> p[N], q[N];
#pragma unroll
  for (int i = 0; i < N; ++i) {
    p[i] = f[i] * au + F[i] * av;
    q[i] = f[i] * AU + F[i] * AV;
  }
  // avx_ort_ps
  Dual<K> pa(0.0), PA(0.0), pA(0.0);
#pragma unroll
  for (int i = 0; i < N; ++i) {
    pa = pa + p[i] * p[i];
    PA = PA + q[i] * q[i];
    pA = pA + p[i] * q[i];
  }
  const Dual<K> s2 = 2.0 * pA, c2 = pa - PA;
  const Dual<K> nn2 = dsqrt(c2 * c2 + s2 * s2);
  const Dual<K> cc2 = c2 / (nn2 + 1e-21);
  const Dual<K> si2 = dsqrt((1.0 - cc2) / 2.0);
  const Dual<K> cq2 = dsqrt((1.0 + cc2) / 2.0);
  const Dual<K> co2 = s2.v > 0.0 ? cq2 : -cq2;
  // avx_stat_ps
  Dual<K> c3(0.0), C3(0.0), ss(0.0), SS(0.0), rq(0.0), RQ(0.0), xs(0.0), XS(0.0);
#pragma unroll
  for (int i = 0; i < N; ++i) {
    const Dual<K> s_ = p[i] * co2 + q[i] * si2;
    const Dual<K> x_ = v00[i] * co2 + v90[i] * si2;
    const Dual<K> S_ = q[i] * co2 - p[i] * si2;
    const Dual<K> X_ = v90[i] * co2 - v00[i] * si2;
    const Dual<K> a = s_ * x_, A = S_ * X_;
    xs = xs + a;
    XS = XS + A;
    c3 = c3 + a * a;
    C3 = C3 + A * A;
    ss = ss + s_ * s_;
    SS = SS + S_ * S_;
    const Dual<K> dd = p[i] - v00[i], DD = q[i] - v90[i];
    rq = rq + dd * dd;
    RQ = RQ + DD * DD;
  }
  if (!(mk.v >= 0.0)) return;  // mk' = mk - 0 < 0: ec = gn = rn = 0
  c3 = c3 / (xs * xs + 1e-9);
  C3 = C3 / (XS * XS + 1e-9);
  o.EC = ss * (1.0 - c3) + SS * (1.0 - C3);
  o.GN = 2.0 * mk;
  o.RN = rq + RQ;
}

// Objective from the sums (sky_scan.py:131-141, no netCC gate).
template <int K>
__device__ Dual<K> objective(const Dual<K>& EE, const Dual<K>& EC, const Dual<K>& GN,
                             const Dual<K>& RN, int M, int n) {
  if (M == 0) return Dual<K>(0.0);
  const Dual<K> Eo = EE / 2.0;
  const Dual<K> No = (GN + RN) / 2.0;
  const Dual<K> CH = No / ((double)n * M + __builtin_sqrt((double)M));
  const Dual<K> cc = CH.v > 1.0 ? CH : Dual<K>(1.0);
  const Dual<K> Co = EC / (EC + No * cc - (double)M * (n - 1));
  const Dual<K> L = Eo.v > 0.0 ? Eo - No : Dual<K>(0.0);
  return L * Co;
}

struct GradArgs {
  const int32_t* clusters; // NULL: block b evaluates cluster b
  int n_ifo;
  const int32_t* npix;
  const int32_t* net;
  const int32_t* K;
  const int64_t* rms_off;
  const int64_t* td_off;
  const float* rms;
  const float2* td;
  const double* geo;       // per network, per ifo: delay vector (3) + response (9)
  const int32_t* net_nifo;
  const double* net_En;
  const double* REG1;      // per cluster
  double reg0;
  const double* theta;     // per cluster
  const double* phi;
  double* value;           // per cluster
  double* grad;            // per cluster (dtheta, dphi)
  float* grad_td;          // NULL or dense, same layout as td (float pairs)
  const double* grad_scale;// NULL or per cluster factor on grad_td
};

// Direction-dependent inputs of one detector: continuous delay and patterns.
__device__ void direction_terms(const double* g, double th, double ph, Dual<2>& tau,
                                Dual<2>& Fp, Dual<2>& Fx) {
  const double st = sin(th), ct = cos(th), sp = sin(ph), cp = cos(ph);
  Dual<2> nx, ny, nz;
  nx.v = st * cp; nx.d[0] = ct * cp; nx.d[1] = -st * sp;
  ny.v = st * sp; ny.d[0] = ct * sp; ny.d[1] = st * cp;
  nz.v = ct;      nz.d[0] = -st;     nz.d[1] = 0.0;
  tau = nx * g[0] + ny * g[1] + nz * g[2];
  // gha = -phi, dec = pi/2 - theta, psi = 0:
  // x = (sin phi, -cos phi, 0), y = (-cos phi cos theta, -sin phi cos theta, sin theta)
  Dual<2> x[3], y[3];
  x[0].v = sp;       x[0].d[0] = 0.0;      x[0].d[1] = cp;
  x[1].v = -cp;      x[1].d[0] = 0.0;      x[1].d[1] = sp;
  x[2] = Dual<2>(0.0);
  y[0].v = -cp * ct; y[0].d[0] = cp * st;  y[0].d[1] = sp * ct;
  y[1].v = -sp * ct; y[1].d[0] = sp * st;  y[1].d[1] = -cp * ct;
  y[2].v = st;       y[2].d[0] = ct;       y[2].d[1] = 0.0;
  const double* D = g + 3;
  Dual<2> Dx[3], Dy[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    Dx[a] = x[0] * D[3 * a] + x[1] * D[3 * a + 1] + x[2] * D[3 * a + 2];
    Dy[a] = y[0] * D[3 * a] + y[1] * D[3 * a + 1] + y[2] * D[3 * a + 2];
  }
  Fp = (x[0] * Dx[0] + x[1] * Dx[1] + x[2] * Dx[2]) - (y[0] * Dy[0] + y[1] * Dy[1] + y[2] * Dy[2]);
  Fx = (x[0] * Dy[0] + x[1] * Dy[1] + x[2] * Dy[2]) + (y[0] * Dx[0] + y[1] * Dx[1] + y[2] * Dx[2]);
}

__device__ __forceinline__ double wave_sum(double x) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
  return x;
}

// Block-wide sum of a dual number (value and every tangent); all threads
// receive the total.  sh holds (kBlock / 64) * (KD + 1) doubles.
template <int KD>
__device__ void block_sum_dual(Dual<KD>& a, double* sh) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  double v = wave_sum(a.v);
  double d[KD];
#pragma unroll
  for (int k = 0; k < KD; ++k) d[k] = wave_sum(a.d[k]);
  __syncthreads();
  if (lane == 0) {
    sh[wave * (KD + 1)] = v;
#pragma unroll
    for (int k = 0; k < KD; ++k) sh[wave * (KD + 1) + 1 + k] = d[k];
  }
  __syncthreads();
  a.v = 0.0;
#pragma unroll
  for (int k = 0; k < KD; ++k) a.d[k] = 0.0;
  for (int w = 0; w < kBlock / 64; ++w) {
    a.v += sh[w * (KD + 1)];
#pragma unroll
    for (int k = 0; k < KD; ++k) a.d[k] += sh[w * (KD + 1) + 1 + k];
  }
  __syncthreads();
}

__device__ int block_sum_int(int v, double* sh) {
  const double x = wave_sum((double)v);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) sh[wave] = x;
  __syncthreads();
  double s = 0.0;
  for (int w = 0; w < kBlock / 64; ++w) s += sh[w];
  __syncthreads();
  return (int)s;
}

template <int N, bool WANT_TD>
__device__ void grad_cluster(const GradArgs& g, int c) {
  __shared__ double sh[kBlock / 64 * 12];
  __shared__ double adj[4];
  const int P = g.npix[c], K = g.K[c], nd = 2 * K + 1, net = g.net[c];
  const double En = g.net_En[net];
  const double* geo = g.geo + (int64_t)net * 8 * 12;
  const float* rms = g.rms + g.rms_off[c];
  const float2* td = g.td + g.td_off[c];
  const double th = g.theta[c], ph = g.phi[c];
  const double rr = g.reg0, RR = g.REG1[c];
  Dual<2> fp0[N], fx0[N], w[N];
  int i0[N];
#pragma unroll
  for (int i = 0; i < N; ++i) {
    Dual<2> tau;
    direction_terms(geo + 12 * i, th, ph, tau, fp0[i], fx0[i]);
    Dual<2> u = tau + (double)K;
    double uv = u.v;
    if (!(uv >= 0.0)) uv = 0.0;  // also maps a NaN direction to an in-bounds tap
    if (uv > nd - 1 - 1e-9) uv = nd - 1 - 1e-9;
    i0[i] = (int)floor(uv);
    if (u.v != uv) u = Dual<2>(uv);  // clamped: the delay no longer moves with direction
    w[i] = u - (double)i0[i];
  }
  PixSums<2> acc;
  acc.EE = Dual<2>(0.0); acc.EC = Dual<2>(0.0); acc.GN = Dual<2>(0.0); acc.RN = Dual<2>(0.0);
  acc.M = 0;
  for (int j = threadIdx.x; j < P; j += kBlock) {
    Dual<2> v00[N], v90[N];
    double r[N];
#pragma unroll
    for (int i = 0; i < N; ++i) {
      const float2 a = td[((int64_t)j * N + i) * nd + i0[i]];
      const float2 b = td[((int64_t)j * N + i) * nd + i0[i] + 1];
      v00[i] = (1.0 - w[i]) * (double)a.x + w[i] * (double)b.x;
      v90[i] = (1.0 - w[i]) * (double)a.y + w[i] * (double)b.y;
      r[i] = rms[(int64_t)j * N + i];
    }
    PixSums<2> o;
    cont_pixel<N, 2>(v00, v90, r, fp0, fx0, rr, RR, En, o);
    acc.EE = acc.EE + o.EE;
    acc.EC = acc.EC + o.EC;
    acc.GN = acc.GN + o.GN;
    acc.RN = acc.RN + o.RN;
    acc.M += o.M;
  }
  block_sum_dual<2>(acc.EE, sh);
  block_sum_dual<2>(acc.EC, sh);
  block_sum_dual<2>(acc.GN, sh);
  block_sum_dual<2>(acc.RN, sh);
  const int M = block_sum_int(acc.M, sh);
  const Dual<2> AA = objective<2>(acc.EE, acc.EC, acc.GN, acc.RN, M, N);
  if (threadIdx.x == 0) {
    g.value[c] = AA.v;
    g.grad[2 * c] = AA.d[0];
    g.grad[2 * c + 1] = AA.d[1];
  }
  if (!WANT_TD) return;
  // ---- amplitude gradient: adjoints of AA with respect to (EE, EC, GN, RN) ----
  if (threadIdx.x == 0) {
    Dual<4> S[4];
    const double vals[4] = {acc.EE.v, acc.EC.v, acc.GN.v, acc.RN.v};
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      S[k] = Dual<4>(vals[k]);
      S[k].d[k] = 1.0;
    }
    const Dual<4> A4 = objective<4>(S[0], S[1], S[2], S[3], M, N);
    const double scale = g.grad_scale ? g.grad_scale[c] : 1.0;
#pragma unroll
    for (int k = 0; k < 4; ++k) adj[k] = scale * A4.d[k];
  }
  __syncthreads();
  float* gtd = g.grad_td + 2 * g.td_off[c];
  Dual<2 * N> fpd[N], fxd[N];
#pragma unroll
  for (int i = 0; i < N; ++i) {
    fpd[i] = Dual<2 * N>(fp0[i].v);
    fxd[i] = Dual<2 * N>(fx0[i].v);
  }
  for (int j = threadIdx.x; j < P; j += kBlock) {
    Dual<2 * N> v00[N], v90[N];
    double r[N];
#pragma unroll
    for (int i = 0; i < N; ++i) {
      const float2 a = td[((int64_t)j * N + i) * nd + i0[i]];
      const float2 b = td[((int64_t)j * N + i) * nd + i0[i] + 1];
      v00[i] = Dual<2 * N>((1.0 - w[i].v) * (double)a.x + w[i].v * (double)b.x);
      v90[i] = Dual<2 * N>((1.0 - w[i].v) * (double)a.y + w[i].v * (double)b.y);
      v00[i].d[i] = 1.0;
      v90[i].d[N + i] = 1.0;
      r[i] = rms[(int64_t)j * N + i];
    }
    PixSums<2 * N> o;
    cont_pixel<N, 2 * N>(v00, v90, r, fpd, fxd, rr, RR, En, o);
#pragma unroll
    for (int i = 0; i < N; ++i) {
#pragma unroll
      for (int ph2 = 0; ph2 < 2; ++ph2) {
        const int k = ph2 * N + i;
        const double gv = adj[0] * o.EE.d[k] + adj[1] * o.EC.d[k] + adj[2] * o.GN.d[k] +
                          adj[3] * o.RN.d[k];
        const int64_t base = 2 * (((int64_t)j * N + i) * nd + i0[i]) + ph2;
        gtd[base] = (float)((1.0 - w[i].v) * gv);
        gtd[base + 2] = (float)(w[i].v * gv);
      }
    }
  }
}

template <int N, bool WANT_TD>
__global__ __launch_bounds__(kBlock) void k_sky_grad(GradArgs g) {
  const int c = g.clusters ? g.clusters[blockIdx.x] : (int)blockIdx.x;
  if (g.net_nifo[g.net[c]] != N) {  // caller grouped clusters wrongly: flag, never read
    if (threadIdx.x == 0) {
      g.value[c] = __builtin_nan("");
      g.grad[2 * c] = __builtin_nan("");
      g.grad[2 * c + 1] = __builtin_nan("");
    }
    return;
  }
  grad_cluster<N, WANT_TD>(g, c);
}

template <int N>
void launch_grad(const GradArgs& g, int blocks, bool want_td, hipStream_t s) {
  if (want_td)
    hipLaunchKernelGGL((k_sky_grad<N, true>), dim3((unsigned)blocks), dim3(kBlock), 0, s, g);
  else
    hipLaunchKernelGGL((k_sky_grad<N, false>), dim3((unsigned)blocks), dim3(kBlock), 0, s, g);
}

}  // namespace

extern "C" {

int pycwb_sky_grad(const pycwb_grad_desc* d, void* stream) {
  if (!d || d->n_clusters < 0 || d->n_ifo < 1 || d->n_ifo > 5) return PYCWB_ERR_ARG;
  if (d->n_clusters == 0) return PYCWB_OK;
  if (!d->n_pix || !d->net || !d->K || !d->rms_off || !d->td_off || !d->rms || !d->td ||
      !d->geo || !d->net_nifo || !d->net_En || !d->REG1 || !d->theta || !d->phi || !d->value ||
      !d->grad)
    return PYCWB_ERR_ARG;
  (void)hipGetLastError();
  GradArgs g;
  g.clusters = d->clusters; g.n_ifo = d->n_ifo;
  g.npix = d->n_pix; g.net = d->net; g.K = d->K; g.rms_off = d->rms_off; g.td_off = d->td_off;
  g.rms = d->rms; g.td = (const float2*)d->td; g.geo = d->geo; g.net_nifo = d->net_nifo;
  g.net_En = d->net_En; g.REG1 = d->REG1; g.reg0 = d->reg0; g.theta = d->theta; g.phi = d->phi;
  g.value = d->value; g.grad = d->grad; g.grad_td = d->grad_td; g.grad_scale = d->grad_scale;
  const bool want_td = d->grad_td != nullptr;
  const hipStream_t s = (hipStream_t)stream;
  switch (d->n_ifo) {
    case 1: launch_grad<1>(g, d->n_clusters, want_td, s); break;
    case 2: launch_grad<2>(g, d->n_clusters, want_td, s); break;
    case 3: launch_grad<3>(g, d->n_clusters, want_td, s); break;
    case 4: launch_grad<4>(g, d->n_clusters, want_td, s); break;
    default: launch_grad<5>(g, d->n_clusters, want_td, s); break;
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    pycwb::record_error("k_sky_grad", hipGetErrorString(e));
    return PYCWB_ERR_HIP;
  }
  return PYCWB_OK;
}

}  // extern "C"
