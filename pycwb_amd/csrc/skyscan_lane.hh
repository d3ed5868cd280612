// Per-lane arithmetic of the coherent sky scan: one lane owns one sky direction
// and sweeps the cluster's pixels in order, carrying every per-direction sum in
// registers.  Each step below restates one pass of PycWB's Numba scan
// (pycwb/modules/likelihoodWP/sky_stat.py, dpf.py) for a single pixel, with the
// float32/float64 typing that Numba's promotion rules give those expressions:
//   int32 (x) float32 -> float64, python float literal -> float64,
//   float32 ** int -> float32 by repeated squaring, loop-carried f32/f64 -> f64.
// All HIP sources are compiled with -ffp-contract=off, IEEE denormals and the
// default correctly rounded f32 divide / sqrt, so every operation below rounds
// exactly like the CPU restatement.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pycwb {

// Thread-local text of the last HIP failure (read by pycwb_last_error).
void record_error(const char* what, const char* why);
const char* last_error();

constexpr int kNumMaps = 11;

// Per-direction sums of one sky sweep.
struct SkySums {
  float EE;      // sum of masked pixel energy (load_data_from_td)
  int M;         // pixels above the network energy threshold
  double Lr;     // reduced likelihood, float64 accumulator (sky_stat.py:398)
  double LL;     // total signal energy, float64 (sky_stat.py:401)
  float EC;      // coherent energy (sky_stat.py:402)
  float GN;      // G-noise correction (sky_stat.py:403)
  float RN;      // residual noise (sky_stat.py:404)
  double Mp;     // pixels with ec > 1e-9 (sky_stat.py:405)
  float sde;     // selected data energy (sky_scan.py:151)
  float pwe;     // energy-weighted |f+|^2 (sky_scan.py:152)
  float cwe;     // energy-weighted |fx|^2 (sky_scan.py:153)
};

__device__ __forceinline__ void sums_init(SkySums& a) {
  a.EE = 0.f; a.M = 0; a.Lr = 0.0; a.LL = 0.0; a.EC = 0.f; a.GN = 0.f; a.RN = 0.f;
  a.Mp = 0.0; a.sde = 0.f; a.pwe = 0.f; a.cwe = 0.f;
}

// Dominant polarisation frame of one pixel (dpf.py:396-450).
template <int N>
struct Dpf {
  float f[N], F[N];
  float fp, fx, ni;
};

template <int N>
__device__ __forceinline__ void dpf_pixel(const float (&r)[N], const float (&fp0)[N],
                                          const float (&fx0)[N], Dpf<N>& d) {
  const float o = 1e-9f;
  float ff = 0.f, FF = 0.f, fF = 0.f;
#pragma unroll
  for (int i = 0; i < N; ++i) {
    d.f[i] = r[i] * fp0[i];
    d.F[i] = r[i] * fx0[i];
  }
#pragma unroll
  for (int i = 0; i < N; ++i) {
    ff += d.f[i] * d.f[i];
    FF += d.F[i] * d.F[i];
    fF += d.F[i] * d.f[i];
  }
  const float s = 2.f * fF;
  const float c = ff - FF;
  const float AP = ff + FF;
  const float nn = __builtin_sqrtf(c * c + s * s);
  const float cc = c / (nn + o);
  d.fp = (AP + nn) / 2.f;
  const float si = __builtin_sqrtf((1.f - cc) / 2.f);
  const float cq = __builtin_sqrtf((1.f + cc) / 2.f);
  const float co = s > 0.f ? cq : -cq;
  float t = 0.f;
#pragma unroll
  for (int i = 0; i < N; ++i) {
    const float a = d.f[i], b = d.F[i];
    d.f[i] = a * co + b * si;
    d.F[i] = b * co - a * si;
  }
#pragma unroll
  for (int i = 0; i < N; ++i) t += d.f[i] * d.F[i];
  t = t / (d.fp + o);
  float fx = 0.f, ni = 0.f;
#pragma unroll
  for (int i = 0; i < N; ++i) {
    d.F[i] -= d.f[i] * t;
    fx += d.F[i] * d.F[i];
    const float f2 = d.f[i] * d.f[i];
    ni += f2 * f2;
  }
  d.fx = fx;
  d.ni = ni / (d.fp * d.fp + o);
}

// Regulator contribution of one pixel: fx / (ni + 1e-9) and [fp > 0]
// (dpf.py:444-450).
template <int N>
__device__ __forceinline__ void regulator_pixel(const float (&r)[N], const float (&fp0)[N],
                                                const float (&fx0)[N], float& NI,
                                                unsigned& NN) {
  Dpf<N> d;
  dpf_pixel<N>(r, fp0, fx0, d);
  NI += d.fx / (d.ni + 1e-9f);
  NN += d.fp > 0.f ? 1u : 0u;
}

// Every intermediate of one (direction, pixel) evaluation that a later
// consumer needs: the sky scan keeps only the sums, the best-sky statistics
// keep the per-pixel arrays.
template <int N>
struct PixelEval {
  float e;         // pixel energy + 1e-12
  int m;           // e > En
  float et;        // masked energy
  Dpf<N> d;        // dominant polarisation frame
  float p[N], q[N];// signal packet (avx_GW_ps)
  float mk;        // updated mask (float32 array of avx_GW_ps)
  float ee, EE;    // orthogonalised packet energies (avx_ort_ps)
  float ec, gn, rn;// coherent energy, G-noise, residual noise (avx_stat_ps)
  double ll, ccd;  // signal energy and correlation factor (float64)
};

// load_data_from_td of one pixel (sky_stat.py:74-87): energy, mask, masked energy.
template <int N>
__device__ __forceinline__ void pixel_energy(const float (&v00)[N], const float (&v90)[N],
                                             double En, float& e, int& m, float& et) {
  float aa = 0.f, AA = 0.f;
#pragma unroll
  for (int i = 0; i < N; ++i) {
    aa += v00[i] * v00[i];
    AA += v90[i] * v90[i];
  }
  e = aa + AA + 1e-12f;
  m = (double)e > En ? 1 : 0;
  et = e * (float)m;
}

// A pixel below the energy threshold (m == 0) with finite energy adds exactly
// nothing to any per-direction sum: fd = Fd = +0, so the signal packet is
// +-0, ll = ec = rn = +0, gn = -0, ccd = +0 and the updated mask is -1.
// Given finite noise weights, antenna patterns and regulators (checked by the
// host), skipping such a pixel leaves every sum bit-identical.
__device__ __forceinline__ bool pixel_contributes(float e, int m) {
  return m != 0 || !isfinite(e);
}

// Everything after load_data_from_td for one pixel.
template <int N>
__device__ __forceinline__ void pixel_rest(const float (&v00)[N], const float (&v90)[N],
                                           const float (&r)[N], const float (&fp0)[N],
                                           const float (&fx0)[N], float rr, float RR, float e,
                                           int m, float et, PixelEval<N>& o) {
  o.e = e;
  o.m = m;
  o.et = et;

  // --- dpf_np_loops_vec ---
  Dpf<N>& d = o.d;
  dpf_pixel<N>(r, fp0, fx0, d);

  // --- avx_GW_ps (sky_stat.py:156-189) ---
  const float eps = 1e-9f;
  float xp = 0.f, XP = 0.f, xx = 0.f, XX = 0.f;
#pragma unroll
  for (int i = 0; i < N; ++i) {
    xp += v00[i] * d.f[i];
    XP += v90[i] * d.f[i];
    xx += v00[i] * d.F[i];
    XX += v90[i] * d.F[i];
  }
  float tf = __builtin_sqrtf(d.ni * (xp * xp + XP * XP) / (et + eps)) * rr - d.fp;
  tf = tf > 0.f ? tf : 0.f;
  const double md = (double)m;
  const double fd = md / (double)(d.fp + tf + eps);
  double h = (double)xp * fd;
  const double H = (double)XP * fd;
  h = h * h + H * H;
  const float Hx = xx * xx + XX * XX;
  double Fd = __builtin_sqrt((double)Hx / (h + (double)eps));
  const float R = 0.1f + RR / (et + eps);
  Fd = Fd * (double)R - (double)d.fx;
  Fd = Fd > 0.0 ? Fd : 0.0;
  Fd = md / (((double)d.fx + Fd) + (double)eps);
  const float au = (float)((double)xp * fd);
  const float AU = (float)((double)XP * fd);
  const float av = (float)((double)xx * Fd);
  const float AV = (float)((double)XX * Fd);
  const double ga = fd * (double)d.fp + Fd * (double)d.fx;
  const float mk = (float)((ga + md) - 1.0);
  o.mk = mk;
#pragma unroll
  for (int i = 0; i < N; ++i) {
    o.p[i] = d.f[i] * au + d.F[i] * av;
    o.q[i] = d.f[i] * AU + d.F[i] * AV;
  }

  // --- avx_ort_ps (sky_stat.py:264-288) ---
  float pa = 0.f, PA = 0.f, pA = 0.f;
#pragma unroll
  for (int i = 0; i < N; ++i) {
    pa += o.p[i] * o.p[i];
    PA += o.q[i] * o.q[i];
    pA += o.p[i] * o.q[i];
  }
  const float s2 = pA * 2.f;
  const float c2 = pa - PA;
  const float et2 = pa + PA + 1e-21f;
  const float nn2 = __builtin_sqrtf(c2 * c2 + s2 * s2);
  o.ee = (et2 + nn2) / 2.f;
  o.EE = (et2 - nn2) / 2.f;
  const float cc2 = c2 / (nn2 + 1e-21f);
  const float si = __builtin_sqrtf((1.f - cc2) / 2.f);
  const float cq = __builtin_sqrtf((1.f + cc2) / 2.f);
  const float co = s2 > 0.f ? cq : -cq;

  // --- avx_stat_ps (sky_stat.py:357-398) ---
  float c = 0.f, C = 0.f, ss = 0.f, SS = 0.f, rq = 0.f, RQ = 0.f, xs = 0.f, XS = 0.f;
#pragma unroll
  for (int i = 0; i < N; ++i) {
    const float s_ = o.p[i] * co + o.q[i] * si;
    const float x_ = v00[i] * co + v90[i] * si;
    const float S_ = o.q[i] * co - o.p[i] * si;
    const float X_ = v90[i] * co - v00[i] * si;
    const float a = s_ * x_;
    const float A = S_ * X_;
    xs += a;
    XS += A;
    c += a * a;
    C += A * A;
    ss += s_ * s_;
    SS += S_ * S_;
    const float dd = o.p[i] - v00[i];
    const float DD = o.q[i] - v90[i];
    rq += dd * dd;
    RQ += DD * DD;
  }
  const double mkd = mk >= 0.f ? 1.0 : 0.0;
  c = c / (xs * xs + eps);
  C = C / (XS * XS + eps);
  o.ll = mkd * (double)(ss + SS);
  const double ssd = (double)ss * (1.0 - (double)c);
  const double SSd = (double)SS * (1.0 - (double)C);
  o.ec = (float)(mkd * (ssd + SSd));
  o.gn = (float)((mkd * 2.0) * (double)mk);
  o.rn = (float)(mkd * (double)(rq + RQ));
  const double a2 = 2.0 * (double)fabsf(o.ec);
  const float A2 = o.rn + o.gn + eps;
  o.ccd = (double)o.ec / (a2 + (double)A2);
}

// load_data_from_td, DPF, avx_GW_ps, avx_ort_ps and avx_stat_ps of one pixel.
template <int N>
__device__ __forceinline__ void pixel_eval(const float (&v00)[N], const float (&v90)[N],
                                           const float (&r)[N], const float (&fp0)[N],
                                           const float (&fx0)[N], float rr, float RR,
                                           double En, PixelEval<N>& o) {
  float e, et;
  int m;
  pixel_energy<N>(v00, v90, En, e, m, et);
  pixel_rest<N>(v00, v90, r, fp0, fx0, rr, RR, e, m, et, o);
}

// One pixel of the scan: pixel_eval plus the per-direction sums
// (sky_stat.py:398-405, sky_scan.py:148-153).  With skip_masked, pixels that
// cannot contribute (pixel_contributes) stop after the energy test.
template <int N>
__device__ __forceinline__ void scan_pixel(const float (&v00)[N], const float (&v90)[N],
                                           const float (&r)[N], const float (&fp0)[N],
                                           const float (&fx0)[N], float rr, float RR,
                                           double En, bool skip_masked, SkySums& acc) {
  float e, et;
  int m;
  pixel_energy<N>(v00, v90, En, e, m, et);
  if (skip_masked && !pixel_contributes(e, m)) return;
  PixelEval<N> o;
  pixel_rest<N>(v00, v90, r, fp0, fx0, rr, RR, e, m, et, o);
  acc.M += o.m;
  acc.EE += o.et;
  acc.Lr += o.ll * o.ccd;
  acc.LL += o.ll;
  acc.EC += o.ec;
  acc.GN += o.gn;
  acc.RN += o.rn;
  acc.Mp += o.ec > 1e-9f ? 1.0 : 0.0;
  if (o.mk > 0.f) {
    acc.sde += o.et;
    acc.pwe += o.d.fp * o.et;
    acc.cwe += o.d.fx * o.et;
  }
}

// Per-direction statistics (sky_scan.py:131-179).
struct SkyStat {
  bool gated;      // skipped: no active pixel, non-finite or sub-threshold Cr
  float maps[kNumMaps];
  float AA;        // nSkyStat value (float32-rounded)
};

__device__ __forceinline__ void sky_epilogue(const SkySums& a, int n, double netCC,
                                             SkyStat& out) {
  const float Eo = a.EE / 2.f;
  const double Cr = (double)2.f * a.Lr / (a.LL + (double)1e-9f);
  const float No = (a.GN + a.RN) / 2.f;
  out.gated = true;
  out.AA = 0.f;
  if (a.M == 0) return;  // 0/0 in the reference; defined here as gated
  const double M = (double)a.M;
  const double CH = (double)No / ((double)((int64_t)n * a.M) + __builtin_sqrt(M));
  const double cc = CH > 1.0 ? CH : 1.0;
  const double Co =
      (double)a.EC / (((double)a.EC + (double)No * cc) - (double)((int64_t)a.M * (n - 1)));
  if (!isfinite(Cr) || Cr < netCC) return;
  out.gated = false;
  const float L = Eo > 0.f ? Eo - No : 0.f;
  const double AA = (double)L * Co;
  const float pwe = a.sde > 0.f ? a.pwe / a.sde : 0.f;
  const float cwe = a.sde > 0.f ? a.cwe / a.sde : 0.f;
  out.maps[0] = __builtin_sqrtf(pwe + cwe);
  out.maps[1] = pwe > 0.f ? __builtin_sqrtf(cwe / pwe) : 0.f;
  out.maps[2] = Eo - No;
  out.maps[3] = No;
  out.maps[4] = a.EC;
  out.maps[5] = (float)Co;
  out.maps[6] = (float)AA;
  out.maps[7] = (float)CH;
  out.maps[8] = (float)cc;
  out.maps[9] = (float)Cr;
  out.maps[10] = (float)a.Mp;
  out.AA = (float)AA;
}

// Argmax key: larger value wins, equal values -> larger valid-list position
// (the reference's forward `>=` scan, sky_scan.py:183-189).  0 means "none".
__device__ __forceinline__ unsigned long long argmax_key(float v, uint32_t pos) {
  if (v != v) return 0ull;         // NaN never satisfies `>=`
  if (v == 0.f) v = 0.f;           // -0.0 compares equal to +0.0
  uint32_t b = __float_as_uint(v);
  b = (b & 0x80000000u) ? ~b : (b | 0x80000000u);
  return ((unsigned long long)b << 32) | (unsigned long long)pos;
}

__device__ __forceinline__ float key_value(unsigned long long k) {
  uint32_t b = (uint32_t)(k >> 32);
  b = (b & 0x80000000u) ? (b & 0x7fffffffu) : ~b;
  return __uint_as_float(b);
}

}  // namespace pycwb
