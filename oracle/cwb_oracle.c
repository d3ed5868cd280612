/*
 * ORACLE -- TEST INFRASTRUCTURE ONLY.
 *
 * CPU restatement of PycWB's Numba likelihood sky scan (pycwb/modules/likelihoodWP)
 * and of the cWB WDM forward transform.  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg load this library, and only as the checker.  The
 * product path (pycwb_amd/) never links or calls it.
 *
 * Two arithmetic modes:
 *   mixed : op-by-op Numba typing of the reference (float32 arrays and sums, the
 *           float64 islands that Numba's promotion rules create, no FMA
 *           contraction, correctly rounded div/sqrt, forward sequential sums).
 *           Build with -ffp-contract=off.
 *   f64   : the same algebra with every intermediate in float64; the yardstick
 *           for "error against an FP64 oracle".
 *
 * The loop structure deliberately follows the reference's separate passes
 * (load_data_from_td -> dpf_np_loops_vec -> avx_GW_ps -> avx_ort_ps ->
 * avx_stat_ps) with per-sky temporaries, so that it is an independent
 * restatement of the fused single-pass GPU kernel.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define NMAPS 11

/* Planted-mistake switch.  0 in every use except tests/test_oracle_pins.py,
 * which sets it to k for one check at a time to show that the closed-form pins
 * catch the k-th plausible transcription slip (a flipped sign, a dropped term
 * or factor, a wrong index or operand).  Each MUT(k) site names its slip. */
int oracle_mutant = 0;
#define MUT(k) (oracle_mutant == (k))

/* ------------------------------------------------------------------------- */
/* dpf_np_loops_vec (likelihoodWP/dpf.py:342-453), mixed precision.          */
/* rms is (P, n) row-major.  f, F are (P, n); fp, fx, si, co, ni are (P).    */
/* Returns sqrt(NI / (NN + 0.01)) in float64 (dpf.py:453).                   */
/* ------------------------------------------------------------------------- */
static double dpf_mixed(const float *Fp0, const float *Fx0, const float *rms,
                        int P, int n, float *f, float *F, float *fp, float *fx,
                        float *si, float *co, float *ni)
{
    const float o = 1e-9f;
    /* dpf.py:396-399: mul_vec is a float32 ufunc */
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < P; ++i) {
            f[i * n + j] = rms[i * n + j] * Fp0[j];
            F[i * n + j] = rms[i * n + j] * Fx0[j];
        }
    /* dpf.py:402-420 */
    for (int i = 0; i < P; ++i) {
        float ff = 0.f, FF = 0.f, fF = 0.f;
        for (int j = 0; j < n; ++j) {
            ff += f[i * n + j] * f[i * n + j];
            FF += F[i * n + j] * F[i * n + j];
            fF += F[i * n + j] * f[i * n + j];
        }
        float s = 2.f * fF;                 /* mul_vec(2, fF) */
        float c = ff - FF;                  /* sub_vec */
        float AP = ff + FF;                 /* add_vec */
        float nn = sqrtf(c * c + s * s);    /* norm_vec(co, si) */
        float cc = c / (nn + o);            /* div_vec(co, nn) */
        fp[i] = (AP + nn) / 2.f;            /* avg_vec */
        si[i] = sqrtf((1.f - cc) / 2.f);    /* sin_from_cc */
        co[i] = s > 0.f ? sqrtf((1.f + cc) / 2.f) : -sqrtf((1.f + cc) / 2.f);
        if (MUT(6)) co[i] = -co[i];                   /* slip: rotation sign */
    }
    /* dpf.py:423-438 */
    for (int i = 0; i < P; ++i) {
        for (int j = 0; j < n; ++j) {
            float a = f[i * n + j], b = F[i * n + j];
            f[i * n + j] = a * co[i] + b * si[i];
            F[i * n + j] = b * co[i] - a * si[i];
        }
        float t = 0.f;
        for (int j = 0; j < n; ++j) t += f[i * n + j] * F[i * n + j];
        t = t / (fp[i] + o);                /* div_vec(fF_new, fp) */
        fx[i] = 0.f;
        ni[i] = 0.f;
        for (int j = 0; j < n; ++j) {
            F[i * n + j] -= f[i * n + j] * t;
            fx[i] += F[i * n + j] * F[i * n + j];
            float f2 = f[i * n + j] * f[i * n + j];   /* float32 ** 4 = (x*x)*(x*x) */
            ni[i] += MUT(5) ? f2 * f[i * n + j] : f2 * f2;   /* slip: f^3 */
        }
    }
    /* dpf.py:441-453 */
    float NI = 0.f;
    uint32_t NN = 0;
    for (int i = 0; i < P; ++i) {
        ni[i] = ni[i] / (fp[i] * fp[i] + o);          /* div_vec(ni, mul_vec(fp, fp)) */
        NI += fx[i] / (ni[i] + o);                    /* div_vec(fx, ni) */
        NN += fp[i] > 0.f ? 1u : 0u;
    }
    return sqrt((double)NI / ((double)NN + 0.01));
}

/* Same algebra, all float64 (f64 oracle mode). */
static double dpf_f64(const double *Fp0, const double *Fx0, const double *rms,
                      int P, int n, double *f, double *F, double *fp, double *fx,
                      double *si, double *co, double *ni)
{
    const double o = 1e-9;
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < P; ++i) {
            f[i * n + j] = rms[i * n + j] * Fp0[j];
            F[i * n + j] = rms[i * n + j] * Fx0[j];
        }
    for (int i = 0; i < P; ++i) {
        double ff = 0, FF = 0, fF = 0;
        for (int j = 0; j < n; ++j) {
            ff += f[i * n + j] * f[i * n + j];
            FF += F[i * n + j] * F[i * n + j];
            fF += F[i * n + j] * f[i * n + j];
        }
        double s = 2.0 * fF, c = ff - FF, AP = ff + FF;
        double nn = sqrt(c * c + s * s);
        double cc = c / (nn + o);
        fp[i] = (AP + nn) / 2.0;
        si[i] = sqrt((1.0 - cc) / 2.0);
        co[i] = s > 0.0 ? sqrt((1.0 + cc) / 2.0) : -sqrt((1.0 + cc) / 2.0);
    }
    for (int i = 0; i < P; ++i) {
        for (int j = 0; j < n; ++j) {
            double a = f[i * n + j], b = F[i * n + j];
            f[i * n + j] = a * co[i] + b * si[i];
            F[i * n + j] = b * co[i] - a * si[i];
        }
        double t = 0;
        for (int j = 0; j < n; ++j) t += f[i * n + j] * F[i * n + j];
        t = t / (fp[i] + o);
        fx[i] = 0;
        ni[i] = 0;
        for (int j = 0; j < n; ++j) {
            F[i * n + j] -= f[i * n + j] * t;
            fx[i] += F[i * n + j] * F[i * n + j];
            double f2 = f[i * n + j] * f[i * n + j];
            ni[i] += f2 * f2;
        }
    }
    double NI = 0;
    unsigned NN = 0;
    for (int i = 0; i < P; ++i) {
        ni[i] = ni[i] / (fp[i] * fp[i] + o);
        NI += fx[i] / (ni[i] + o);
        NN += fp[i] > 0.0;
    }
    return sqrt(NI / ((double)NN + 0.01));
}

/* ------------------------------------------------------------------------- */
/* calculate_dpf (likelihoodWP/dpf.py:7-34).                                 */
/* FP, FX: (S, n) f32; rms (P, n) f32.  Returns the f64 regulator.           */
/* If quality_out != NULL the per-valid-sky quality aa[k] is stored there.   */
/* ------------------------------------------------------------------------- */
double oracle_calculate_dpf(int n_ifo, int n_pix, int n_sky, const float *FP,
                            const float *FX, const float *rms, double gamma_regulator,
                            double En, const int64_t *valid, int64_t n_valid,
                            double *quality_out, int n_threads)
{
    (void)n_sky;
    int64_t count = 0;
#ifdef _OPENMP
    if (n_threads > 0) omp_set_num_threads(n_threads);
#pragma omp parallel reduction(+ : count)
#endif
    {
        float *f = malloc(sizeof(float) * n_pix * n_ifo * 2 + sizeof(float) * n_pix * 5);
        float *F = f + n_pix * n_ifo, *fp = F + n_pix * n_ifo, *fx = fp + n_pix,
              *si = fx + n_pix, *co = si + n_pix, *ni = co + n_pix;
#ifdef _OPENMP
#pragma omp for schedule(static)
#endif
        for (int64_t k = 0; k < n_valid; ++k) {
            int64_t s = valid[k];
            double q = dpf_mixed(FP + s * n_ifo, FX + s * n_ifo, rms, n_pix, n_ifo,
                                 f, F, fp, fx, si, co, ni);
            if (quality_out) quality_out[k] = q;
            count += q > gamma_regulator;
        }
        free(f);
    }
    double FF = (double)n_valid, ff = (double)count;
    return (FF * FF / (ff * ff + 1.e-9) - 1) * En;
}

/* Per-pixel float32 DPF terms of the listed directions (dpf_np_loops_vec,
 * dpf.py:342-453): fp, fx and ni as the mixed-precision oracle computes them,
 * each (n_dirs, P). */
void oracle_dpf_terms(int n_ifo, int n_pix, const float *FP, const float *FX, const float *rms,
                      const int64_t *dirs, int64_t n_dirs, float *fp_out, float *fx_out,
                      float *ni_out, int n_threads)
{
#ifdef _OPENMP
    if (n_threads > 0) omp_set_num_threads(n_threads);
#pragma omp parallel
#endif
    {
        float *f = malloc(sizeof(float) * n_pix * n_ifo * 2 + sizeof(float) * n_pix * 5);
        float *F = f + n_pix * n_ifo, *fp = F + n_pix * n_ifo, *fx = fp + n_pix,
              *si = fx + n_pix, *co = si + n_pix, *ni = co + n_pix;
#ifdef _OPENMP
#pragma omp for schedule(static)
#endif
        for (int64_t k = 0; k < n_dirs; ++k) {
            int64_t s = dirs[k];
            dpf_mixed(FP + s * n_ifo, FX + s * n_ifo, rms, n_pix, n_ifo, f, F, fp, fx, si, co,
                      ni);
            memcpy(fp_out + k * n_pix, fp, sizeof(float) * n_pix);
            memcpy(fx_out + k * n_pix, fx, sizeof(float) * n_pix);
            memcpy(ni_out + k * n_pix, ni, sizeof(float) * n_pix);
        }
        free(f);
    }
}

double oracle_calculate_dpf_f64(int n_ifo, int n_pix, int n_sky, const double *FP,
                                const double *FX, const double *rms,
                                double gamma_regulator, double En,
                                const int64_t *valid, int64_t n_valid, int n_threads)
{
    (void)n_sky;
    int64_t count = 0;
#ifdef _OPENMP
    if (n_threads > 0) omp_set_num_threads(n_threads);
#pragma omp parallel reduction(+ : count)
#endif
    {
        double *f = malloc(sizeof(double) * (n_pix * n_ifo * 2 + n_pix * 5));
        double *F = f + n_pix * n_ifo, *fp = F + n_pix * n_ifo, *fx = fp + n_pix,
               *si = fx + n_pix, *co = si + n_pix, *ni = co + n_pix;
#ifdef _OPENMP
#pragma omp for schedule(static)
#endif
        for (int64_t k = 0; k < n_valid; ++k) {
            int64_t s = valid[k];
            double q = dpf_f64(FP + s * n_ifo, FX + s * n_ifo, rms, n_pix, n_ifo,
                               f, F, fp, fx, si, co, ni);
            count += q > gamma_regulator;
        }
        free(f);
    }
    double FF = (double)n_valid, ff = (double)count;
    return (FF * FF / (ff * ff + 1.e-9) - 1) * En;
}

/* ------------------------------------------------------------------------- */
/* Per-sky temporaries of the scan (the reference allocates them inside the  */
/* prange body, sky_scan.py:100-101).                                         */
/* ------------------------------------------------------------------------- */
typedef struct {
    float *v00, *v90;             /* (n, P) */
    float *et;                    /* (P) masked total energy */
    int32_t *mask;                /* (P) */
    float *f, *F;                 /* (P, n) */
    float *fp, *fx, *si, *co, *ni;/* (P) */
    float *p, *q;                 /* (n, P) signal packet */
    float *mk;                    /* (P) updated mask (float32 array) */
    float *au, *AU, *av, *AV;     /* (P) */
    float *si2, *co2, *ee, *EE;   /* (P) */
    float *ec, *gn, *rn;          /* (P) */
} scan_tmp;

static scan_tmp tmp_alloc(int n, int P)
{
    scan_tmp t;
    size_t nf = (size_t)n * P * 6 + (size_t)P * 20;
    float *b = calloc(nf, sizeof(float));
    t.v00 = b; b += n * P;
    t.v90 = b; b += n * P;
    t.f = b; b += n * P;
    t.F = b; b += n * P;
    t.p = b; b += n * P;
    t.q = b; b += n * P;
    t.et = b; b += P;
    t.mask = (int32_t *)b; b += P;
    t.fp = b; b += P; t.fx = b; b += P; t.si = b; b += P; t.co = b; b += P; t.ni = b; b += P;
    t.mk = b; b += P;
    t.au = b; b += P; t.AU = b; b += P; t.av = b; b += P; t.AV = b; b += P;
    t.si2 = b; b += P; t.co2 = b; b += P; t.ee = b; b += P; t.EE = b; b += P;
    t.ec = b; b += P; t.gn = b; b += P; t.rn = b; b += P;
    return t;
}

static void tmp_free(scan_tmp *t) { free(t->v00); }

/* load_data_from_td (sky_stat.py:36-89): returns EE/2 (f32), NN via *M. */
static float load_data_mixed(const float *v00, const float *v90, int n, int P, double En,
                             float *et, int32_t *mask, int64_t *M)
{
    float EE = 0.f;
    int64_t NN = 0;
    for (int i = 0; i < P; ++i) {
        float aa = 0.f, AA = 0.f;
        for (int j = 0; j < n; ++j) {
            aa += v00[j * P + i] * v00[j * P + i];
            AA += v90[j * P + i] * v90[j * P + i];
        }
        et[i] = aa + AA + 1e-12f;
        mask[i] = (double)et[i] > En;            /* float32 widened vs float64 En */
        NN += mask[i];
        et[i] = (float)((double)et[i] * (double)mask[i]);  /* f32 * int32 -> f64 -> f32 */
        EE += et[i];
    }
    *M = NN;
    return MUT(7) ? EE : EE / 2.f;                   /* slip: dropped 1/2 */
}

/* avx_GW_ps (sky_stat.py:92-219).  Returns NN (number of mask pixels). */
static int64_t gw_mixed(const float *v00, const float *v90, const float *f, const float *F,
                        const float *fp, const float *fx, const float *ni, const float *et,
                        const int32_t *mask, const float *reg, int n, int P,
                        float *p, float *q, float *mk, float *au, float *AU, float *av,
                        float *AV)
{
    const float o = 1e-9f;
    const float rr = reg[0], RR = reg[1];
    int64_t NN = 0;
    for (int i = 0; i < P; ++i) {
        float xp = 0.f, XP = 0.f, xx = 0.f, XX = 0.f;
        for (int j = 0; j < n; ++j) {
            xp += v00[j * P + i] * f[i * n + j];
            XP += v90[j * P + i] * f[i * n + j];
            xx += v00[j * P + i] * F[i * n + j];
            XX += v90[j * P + i] * F[i * n + j];
        }
        float tf = sqrtf(ni[i] * (xp * xp + XP * XP) / (et[i] + o)) * rr - fp[i];
        tf = tf > 0.f ? tf : 0.f;
        /* int32 / float32 -> float64 */
        double fd = (double)mask[i] / (double)(fp[i] + tf + o);
        double h = (double)xp * fd;
        double H = (double)XP * fd;
        h = h * h + H * H;
        float Hx = xx * xx + XX * XX;
        double Fd = sqrt((double)Hx / (h + (double)o));
        float R = MUT(1) ? 0.1f - RR / (et[i] + o) : 0.1f + RR / (et[i] + o); /* slip: sign */
        Fd = Fd * (double)R - (double)fx[i];
        Fd = Fd > 0.0 ? Fd : 0.0;
        Fd = (double)mask[i] / ((double)fx[i] + Fd + (double)o);
        au[i] = (float)((double)xp * fd);
        AU[i] = (float)((double)XP * fd);
        av[i] = (float)((double)xx * Fd);
        AV[i] = (float)((double)XX * Fd);
        double a = fd * (double)fp[i] + Fd * (double)fx[i];
        NN += mask[i];
        mk[i] = (float)(a + (double)mask[i] - 1.0);
        for (int j = 0; j < n; ++j) {
            p[j * P + i] = MUT(2) ? f[i * n + j] * au[i] - F[i * n + j] * av[i]  /* slip */
                                  : f[i * n + j] * au[i] + F[i * n + j] * av[i];
            q[j * P + i] = f[i * n + j] * AU[i] + F[i * n + j] * AV[i];
        }
    }
    return NN;
}

/* avx_ort_ps (sky_stat.py:222-294).  Returns e+E in float64 (unused by scan). */
static double ort_mixed(const float *v00, const float *v90, const float *mask, int n, int P,
                        float *si, float *co, float *ee, float *EE)
{
    const float o = 1e-21f;
    double e = 0.0, E = 0.0;
    for (int i = 0; i < P; ++i) {
        float aa = 0.f, AA = 0.f, aA = 0.f;
        for (int j = 0; j < n; ++j) {
            aa += v00[j * P + i] * v00[j * P + i];
            AA += v90[j * P + i] * v90[j * P + i];
            aA += v00[j * P + i] * v90[j * P + i];
        }
        si[i] = aA * 2.f;
        co[i] = aa - AA;
        float et = aa + AA + o;
        float cc = co[i] * co[i];
        float ss = si[i] * si[i];
        float nn = sqrtf(cc + ss);
        ee[i] = (et + nn) / 2.f;
        EE[i] = (et - nn) / 2.f;
        cc = co[i] / (nn + o);
        int64_t sgn = si[i] > 0.f ? 1 : -1;
        si[i] = sqrtf((1.f - cc) / 2.f);
        co[i] = sqrtf((1.f + cc) / 2.f);
        co[i] = (float)((double)co[i] * (double)sgn);
        int mkk = mask[i] > 0.f;
        e += (double)mkk * (double)ee[i];
        E += (double)mkk * (double)EE[i];
    }
    return e + E;
}

/* avx_stat_ps (sky_stat.py:297-410). */
typedef struct { double Cr; float EC; double Mp; float No; } stat_out;

static stat_out stat_mixed(const float *v00, const float *v90, const float *s, const float *S,
                           const float *si, const float *co, const float *mask, int n, int P,
                           float *ec, float *gn, float *rn)
{
    const float o = 1e-9f;
    double LL = 0.0, Lr = 0.0, NN = 0.0;   /* loop-carried f32/f64 unification -> f64 */
    float EC = 0.f, GN = 0.f, RN = 0.f;
    for (int i = 0; i < P; ++i) {
        float c = 0.f, C = 0.f, ss = 0.f, SS = 0.f, rr = 0.f, RR = 0.f, xs = 0.f, XS = 0.f;
        for (int j = 0; j < n; ++j) {
            float s_ = s[j * P + i] * co[i] + S[j * P + i] * si[i];
            float x_ = v00[j * P + i] * co[i] + v90[j * P + i] * si[i];
            float S_ = S[j * P + i] * co[i] - s[j * P + i] * si[i];
            float X_ = v90[j * P + i] * co[i] - v00[j * P + i] * si[i];
            float a = s_ * x_;
            float A = S_ * X_;
            xs += a;
            XS += A;
            c += a * a;
            C += A * A;
            ss += s_ * s_;
            SS += S_ * S_;
            float d = s[j * P + i] - (MUT(3) ? v90 : v00)[j * P + i];   /* slip: index */
            float D = S[j * P + i] - v90[j * P + i];
            rr += d * d;
            RR += D * D;
        }
        int64_t mk = mask[i] >= 0.f ? 1 : 0;
        c = c / ((MUT(4) ? XS * XS : xs * xs) + o);   /* slip: operand */
        C = C / (XS * XS + o);
        double ll = (double)mk * (double)(ss + SS);
        double ssd = (double)ss * (1.0 - (double)c);
        double SSd = (double)SS * (1.0 - (double)C);
        ec[i] = (float)((double)mk * (ssd + SSd));
        gn[i] = (float)(((double)mk * (MUT(9) ? 1.0 : 2.0)) * (double)mask[i]);  /* slip */
        rn[i] = (float)((double)mk * (double)(rr + RR));
        double a = 2.0 * fabs((double)ec[i]);
        float A = rn[i] + gn[i] + o;
        double cc = (double)ec[i] / (a + (double)A);
        Lr += ll * cc;
        int64_t mm = ec[i] > o ? 1 : 0;
        LL += ll;
        EC += ec[i];
        GN += gn[i];
        RN += rn[i];
        NN += (double)mm;
    }
    stat_out r;
    r.Cr = (double)2.f * Lr / (LL + (double)o);
    r.EC = EC;
    r.Mp = NN;
    r.No = (GN + RN) / 2.f;
    return r;
}

/* ------------------------------------------------------------------------- */
/* scan_sky_for_best_fit (likelihoodWP/sky_scan.py:17-192), mixed precision. */
/* td00/td90: (n_delay, n, P); ml: (n, S) int32; FP/FX (S, n); rms (P, n).   */
/* maps: (11, S) float32, tuple order of sky_scan.py:191-192.                */
/* best_stat (optional, S floats): -1e12 prefilled best-stat array.          */
/* Directions with zero active pixels are gated out (the reference divides   */
/* by zero there and raises ZeroDivisionError).                              */
/* ------------------------------------------------------------------------- */
void oracle_scan_sky(int n_ifo, int n_pix, int n_sky, int n_delay, const float *FP,
                     const float *FX, const float *rms, const float *td00,
                     const float *td90, const int32_t *ml, const float *REG,
                     double netCC, double delta_regulator, double En,
                     const int64_t *valid, int64_t n_valid, float *maps,
                     float *best_stat, int64_t *l_max_out, double *stat_max_out,
                     int n_threads)
{
    const int n = n_ifo, P = n_pix, S = n_sky;
    const int offset = n_delay / 2;
    float *bs = best_stat ? best_stat : malloc(sizeof(float) * S);
    memset(maps, 0, sizeof(float) * NMAPS * (size_t)S);
    for (int s = 0; s < S; ++s) bs[s] = -1.e12f;
#ifdef _OPENMP
    if (n_threads > 0) omp_set_num_threads(n_threads);
#pragma omp parallel
#endif
    {
        scan_tmp t = tmp_alloc(n, P);
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 64)
#endif
        for (int64_t k = 0; k < n_valid; ++k) {
            int64_t l = valid[k];
            for (int i = 0; i < n; ++i) {
                int64_t row = (int64_t)ml[(int64_t)i * S + l] + offset;
                memcpy(t.v00 + i * P, td00 + (row * n + i) * P, sizeof(float) * P);
                memcpy(t.v90 + i * P, td90 + (row * n + i) * P, sizeof(float) * P);
            }
            int64_t M0;
            float Eo = load_data_mixed(t.v00, t.v90, n, P, En, t.et, t.mask, &M0);
            dpf_mixed(FP + l * n, FX + l * n, rms, P, n, t.f, t.F, t.fp, t.fx, t.si, t.co,
                      t.ni);
            int64_t M = gw_mixed(t.v00, t.v90, t.f, t.F, t.fp, t.fx, t.ni, t.et, t.mask, REG,
                                 n, P, t.p, t.q, t.mk, t.au, t.AU, t.av, t.AV);
            ort_mixed(t.p, t.q, t.mk, n, P, t.si2, t.co2, t.ee, t.EE);
            stat_out st = stat_mixed(t.v00, t.v90, t.p, t.q, t.si2, t.co2, t.mk, n, P, t.ec,
                                     t.gn, t.rn);
            if (M == 0) continue; /* defined behaviour: gate out (reference: 0/0 raises) */
            double CH = (double)st.No / ((double)(n * M) + (MUT(8) ? 0.0 : sqrt((double)M)));
            double cc = CH > 1.0 ? CH : 1.0;
            double Co = (double)st.EC / (((double)st.EC + (double)st.No * cc) -
                                         (double)(M * (n - 1)));
            if (!isfinite(st.Cr) || st.Cr < netCC) continue;
            float L = Eo > 0.f ? Eo - st.No : 0.f;
            double AA = (double)L * Co;
            (void)delta_regulator; /* probability map is not returned by the reference */
            float sde = 0.f, pwe = 0.f, cwe = 0.f;
            for (int j = 0; j < P; ++j) {
                if (t.mk[j] <= 0.f) continue;
                sde += t.et[j];
                pwe += t.fp[j] * t.et[j];
                cwe += t.fx[j] * t.et[j];
            }
            pwe = sde > 0.f ? pwe / sde : 0.f;
            cwe = sde > 0.f ? cwe / sde : 0.f;
            maps[0 * (size_t)S + l] = sqrtf(pwe + cwe);
            maps[1 * (size_t)S + l] = pwe > 0.f ? sqrtf(cwe / pwe) : 0.f;
            maps[2 * (size_t)S + l] = Eo - st.No;
            maps[3 * (size_t)S + l] = st.No;
            maps[4 * (size_t)S + l] = st.EC;
            maps[5 * (size_t)S + l] = (float)Co;
            maps[6 * (size_t)S + l] = (float)AA;
            maps[7 * (size_t)S + l] = (float)CH;
            maps[8 * (size_t)S + l] = (float)cc;
            maps[9 * (size_t)S + l] = (float)st.Cr;
            maps[10 * (size_t)S + l] = (float)st.Mp;
            bs[l] = (float)AA;
        }
        tmp_free(&t);
    }
    /* sky_scan.py:183-189: last index with best >= running max (starts at 0) */
    double stat_max = 0.0;
    int64_t l_max = n_valid > 0 ? valid[0] : 0;
    for (int64_t k = 0; k < n_valid; ++k) {
        int64_t l = valid[k];
        if ((double)bs[l] >= stat_max) {
            stat_max = (double)bs[l];
            l_max = l;
        }
    }
    *l_max_out = l_max;
    *stat_max_out = stat_max;
    if (!best_stat) free(bs);
}

/* ------------------------------------------------------------------------- */
/* f64 mode of the scan: same passes, float64 throughout.  Inputs are the    */
/* same float32 arrays widened exactly; REG is the unrounded float64 pair.   */
/* Outputs maps as float64 (11, S).                                          */
/* ------------------------------------------------------------------------- */
void oracle_scan_sky_f64(int n_ifo, int n_pix, int n_sky, int n_delay, const float *FP,
                         const float *FX, const float *rms, const float *td00,
                         const float *td90, const int32_t *ml, const double *REG,
                         double netCC, double En, const int64_t *valid, int64_t n_valid,
                         double *maps, int64_t *l_max_out, double *stat_max_out,
                         int n_threads)
{
    const int n = n_ifo, P = n_pix, S = n_sky;
    const int offset = n_delay / 2;
    double *bs = malloc(sizeof(double) * S);
    memset(maps, 0, sizeof(double) * NMAPS * (size_t)S);
    for (int s = 0; s < S; ++s) bs[s] = -1.e12;
    double *rmsd = malloc(sizeof(double) * n * P);
    for (int i = 0; i < n * P; ++i) rmsd[i] = rms[i];
#ifdef _OPENMP
    if (n_threads > 0) omp_set_num_threads(n_threads);
#pragma omp parallel
#endif
    {
        size_t nb = (size_t)n * P * 6 + (size_t)P * 20 + 2 * n;
        double *b = malloc(sizeof(double) * nb);
        double *v00 = b, *v90 = v00 + n * P, *f = v90 + n * P, *F = f + n * P,
               *p = F + n * P, *q = p + n * P, *et = q + n * P, *mask = et + P,
               *fp = mask + P, *fx = fp + P, *si = fx + P, *co = si + P, *ni = co + P,
               *mk = ni + P, *si2 = mk + P, *co2 = si2 + P, *Fp0 = co2 + P, *Fx0 = Fp0 + n;
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 64)
#endif
        for (int64_t k = 0; k < n_valid; ++k) {
            int64_t l = valid[k];
            for (int i = 0; i < n; ++i) {
                int64_t row = (int64_t)ml[(int64_t)i * S + l] + offset;
                for (int j = 0; j < P; ++j) {
                    v00[i * P + j] = td00[(row * n + i) * P + j];
                    v90[i * P + j] = td90[(row * n + i) * P + j];
                }
                Fp0[i] = FP[l * n + i];
                Fx0[i] = FX[l * n + i];
            }
            /* load_data_from_td */
            double EE = 0;
            int64_t M = 0;
            for (int j = 0; j < P; ++j) {
                double aa = 0, AA = 0;
                for (int i = 0; i < n; ++i) {
                    aa += v00[i * P + j] * v00[i * P + j];
                    AA += v90[i * P + j] * v90[i * P + j];
                }
                et[j] = aa + AA + 1e-12;
                mask[j] = et[j] > En;
                M += (int64_t)mask[j];
                et[j] *= mask[j];
                EE += et[j];
            }
            double Eo = EE / 2.0;
            dpf_f64(Fp0, Fx0, rmsd, P, n, f, F, fp, fx, si, co, ni);
            /* avx_GW_ps */
            const double o = 1e-9;
            for (int j = 0; j < P; ++j) {
                double xp = 0, XP = 0, xx = 0, XX = 0;
                for (int i = 0; i < n; ++i) {
                    xp += v00[i * P + j] * f[j * n + i];
                    XP += v90[i * P + j] * f[j * n + i];
                    xx += v00[i * P + j] * F[j * n + i];
                    XX += v90[i * P + j] * F[j * n + i];
                }
                double tf = sqrt(ni[j] * (xp * xp + XP * XP) / (et[j] + o)) * REG[0] - fp[j];
                tf = tf > 0 ? tf : 0;
                tf = mask[j] / (fp[j] + tf + o);
                double h = xp * tf, H = XP * tf;
                h = h * h + H * H;
                H = xx * xx + XX * XX;
                double Fd = sqrt(H / (h + o));
                double R = 0.1 + REG[1] / (et[j] + o);
                Fd = Fd * R - fx[j];
                Fd = Fd > 0 ? Fd : 0;
                Fd = mask[j] / (fx[j] + Fd + o);
                double au = xp * tf, AU = XP * tf, av = xx * Fd, AV = XX * Fd;
                double a = tf * fp[j] + Fd * fx[j];
                mk[j] = a + mask[j] - 1.0;
                for (int i = 0; i < n; ++i) {
                    p[i * P + j] = f[j * n + i] * au + F[j * n + i] * av;
                    q[i * P + j] = f[j * n + i] * AU + F[j * n + i] * AV;
                }
            }
            /* avx_ort_ps */
            for (int j = 0; j < P; ++j) {
                double aa = 0, AA = 0, aA = 0;
                for (int i = 0; i < n; ++i) {
                    aa += p[i * P + j] * p[i * P + j];
                    AA += q[i * P + j] * q[i * P + j];
                    aA += p[i * P + j] * q[i * P + j];
                }
                double s_ = aA * 2, c_ = aa - AA;
                double nn = sqrt(c_ * c_ + s_ * s_);
                double cc = c_ / (nn + 1e-21);
                si2[j] = sqrt((1 - cc) / 2);
                co2[j] = sqrt((1 + cc) / 2) * (s_ > 0 ? 1.0 : -1.0);
            }
            /* avx_stat_ps */
            double LL = 0, Lr = 0, EC = 0, GN = 0, RN = 0, NN = 0;
            for (int j = 0; j < P; ++j) {
                double c = 0, C = 0, ss = 0, SS = 0, rr = 0, RR = 0, xs = 0, XS = 0;
                for (int i = 0; i < n; ++i) {
                    double s_ = p[i * P + j] * co2[j] + q[i * P + j] * si2[j];
                    double x_ = v00[i * P + j] * co2[j] + v90[i * P + j] * si2[j];
                    double S_ = q[i * P + j] * co2[j] - p[i * P + j] * si2[j];
                    double X_ = v90[i * P + j] * co2[j] - v00[i * P + j] * si2[j];
                    double a = s_ * x_, A = S_ * X_;
                    xs += a; XS += A; c += a * a; C += A * A;
                    ss += s_ * s_; SS += S_ * S_;
                    double d = p[i * P + j] - v00[i * P + j], D = q[i * P + j] - v90[i * P + j];
                    rr += d * d; RR += D * D;
                }
                double mkk = mk[j] >= 0 ? 1.0 : 0.0;
                c = c / (xs * xs + o);
                C = C / (XS * XS + o);
                double ll = mkk * (ss + SS);
                double ec = mkk * (ss * (1 - c) + SS * (1 - C));
                double gn = mkk * 2.0 * mk[j];
                double rn = mkk * (rr + RR);
                double cc = ec / (2 * fabs(ec) + rn + gn + o);
                Lr += ll * cc;
                LL += ll; EC += ec; GN += gn; RN += rn; NN += ec > o;
            }
            double Cr = 2 * Lr / (LL + o), No = (GN + RN) / 2;
            if (M == 0) continue;
            double CH = No / (n * M + sqrt((double)M));
            double ccn = CH > 1 ? CH : 1;
            double Co = EC / (EC + No * ccn - M * (n - 1));
            if (!isfinite(Cr) || Cr < netCC) continue;
            double L = Eo > 0 ? Eo - No : 0;
            double AA = L * Co;
            double sde = 0, pwe = 0, cwe = 0;
            for (int j = 0; j < P; ++j) {
                if (mk[j] <= 0) continue;
                sde += et[j]; pwe += fp[j] * et[j]; cwe += fx[j] * et[j];
            }
            pwe = sde > 0 ? pwe / sde : 0;
            cwe = sde > 0 ? cwe / sde : 0;
            maps[0 * (size_t)S + l] = sqrt(pwe + cwe);
            maps[1 * (size_t)S + l] = pwe > 0 ? sqrt(cwe / pwe) : 0;
            maps[2 * (size_t)S + l] = Eo - No;
            maps[3 * (size_t)S + l] = No;
            maps[4 * (size_t)S + l] = EC;
            maps[5 * (size_t)S + l] = Co;
            maps[6 * (size_t)S + l] = AA;
            maps[7 * (size_t)S + l] = CH;
            maps[8 * (size_t)S + l] = ccn;
            maps[9 * (size_t)S + l] = Cr;
            maps[10 * (size_t)S + l] = NN;
            bs[l] = AA;
        }
        free(b);
    }
    double stat_max = 0.0;
    int64_t l_max = n_valid > 0 ? valid[0] : 0;
    for (int64_t k = 0; k < n_valid; ++k) {
        int64_t l = valid[k];
        if (bs[l] >= stat_max) { stat_max = bs[l]; l_max = l; }
    }
    *l_max_out = l_max;
    *stat_max_out = stat_max;
    free(bs);
    free(rmsd);
}

/* ========================================================================= */
/* Best-sky statistics: compute_statistics_at_sky_position                   */
/* (likelihoodWP/sky_statistics.py:26-290) with the packet operations of     */
/* likelihoodWP/packet_ops.py.  Numba kernels keep their float32/float64     */
/* typing; the NumPy helpers (avx_noise_ps, avx_setAMP_ps, xtalk sums) run   */
/* in float64 with sequential sums.                                          */
/*                                                                           */
/* xtalk CSR: lookup (P, 2) int32 [start, end), xt (m, 8) float32 with col 0 */
/* = neighbour pixel index and cols 4..7 = coefficients.                     */
/* scal[16] = Gn, Ec, Dc, Rc, Eh, Es, Np, Em, Lm, norm, cc, rho, xrho, Lo,   */
/*            Eo, N_pix_effective (all float64; the dataclass stores f32).   */
/* arrays (may be NULL): pd, pD, ps, pS (n, P) f64; nul00, nul90 (n, P) f32; */
/* pmask, gnc, ecv, ee, EE (P) f32; S_snr (n) f64.                           */
/* ========================================================================= */

/* avx_packet_ps (packet_ops.py:658-753), numba: f32 arrays, f64 islands. */
static double packet_mixed(const float *v00, const float *v90, const float *mask, int n, int P,
                           float *out00, float *out90, float *E, float *si, float *co,
                           float *a_save, float *A_save)
{
    const double o = 1.e-9;   /* python float */
    for (int j = 0; j < n; ++j) {
        float aa = 0.f, AA = 0.f, aA = 0.f;
        for (int i = 0; i < P; ++i) {
            float mk = mask[i] > 0.f ? 1.f : 0.f;
            aa += mk * (v00[j * P + i] * v00[j * P + i]);
            AA += mk * (v90[j * P + i] * v90[j * P + i]);
            aA += mk * (v00[j * P + i] * v90[j * P + i]);
        }
        float _si = 2.f * aA;
        float _co = aa - AA;
        double _x = (double)(aa + AA) + o;
        float _cc = _co * _co;
        float _ss = _si * _si;
        float _nn = sqrtf(_cc + _ss);
        float a = (float)sqrt((_x + (double)_nn) / 2.0);
        float A = (float)sqrt(fabs((_x + (MUT(10) ? 0.0 : -(double)_nn)) / 2.0)); /* slip */
        double cc = (double)_co / ((double)_nn + o);
        float sgn = _si > 0.f ? 1.f : -1.f;
        si[j] = (float)sqrt(((double)1.f - cc) / 2.0);
        co[j] = (float)(sqrt(((double)1.f + cc) / 2.0) * (double)sgn);
        float s = a + A;
        E[j] = (s * s) / 2.f;
        a_save[j] = a;
        A_save[j] = A;
        float ai = (float)(1.0 / ((double)a + o));
        float Ai = (float)(1.0 / ((double)A + o));
        for (int i = 0; i < P; ++i) {
            float mk = mask[i] > 0.f ? 1.f : 0.f;
            float _a = v00[j * P + i] * co[j] + v90[j * P + i] * si[j];
            float _A = v90[j * P + i] * co[j] - v00[j * P + i] * si[j];
            out00[j * P + i] = mk * _a * ai;
            out90[j * P + i] = mk * _A * Ai;
        }
    }
    double Ep = 0.0;
    for (int j = 0; j < n; ++j) Ep += (double)E[j];
    return Ep / 2.0;
}

/* packet_norm_numpy (packet_ops.py:813-901): returns per-ifo norm (clamped
 * >= 2) in norm[n], halo noise rn[P], q_norm (n, P); p, q given in f64. */
static void packet_norm_f64(const double *p, const double *q, const int32_t *lookup,
                            const float *xt, const float *mk, int n, int P, double *norm,
                            double *rn, double *q_norm)
{
    const double o = 1.e-12;
    double *x = malloc(sizeof(double) * 4 * n);
    for (int j = 0; j < n; ++j) norm[j] = 0.0;
    for (int i = 0; i < n * P; ++i) q_norm[i] = 0.0;
    for (int i = 0; i < P; ++i) {
        rn[i] = 0.0;
        if (mk[i] <= 0.f) continue;
        for (int k = 0; k < 4 * n; ++k) x[k] = 0.0;
        for (int e = lookup[2 * i]; e < lookup[2 * i + 1]; ++e) {
            int nb = (int)xt[8 * e];
            for (int j = 0; j < n; ++j) {
                x[0 * n + j] += p[j * P + nb] * (double)xt[8 * e + 4];
                x[1 * n + j] += p[j * P + nb] * (double)xt[8 * e + 5];
                x[2 * n + j] += q[j * P + nb] * (double)xt[8 * e + 6];
                x[3 * n + j] += q[j * P + nb] * (double)xt[8 * e + 7];
            }
        }
        double s = 0.0;
        for (int j = 0; j < n; ++j) {
            double pi = p[j * P + i], qi = q[j * P + i];
            double t = x[0 * n + j] * pi + x[1 * n + j] * qi + x[2 * n + j] * pi +
                       x[3 * n + j] * qi;
            double tc = t < 0 ? 0 : t;
            norm[j] += tc;
            double e = (pi * pi + qi * qi) / (tc + o);
            q_norm[j * P + i] = e >= 1 ? 0 : e;
            double u = x[0 * n + j] + x[2 * n + j], v = x[1 * n + j] + x[3 * n + j];
            s += u * u + v * v;
        }
        rn[i] = s;
    }
    for (int j = 0; j < n; ++j) norm[j] = norm[j] < 2.0 ? 2.0 : norm[j];
    free(x);
}

/* xtalk_energy_sum_numpy (packet_ops.py:756-810) */
static double xtalk_sum_f64(const double *p, const double *q, const int32_t *lookup,
                            const float *xt, const float *mk, int n, int P)
{
    double *g = calloc(n, sizeof(double));
    double x[4];
    for (int i = 0; i < P; ++i) {
        if (mk[i] <= 0.f) continue;
        for (int j = 0; j < n; ++j) {
            x[0] = x[1] = x[2] = x[3] = 0.0;
            for (int e = lookup[2 * i]; e < lookup[2 * i + 1]; ++e) {
                int nb = (int)xt[8 * e];
                x[0] += p[j * P + nb] * (double)xt[8 * e + 4];
                x[1] += p[j * P + nb] * (double)xt[8 * e + 5];
                x[2] += q[j * P + nb] * (double)xt[8 * e + 6];
                x[3] += q[j * P + nb] * (double)xt[8 * e + 7];
            }
            double pi = p[j * P + i], qi = q[j * P + i];
            double t = x[0] * pi + x[1] * qi + x[2] * pi + x[3] * qi;
            if (t > 0.0) g[j] += t;
        }
    }
    double s = 0.0;
    for (int j = 0; j < n; ++j) s += g[j];
    free(g);
    return s;
}

void oracle_sky_statistics(int n_ifo, int n_pix, int n_sky, int n_delay, int64_t l,
                           const float *FP, const float *FX, const float *rms,
                           const float *td00, const float *td90, const int32_t *ml,
                           const float *REG, double En, const int32_t *lookup, const float *xt,
                           int xgb_rho_mode, double *scal, double *pd, double *pD, double *ps,
                           double *pS, float *nul00, float *nul90, float *pmask, float *gnc,
                           float *ecv, float *eev, float *EEv, double *S_snr, float *pol)
{
    const int n = n_ifo, P = n_pix, S = n_sky, offset = n_delay / 2;
    scan_tmp t = tmp_alloc(n, P);
    for (int i = 0; i < n; ++i) {
        int64_t row = (int64_t)ml[(int64_t)i * S + l] + offset;
        memcpy(t.v00 + i * P, td00 + (row * n + i) * P, sizeof(float) * P);
        memcpy(t.v90 + i * P, td90 + (row * n + i) * P, sizeof(float) * P);
    }
    int64_t M0;
    load_data_mixed(t.v00, t.v90, n, P, En, t.et, t.mask, &M0);
    dpf_mixed(FP + l * n, FX + l * n, rms, P, n, t.f, t.F, t.fp, t.fx, t.si, t.co, t.ni);
    gw_mixed(t.v00, t.v90, t.f, t.F, t.fp, t.fx, t.ni, t.et, t.mask, REG, n, P, t.p, t.q, t.mk,
             t.au, t.AU, t.av, t.AV);
    ort_mixed(t.p, t.q, t.mk, n, P, t.si2, t.co2, t.ee, t.EE);
    stat_mixed(t.v00, t.v90, t.p, t.q, t.si2, t.co2, t.mk, n, P, t.ec, t.gn, t.rn);

    float *d00 = malloc(sizeof(float) * n * P * 4), *d90 = d00 + n * P, *s00 = d90 + n * P,
          *s90 = s00 + n * P;
    float dE[8], dsi[8], dco[8], da[8], dA[8], sE[8], ssi[8], sco[8], sa[8], sA[8];
    double Eo = packet_mixed(t.v00, t.v90, t.mk, n, P, d00, d90, dE, dsi, dco, da, dA);
    double Lo = packet_mixed(t.p, t.q, t.mk, n, P, s00, s90, sE, ssi, sco, sa, sA);

    double *dd00 = malloc(sizeof(double) * n * P * 8);
    double *dd90 = dd00 + n * P, *qn = dd90 + n * P, *pn = qn + n * P, *sd00 = pn + n * P,
           *sd90 = sd00 + n * P, *ap00 = sd90 + n * P, *ap90 = ap00 + n * P;
    double *rnx = malloc(sizeof(double) * P);
    double norm[8], snorm[8];
    for (int i = 0; i < n * P; ++i) { dd00[i] = d00[i]; dd90[i] = d90[i]; }
    packet_norm_f64(dd00, dd90, lookup, xt, t.mk, n, P, norm, rnx, qn);
    /* gw_norm_numpy (packet_ops.py:904-948): S_snr = 2 * E_signal / norm,
     * p_norm = q_norm where ec > 0. */
    for (int j = 0; j < n; ++j) {
        snorm[j] = (double)sE[j] * 2 / norm[j];
        if (S_snr) S_snr[j] = snorm[j];
        for (int i = 0; i < P; ++i) pn[j * P + i] = t.ec[i] > 0.f ? qn[j * P + i] : 0.0;
    }
    /* avx_noise_ps (packet_ops.py:93-153), float64 */
    double EC = 0, NC = 0, ES = 0, SC = 0, RC = 0, GN = 0, NS = 0, EH = 0;
    for (int i = 0; i < P; ++i) {
        double ns = 0, nx = 0;
        for (int j = 0; j < n; ++j) { ns += pn[j * P + i]; nx += qn[j * P + i]; }
        ns /= n;
        nx /= n;
        double mk = t.mk[i] > 0.f ? 1.0 : 0.0;
        double core = mk * (nx > 0 ? 1.0 : 0.0);
        EC += core * (double)t.ec[i];
        NC += core;
        ES += (mk - core) * rnx[i];
        double gn = t.gn[i];
        double rc = gn < 2.0 ? 1.0 : 0.0;
        double nn = gn * (1.0 - rc);
        rc = rc + nn * (MUT(11) ? 1.0 : 0.5);                        /* slip: factor */
        rc = (double)t.ec[i] / (rc + 1e-9);
        double sig = mk * (ns > 0 ? 1.0 : 0.0);
        t.gn[i] = (float)(mk * gn * nx);
        SC += sig * (double)t.ec[i];
        RC += sig * rc;
        GN += sig * (double)t.gn[i];
        NS += sig;
        EH += (mk - sig) * (double)t.et[i];
    }
    double Gn = GN, Ec = EC, Dc = SC - EC, Rc = RC / (SC + 0.01), Eh = EH / 2, Es = ES / 2;
    /* avx_setAMP_ps (packet_ops.py:234-285) for data and signal packets */
    double Np_eff = 0;
    for (int pass = 0; pass < 2; ++pass) {
        const float *p = pass ? s00 : d00, *q = pass ? s90 : d90;
        const double *nrm = pass ? pn : qn;
        const float *a = pass ? sa : da, *A = pass ? sA : dA, *si = pass ? ssi : dsi,
                    *co = pass ? sco : dco;
        double *o00 = pass ? ap00 : sd00, *o90 = pass ? ap90 : sd90;
        double Npp = 0;
        for (int i = 0; i < P; ++i) {
            float mk = 0.5f * (t.mk[i] > 0.f ? 1.f : 0.f);
            double col = 0;
            for (int j = 0; j < n; ++j) {
                double na = (double)((a[j] + A[j]) * mk) * nrm[j * P + i];
                float rp = p[j * P + i] * co[j] - q[j * P + i] * si[j];
                float rq = q[j * P + i] * co[j] + p[j * P + i] * si[j];
                o00[j * P + i] = na * (double)rp;
                o90[j * P + i] = na * (double)rq;
                col += nrm[j * P + i];
            }
            Npp += col * (double)mk;
        }
        if (!pass) Np_eff = Npp * 4 / n - (MUT(12) ? 0 : 1);     /* slip: dropped -1 */
    }
    /* avx_loadNULL_ps: f32(data) - f32(signal) */
    float *n00 = malloc(sizeof(float) * n * P * 2), *n90 = n00 + n * P;
    double *dn00 = malloc(sizeof(double) * n * P * 2), *dn90 = dn00 + n * P;
    for (int i = 0; i < n * P; ++i) {
        n00[i] = (float)sd00[i] - (float)ap00[i];
        n90[i] = (float)sd90[i] - (float)ap90[i];
        dn00[i] = n00[i];
        dn90[i] = n90[i];
    }
    double Em = xtalk_sum_f64(sd00, sd90, lookup, xt, t.mk, n, P);
    double Np = xtalk_sum_f64(dn00, dn90, lookup, xt, t.mk, n, P);
    double Lm = MUT(13) ? Em - Np + Gn : Em - Np - Gn;        /* slip: sign */
    double nrm = Em > 0 ? (Eo - Eh) / Em : 1.e9;
    if (nrm < 1) nrm = 1;
    Ec /= nrm;
    Dc /= nrm;
    double chi = (Np + Gn) / (Np_eff * n);
    double ccf = chi > 1 ? chi : 1, rho, xrho = 0;
    if (!xgb_rho_mode) {
        rho = Ec > 0 ? sqrt(Ec * Rc / 2.) : 0;
    } else {
        double pen = chi;
        rho = Ec > 0 ? sqrt(Ec / (1 + pen * ((pen > 1 ? pen : 1) - 1))) : 0;
        xrho = Ec > 0 ? sqrt(Ec * Rc / 2.) : 0;
    }
    /* avx_pol_ps (packet_ops.py:564-633), first call: p00_POL / p90_POL as
     * (radius, angle) per pixel, float64 NumPy arithmetic with float32 results;
     * the row sums over detectors start from 0.0 in detector order (NumPy's
     * reduction below eight terms; the 8-term pairing for n = 8). */
    if (pol) {
        for (int i = 0; i < P; ++i) {
            double m = t.mk[i] > 0.f ? 1.0 : 0.0, tt[4][8];
            for (int j = 0; j < n; ++j) {
                double pj = (double)t.v00[j * P + i] * m, qj = (double)t.v90[j * P + i] * m;
                tt[0][j] = (double)t.f[i * n + j] * pj;
                tt[1][j] = (double)t.f[i * n + j] * qj;
                tt[2][j] = (double)t.F[i * n + j] * pj;
                tt[3][j] = (double)t.F[i * n + j] * qj;
            }
            double sum[4];
            for (int k = 0; k < 4; ++k) {
                if (n < 8) {
                    sum[k] = 0.0;
                    for (int j = 0; j < n; ++j) sum[k] += tt[k][j];
                } else {
                    sum[k] = ((tt[k][0] + tt[k][1]) + (tt[k][2] + tt[k][3])) +
                             ((tt[k][4] + tt[k][5]) + (tt[k][6] + tt[k][7]));
                }
            }
            double xp = sum[0], XP = sum[1], xx = sum[2], XX = sum[3];
            double fp = t.fp[i], fx = t.fx[i];
            double sfp = sqrt(fp) + 1e-9, sfx = sqrt(fx) + 1e-9;
            double rp = xp * xp / (fp + 1e-9) + xx * xx / (fx + 1e-9);
            double RP = XP * XP / (fp + 1e-9) + XX * XX / (fx + 1e-9);
            pol[i] = (float)sqrt(rp);
            pol[P + i] = (float)atan2(xx / sfx, xp / sfp);
            pol[2 * P + i] = (float)sqrt(RP);
            pol[3 * P + i] = (float)atan2(XX / sfx, XP / sfp);
        }
    }
    double v[16] = {Gn, Ec, Dc, Rc, Eh, Es, Np, Em, Lm, nrm, ccf, rho, xrho, Lo, Eo, Np_eff};
    memcpy(scal, v, sizeof(v));
    if (pd) memcpy(pd, sd00, sizeof(double) * n * P);
    if (pD) memcpy(pD, sd90, sizeof(double) * n * P);
    if (ps) memcpy(ps, ap00, sizeof(double) * n * P);
    if (pS) memcpy(pS, ap90, sizeof(double) * n * P);
    if (nul00) memcpy(nul00, n00, sizeof(float) * n * P);
    if (nul90) memcpy(nul90, n90, sizeof(float) * n * P);
    if (pmask) memcpy(pmask, t.mk, sizeof(float) * P);
    if (gnc) memcpy(gnc, t.gn, sizeof(float) * P);
    if (ecv) memcpy(ecv, t.ec, sizeof(float) * P);
    if (eev) memcpy(eev, t.ee, sizeof(float) * P);
    if (EEv) memcpy(EEv, t.EE, sizeof(float) * P);
    free(n00); free(dn00); free(dd00); free(rnx); free(d00);
    tmp_free(&t);
}

/* Exported single-purpose wrappers for the reference's packet-op contract
 * tests (packet_ops.py:756-901). */
void oracle_packet_norm(const double *p, const double *q, const int32_t *lookup, const float *xt,
                        const float *mk, int n, int P, const float *q_E, double *det_snr,
                        double *norm, double *rn, double *q_norm)
{
    packet_norm_f64(p, q, lookup, xt, mk, n, P, norm, rn, q_norm);
    for (int j = 0; j < n; ++j) det_snr[j] = (double)q_E[j] * 2.0 / norm[j];
}

double oracle_xtalk_sum(const double *p, const double *q, const int32_t *lookup, const float *xt,
                        const float *mk, int n, int P)
{
    return xtalk_sum_f64(p, q, lookup, xt, mk, n, P);
}
