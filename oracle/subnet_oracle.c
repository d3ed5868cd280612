/*
 * ORACLE -- TEST INFRASTRUCTURE ONLY.
 *
 * CPU restatement of PycWB's sub-network cut (pycwb/modules/super_cluster_native/
 * sub_net_cut.py): the sky loop optimze_sky_loc_from_td (:404-517) and its older
 * twin optimze_sky_loc (:291-401, which takes the per-detector energies as a
 * separate table), the reduced-pixel DPF _dpf_np_loops_vec_into (:520-565),
 * sse_like_ps (:728-742), the greedy multi-resolution pass sse_MRA_ps (:746-825)
 * and mra_statistics_from_td (:659-724).  Only tests/ and bench.py's cpu_baseline
 * leg load this library, and only as the checker.
 *
 * Arithmetic follows Numba's typing of the reference op by op: float32 arrays,
 * float32 scalars and sequential float32 sums; an int64 or float64 operand
 * promotes the operation to float64 (the `2 * m * (Eo - Ln) / Eo` term, the
 * subcut test, AA, the suball and submra ratios, rHo); no FMA contraction
 * (build with -ffp-contract=off); correctly rounded division and sqrt.
 * sse_like_ps's np.dot of two float32 vectors is restated as a forward float32
 * sum of float32 products.  (The reference's np.dot is the BLAS library's sdot,
 * whose accumulation order is the library's; parity with the reference there
 * is to float32 rounding, not bits.)
 *
 * Layouts (the reference's): td00/td90 (nd, n, P) float32 with the delay
 * offset nd / 2; ml (n, S) int32; FP/FX (S, n) float32; rms (P, n) float32;
 * xtalk (m, 8) float32 rows (column 0 the neighbour's pixel index, columns 4..7
 * the coefficients) and lookup (P, 2) int32 [start, end) rows.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define TD(a, d, i, j) (a)[((int64_t)(d) * n + (i)) * P + (j)]

/* _dpf_np_loops_vec_into (sub_net_cut.py:520-565): DPF antenna vectors f, F
 * (m, n) of m reduced pixels for one direction (Fp0, Fx0 (n)). */
static void dpf_into(const float *Fp0, const float *Fx0, const float *rms, int m, int n,
                     float *f, float *F, float *si, float *co, float *fp)
{
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < m; ++i) {
            f[i * n + j] = rms[i * n + j] * Fp0[j];
            F[i * n + j] = rms[i * n + j] * Fx0[j];
        }
    for (int i = 0; i < m; ++i) {
        float ff = 0.0f, FF = 0.0f, fF = 0.0f;
        for (int j = 0; j < n; ++j) {
            ff += f[i * n + j] * f[i * n + j];
            FF += F[i * n + j] * F[i * n + j];
            fF += F[i * n + j] * f[i * n + j];
        }
        float s2 = 2.0f * fF;                      /* mul_vec */
        float c2 = ff - FF;                        /* sub_vec */
        float AP = ff + FF;                        /* add_vec */
        float nn = sqrtf(c2 * c2 + s2 * s2);       /* norm_vec */
        float cc = c2 / (nn + 1e-9f);              /* div_vec */
        fp[i] = (AP + nn) / 2.0f;                  /* avg_vec */
        si[i] = sqrtf((1.0f - cc) / 2.0f);         /* sin_from_cc */
        co[i] = s2 > 0.0f ? sqrtf((1.0f + cc) / 2.0f) : -sqrtf((1.0f + cc) / 2.0f);
    }
    for (int i = 0; i < m; ++i) {
        for (int j = 0; j < n; ++j) {
            float a = f[i * n + j], b = F[i * n + j];
            f[i * n + j] = a * co[i] + b * si[i];
            F[i * n + j] = b * co[i] - a * si[i];
        }
        float g = 0.0f;
        for (int j = 0; j < n; ++j) g += f[i * n + j] * F[i * n + j];
        g = g / (fp[i] + 1e-9f);
        for (int j = 0; j < n; ++j) F[i * n + j] -= f[i * n + j] * g;
    }
}

static float dot32(const float *a, const float *b, int n)
{
    float s = 0.0f;
    for (int i = 0; i < n; ++i) s += a[i] * b[i];
    return s;
}

/* sse_like_ps (sub_net_cut.py:728-742). */
static float sse_like(const float *fp, const float *fx, const float *am, const float *AM, int n)
{
    float xp = dot32(fp, am, n), XP = dot32(fp, AM, n);
    float xx = dot32(fx, am, n), XX = dot32(fx, AM, n);
    float gp = dot32(fp, fp, n) + 1e-12f;
    float gx = dot32(fx, fx, n) + 1e-12f;
    xp = xp * xp + XP * XP;
    xx = xx * xx + XX * XX;
    return xp / gp + xx / gx;
}

typedef struct {
    float *rms, *v00, *v90, *f, *F, *si, *co, *fp;
} reduced_t;

static reduced_t reduced_alloc(int P, int n)
{
    reduced_t r;
    size_t a = (size_t)P * n;
    r.rms = malloc(sizeof(float) * (6 * a + 3 * (size_t)P + 8));
    r.v00 = r.rms + a;
    r.v90 = r.v00 + a;
    r.f = r.v90 + a;
    r.F = r.f + a;
    r.si = r.F + a + a;
    r.co = r.si + P;
    r.fp = r.co + P;
    return r;
}

/* Lo over the reduced pixels of one direction (sub_net_cut.py:493-500). */
static float reduced_likelihood(const float *Fp0, const float *Fx0, reduced_t *r, int m, int n)
{
    dpf_into(Fp0, Fx0, r->rms, m, n, r->f, r->F, r->si, r->co, r->fp);
    float Lo = 0.0f;
    for (int j = 0; j < m; ++j)
        Lo += sse_like(r->f + j * n, r->F + j * n, r->v00 + j * n, r->v90 + j * n, n);
    return Lo;
}

/* Shared tail of both sky loops: the per-direction statistics and the running
 * maximum `AA > AA_max` (sub_net_cut.py:502-515).  out: l_max, stat, Em, Am, lm,
 * Vm, suball, EE. */
static void sky_tail(float aa, float Eo, float Ls, float Ln, float Lo, int64_t m, int l,
                     double *AA_max, double *out)
{
    double AA = (double)aa / ((double)(fabsf(aa) + fabsf(Eo - Lo))
                              + (double)(2 * m) * (double)(Eo - Ln) / (double)Eo);
    float ee = Ls * Eo / (Eo - Ls);
    double em = (double)fabsf(Eo - Lo) + (double)(2 * m);
    double ee2 = (double)ee / ((double)ee + em);
    double aa2 = ((double)aa - (double)m) / ((double)aa + (double)m);
    if (AA > *AA_max) {
        *AA_max = AA;
        out[0] = l;
        out[1] = AA;
        out[2] = Eo;
        out[3] = aa2;
        out[4] = l;
        out[5] = (double)m;
        out[6] = ee2;
        out[7] = em;
    }
}

static int subcut_rejects(float aa, int64_t m_cut, double subcut)
{
    return subcut >= 0.0 &&
           ((double)aa - (double)m_cut) / ((double)aa + (double)m_cut + (double)1e-16f) < subcut;
}

/* optimze_sky_loc_from_td (sub_net_cut.py:404-517). */
void subnet_oracle_sky_loc_from_td(int n, int P, int S, const float *FP, const float *FX,
                                   const float *rms, const float *td00, const float *td90,
                                   int nd, const int32_t *ml, float En, double e2or,
                                   double subcut, double *out)
{
    const float Es = (float)(2.0 * e2or);
    const int off = nd / 2;
    float *rNRG = malloc(sizeof(float) * (P + 1));
    reduced_t r = reduced_alloc(P, n);
    double AA_max = 0.0;
    for (int k = 0; k < 8; ++k) out[k] = 0.0;
    for (int l = 0; l < S; ++l) {
        float m_gt = 0.0f, Eo_gt = 0.0f, Ls_gt = 0.0f, Ln_gt = 0.0f;
        float Eo_ge = 0.0f, Ls_ge = 0.0f, Ln_ge = 0.0f;
        for (int j = 0; j < P; ++j) {
            float rE = 0.0f, em = 0.0f;
            for (int i = 0; i < n; ++i) {
                int d = ml[(int64_t)i * S + l] + off;
                float a = TD(td00, d, i, j), b = TD(td90, d, i, j);
                float de = a * a + b * b;
                rE += de;
                if (de > em) em = de;
            }
            rNRG[j] = rE;
            float sub = rE - em;
            if (rE >= En) {
                Eo_ge += rE;
                Ls_ge += sub;
                if (sub > Es) Ln_ge += rE;
            }
            if (rE > En) {
                m_gt += 1.0f;
                Eo_gt += rE;
                Ls_gt += sub;
                if (sub > Es) Ln_gt += rE;
            }
        }
        float Eo_cut = Eo_gt + 0.01f;
        int64_t m_cut = (int64_t)(2.0 * (double)m_gt + 0.01);
        float aa = Ls_gt * Ln_gt / (Eo_cut - Ls_gt);
        if (subcut_rejects(aa, m_cut, subcut)) continue;
        int64_t m = 0;
        for (int j = 0; j < P; ++j) {
            if (rNRG[j] < En) continue;
            for (int i = 0; i < n; ++i) {
                int d = ml[(int64_t)i * S + l] + off;
                r.rms[m * n + i] = rms[(int64_t)j * n + i];
                r.v00[m * n + i] = TD(td00, d, i, j);
                r.v90[m * n + i] = TD(td90, d, i, j);
            }
            m += 1;
        }
        float Eo = Eo_ge, Ls = Ls_ge, Ln = Ln_ge;
        if (Eo <= 0.0f) continue;
        float Lo = reduced_likelihood(FP + (int64_t)l * n, FX + (int64_t)l * n, &r, (int)m, n);
        sky_tail(aa, Eo, Ls, Ln, Lo, m, l, &AA_max, out);
    }
    free(rNRG);
    free(r.rms);
}

/* optimze_sky_loc (sub_net_cut.py:291-401): the same statistic from a separate
 * per-detector energy table te (nd, n, P), with the sub-network energy taken as
 * min_i (E - e_i) and the reduced pixels selected by recomputing E (< En). */
void subnet_oracle_sky_loc(int n, int P, int S, const float *FP, const float *FX,
                           const float *rms, const float *td00, const float *td90,
                           const float *te, int nd, const int32_t *ml, float En, double e2or,
                           double subcut, double *out)
{
    const float Es = (float)(2.0 * e2or);
    const int off = nd / 2;
    float *rNRG = malloc(sizeof(float) * (P + 1));
    float *pNRG = malloc(sizeof(float) * (P + 1));
    reduced_t r = reduced_alloc(P, n);
    double AA_max = 0.0;
    for (int k = 0; k < 8; ++k) out[k] = 0.0;
    for (int l = 0; l < S; ++l) {
        float m = 0.0f, Eo = 0.0f, Ls = 0.0f, Ln = 0.0f;
        for (int j = 0; j < P; ++j) {
            float rE = 0.0f;
            for (int i = 0; i < n; ++i) rE += TD(te, ml[(int64_t)i * S + l] + off, i, j);
            rNRG[j] = rE;
            float msk = rNRG[j] > En ? 1.0f : 0.0f;
            m += msk;
            pNRG[j] = rNRG[j] * msk;
            Eo += pNRG[j];
            for (int i = 0; i < n; ++i) {
                float t = rNRG[j] - TD(te, ml[(int64_t)i * S + l] + off, i, j);
                pNRG[j] = t < pNRG[j] ? t : pNRG[j];
            }
            Ls += pNRG[j];
            msk = pNRG[j] > Es ? 1.0f : 0.0f;
            Ln += rNRG[j] * msk;
        }
        Eo = Eo + 0.01f;
        int64_t m_cut = (int64_t)(2.0 * (double)m + 0.01);
        float aa = Ls * Ln / (Eo - Ls);
        if (subcut_rejects(aa, m_cut, subcut)) continue;
        int64_t mm = 0;
        float Ls2 = 0.0f, Ln2 = 0.0f, Eo2 = 0.0f;
        for (int j = 0; j < P; ++j) {
            float ee = 0.0f;
            for (int i = 0; i < n; ++i) {
                int d = ml[(int64_t)i * S + l] + off;
                float a = TD(td00, d, i, j), b = TD(td90, d, i, j);
                ee += a * a + b * b;
            }
            if (ee < En) continue;
            float em = 0.0f;
            for (int i = 0; i < n; ++i) {
                int d = ml[(int64_t)i * S + l] + off;
                float a = TD(td00, d, i, j), b = TD(td90, d, i, j);
                r.rms[mm * n + i] = rms[(int64_t)j * n + i];
                r.v00[mm * n + i] = a;
                r.v90[mm * n + i] = b;
                float e = a * a + b * b;
                if (e > em) em = e;
            }
            mm += 1;
            Ls2 += ee - em;
            Eo2 += ee;
            if (ee - em > Es) Ln2 += ee;
        }
        if (Eo2 <= 0.0f) continue;
        float Lo = reduced_likelihood(FP + (int64_t)l * n, FX + (int64_t)l * n, &r, (int)mm, n);
        sky_tail(aa, Eo2, Ls2, Ln2, Lo, mm, l, &AA_max, out);
    }
    free(rNRG);
    free(pNRG);
    free(r.rms);
}

/* sse_MRA_ps (sub_net_cut.py:746-825): greedy extraction of K principal
 * pixels.  a00/a90 (n, P) are the working amplitudes (modified), rNRG (P) the
 * working energies (modified); amp/AMP (n, P) receive the components.
 * Returns the number of extracted components. */
static int sse_mra(float Eo, int64_t K, float *rNRG, float *a00, float *a90, const float *xt,
                   const int32_t *lookup, int n, int P, float *amp, float *AMP)
{
    float EE = 0.0f;
    float mam[16], mAM[16];
    memset(amp, 0, sizeof(float) * (size_t)n * P);
    memset(AMP, 0, sizeof(float) * (size_t)n * P);
    int64_t k = 0;
    while (k < K) {
        int m = 0; /* np.argmax: first index of the maximum */
        for (int j = 1; j < P; ++j)
            if (rNRG[j] > rNRG[m]) m = j;
        if (rNRG[m] <= Eo) break;
        float E = 0.0f;
        for (int i = 0; i < n; ++i)
            E += a00[i * P + m] * a00[i * P + m] + a90[i * P + m] * a90[i * P + m];
        EE += E;
        if ((double)(E / EE) < 0.01) break;
        for (int i = 0; i < n; ++i) {
            mam[i] = a00[i * P + m];
            mAM[i] = a90[i * P + m];
        }
        for (int i = 0; i < n; ++i) {
            amp[i * P + m] += mam[i];
            AMP[i * P + m] += mAM[i];
        }
        for (int e = lookup[2 * m]; e < lookup[2 * m + 1]; ++e) {
            const float *c = xt + (int64_t)e * 8;
            int q = (int)c[0];
            if (rNRG[q] > Eo) {
                rNRG[q] = 0.0f;
                for (int i = 0; i < n; ++i) {
                    a00[i * P + q] -= mam[i] * c[4] + mAM[i] * c[5];
                    a90[i * P + q] -= mam[i] * c[6] + mAM[i] * c[7];
                    rNRG[q] += a00[i * P + q] * a00[i * P + q] + a90[i * P + q] * a90[i * P + q];
                }
            }
        }
        k += 1;
    }
    return (int)k;
}

/* mra_statistics_from_td (sub_net_cut.py:659-724) at direction l.
 * out: submra, rHo, Eo, Lo, Ls, m.  amp_out/AMP_out (n, P), if not NULL,
 * receive the extracted components; the return value is their count. */
int subnet_oracle_mra_from_td(int n, int P, int S, const float *FP, const float *FX,
                              const float *rms, const float *td00, const float *td90, int nd,
                              const int32_t *ml, float En, double e2or, const float *xt,
                              const int32_t *lookup, int l, double *out, float *amp_out,
                              float *AMP_out)
{
    const float Es = (float)(2.0 * e2or);
    const int off = nd / 2;
    size_t a = (size_t)n * P;
    float *rNRG = malloc(sizeof(float) * (P + 1));
    float *w = malloc(sizeof(float) * (4 * a + 4));
    float *v00 = w, *v90 = w + a, *amp = w + 2 * a, *AMP = w + 3 * a;
    float cnt = 0.0f;
    for (int j = 0; j < P; ++j) {
        float rE = 0.0f;
        for (int i = 0; i < n; ++i) {
            int d = ml[(int64_t)i * S + l] + off;
            float x = TD(td00, d, i, j), y = TD(td90, d, i, j);
            rE += x * x + y * y;
            v00[i * P + j] = x;
            v90[i * P + j] = y;
        }
        rNRG[j] = rE;
        cnt += rNRG[j] > En ? 1.0f : 0.0f;
    }
    int ncomp = sse_mra(En, (int64_t)cnt, rNRG, v00, v90, xt, lookup, n, P, amp, AMP);
    reduced_t r = reduced_alloc(P, n);
    int64_t m = 0;
    float Ls = 0.0f, Ln = 0.0f, Eo = 0.0f;
    for (int j = 0; j < P; ++j) {
        float ee = 0.0f;
        for (int i = 0; i < n; ++i)
            ee += amp[i * P + j] * amp[i * P + j] + AMP[i * P + j] * AMP[i * P + j];
        if (ee < En) continue;
        float em = 0.0f;
        for (int i = 0; i < n; ++i) {
            r.rms[m * n + i] = rms[(int64_t)j * n + i];
            r.v00[m * n + i] = amp[i * P + j];
            r.v90[m * n + i] = AMP[i * P + j];
            float e = amp[i * P + j] * amp[i * P + j] + AMP[i * P + j] * AMP[i * P + j];
            if (e > em) em = e;
        }
        m += 1;
        Ls += ee - em;
        Eo += ee;
        if (ee - em > Es) Ln += ee;
    }
    float Lo = reduced_likelihood(FP + (int64_t)l * n, FX + (int64_t)l * n, &r, (int)m, n);
    float sm = Ls * Eo / (Eo - Ls + 1e-16f);
    double submra = (double)sm / ((double)(fabsf(sm) + fabsf(Eo - Lo)) + (double)(2 * (m + 6)));
    double rHo = sqrt((double)(Lo * Lo) / ((double)(Eo) + (double)(2 * m) + (double)1e-16f) / 2.0);
    out[0] = submra;
    out[1] = rHo;
    out[2] = Eo;
    out[3] = Lo;
    out[4] = Ls;
    out[5] = (double)m;
    if (amp_out) memcpy(amp_out, amp, sizeof(float) * a);
    if (AMP_out) memcpy(AMP_out, AMP, sizeof(float) * a);
    free(rNRG);
    free(w);
    free(r.rms);
    return ncomp;
}

/* The reduced-pixel likelihood alone (for tests of the DPF and sse_like_ps):
 * Lo of m pixels with rms (m, n), v00/v90 (m, n) at antenna vectors Fp0/Fx0. */
float subnet_oracle_reduced_likelihood(const float *Fp0, const float *Fx0, const float *rms,
                                       const float *v00, const float *v90, int m, int n)
{
    reduced_t r = reduced_alloc(m > 0 ? m : 1, n);
    memcpy(r.rms, rms, sizeof(float) * (size_t)m * n);
    memcpy(r.v00, v00, sizeof(float) * (size_t)m * n);
    memcpy(r.v90, v90, sizeof(float) * (size_t)m * n);
    float Lo = reduced_likelihood(Fp0, Fx0, &r, m, n);
    free(r.rms);
    return Lo;
}
