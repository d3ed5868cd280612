/*
 * ORACLE -- TEST INFRASTRUCTURE ONLY.
 *
 * CPU restatement of PycWB's time-delay max-energy map, pattern path
 * (pycwb/modules/coherence_native/time_delay_numba.py:183-288 with the packet
 * energy _wdm_packet_energy_nb, :49-104, and its parameters
 * _compute_packet_energy_params, time_delay_packet.py:8-64): the quadrature
 * t2w (wdm_oracle_t2w) of the series and of its copies shifted left and right
 * by k = d, 2d, ... <= max_delay samples (the other end keeps the original
 * samples), the packet energy of every map, the pixel-wise maximum, and the
 * lowest layer (two for patterns 5, 6, 9) zeroed.  Only tests/ load this
 * library, and only as the checker.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

int64_t wdm_oracle_t2w(const double *x, int64_t L, int M, const double *filt, int m_H, int MM,
                       double *map00, double *map90);

/* time_delay_packet.py:8-64 */
void maxenergy_oracle_params(int M1, int64_t T, int pattern, double edge, int wavelet_rate,
                             double f_low, double f_high, double df, int64_t *jb, int64_t *je,
                             int *mL, int *mH, double *mean, int64_t *p)
{
    int64_t J = (int64_t)M1 * T;
    int64_t b = (int64_t)(edge * (double)wavelet_rate / 4.0) * M1;
    if (b < 4 * (int64_t)M1) b = 4 * (int64_t)M1;
    int64_t e = J - b;
    if (b < 0) b = 0;
    if (e > J) e = J;
    int lo = (int)(f_low / df + 0.1), hi = (int)(f_high / df + 0.1);
    if (lo < 0) lo = 0;
    if (hi > M1 - 1) hi = M1 - 1;
    int pt = pattern < 0 ? -pattern : pattern;
    double mn = 1.0;
    if (pt == 1 || pt == 3 || pt == 4) { mn = 3.0; lo += 1; hi -= 1; }
    else if (pt == 2) { mn = 3.0; }
    else if (pt == 5 || pt == 6) { mn = 5.0; lo += 2; hi -= 2; }
    else if (pt == 7 || pt == 8) { mn = 5.0; lo += 1; hi -= 1; }
    else if (pt == 9) { mn = 9.0; lo += 1; hi -= 1; }
    for (int k = 0; k < 9; ++k) p[k] = 0;
    int64_t M = M1;
    if (pt == 1) { p[1] = 1; p[2] = -1; }
    else if (pt == 2) { p[1] = M; p[2] = -M; }
    else if (pt == 3) { p[1] = M + 1; p[2] = -M - 1; }
    else if (pt == 4) { p[1] = -M + 1; p[2] = M - 1; }
    else if (pt == 5) { p[1] = M + 1; p[2] = -M - 1; p[3] = 2 * M + 2; p[4] = -2 * M - 2; }
    else if (pt == 6) { p[1] = -M + 1; p[2] = M - 1; p[3] = -2 * M + 2; p[4] = 2 * M - 2; }
    else if (pt == 7) { p[1] = 1; p[2] = -1; p[3] = M; p[4] = -M; }
    else if (pt == 8) { p[1] = M + 1; p[2] = -M + 1; p[3] = M - 1; p[4] = -M - 1; }
    else if (pt == 9) {
        int64_t q[8] = {1, -1, M, -M, M + 1, M - 1, -M + 1, -M - 1};
        for (int k = 0; k < 8; ++k) p[k + 1] = q[k];
    }
    *jb = b; *je = e; *mL = lo; *mH = hi; *mean = mn;
}

/* _wdm_packet_energy_nb (time_delay_numba.py:49-104); maps time-major
 * (T, M1); energy time-major. */
static void packet_energy(const double *re, const double *im, int M1, int64_t T, int64_t jb,
                          int64_t je, int mL, int mH, double mean, const int64_t *p,
                          double *energy)
{
    int64_t J = (int64_t)M1 * T;
    for (int64_t j = 0; j < J; ++j) {
        int m = (int)(j % M1);
        energy[j] = 0.0;
        if (j < jb || j >= je) {
            if (re[j] > 0.0) energy[j] = re[j];
            continue;
        }
        if (m < mL || m > mH) continue;
        double ss = 0.0, ee = 0.0, EE = 0.0;
        for (int n = 1; n < 9; ++n) {
            int64_t idx = j + p[n];
            if (idx < 0) idx = 0;
            else if (idx >= J) idx = J - 1;
            ss += re[idx] * im[idx];
            ee += re[idx] * re[idx];
            EE += im[idx] * im[idx];
        }
        ss += re[j] * im[j] * (mean - 8.0);
        ee += re[j] * re[j] * (mean - 8.0);
        EE += im[j] * im[j] * (mean - 8.0);
        double cc = ee - EE, s2 = 2.0 * ss;
        double nn2 = cc * cc + s2 * s2;
        double nn = nn2 > 0.0 ? sqrt(nn2) : 0.0;
        double se = ee + EE;
        if (se < nn) nn = se;
        double h1 = (se + nn) / 2.0, h2 = (se - nn) / 2.0;
        double a = (h1 > 0.0 ? sqrt(h1) : 0.0) + (h2 > 0.0 ? sqrt(h2) : 0.0);
        energy[j] = mean == 1.0 ? se / 2.0 : a * a / 4.0;
    }
}

/* out: (T, M1) time-major maximum energy; returns T (or -1). */
int64_t maxenergy_oracle(const double *ts, int64_t L, int M, const double *filt, int m_H,
                         int64_t max_delay, int downsample, int pattern, double edge,
                         int wavelet_rate, double f_low, double f_high, double df, double *out)
{
    int M1 = M + 1;
    int64_t T = (L + (L % M ? M - L % M : 0)) / M;
    int64_t J = (int64_t)M1 * T;
    int64_t jb, je, p[9];
    int mL, mH;
    double mean;
    maxenergy_oracle_params(M1, T, pattern, edge, wavelet_rate, f_low, f_high, df, &jb, &je,
                            &mL, &mH, &mean, p);
    double *re = malloc(sizeof(double) * J), *im = malloc(sizeof(double) * J);
    double *en = malloc(sizeof(double) * J), *xx = malloc(sizeof(double) * L);
    wdm_oracle_t2w(ts, L, M, filt, m_H, -1, re, im);
    packet_energy(re, im, M1, T, jb, je, mL, mH, mean, p, out);
    for (int64_t k = downsample; k <= max_delay && k < L; k += downsample) {
        for (int side = 0; side < 2; ++side) {
            memcpy(xx, ts, sizeof(double) * L);
            if (side == 0) memcpy(xx, ts + k, sizeof(double) * (L - k));   /* left */
            else memcpy(xx + k, ts, sizeof(double) * (L - k));             /* right */
            wdm_oracle_t2w(xx, L, M, filt, m_H, -1, re, im);
            packet_energy(re, im, M1, T, jb, je, mL, mH, mean, p, en);
            for (int64_t j = 0; j < J; ++j)
                if (en[j] > out[j]) out[j] = en[j];
        }
    }
    int pt = pattern < 0 ? -pattern : pattern;
    for (int64_t t = 0; t < T; ++t) {
        out[t * M1] = 0.0;
        if ((pt == 5 || pt == 6 || pt == 9) && M1 > 2) out[t * M1 + 1] = 0.0;
    }
    free(re); free(im); free(en); free(xx);
    return T;
}
