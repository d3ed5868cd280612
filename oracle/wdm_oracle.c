/*
 * ORACLE -- TEST INFRASTRUCTURE ONLY.
 *
 * CPU restatement of the cWB WDM forward transform used by PycWB
 * (third-party wdm-wavelet>=0.3.1, un-vendored; the source-level spec is
 * cwb-core/WDM.cc).  Everything here is float64.
 *
 *  - wdm_oracle_fourier : Fourier coefficients of the Meyer-type band edge
 *    cos(pi/2 * I_x(beta, beta)) (regularised incomplete beta), the quantity
 *    that cwb-core/FourierCoefficients.icc tabulates as Cos[beta][j]:
 *        a_0 = int_0^1 g(x) dx,  a_j = sqrt(2) int_0^1 g(x) cos(j pi x) dx.
 *    Computed here by Gauss-Legendre quadrature (independent of the table,
 *    which tests use only as known answers).
 *  - wdm_oracle_filter  : WDM<T>::getFilter + the length rule of the
 *    constructor (cwb-core/WDM.cc:108-173, 498-548).
 *  - wdm_oracle_t2w     : WDM<T>::t2w (cwb-core/WDM.cc:1385-1534) with mirror
 *    padding (also coherence_native/tf_batch_generation.py:37-61), the
 *    symmetric window fold, a direct DFT of length 2M (textbook routine in
 *    place of FFTW) and the 00/90 phase assignment.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

/* Number of Fourier terms per beta order, as sized in FourierCoefficients.icc. */
static const int kCosSize[8] = {0, 0, 401, 501, 151, 126, 81, 61};

int wdm_oracle_cos_size(int beta) { return (beta >= 2 && beta <= 7) ? kCosSize[beta] : 0; }

static double binom(int n, int k)
{
    double r = 1;
    for (int i = 1; i <= k; ++i) r = r * (n - k + i) / i;
    return r;
}

/* regularised incomplete beta I_x(b, b) for integer b (a polynomial) */
static double ibeta_sym(int b, double x)
{
    int m = 2 * b - 1;
    double s = 0;
    for (int k = b; k <= m; ++k) s += binom(m, k) * pow(x, k) * pow(1 - x, m - k);
    return s;
}

/* Gauss-Legendre nodes/weights on [0,1] (Newton on P_N). */
static void gauss_legendre01(int N, double *x, double *w)
{
    for (int i = 0; i < (N + 1) / 2; ++i) {
        double z = cos(M_PI * (i + 0.75) / (N + 0.5)), pp = 0;
        for (int it = 0; it < 100; ++it) {
            double p1 = 1, p2 = 0;
            for (int j = 1; j <= N; ++j) {
                double p3 = p2;
                p2 = p1;
                p1 = ((2.0 * j - 1) * z * p2 - (j - 1.0) * p3) / j;
            }
            pp = N * (z * p1 - p2) / (z * z - 1);
            double dz = p1 / pp;
            z -= dz;
            if (fabs(dz) < 1e-16) break;
        }
        x[i] = 0.5 * (1 - z);
        x[N - 1 - i] = 0.5 * (1 + z);
        w[i] = w[N - 1 - i] = 1.0 / ((1 - z * z) * pp * pp);
    }
}

int wdm_oracle_fourier(int beta, double *out)
{
    int nf = wdm_oracle_cos_size(beta);
    if (!nf) return 0;
    const int NG = 2048;
    double *x = malloc(sizeof(double) * NG), *w = malloc(sizeof(double) * NG);
    double *g = malloc(sizeof(double) * NG);
    gauss_legendre01(NG, x, w);
    for (int i = 0; i < NG; ++i) g[i] = cos(M_PI / 2 * ibeta_sym(beta, x[i]));
    for (int j = 0; j < nf; ++j) {
        double s = 0;
        for (int i = 0; i < NG; ++i) s += w[i] * g[i] * (j ? cos(j * M_PI * x[i]) : 1.0);
        out[j] = j ? sqrt(2.0) * s : s;
    }
    free(x); free(w); free(g);
    return nf;
}

/* WDM<T>::getFilter(n) followed by the constructor's length rule.
 * Writes up to max_taps filter coefficients into `filter`; returns m_H
 * (filter length) or -1 when max_taps is too small. */
int wdm_oracle_filter(int M, int K, int beta, int precision, double *filter, int max_taps)
{
    if (beta < 2) beta = 2;
    if (beta > 7) beta = 7;
    int nf = kCosSize[beta];
    double *fourier = malloc(sizeof(double) * nf);
    wdm_oracle_fourier(beta, fourier);
    const double Pi = M_PI;
    double B = Pi / K;
    double A = (K - M) * Pi / 2. / K / M;
    double K2 = (double)K * K;
    double gNorm = sqrt(2.0 * M) / Pi;
    double residual, prec = pow(10., -(double)precision);
    int M2 = 2 * M, N = 1;
    filter[0] = (A + fourier[0] * B) * gNorm;
    fourier[0] /= sqrt(2);
    residual = 1 - filter[0] * filter[0];
    for (;;) {
        int i = N;
        if (i >= max_taps) { free(fourier); return -1; }
        double di = i, i2 = di * di, sumEven = 0, sumOdd = 0;
        for (int j = 0; j < nf; ++j) {
            if (i % K == 0 && j == i / K) continue;
            if (j & 1) sumOdd += di / (i2 - j * j * K2) * fourier[j];
            else sumEven += di / (i2 - j * j * K2) * fourier[j];
        }
        double intAB = 0;
        if (i % K == 0 && i / K < nf) intAB = fourier[i / K] * B / 2 * cos(i * A);
        intAB += 2 * (sumEven * sin(i * B / 2) * cos(i * Pi / (2 * M)) -
                      sumOdd * sin(i * Pi / (2 * M)) * cos(i * B / 2));
        filter[i] = gNorm * (sin(i * A) / i + sqrt(2) * intAB);
        /* constructor loop: residual -= 2*f[N]^2; N++; while(...) */
        residual -= 2 * filter[N] * filter[N];
        N++;
        if (!(residual > prec || (N - 1) % M2 || N / M2 < 3)) break;
    }
    free(fourier);
    return N;
}

/* Forward transform of one series x[0..L).  stride MM (<=0 means M);
 * MM < 0 (the reference's KK < 0) writes the quadrature maps map00/map90,
 * otherwise the power map into map00; layout (N, M+1); returns N. */
int64_t wdm_oracle_t2w(const double *x, int64_t L, int M, const double *filt, int m_H,
                       int MM, double *map00, double *map90)
{
    int stride = MM <= 0 ? M : MM;
    int64_t nW = L + (L % stride ? stride - L % stride : 0);
    int64_t m = nW + 2 * (int64_t)m_H;
    double *ts = calloc(m, sizeof(double));
    for (int64_t n = 0; n <= m_H && n < L; ++n) ts[m_H - n] = x[n];
    for (int64_t n = 0; n < L; ++n) ts[m_H + n] = x[n];
    for (int64_t n = 0; n < m - m_H - L; ++n) ts[n + m_H + L] = L - n - 1 >= 0 ? x[L - n - 1] : 0.0;
    int64_t N = nW / stride;
    int M2 = 2 * M;
    double *re = malloc(sizeof(double) * M2);
    double *cs = malloc(sizeof(double) * M2), *sn = malloc(sizeof(double) * M2);
    for (int k = 0; k < M2; ++k) {
        cs[k] = cos(2 * M_PI * k / M2);
        sn[k] = sin(2 * M_PI * k / M2);
    }
    const double sqrt2 = sqrt(2.0);
    for (int64_t n = 0; n < N; ++n) {
        const double *pTS = ts + m_H + n * stride;
        for (int k = 0; k < M2; ++k) re[k] = 0;
        /* symmetric fold: t in [-(m_H-1), m_H-1], weight filt[|t|] */
        for (int t = -(m_H - 1); t <= m_H - 1; ++t) {
            int k = ((t % M2) + M2) % M2;
            re[k] += pTS[t] * filt[t < 0 ? -t : t];
        }
        int odd = (int)(n & 1);
        for (int f = 0; f <= M; ++f) {
            double R = 0, I = 0;
            for (int k = 0; k < M2; ++k) {
                int a = (int)(((int64_t)f * k) % M2);
                R += re[k] * cs[a];
                I -= re[k] * sn[a];   /* forward DFT, exp(-i 2 pi f k / 2M) */
            }
            if (f == 0 || f == M) { R = R / sqrt2; I = R; }
            double v00, v90;
            if ((f + odd) & 1) { v00 = sqrt2 * I; v90 = sqrt2 * R; }
            else { v00 = sqrt2 * R; v90 = -sqrt2 * I; }
            if (MM >= 0) { v00 = R * R + I * I; v90 = 0; }
            map00[n * (M + 1) + f] = v00;
            if (map90) map90[n * (M + 1) + f] = v90;
        }
    }
    free(ts); free(re); free(cs); free(sn);
    return N;
}

/* WDM<T>::w2t / w2tQ (cwb-core/WDM.cc:1536-1710): inverse transform of an
 * (N, M+1) map (map: 00 coefficients for quad = 0, 90 coefficients for
 * quad = 1), n_out output samples.  Per time bin the half spectrum is built
 * with the reference's phase rules, the complex-to-real inverse transform of
 * length 2M is a direct sum (textbook routine in place of FFTW's c2r,
 * unnormalised: r[k] = X0 + (-1)^k XM + 2 sum_j Re(X_j e^{+i pi j k / M})),
 * and the filter-weighted overlap-add runs over the bins in order. */
int wdm_oracle_w2t(const double *map, int64_t N, int M, const double *filt, int m_H, int quad,
                   int64_t n_out, double *out)
{
    int M2 = 2 * M;
    int64_t nTS = (int64_t)M * N + 2 * (int64_t)m_H;
    double *ts = calloc((size_t)nTS, sizeof(double));
    double *tmp = ts + m_H;
    double *reX = malloc(sizeof(double) * M2), *imX = malloc(sizeof(double) * M2);
    double *r = malloc(sizeof(double) * M2);
    double *cs = malloc(sizeof(double) * M2), *sn = malloc(sizeof(double) * M2);
    for (int k = 0; k < M2; ++k) {
        cs[k] = cos(M_PI * k / M);
        sn[k] = sin(M_PI * k / M);
    }
    const double sqrt2 = sqrt(2.0);
    for (int64_t n = 0; n < N; ++n) {
        const double *col = map + n * (M + 1);
        for (int j = 0; j < M2; ++j) reX[j] = imX[j] = 0;
        if (!quad) {
            for (int j = 1; j < M; ++j) {
                if ((n + j) & 1) imX[j] = col[j] / sqrt2;
                else if (j & 1) reX[j] = -col[j] / sqrt2;
                else reX[j] = col[j] / sqrt2;
            }
            if ((n & 1) == 0) reX[0] = col[0];
            if (((n + M) & 1) == 0) reX[M] = (M & 1) ? -col[M] : col[M];
        } else {
            for (int j = 1; j < M; ++j) {
                if ((n + j) & 1) reX[j] = col[j] / sqrt2;
                else if (j & 1) imX[j] = col[j] / sqrt2;
                else imX[j] = -col[j] / sqrt2;
            }
            if ((n + M) & 1) reX[M] = col[M];
            if (n & 1) reX[0] = col[0];
        }
        for (int k = 0; k < M2; ++k) {
            double s = reX[0] + ((k & 1) ? -reX[M] : reX[M]);
            for (int j = 1; j < M; ++j) {
                int q = (int)(((int64_t)j * k) % M2);
                s += 2.0 * (reX[j] * cs[q] - imX[j] * sn[q]);
            }
            r[k] = s;
        }
        double *x = tmp + n * M;
        int m = M * (int)(n & 1), mm = m;
        x[0] += filt[0] * r[m];
        for (int j = 1; j < m_H; ++j) {
            ++m;
            if (m == M2) m = 0;
            x[j] += filt[j] * r[m];
            if (mm == 0) mm = M2 - 1;
            else --mm;
            x[-j] += filt[j] * r[mm];
        }
    }
    for (int64_t n = 0; n < n_out; ++n) out[n] = n < nTS - m_H ? tmp[n] : 0.0;
    free(ts); free(reX); free(imX); free(r); free(cs); free(sn);
    return 0;
}
