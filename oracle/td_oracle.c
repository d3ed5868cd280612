/*
 * ORACLE -- TEST INFRASTRUCTURE ONLY.
 *
 * CPU restatement, in float64, of the time-delay (TD) amplitude extraction
 * that produces the likelihood's td00/td90 tables:
 *
 *  - td_oracle_fourier : cosine-series coefficients of cos^2(pi/2 nu) and of
 *    sin(pi nu), nu = I_x(beta, beta) -- the Cos2 and SinCos tables of
 *    cwb-core/FourierCoefficients.icc (tests use the table as known answers);
 *  - td_oracle_filters : WDM<T>::getTDFilter1, getTDFilter2 and setTDFilter
 *    (cwb-core/WDM.cc:551-700) as dense T0/Tx tables (2J+1, 2n+1);
 *  - td_oracle_vectors : batch_get_td_vecs / _get_pixel_amplitude_nb
 *    (pycwb/utils/td_vector_kernels.py:16-199): per pixel and delay the
 *    same-band and the two cross-band filter sums over the zero-padded TF
 *    planes, with the whole-pixel shift and quadrature swap of getTDamp.
 *    Products float32 x float64 and sums in float64, output float32.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>

static const int kCos2Size[8] = {0, 0, 101, 61, 46, 36, 36, 31};
static const int kSinCosSize[8] = {0, 0, 601, 801, 151, 151, 81, 81};

static double tbinom(int n, int k)
{
    double r = 1;
    for (int i = 1; i <= k; ++i) r = r * (n - k + i) / i;
    return r;
}

static double tibeta(int b, double x)
{
    int m = 2 * b - 1;
    double s = 0;
    for (int k = b; k <= m; ++k) s += tbinom(m, k) * pow(x, k) * pow(1 - x, m - k);
    return s;
}

/* Gauss-Legendre on [0, 1] by Newton iteration on P_N. */
static void tgl01(int N, double *x, double *w)
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

int td_oracle_size(int beta, int which)
{
    if (beta < 2 || beta > 7) return 0;
    return which == 0 ? kCos2Size[beta] : kSinCosSize[beta];
}

/* which = 0: cos^2(pi/2 nu); which = 1: sin(pi nu). */
int td_oracle_fourier(int beta, int which, double *out)
{
    int nf = td_oracle_size(beta, which);
    if (!nf) return 0;
    const int NG = 2048;
    double *x = malloc(sizeof(double) * NG), *w = malloc(sizeof(double) * NG);
    double *f = malloc(sizeof(double) * NG);
    tgl01(NG, x, w);
    for (int i = 0; i < NG; ++i) {
        double nu = tibeta(beta, x[i]);
        double c = cos(M_PI / 2 * nu);
        f[i] = which == 0 ? c * c : sin(M_PI * nu);
    }
    for (int j = 0; j < nf; ++j) {
        /* symmetric about x = 1/2: these coefficients vanish exactly */
        if ((which == 0 && j > 0 && !(j & 1)) || (which == 1 && (j & 1))) {
            out[j] = 0.0;
            continue;
        }
        double s = 0;
        for (int i = 0; i < NG; ++i) s += w[i] * f[i] * (j ? cos(j * M_PI * x[i]) : 1.0);
        out[j] = j ? sqrt(2.0) * s : s;
    }
    free(x); free(w); free(f);
    return nf;
}

/* getTDFilter1 (WDM.cc:551-594) */
static void td_filter1(int M, int K, int beta, int n, int L, double *out)
{
    int nf = td_oracle_size(beta, 0);
    double *four = malloc(sizeof(double) * nf);
    td_oracle_fourier(beta, 0, four);
    double B = M_PI / K, A = (K - M) * M_PI / 2. / K / M, K2 = (double)K * K;
    double gNorm = 2.0 * M / M_PI;
    out[0] = (A + four[0] * B) * gNorm;
    four[0] /= sqrt(2.0);
    for (int i = 1; i < n * L; ++i) {
        double di = i * (1. / L), i2 = di * di, se = 0, so = 0;
        for (int j = 0; j < nf; ++j) {
            if (j * K * L == i) continue;
            if (j & 1) so += di / (i2 - j * j * K2) * four[j];
            else se += di / (i2 - j * j * K2) * four[j];
        }
        double ab = 0;
        if (i % (K * L) == 0 && i / (K * L) < nf) ab = four[i / (K * L)] * B / 2 * cos(di * A);
        ab += 2 * (se * sin(di * B / 2) * cos(di * M_PI / (2 * M)) -
                   so * sin(di * M_PI / (2 * M)) * cos(di * B / 2));
        out[i] = gNorm * (sin(di * A) / di + sqrt(2.0) * ab);
    }
    free(four);
}

/* getTDFilter2 (WDM.cc:596-636) */
static void td_filter2(int M, int K, int beta, int n, int L, double *out)
{
    int nf = td_oracle_size(beta, 1);
    double *four = malloc(sizeof(double) * nf);
    td_oracle_fourier(beta, 1, four);
    double B = M_PI / K, K2 = (double)K * K, gNorm = M / M_PI;
    out[0] = four[0] * B * gNorm;
    four[0] /= sqrt(2.0);
    for (int i = 1; i < n * L; ++i) {
        double di = i * (1. / L), i2 = di * di, sum = 0;
        for (int j = 0; j < nf; j += 2) {  /* nf is odd: j <= nf stays inside */
            if (j * K * L == i) continue;
            sum += di / (i2 - j * j * K2) * four[j];
        }
        sum *= 2 * sin(di * B / 2);
        if (i % (2 * K * L) == 0 && i / (K * L) < nf) {
            if ((i / (2 * K * L)) & 1) sum -= four[i / K / L] * B / 2;
            else sum += four[i / K / L] * B / 2;
        }
        out[i] = gNorm * sqrt(2.0) * sum;
    }
    free(four);
}

/* setTDFilter (WDM.cc:638-700): T0, Tx (2J+1, 2n+1), J = M L. */
void td_oracle_filters(int M, int K, int beta, int nc, int L, double *T0, double *Tx)
{
    int n = M * (nc + 1) + 1, J = M * L, W = 2 * nc + 1;
    double *f1 = malloc(sizeof(double) * n * L), *f2 = malloc(sizeof(double) * n * L);
    td_filter1(M, K, beta, n, L, f1);
    td_filter2(M, K, beta, n, L, f2);
    for (int i = -J; i <= J; ++i) {
        double *t0 = T0 + (int64_t)(i + J) * W + nc, *tx = Tx + (int64_t)(i + J) * W + nc;
        t0[0] = f1[abs(i)];
        tx[0] = f2[abs(i)] / 2.;
        for (int j = 1; j <= nc; ++j) {
            t0[j] = f1[i + J * j];
            t0[-j] = f1[J * j - i];
            tx[j] = f2[i + J * j] / 2.;
            tx[-j] = f2[J * j - i] / 2.;
        }
        for (int j = 1; j <= nc; ++j) switch (j % 4) {
            case 1: tx[-j] *= -1; break;
            case 2: tx[j] *= -1; tx[-j] *= -1; break;
            case 3: tx[j] *= -1; break;
        }
    }
    free(f1); free(f2);
}

/* one band's even / odd filter sums */
/* rows outside the planes (whole-bin shifts at the map's ends) read as zero;
 * the reference kernel does not guard them */
static int64_t g_rows;

static void band_sums(const float *plane, int M1, int n, int m, const double *T, int W,
                      double *se, double *so)
{
    double e = 0, o = 0;
    for (int k = 0; k < W; ++k) {
        int64_t row = (int64_t)n + k;
        double pv = (row >= 0 && row < g_rows) ? (double)plane[row * M1 + m] : 0.0;
        double v = pv * T[k];
        if (k % 2 == 0) e += v;
        else o += v;
    }
    *se = e;
    *so = o;
}

/* _get_pixel_amplitude_nb (td_vector_kernels.py:16-129) */
static double pixel_amp(int n, int m, int dT, const float *plane, const double *T0,
                        const double *Tx, int M, int nc, int J, int quad)
{
    const int W = 2 * nc + 1, M1 = M + 1;
    int odd_n = n & 1, odd_mn = (m + n) & 1, cancel_even = odd_n ^ (quad ? 1 : 0);
    int edge = (m == 0) || (m == M);
    const double *t0 = T0 + (int64_t)(dT + J) * W, *tx = Tx + (int64_t)(dT + J) * W;
    double dt = (double)dT * (double)M / (double)J;
    double se, so;
    band_sums(plane, M1, n, m, t0, W, &se, &so);
    if (edge && cancel_even) se = 0;
    if (edge && !cancel_even) so = 0;
    double ph = m * M_PI * dt / M, r;
    if (odd_n) r = se * cos(ph) - so * sin(ph);
    else r = se * cos(ph) + so * sin(ph);
    if (m > 0) {
        double le, lo;
        band_sums(plane, M1, n, m - 1, tx, W, &le, &lo);
        if (m == 1 && cancel_even) le = 0;
        if (m == 1 && !cancel_even) lo = 0;
        double lp = (2.0 * m - 1.0) * dt * M_PI / (2.0 * M);
        double t = odd_n ? (le - lo) * sin(lp) : (le + lo) * sin(lp);
        if (m == 1 || m == M) t *= sqrt(2.0);
        r = odd_mn ? r - t : r + t;
    }
    if (m < M) {
        double he, ho;
        band_sums(plane, M1, n, m + 1, tx, W, &he, &ho);
        if (m == M - 1 && cancel_even) he = 0;
        if (m == M - 1 && !cancel_even) ho = 0;
        double hp = (2.0 * m + 1.0) * dt * M_PI / (2.0 * M);
        double t = odd_n ? (he + ho) * sin(hp) : (he - ho) * sin(hp);
        if (m == 0 || m == M - 1) t *= sqrt(2.0);
        r = odd_mn ? r - t : r + t;
    }
    if (edge && cancel_even) return 0.0;
    return r;
}

/* batch_get_td_vecs (td_vector_kernels.py:131-199): out (P, 4K+2) float32 rows
 * [a00(-K..K), a90(-K..K)]; planes (n_time + 2 nc, M + 1) float32. */
void td_oracle_vectors(const int32_t *pix, int64_t P, const float *p00, const float *p90,
                       int64_t n_rows, const double *T0, const double *Tx, int M, int nc, int K,
                       int J, float *out)
{
    const int M1 = M + 1, half = 2 * K + 1;
    g_rows = n_rows;
    for (int64_t p = 0; p < P; ++p) {
        int n = pix[p] / M1, m = pix[p] % M1;
        for (int ki = 0; ki < half; ++ki) {
            int dT = ki - K;
            int sh = dT >= 0 ? dT / J : -((-dT) / J);
            int sub = dT - sh * J, ne = n - sh;
            double a00, a90;
            if (sh % 2 != 0) {
                if ((n + m) % 2 != 0) {
                    a00 = -pixel_amp(ne, m, sub, p90, T0, Tx, M, nc, J, 1);
                    a90 = pixel_amp(ne, m, sub, p00, T0, Tx, M, nc, J, 0);
                } else {
                    a00 = pixel_amp(ne, m, sub, p90, T0, Tx, M, nc, J, 1);
                    a90 = -pixel_amp(ne, m, sub, p00, T0, Tx, M, nc, J, 0);
                }
            } else {
                a00 = pixel_amp(ne, m, sub, p00, T0, Tx, M, nc, J, 0);
                a90 = pixel_amp(ne, m, sub, p90, T0, Tx, M, nc, J, 1);
            }
            out[p * (4 * K + 2) + ki] = (float)a00;
            out[p * (4 * K + 2) + half + ki] = (float)a90;
        }
    }
}
