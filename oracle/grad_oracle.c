/*
 * ORACLE -- TEST INFRASTRUCTURE ONLY.
 *
 * CPU float64 restatement of the continuous-direction sky statistic whose
 * gradient the GPU computes, with derivatives by the complex-step method
 * (Squire & Trapp, SIAM Review 40 (1998) 110): for a real-analytic f,
 * f'(x) = Im f(x + i h) / h with no subtractive cancellation, so h = 1e-30
 * gives the derivative to rounding.  Branch tests (masks, max(., 0), the DPF
 * sign choice, the interpolation cell) read real parts only, which freezes
 * them at the evaluation point exactly as the GPU's tangent propagation does,
 * and sqrt uses its first-order complex-step form (derivative 0 at 0).
 *
 * Statistic (per cluster, direction theta = colatitude, phi = longitude):
 *   n      = (sin th cos ph, sin th sin ph, cos th)       (detector.py:1294-1297)
 *   tau_i  = g_i . n, g_i = (R_i - R_ref) rate / c        (detector.py:1299-1303,
 *            the continuous value before rint)
 *   v_i    = td at tap tau_i + K, linearly interpolated, clamped to the table
 *   F+, Fx = response-tensor contraction at gha = -ph, dec = pi/2 - th, psi = 0
 *            (detector.py:272-378)
 *   AA     = L * Co of the float64 scan chain (sky_stat.py:74-405, dpf.py:396-450,
 *            sky_scan.py:131-141) with REG fixed and no netCC gate.
 * The pixel chain below is written pass by pass like oracle_scan_sky_f64, so
 * that on integer taps with table patterns it reproduces that function's AA.
 */
#include <complex.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>

typedef double complex cd;

static cd cs_sqrt(cd z)
{
    double r = sqrt(creal(z));
    return r > 0 ? r + I * (cimag(z) / (2.0 * r)) : 0.0;
}

static cd cs_pos(cd z) { return creal(z) > 0 ? z : 0.0; }

/* Direction terms of one detector: continuous delay (taps) and patterns. */
static void cs_terms(const double *g, cd th, cd ph, cd *tau, cd *fp, cd *fx)
{
    cd st = csin(th), ct = ccos(th), sp = csin(ph), cp = ccos(ph);
    cd nv[3] = {st * cp, st * sp, ct};
    *tau = nv[0] * g[0] + nv[1] * g[1] + nv[2] * g[2];
    /* gha = -ph, dec = pi/2 - th: x = (-sin gha, -cos gha, 0),
     * y = (-cos gha sin dec, sin gha sin dec, cos dec) */
    cd x[3] = {sp, -cp, 0.0};
    cd y[3] = {-cp * ct, -sp * ct, st};
    const double *D = g + 3;
    cd dx[3], dy[3];
    for (int a = 0; a < 3; ++a) {
        dx[a] = D[3 * a] * x[0] + D[3 * a + 1] * x[1] + D[3 * a + 2] * x[2];
        dy[a] = D[3 * a] * y[0] + D[3 * a + 1] * y[1] + D[3 * a + 2] * y[2];
    }
    *fp = x[0] * dx[0] + x[1] * dx[1] + x[2] * dx[2] - (y[0] * dy[0] + y[1] * dy[1] + y[2] * dy[2]);
    *fx = x[0] * dy[0] + x[1] * dy[1] + x[2] * dy[2] + y[0] * dx[0] + y[1] * dx[1] + y[2] * dx[2];
}

/*
 * AA from delays and patterns.  rms (P, n) float32; td (P, n, 2K+1) float
 * pairs {00, 90}.  Amplitude entry td_index (flat float index, -1 for none)
 * is offset by td_delta.  Pixel arrays are laid out (P) like the reference's
 * passes.
 */
static cd cs_objective(int n, int P, int K, const float *rms, const float *td, const cd *tau,
                       const cd *Fp0, const cd *Fx0, double reg0, double reg1, double En,
                       int64_t td_index, cd td_delta)
{
    const int nd = 2 * K + 1;
    const double o = 1e-9;
    cd *buf = malloc(sizeof(cd) * (size_t)P * (6 * n + 5));
    cd *v00 = buf, *v90 = v00 + (size_t)n * P, *p = v90 + (size_t)n * P, *q = p + (size_t)n * P;
    cd *f = q + (size_t)n * P, *F = f + (size_t)n * P;
    cd *et = F + (size_t)n * P, *fp = et + P, *fx = fp + P, *ni = fx + P, *mk = ni + P;
    int *mask = malloc(sizeof(int) * (size_t)P);

    /* delayed amplitudes by linear interpolation */
    for (int i = 0; i < n; ++i) {
        cd u = tau[i] + K;
        double uv = creal(u);
        double lo = 0.0, hi = nd - 1 - 1e-9;
        if (!(uv >= lo) || uv > hi) u = uv > hi ? hi : lo;
        int i0 = (int)floor(creal(u));
        cd w = u - i0;
        for (int j = 0; j < P; ++j) {
            int64_t e0 = 2 * (((int64_t)j * n + i) * nd + i0);
            cd a00 = td[e0], a90 = td[e0 + 1], b00 = td[e0 + 2], b90 = td[e0 + 3];
            if (td_index == e0) a00 += td_delta;
            if (td_index == e0 + 1) a90 += td_delta;
            if (td_index == e0 + 2) b00 += td_delta;
            if (td_index == e0 + 3) b90 += td_delta;
            v00[i * P + j] = (1.0 - w) * a00 + w * b00;
            v90[i * P + j] = (1.0 - w) * a90 + w * b90;
        }
    }
    /* load_data_from_td */
    cd EE = 0;
    int64_t M = 0;
    for (int j = 0; j < P; ++j) {
        cd aa = 0, AA = 0;
        for (int i = 0; i < n; ++i) {
            aa += v00[i * P + j] * v00[i * P + j];
            AA += v90[i * P + j] * v90[i * P + j];
        }
        et[j] = aa + AA + 1e-12;
        mask[j] = creal(et[j]) > En;
        M += mask[j];
        if (!mask[j]) et[j] = 0;
        EE += et[j];
    }
    /* dpf_np_loops_vec */
    for (int j = 0; j < P; ++j) {
        cd ff = 0, FF = 0, fF = 0;
        for (int i = 0; i < n; ++i) {
            f[j * n + i] = (double)rms[j * n + i] * Fp0[i];
            F[j * n + i] = (double)rms[j * n + i] * Fx0[i];
            ff += f[j * n + i] * f[j * n + i];
            FF += F[j * n + i] * F[j * n + i];
            fF += F[j * n + i] * f[j * n + i];
        }
        cd s = 2.0 * fF, c = ff - FF, AP = ff + FF;
        cd nn = cs_sqrt(c * c + s * s);
        cd cc = c / (nn + o);
        fp[j] = (AP + nn) / 2.0;
        cd si = cs_sqrt((1.0 - cc) / 2.0);
        cd co = cs_sqrt((1.0 + cc) / 2.0);
        if (!(creal(s) > 0)) co = -co;
        cd t = 0;
        for (int i = 0; i < n; ++i) {
            cd a = f[j * n + i], b = F[j * n + i];
            f[j * n + i] = a * co + b * si;
            F[j * n + i] = b * co - a * si;
            t += f[j * n + i] * F[j * n + i];
        }
        t = t / (fp[j] + o);
        cd xx = 0, nq = 0;
        for (int i = 0; i < n; ++i) {
            F[j * n + i] -= f[j * n + i] * t;
            xx += F[j * n + i] * F[j * n + i];
            cd f2 = f[j * n + i] * f[j * n + i];
            nq += f2 * f2;
        }
        fx[j] = xx;
        ni[j] = nq / (fp[j] * fp[j] + o);
    }
    /* avx_GW_ps (masked pixels: packet 0, mk = -1) */
    for (int j = 0; j < P; ++j) {
        if (!mask[j]) {
            mk[j] = -1.0;
            for (int i = 0; i < n; ++i) p[i * P + j] = q[i * P + j] = 0;
            continue;
        }
        cd xp = 0, XP = 0, xx = 0, XX = 0;
        for (int i = 0; i < n; ++i) {
            xp += v00[i * P + j] * f[j * n + i];
            XP += v90[i * P + j] * f[j * n + i];
            xx += v00[i * P + j] * F[j * n + i];
            XX += v90[i * P + j] * F[j * n + i];
        }
        cd tf = cs_pos(cs_sqrt(ni[j] * (xp * xp + XP * XP) / (et[j] + o)) * reg0 - fp[j]);
        cd fd = 1.0 / (fp[j] + tf + o);
        cd h = xp * fd, H = XP * fd;
        h = h * h + H * H;
        cd Hx = xx * xx + XX * XX;
        cd R = 0.1 + reg1 / (et[j] + o);
        cd Fd = cs_pos(cs_sqrt(Hx / (h + o)) * R - fx[j]);
        Fd = 1.0 / (fx[j] + Fd + o);
        cd au = xp * fd, AU = XP * fd, av = xx * Fd, AV = XX * Fd;
        mk[j] = fd * fp[j] + Fd * fx[j];
        for (int i = 0; i < n; ++i) {
            p[i * P + j] = f[j * n + i] * au + F[j * n + i] * av;
            q[i * P + j] = f[j * n + i] * AU + F[j * n + i] * AV;
        }
    }
    /* avx_ort_ps + avx_stat_ps */
    cd EC = 0, GN = 0, RN = 0;
    for (int j = 0; j < P; ++j) {
        if (!(creal(mk[j]) >= 0)) continue;
        cd aa = 0, AA = 0, aA = 0;
        for (int i = 0; i < n; ++i) {
            aa += p[i * P + j] * p[i * P + j];
            AA += q[i * P + j] * q[i * P + j];
            aA += p[i * P + j] * q[i * P + j];
        }
        cd s_ = aA * 2.0, c_ = aa - AA;
        cd nn = cs_sqrt(c_ * c_ + s_ * s_);
        cd cc = c_ / (nn + 1e-21);
        cd si = cs_sqrt((1.0 - cc) / 2.0);
        cd co = cs_sqrt((1.0 + cc) / 2.0);
        if (!(creal(s_) > 0)) co = -co;
        cd c = 0, C = 0, ss = 0, SS = 0, rr = 0, RR = 0, xs = 0, XS = 0;
        for (int i = 0; i < n; ++i) {
            cd s = p[i * P + j] * co + q[i * P + j] * si;
            cd x = v00[i * P + j] * co + v90[i * P + j] * si;
            cd S = q[i * P + j] * co - p[i * P + j] * si;
            cd X = v90[i * P + j] * co - v00[i * P + j] * si;
            cd a = s * x, A = S * X;
            xs += a; XS += A; c += a * a; C += A * A;
            ss += s * s; SS += S * S;
            cd d = p[i * P + j] - v00[i * P + j], D = q[i * P + j] - v90[i * P + j];
            rr += d * d; RR += D * D;
        }
        c = c / (xs * xs + o);
        C = C / (XS * XS + o);
        EC += ss * (1.0 - c) + SS * (1.0 - C);
        GN += 2.0 * mk[j];
        RN += rr + RR;
    }
    free(buf);
    free(mask);
    if (M == 0) return 0.0;
    cd Eo = EE / 2.0, No = (GN + RN) / 2.0;
    cd CH = No / ((double)n * M + sqrt((double)M));
    cd cc = creal(CH) > 1 ? CH : 1.0;
    cd Co = EC / (EC + No * cc - (double)M * (n - 1));
    cd L = creal(Eo) > 0 ? Eo - No : 0.0;
    return L * Co;
}

/* Continuous delays and patterns of n detectors at (theta, phi), float64. */
void oracle_direction_terms(int n, const double *geo, double theta, double phi, double *tau,
                            double *fp, double *fx)
{
    for (int i = 0; i < n; ++i) {
        cd t, a, b;
        cs_terms(geo + 12 * i, theta, phi, &t, &a, &b);
        tau[i] = creal(t);
        fp[i] = creal(a);
        fx[i] = creal(b);
    }
}

/*
 * Statistic at theta + i*theta_im, phi + i*phi_im with amplitude entry
 * td_index offset by td_re + i*td_im.  out[0] = real part, out[1] = imaginary
 * part (the complex step: derivative = out[1] / step).
 */
void oracle_sky_objective(int n, int P, int K, const float *rms, const float *td,
                          const double *geo, double reg0, double reg1, double En, double theta,
                          double theta_im, double phi, double phi_im, int64_t td_index,
                          double td_re, double td_im, double *out)
{
    cd tau[8], fp[8], fx[8];
    for (int i = 0; i < n; ++i)
        cs_terms(geo + 12 * i, theta + I * theta_im, phi + I * phi_im, &tau[i], &fp[i], &fx[i]);
    cd v = cs_objective(n, P, K, rms, td, tau, fp, fx, reg0, reg1, En, td_index, td_re + I * td_im);
    out[0] = creal(v);
    out[1] = cimag(v);
}

/* Statistic at given delays (taps) and patterns, float64 (no derivative). */
double oracle_sky_objective_at(int n, int P, int K, const float *rms, const float *td,
                               const double *tau, const double *fp, const double *fx,
                               double reg0, double reg1, double En)
{
    cd t[8], a[8], b[8];
    for (int i = 0; i < n; ++i) {
        t[i] = tau[i];
        a[i] = fp[i];
        b[i] = fx[i];
    }
    return creal(cs_objective(n, P, K, rms, td, t, a, b, reg0, reg1, En, -1, 0.0));
}
