/*
 * ORACLE -- TEST INFRASTRUCTURE ONLY.
 *
 * CPU restatement of PycWB's connected-component clustering of the selected
 * network pixels (pycwb/modules/coherence_native/clustering.py:18-157):
 *   cluster_oracle_labels   _label_components_grid (kernels.py:24-105): a
 *                           serial union-find over a dense (n_freq, n_time)
 *                           index map holding the last pixel of every cell;
 *                           roots are merged smaller-index-wins, labels are
 *                           handed out 1, 2, ... in pixel order of the roots
 *   cluster_oracle_np_sum   numpy's add.reduce of a contiguous float64 array
 *                           (the reference's np.sum / np.mean): 8192-element
 *                           chunks added to 0.0 in order, each chunk summed
 *                           pairwise (numpy loops_utils.h.src pairwise_sum:
 *                           below 8 values a plain loop, up to 128 values eight
 *                           strided accumulators, else split at n/2 rounded
 *                           down to a multiple of 8)
 *   cluster_oracle_subnet   _subnet_subrho_numba (kernels.py:141-191)
 *   cluster_oracle_run      cluster_pixels for one payload: labels, the groups
 *                           in label order with their pixels in index order,
 *                           and each group's energy, c_time, c_freq, subnet,
 *                           subrho (clustering.py:67-135)
 * Only tests/ load this library, and only as the checker.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

static int64_t uf_find(int64_t *parent, int64_t x)
{
    while (parent[x] != x) {
        parent[x] = parent[parent[x]];
        x = parent[x];
    }
    return x;
}

static void uf_union(int64_t *parent, int64_t a, int64_t b)
{
    int64_t ra = uf_find(parent, a), rb = uf_find(parent, b);
    if (ra == rb) return;
    if (ra < rb) parent[rb] = ra;
    else parent[ra] = rb;
}

/* kernels.py:24-105; labels 1-based. */
void cluster_oracle_labels(const int64_t *f, const int64_t *t, int64_t n, int n_freq, int n_time,
                           int kf, int kt, int64_t *labels)
{
    if (n <= 0) return;
    int64_t *parent = malloc(sizeof(int64_t) * n);
    int64_t *map = malloc(sizeof(int64_t) * (int64_t)n_freq * n_time);
    for (int64_t i = 0; i < n; ++i) parent[i] = i;
    for (int64_t c = 0; c < (int64_t)n_freq * n_time; ++c) map[c] = -1;
    for (int64_t i = 0; i < n; ++i) {
        if (f[i] < 0 || f[i] >= n_freq || t[i] < 0 || t[i] >= n_time) continue;
        int64_t *cell = map + f[i] * n_time + t[i];
        if (*cell >= 0) uf_union(parent, i, *cell);
        *cell = i;
    }
    for (int64_t i = 0; i < n; ++i) {
        if (f[i] < 0 || f[i] >= n_freq || t[i] < 0 || t[i] >= n_time) continue;
        int64_t f0 = f[i] - kf < 0 ? 0 : f[i] - kf, f1 = f[i] + kf + 1 > n_freq ? n_freq : f[i] + kf + 1;
        int64_t t0 = t[i] - kt < 0 ? 0 : t[i] - kt, t1 = t[i] + kt + 1 > n_time ? n_time : t[i] + kt + 1;
        for (int64_t ff = f0; ff < f1; ++ff)
            for (int64_t tt = t0; tt < t1; ++tt) {
                int64_t j = map[ff * n_time + tt];
                if (j >= 0 && j < i) uf_union(parent, i, j);
            }
    }
    int64_t *root_label = malloc(sizeof(int64_t) * n);
    for (int64_t i = 0; i < n; ++i) root_label[i] = -1;
    int64_t next = 1;
    for (int64_t i = 0; i < n; ++i) {
        int64_t r = uf_find(parent, i);
        if (root_label[r] < 0) root_label[r] = next++;
        labels[i] = root_label[r];
    }
    free(root_label); free(map); free(parent);
}

typedef double (*getter)(const void *ctx, int64_t i);

static double pw_sum(getter g, const void *ctx, int64_t s, int64_t n)
{
    if (n < 8) {
        double r = 0.0;
        for (int64_t i = 0; i < n; ++i) r += g(ctx, s + i);
        return r;
    }
    if (n <= 128) {
        double r[8];
        int64_t i;
        for (int k = 0; k < 8; ++k) r[k] = g(ctx, s + k);
        for (i = 8; i < n - (n % 8); i += 8)
            for (int k = 0; k < 8; ++k) r[k] += g(ctx, s + i + k);
        double res = ((r[0] + r[1]) + (r[2] + r[3])) + ((r[4] + r[5]) + (r[6] + r[7]));
        for (; i < n; ++i) res += g(ctx, s + i);
        return res;
    }
    int64_t n2 = n / 2;
    n2 -= n2 % 8;
    return pw_sum(g, ctx, s, n2) + pw_sum(g, ctx, s + n2, n - n2);
}

static double np_reduce(getter g, const void *ctx, int64_t n)
{
    double r = 0.0;
    for (int64_t s = 0; s < n; s += 8192) r += pw_sum(g, ctx, s, n - s < 8192 ? n - s : 8192);
    return r;
}

static double get_plain(const void *ctx, int64_t i) { return ((const double *)ctx)[i]; }

double cluster_oracle_np_sum(const double *x, int64_t n) { return np_reduce(get_plain, x, n); }

/* kernels.py:141-191; asnr and noise_rms (n_pix, n_ifo) row-major. */
void cluster_oracle_subnet(const double *asnr, const double *nrms, int64_t n_pix, int n_ifo,
                           double n_sub, double *subnet, double *subrho)
{
    double rho = 0.0, e_sub_total = 0.0, e_max_sum = 0.0, acc = 0.0;
    for (int64_t p = 0; p < n_pix; ++p) {
        double amp_sum = 0.0, e_max = 0.0, e_tot = 0.0, nsd = 0.0, msd = 0.0;
        for (int d = 0; d < n_ifo; ++d) {
            double amp = asnr[p * n_ifo + d];
            double x = amp * amp;
            double v = nrms[p * n_ifo + d] * nrms[p * n_ifo + d];
            amp_sum += fabs(amp);
            if (x > e_max) { e_max = x; msd = v; }
            e_tot += x;
            if (v > 0.0) nsd += 1.0 / v;
        }
        double a_fac = amp_sum > 0.0 ? e_tot / (amp_sum * amp_sum) : 1.0;
        rho += (1.0 - a_fac) * (e_tot - n_sub * 2.0);
        double y = e_tot - e_max;
        double x_corr = y * (1.0 + y / (e_max + 1.0e-5));
        if (msd > 0.0) nsd -= 1.0 / msd;
        double v_corr = (2.0 * e_max - e_tot) * msd * nsd / 10.0;
        e_sub_total += e_tot - e_max;
        e_max_sum += e_max;
        double a_cut = (x_corr + n_sub) != 0.0 ? x_corr / (x_corr + n_sub) : 0.0;
        double denom = x_corr + (v_corr > 0.0 ? v_corr : 1.0e-5);
        acc += (e_tot * x_corr / denom) * (a_cut > 0.5 ? a_cut : 0.0);
    }
    *subnet = acc / (e_max_sum + e_sub_total + 0.01);
    *subrho = rho >= 0.0 ? sqrt(rho) : NAN;
}

struct member_ctx { const double *v; const int64_t *idx; };
static double get_member(const void *ctx, int64_t i)
{
    const struct member_ctx *c = ctx;
    return c->v[c->idx[i]];
}
struct time_ctx { const int64_t *t; const int64_t *idx; double a, b; };
/* t * a - b, as numpy's element-wise (t_group * dt - dt / 2) */
static double get_time(const void *ctx, int64_t i)
{
    const struct time_ctx *c = ctx;
    return (double)c->t[c->idx[i]] * c->a - c->b;
}
static double get_scaled(const void *ctx, int64_t i)
{
    const struct time_ctx *c = ctx;
    return (double)c->t[c->idx[i]] * c->a;
}

/*
 * clustering.py:18-157 for one payload.  labels (n); members (n) pixel indices
 * grouped by cluster; cluster_off (n + 1); per cluster (n capacity): energy,
 * c_time, c_freq, subnet, subrho.  Returns the number of clusters.
 */
int64_t cluster_oracle_run(const int64_t *f, const int64_t *t, const double *energy,
                           const double *det_energy, int64_t n, int n_ifo, int n_freq, int n_time,
                           int kf, int kt, double rate, double n_sub, int64_t *labels,
                           int64_t *members, int64_t *cluster_off, double *c_energy,
                           double *c_time, double *c_freq, double *c_subnet, double *c_subrho)
{
    if (n <= 0) { cluster_off[0] = 0; return 0; }
    if (kf < 1) kf = 1;
    if (kt < 1) kt = 1;
    cluster_oracle_labels(f, t, n, n_freq, n_time, kf, kt, labels);
    int64_t nc = 0;
    for (int64_t i = 0; i < n; ++i) if (labels[i] > nc) nc = labels[i];
    for (int64_t c = 0; c <= nc; ++c) cluster_off[c] = 0;
    for (int64_t i = 0; i < n; ++i) cluster_off[labels[i]] += 1;
    for (int64_t c = 1; c <= nc; ++c) cluster_off[c] += cluster_off[c - 1];
    int64_t *fill = calloc(nc + 1, sizeof(int64_t));
    for (int64_t i = 0; i < n; ++i) {
        int64_t c = labels[i] - 1;
        members[cluster_off[c] + fill[c]++] = i;
    }
    free(fill);
    double dt = rate > 0.0 ? 1.0 / rate : 1.0;
    double *asnr = malloc(sizeof(double) * (n_ifo > 0 ? n_ifo : 1) * n);
    double *ones = malloc(sizeof(double) * (n_ifo > 0 ? n_ifo : 1) * n);
    for (int64_t c = 0; c < nc; ++c) {
        int64_t a = cluster_off[c], m = cluster_off[c + 1] - a;
        const int64_t *idx = members + a;
        struct member_ctx me = {energy, idx};
        c_energy[c] = np_reduce(get_member, &me, m);
        struct time_ctx tc = {t, idx, dt, dt / 2.0};
        c_time[c] = np_reduce(get_time, &tc, m) / (double)m;
        struct time_ctx fc = {f, idx, rate / 2.0, 0.0};
        c_freq[c] = np_reduce(get_scaled, &fc, m) / (double)m;
        if (n_ifo > 1) {
            for (int64_t k = 0; k < m; ++k)
                for (int d = 0; d < n_ifo; ++d) {
                    asnr[k * n_ifo + d] = sqrt(det_energy[idx[k] * n_ifo + d]);
                    ones[k * n_ifo + d] = 1.0;
                }
            cluster_oracle_subnet(asnr, ones, m, n_ifo, n_sub, c_subnet + c, c_subrho + c);
        } else {
            c_subnet[c] = 0.0;
            c_subrho[c] = 0.0;
        }
    }
    free(asnr); free(ones);
    return nc;
}
