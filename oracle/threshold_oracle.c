/*
 * ORACLE -- TEST INFRASTRUCTURE ONLY.
 *
 * CPU restatement of the array work of PycWB's compute_threshold
 * (pycwb/modules/coherence_native/veto_threshold.py:98-222): the detectors'
 * energy maps summed in detector order and clipped to [0, hi] in one of the
 * reference's two orders (time_major = 0: frequency rows over the time bins
 * [begin, begin + count / n_freq), :173-176; time_major = 1: the time-major
 * flat indices [begin, begin + count), :136-157), the counts above 1e-4 and
 * 1e-3, the values sorted ascending (np.sort), and the sums of the values above
 * 1e-3 and of their logs in numpy's add.reduce order (cluster_oracle_np_sum).
 * Only tests/ load this library, and only as the checker.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

double cluster_oracle_np_sum(const double *x, int64_t n);

void threshold_oracle_values(const double *maps, int n_ifo, int n_freq, int64_t n_time,
                             int time_major, int64_t begin, int64_t count, double hi,
                             double *values, int64_t *counts)
{
    int64_t ncol = time_major ? 0 : count / n_freq;
    counts[0] = counts[1] = 0;
    for (int64_t k = 0; k < count; ++k) {
        int64_t f, t;
        if (time_major) {
            t = (begin + k) / n_freq;
            f = (begin + k) % n_freq;
        } else {
            f = k / ncol;
            t = begin + k % ncol;
        }
        /* np.sum(axis=0) starts from 0.0 (so -0.0 sums to +0.0); the time-major
         * path's w = flat0.copy(); w += ... starts from the first map */
        double w = time_major ? maps[f * n_time + t] : 0.0 + maps[f * n_time + t];
        for (int d = 1; d < n_ifo; ++d) w = w + maps[((int64_t)d * n_freq + f) * n_time + t];
        double v = w < 0.0 ? 0.0 : w;   /* np.clip keeps -0.0 and NaN */
        v = v > hi ? hi : v;
        values[k] = v;
        counts[0] += v > 1.0e-4;
        counts[1] += v > 1.0e-3;
    }
}

static int cmp_double(const void *a, const void *b)
{
    double x = *(const double *)a, y = *(const double *)b;
    return (x > y) - (x < y);
}

void threshold_oracle_sorted(const double *values, int64_t count, double *out)
{
    memcpy(out, values, sizeof(double) * count);
    qsort(out, count, sizeof(double), cmp_double);
}

/* np.sum(positive), np.sum(np.log(positive)) with positive = values > 1e-3 in order */
int64_t threshold_oracle_positive_sums(const double *values, int64_t count, double *sum,
                                       double *log_sum)
{
    double *p = calloc(count > 0 ? count : 1, sizeof(double));
    int64_t m = 0;
    for (int64_t k = 0; k < count; ++k)
        if (values[k] > 1.0e-3) p[m++] = values[k];
    *sum = cluster_oracle_np_sum(p, m);
    for (int64_t k = 0; k < m; ++k) p[k] = log(p[k]);
    *log_sum = cluster_oracle_np_sum(p, m);
    free(p);
    return m;
}
