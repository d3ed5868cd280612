/*
 * ORACLE -- TEST INFRASTRUCTURE ONLY.
 *
 * CPU restatement of PycWB's network-pixel selection for one lag
 * (pycwb/modules/coherence_native/selection.py:144-243 with the kernels
 * _align_threshold_map_numba, kernels.py:231-302, _candidate_passes_support_numba,
 * :107-137, and _select_candidates_numba, :305-368).  Only tests/ load this
 * library, and only as the checker.
 *
 * energy: (n_ifo, n_freq, n_time) float64; shift (n_ifo) circular shifts in
 * bins; veto (n_time) int16 or NULL (1 keep, 0 reject).  Outputs: live
 * (n_time) uint8, mask (n_freq, n_time) uint8 and, in (frequency, time)
 * row-major order, the selected pixels' frequency and time indices, summed raw
 * energy, per-detector energy (clipped at 0) and per-detector map index
 * (time * n_freq + frequency).  Returns the pixel count, or -1 when cap is
 * too small.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define E(d, f, t) energy[((int64_t)(d) * n_freq + (f)) * n_time + (t)]
#define CB(f, t) comb[(int64_t)(f) * n_time + (t)]

int64_t select_oracle(int n_ifo, int n_freq, int64_t n_time, const double *energy,
                      const int64_t *shift, const int16_t *veto, int edge_bins, int ib, int ie,
                      double eo, uint8_t *live, uint8_t *mask, int64_t *freq, int64_t *tidx,
                      double *value, double *det_energy, int64_t *det_index, int64_t cap)
{
    const int64_t vs = edge_bins, vstop = n_time - edge_bins, nn = vstop - vs;
    const double em = 2.0 * eo, eh = em * em;
    double *comb = calloc((size_t)n_freq * n_time, sizeof(double));
    memset(live, 0, (size_t)n_time);
    memset(mask, 0, (size_t)n_freq * n_time);
    for (int64_t t = vs; t < vs + nn; ++t) {
        int64_t u = t - vs;
        int ok = 1;
        if (veto)
            for (int d = 0; d < n_ifo; ++d)
                if (veto[vs + (u + shift[d]) % nn] == 0) {
                    ok = 0;
                    break;
                }
        live[t] = (uint8_t)ok;
    }
    for (int f = ib; f < n_freq; ++f)
        for (int64_t t = vs; t < vstop; ++t) {
            if (veto && !live[t]) continue;
            int64_t u = t - vs;
            double v = 0.0;
            for (int d = 0; d < n_ifo; ++d) v += E(d, f, vs + (u + shift[d]) % nn);
            if (v < eo) continue;
            CB(f, t) = v > em ? em + 0.1 : v;
        }
    const int ii = n_freq - 2;
    const int64_t margin = edge_bins > 2 ? edge_bins : 2;
    int fe = ie > ib ? ie : ib;
    if (fe > n_freq - 1) fe = n_freq - 1;
    int64_t n = 0;
    for (int f = ib; f < fe; ++f)
        for (int64_t t = margin; t < n_time - margin; ++t) {
            double e = CB(f, t);
            if (e < eo) continue;
            double ct = CB(f + 1, t) + CB(f, t + 1) + CB(f + 1, t + 1);
            double cb = CB(f - 1, t) + CB(f, t - 1) + CB(f - 1, t - 1);
            double ht = CB(f + 1, t + 2);
            if (f < ii) ht += CB(f + 2, t + 2) + CB(f + 2, t + 1);
            double hb = CB(f - 1, t - 2);
            if (f >= 2) hb += CB(f - 2, t - 2) + CB(f - 2, t - 1);
            if ((ct + cb) * e < eh && (ct + ht) * e < eh && (cb + hb) * e < eh && e < em)
                continue;
            mask[(int64_t)f * n_time + t] = 1;
            if (n >= cap) {
                free(comb);
                return -1;
            }
            freq[n] = f;
            tidx[n] = t;
            double raw = 0.0;
            for (int d = 0; d < n_ifo; ++d) {
                int64_t dt = (nn > 0 && t >= vs && t < vstop) ? vs + (t - vs + shift[d]) % nn : t;
                double x = E(d, f, dt);
                raw += x;
                det_energy[n * n_ifo + d] = x > 0.0 ? x : 0.0;
                det_index[n * n_ifo + d] = dt * n_freq + f;
            }
            value[n] = raw;
            n++;
        }
    free(comb);
    return n;
}
