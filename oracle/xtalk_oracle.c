/*
 * ORACLE -- TEST INFRASTRUCTURE ONLY.
 *
 * CPU restatement of PycWB's cross-talk lookups (pycwb/modules/xtalk/monster.py):
 *   getXTalk              (:157-192)  coefficients of one pixel pair;
 *   getXTalk_pixels_numba (:195-237)  a cluster's CSR from every ordered pair;
 *   getXTalk_pixels_fast  (:240-440)  the same CSR from each pixel's catalog
 *                                      entries, in the reference's entry order.
 * Only tests/ load this library, and only as the checker.
 *
 * Catalog layout (load_catalog, :63-153): layers (nRes) int32; table
 * (nRes, nRes, maxL+1, 2, 2) int32 [start, end) entry ranges for
 * (r1 >= r2, frequency index, odd time); coeff (E, 5) float32 rows
 * (relative index, c1, c2, c3, c4).  A pixel is (nLayer = layers + 1, time =
 * WDM map index t (nLayer) + f).
 * CSR rows (8 float32): neighbour index, c1^2+c2^2, c3^2+c4^2, the sum of all
 * four squares, then c1, c3, c2, c4 (a pixel pair seen from the coarser pixel
 * swaps c2 and c3).
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define TAB(r1, r2, k, odd, s) \
    table[((((int64_t)(r1) * nRes + (r2)) * (maxL + 1) + (k)) * 2 + (odd)) * 2 + (s)]

/* Resolution index of a pixel's nLayer: the last r with layers[r] + 1 == nLayer
 * (both the pairwise lookup's loop and the fast path's layer_to_r keep the last). */
static int res_of(int nLayer, const int32_t *layers, int nRes)
{
    int r = -1;
    for (int i = 0; i < nRes; ++i)
        if (layers[i] == nLayer - 1) r = i;
    return r;
}

/* getXTalk (monster.py:157-192): out = (c1, c2, c3, c4), or 3s when the pair
 * has no catalog entry.  Returns 0, or -1 when a resolution is missing. */
int xtalk_oracle_get(int nLayer1, int64_t indx1, int nLayer2, int64_t indx2,
                     const int32_t *layers, int nRes, int maxL, const int32_t *table,
                     const float *coeff, float *out)
{
    int r1 = res_of(nLayer1, layers, nRes), r2 = res_of(nLayer2, layers, nRes);
    if (r1 < 0 || r2 < 0) return -1;
    int swap = r1 < r2;
    if (swap) {
        int t = r1; r1 = r2; r2 = t;
        int64_t u = indx1; indx1 = indx2; indx2 = u;
    }
    int64_t layer1 = layers[r1];
    int64_t time1 = indx1 / (layer1 + 1), freq1 = indx1 % (layer1 + 1);
    int64_t odd = time1 & 1;
    int64_t index = indx2 - (time1 - odd) * (layer1 / layers[r2]) * (layers[r2] + 1);
    out[0] = out[1] = out[2] = out[3] = 3.0f;
    for (int64_t e = TAB(r1, r2, freq1, odd, 0); e < TAB(r1, r2, freq1, odd, 1); ++e) {
        if (index == (int64_t)coeff[e * 5]) {
            out[0] = coeff[e * 5 + 1];
            out[1] = coeff[e * 5 + 2];
            out[2] = coeff[e * 5 + 3];
            out[3] = coeff[e * 5 + 4];
            if (swap) {
                float t = out[1]; out[1] = out[2]; out[2] = t;
            }
            break;
        }
    }
    return 0;
}

static void put_row(float *row, float j, float a, float b, float c, float d)
{
    row[0] = j;
    row[1] = a * a + b * b;
    row[2] = c * c + d * d;
    row[3] = row[1] + row[2];
    row[4] = a;
    row[5] = c;
    row[6] = b;
    row[7] = d;
}

/* getXTalk_pixels_numba (monster.py:195-237).  nlayer/time (P); lookup (P, 2);
 * xt receives at most cap rows.  Returns the number of rows (or -1). */
int64_t xtalk_oracle_pixels_pairwise(int P, const int32_t *nlayer, const int64_t *time,
                                     const int32_t *layers, int nRes, int maxL,
                                     const int32_t *table, const float *coeff,
                                     int32_t *lookup, float *xt, int64_t cap)
{
    int64_t n = 0;
    float o[4];
    for (int i = 0; i < P; ++i) {
        lookup[2 * i] = (int32_t)n;
        for (int j = 0; j < P; ++j) {
            if (xtalk_oracle_get(nlayer[i], time[i], nlayer[j], time[j], layers, nRes, maxL,
                                 table, coeff, o) != 0)
                return -1;
            if (o[0] > 2.0f) continue;
            if (n >= cap) return -1;
            /* columns 4..7 = tmpOvlp[0], [2], [1], [3] */
            put_row(xt + 8 * n, (float)j, o[0], o[1], o[2], o[3]);
            n++;
        }
        lookup[2 * i + 1] = (int32_t)n;
    }
    return n;
}

typedef struct {
    int64_t key;
    int idx;
} keyed_t;

static int cmp_keyed(const void *a, const void *b)
{
    const keyed_t *x = a, *y = b;
    if (x->key != y->key) return x->key < y->key ? -1 : 1;
    return x->idx - y->idx;
}

/* The (nLayer, time) -> pixel map of the fast path: a later pixel with the
 * same key overwrites an earlier one (dictionary assignment). */
static int find_pixel(const keyed_t *map, int P, int64_t key)
{
    int lo = 0, hi = P;
    while (lo < hi) {
        int mid = (lo + hi) / 2;
        if (map[mid].key < key) lo = mid + 1; else hi = mid;
    }
    int found = -1;
    while (lo < P && map[lo].key == key) found = map[lo++].idx;
    return found;
}

#define KEY(nl, t) ((int64_t)(nl) * ((int64_t)1 << 32) + (int64_t)(t))

typedef struct {
    int64_t src, seq;
    float row[8];
} entry_t;

static int cmp_entry(const void *a, const void *b)
{
    const entry_t *x = a, *y = b;
    if (x->src != y->src) return x->src < y->src ? -1 : 1;
    if (x->seq != y->seq) return x->seq < y->seq ? -1 : 1;
    return 0;
}

/* getXTalk_pixels_fast (monster.py:240-440).  Entries of row j appear in the
 * order the reference's fill loop writes them: by the pixel i whose catalog
 * walk produced them (i == j for j's own entries), then by that walk's
 * (r_j, entry) order.  Returns the number of rows (or -1). */
int64_t xtalk_oracle_pixels_fast(int P, const int32_t *nlayer, const int64_t *time,
                                 const int32_t *layers, int nRes, int maxL,
                                 const int32_t *table, const float *coeff, int64_t n_coeff,
                                 int32_t *lookup, float *xt, int64_t cap)
{
    if (P == 0) return 0;
    keyed_t *map = malloc(sizeof(keyed_t) * P);
    for (int i = 0; i < P; ++i) {
        map[i].key = KEY(nlayer[i], time[i]);
        map[i].idx = i;
    }
    qsort(map, P, sizeof(keyed_t), cmp_keyed);
    int64_t n_ent = 0, cap_ent = 1024;
    entry_t *ent = malloc(sizeof(entry_t) * cap_ent);
    int64_t *row_of = malloc(sizeof(int64_t) * cap_ent);
    {
        for (int i = 0; i < P; ++i) {
            int ri = res_of(nlayer[i], layers, nRes);
            if (ri < 0) continue;
            int64_t layer_i = layers[ri];
            int64_t t1 = time[i] / (layer_i + 1), f1 = time[i] - t1 * (layer_i + 1);
            int64_t odd = t1 & 1;
            for (int rj = 0; rj <= ri; ++rj) {
                int64_t layer_j = layers[rj];
                int64_t base = (t1 - odd) * (layer_i / layer_j) * (layer_j + 1);
                for (int64_t e = TAB(ri, rj, f1, odd, 0); e < TAB(ri, rj, f1, odd, 1); ++e) {
                    int64_t indx2 = (int64_t)coeff[e * 5] + base;
                    if (indx2 < 0) continue;
                    int j = find_pixel(map, P, KEY(layer_j + 1, indx2));
                    if (j < 0) continue;
                    float c1 = coeff[e * 5 + 1], c2 = coeff[e * 5 + 2];
                    float c3 = coeff[e * 5 + 3], c4 = coeff[e * 5 + 4];
                    for (int k = 0; k < (rj < ri ? 2 : 1); ++k) {
                        if (n_ent == cap_ent) {
                            cap_ent *= 2;
                            ent = realloc(ent, sizeof(entry_t) * cap_ent);
                            row_of = realloc(row_of, sizeof(int64_t) * cap_ent);
                        }
                        entry_t *x = ent + n_ent;
                        x->src = i;
                        x->seq = (int64_t)rj * n_coeff + e;
                        if (k == 0) {
                            row_of[n_ent] = i;
                            x->row[0] = (float)j;
                            x->row[1] = c1 * c1 + c2 * c2;
                            x->row[2] = c3 * c3 + c4 * c4;
                            x->row[3] = c1 * c1 + c2 * c2 + c3 * c3 + c4 * c4;
                            x->row[4] = c1; x->row[5] = c3; x->row[6] = c2; x->row[7] = c4;
                        } else {
                            row_of[n_ent] = j;
                            x->row[0] = (float)i;
                            x->row[1] = c1 * c1 + c3 * c3;
                            x->row[2] = c2 * c2 + c4 * c4;
                            x->row[3] = c1 * c1 + c2 * c2 + c3 * c3 + c4 * c4;
                            x->row[4] = c1; x->row[5] = c2; x->row[6] = c3; x->row[7] = c4;
                        }
                        n_ent++;
                    }
                }
            }
        }
    }
    /* bucket by row, then order each row by (src, seq) */
    int64_t *cnt = calloc((size_t)P + 1, sizeof(int64_t));
    for (int64_t k = 0; k < n_ent; ++k) cnt[row_of[k] + 1]++;
    for (int i = 0; i < P; ++i) cnt[i + 1] += cnt[i];
    if (n_ent > cap) {
        free(map); free(ent); free(row_of); free(cnt);
        return -1;
    }
    entry_t *sorted = malloc(sizeof(entry_t) * (n_ent + 1));
    int64_t *fill = calloc((size_t)P, sizeof(int64_t));
    for (int64_t k = 0; k < n_ent; ++k) {
        int64_t r = row_of[k];
        sorted[cnt[r] + fill[r]++] = ent[k];
    }
    for (int i = 0; i < P; ++i) {
        qsort(sorted + cnt[i], (size_t)(cnt[i + 1] - cnt[i]), sizeof(entry_t), cmp_entry);
        lookup[2 * i] = (int32_t)cnt[i];
        lookup[2 * i + 1] = (int32_t)cnt[i + 1];
    }
    for (int64_t k = 0; k < n_ent; ++k) memcpy(xt + 8 * k, sorted[k].row, sizeof(float) * 8);
    free(map); free(ent); free(row_of); free(cnt); free(sorted); free(fill);
    return n_ent;
}
