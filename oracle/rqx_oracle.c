/*
 * C/OpenMP port of the reference's SQD matvec, cache strategy (1, 0)
 * (rqutils/sqd.py:347-351 builds the source-index cache; apply_h_x_cached,
 * sqd.py:694-709, recomposes each group's diagonal and gathers per matvec).
 *
 * TEST INFRASTRUCTURE ONLY: timed as the CPU baseline and used as a checker;
 * never linked into the product library.
 *
 * States are uint64 (bit q = qubit q); by the index-parity lemma the ranks
 * equal those of the reference's packed-byte rows.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
    uint64_t key;
    int64_t idx;
} pair_t;

/* stable LSD radix sort of pairs by key (8 passes of 8 bits) */
static void radix_sort_pairs(pair_t *a, pair_t *tmp, int64_t n) {
    for (int pass = 0; pass < 8; ++pass) {
        const int shift = 8 * pass;
        int64_t count[257];
        memset(count, 0, sizeof(count));
        for (int64_t i = 0; i < n; ++i) count[((a[i].key >> shift) & 0xff) + 1]++;
        int all_same = 0;
        for (int b = 1; b <= 256; ++b)
            if (count[b] == n) all_same = 1;
        if (all_same) continue;
        for (int b = 0; b < 256; ++b) count[b + 1] += count[b];
        for (int64_t i = 0; i < n; ++i) tmp[count[(a[i].key >> shift) & 0xff]++] = a[i];
        memcpy(a, tmp, (size_t)n * sizeof(pair_t));
    }
}

/*
 * get_xsource (sqd.py:514-540): stable sort of [S; S^x] with row indices,
 * match adjacent equal keys, then pick the entries of S in sorted order.
 * out[i] = rank of keys[i]^x in keys, or -1.  Returns 0, or -1 on allocation failure.
 */
int oracle_xsource(const uint64_t *keys, int64_t n, uint64_t x, int32_t *out) {
    const int64_t m = 2 * n;
    pair_t *joined = (pair_t *)malloc((size_t)m * sizeof(pair_t));
    pair_t *tmp = (pair_t *)malloc((size_t)m * sizeof(pair_t));
    if (!joined || !tmp) {
        free(joined);
        free(tmp);
        return -1;
    }
    for (int64_t i = 0; i < n; ++i) {
        joined[i].key = keys[i];
        joined[i].idx = i;
        joined[n + i].key = keys[i] ^ x;
        joined[n + i].idx = n + i;
    }
    radix_sort_pairs(joined, tmp, m);
    /* position of each S entry in sorted order: S entries appear in index order */
    int64_t s_seen = 0;
    for (int64_t k = 0; k < m; ++k) {
        if (joined[k].idx >= n) continue;
        int32_t src = -1;
        if (k + 1 < m && joined[k + 1].key == joined[k].key) src = (int32_t)(joined[k + 1].idx - n);
        out[s_seen++] = src;
    }
    free(joined);
    free(tmp);
    return 0;
}

/*
 * One group of apply_h_x_cached (sqd.py:704-707): y += C(s) * v[src],
 * C(s) = sum_k c_k (1 - 2 (popcount(s & z_k) & 1)) (get_diagonal, sqd.py:603-628).
 * c, v, y complex interleaved doubles.
 */
void oracle_apply_group(const uint64_t *keys, int64_t n, const int32_t *src, const uint64_t *z,
                        const double *c, int64_t k, const double *v, double *y) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
        const int32_t s = src[i];
        if (s < 0) continue;
        double dr = 0.0, di = 0.0;
        const uint64_t b = keys[i];
        for (int64_t t = 0; t < k; ++t) {
            const double sg = (__builtin_popcountll(b & z[t]) & 1) ? -1.0 : 1.0;
            dr += sg * c[2 * t];
            di += sg * c[2 * t + 1];
        }
        const double vr = v[2 * s], vi = v[2 * s + 1];
        y[2 * i] += dr * vr - di * vi;
        y[2 * i + 1] += dr * vi + di * vr;
    }
}

/*
 * The same source indices by binary search (OpenMP over rows): the untimed setup
 * of the CPU baseline at sizes where the sort-merge of every group would take
 * minutes (the reference builds this cache once, sqd.py:347-351).
 */
void oracle_xsource_bsearch(const uint64_t *keys, int64_t n, uint64_t x, int32_t *out) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
        const uint64_t t = keys[i] ^ x;
        int64_t lo = 0, hi = n;
        while (lo < hi) {
            const int64_t mid = lo + (hi - lo) / 2;
            if (keys[mid] < t) lo = mid + 1;
            else hi = mid;
        }
        out[i] = (lo < n && keys[lo] == t) ? (int32_t)lo : -1;
    }
}

/*
 * oracle_apply_group in single precision: complex64 vectors (interleaved floats),
 * coefficients c as complex interleaved doubles rounded to float, sums in float.
 */
void oracle_apply_group_c64(const uint64_t *keys, int64_t n, const int32_t *src, const uint64_t *z,
                            const double *c, int64_t k, const float *v, float *y) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
        const int32_t s = src[i];
        if (s < 0) continue;
        float dr = 0.0f, di = 0.0f;
        const uint64_t b = keys[i];
        for (int64_t t = 0; t < k; ++t) {
            const float sg = (__builtin_popcountll(b & z[t]) & 1) ? -1.0f : 1.0f;
            dr += sg * (float)c[2 * t];
            di += sg * (float)c[2 * t + 1];
        }
        const float vr = v[2 * s], vi = v[2 * s + 1];
        y[2 * i] += dr * vr - di * vi;
        y[2 * i + 1] += dr * vi + di * vr;
    }
}

int oracle_num_threads(void) {
    int t = 1;
#pragma omp parallel
    {
#pragma omp single
        {
#ifdef _OPENMP
            extern int omp_get_num_threads(void);
            t = omp_get_num_threads();
#endif
        }
    }
    return t;
}
