/*
 * CPU oracle for the LOF forward path.
 *
 * TEST INFRASTRUCTURE ONLY. Results of the GPU path are compared with this
 * file's; it is loaded by the test-suite, by the smoke check and by the
 * benchmark's CPU-baseline leg, and by nothing else. The product path (culof_amd/) never links, loads or calls
 * it, and has no CPU fallback of any kind.
 *
 * Every function restates one stage of the reference algorithm in plain C,
 * cited by file:line of the reference sources:
 *
 *   oracle_center        src/cuda/preprocess.cu:31-84   (FP64 centring, norms)
 *   oracle_key_row       src/cuda/knn.cu:150-151 (squared distance, diagonal and
 *                        NaN -> +inf), in difference form; see "Selection keys"
 *   oracle_gram_key_row  src/cuda/pipeline.cu:182-188 + src/cuda/knn.cu:150-151
 *   oracle_select_row    src/cuda/knn.cu:136-231        (exact k-th key, gather order)
 *   oracle_lrd           src/cuda/lof.cu:32-46
 *   oracle_scores        src/cuda/lof.cu:49-62
 *   oracle_lof           src/cuda/pipeline.cu:131-207   (the whole forward)
 *   oracle_lof_fp64      tests/cpp/test_culof.cu:39-80  (FP64 brute-force LOF)
 *   oracle_knn_ties      the FP64 tie-break form (no reference code; its sets
 *                        are pinned to oracle_lof_fp64's)
 *
 * Where the reference leaves an arithmetic order unspecified (cuBLAS), this
 * oracle fixes one: single-precision FMA chains in ascending feature order
 * starting from +0. The HIP kernels use the same chains, so keys, neighbour
 * sets and scores compare bit for bit.
 *
 * Pinning (what checks this file, in tests/test_oracle.py): scikit-learn
 * 1.7.2 FP64 golden vectors in tests/golden/, the scikit-learn docstring
 * known-answer test, the equally-spaced circle (LOF == 1 by symmetry), n=2,
 * a brute-force sort for the selection rule, integer-exact sums for the
 * centring order, and planted mutations (the `mutation` argument) that each of
 * those checks must catch.
 */

#include <math.h>
#include <stddef.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define REDUCE_LANES 256

/* ---------------------------------------------------------------------------
 * Centring and norms: src/cuda/preprocess.cu:31-84.
 *
 * The reference reduces each feature with 256 partial sums, partial p holding
 * rows p, p+256, p+512, ... in ascending order, followed by a halving tree
 * over the partials (preprocess.cu:35-45). That order is restated exactly so
 * the FP64 mean is bit-identical to the reference's.
 * ------------------------------------------------------------------------- */

static double tree_reduce(double* part) {
    for (int off = REDUCE_LANES / 2; off > 0; off >>= 1) {
        for (int p = 0; p < off; ++p) part[p] += part[p + off];
    }
    return part[0];
}

void oracle_center(const float* x, int n, int d, int normalize, float* xc, float* norms,
                   double* mean_out, double* inv_std_out) {
    double part[REDUCE_LANES];
    for (int f = 0; f < d; ++f) {
        for (int p = 0; p < REDUCE_LANES; ++p) {
            double s = 0.0;
            for (int i = p; i < n; i += REDUCE_LANES) s += (double)x[(size_t)i * d + f];
            part[p] = s;
        }
        const double mean = tree_reduce(part) / n;
        double inv_std = 1.0;
        if (normalize) {
            /* Population variance; preprocess.cu:48-65. The squared
             * deviation is accumulated with a fused multiply-add. */
            for (int p = 0; p < REDUCE_LANES; ++p) {
                double s = 0.0;
                for (int i = p; i < n; i += REDUCE_LANES) {
                    const double dv = (double)x[(size_t)i * d + f] - mean;
                    s = fma(dv, dv, s);
                }
                part[p] = s;
            }
            const double sd = sqrt(tree_reduce(part) / n);
            inv_std = (sd > 1e-12) ? 1.0 / sd : 1.0;
        }
        if (mean_out) mean_out[f] = mean;
        if (inv_std_out) inv_std_out[f] = inv_std;
        for (int i = 0; i < n; ++i) {
            const size_t o = (size_t)i * d + f;
            xc[o] = (float)(((double)x[o] - mean) * inv_std);
        }
    }
    /* preprocess.cu:75-84: FP64 sum of exact float*float products. */
    for (int i = 0; i < n; ++i) {
        double s = 0.0;
        const float* row = xc + (size_t)i * d;
        for (int t = 0; t < d; ++t) s += (double)row[t] * (double)row[t];
        norms[i] = (float)s;
    }
}

/* ---------------------------------------------------------------------------
 * Selection keys.
 *
 * The reference keys its selection by the Gram expansion
 * |x_i|^2 + |x_j|^2 - 2 x_i.x_j (pipeline.cu:182-188, knn.cu:150-151), with the
 * diagonal and a NaN expansion mapped to +inf. In FP32 that expansion cancels:
 * its error is about d u (|x_i|^2 + |x_j|^2), 1e-4 relative to near-neighbour
 * distances on clustered data. Selection here is keyed by the squared
 * distance in difference form instead,
 *     e(i, j) = sum_t (a_t - b_t)^2  as an FP32 FMA chain in ascending t,
 * whose error is about d u e(i, j), and uses the Gram expansion (over the
 * centred rows) only as a screen with a rigorous error bound
 * (oracle_gram_key_row restates it for the bound test).
 *
 * Key rows `a`, `b` are the input rows themselves, or the z-scored rows when
 * normalising (distances are translation-invariant, so centring is not
 * applied to the key rows; it would only add rounding).
 * ------------------------------------------------------------------------- */

static float key_of(const float* xi, const float* xj, int d) {
    float e = 0.0f;
    for (int t = 0; t < d; ++t) {
        const float diff = xi[t] - xj[t];
        e = fmaf(diff, diff, e);
    }
    return isnan(e) ? INFINITY : e;
}

void oracle_key_row(const float* xs, int n, int d, int i, float* keys) {
    const float* xi = xs + (size_t)i * d;
    for (int j = 0; j < n; ++j) keys[j] = (j == i) ? INFINITY : key_of(xi, xs + (size_t)j * d, d);
}

/* The reference's Gram-expansion key over centred rows (knn.cu:150-151):
 * g = FP32 FMA chain of x_i.x_j, raw = fma(-2, g, n_i + n_j), NaN and the
 * diagonal -> +inf, negatives -> +0. */
void oracle_gram_key_row(const float* xc, const float* norms, int n, int d, int i, float* keys) {
    const float* xi = xc + (size_t)i * d;
    for (int j = 0; j < n; ++j) {
        if (j == i) {
            keys[j] = INFINITY;
            continue;
        }
        const float* xj = xc + (size_t)j * d;
        float g = 0.0f;
        for (int t = 0; t < d; ++t) g = fmaf(xi[t], xj[t], g);
        const float raw = fmaf(-2.0f, g, norms[i] + norms[j]);
        keys[j] = isnan(raw) ? INFINITY : (raw > 0.0f ? raw : 0.0f);
    }
}

static uint32_t bits_of(float f) {
    uint32_t u;
    memcpy(&u, &f, sizeof u);
    return u;
}

static float float_of(uint32_t u) {
    float f;
    memcpy(&f, &u, sizeof f);
    return f;
}

/* k-th smallest (1-based) of v[0..n) by quickselect; v is permuted. */
static uint32_t kth_smallest_u32(uint32_t* v, int n, int k) {
    int lo = 0, hi = n - 1;
    const int target = k - 1;
    while (lo < hi) {
        const int mid = lo + (hi - lo) / 2;
        /* median of three as pivot */
        uint32_t a = v[lo], b = v[mid], c = v[hi];
        uint32_t pivot = (a < b) ? ((b < c) ? b : (a < c ? c : a)) : ((a < c) ? a : (b < c ? c : b));
        int i = lo, j = hi;
        while (i <= j) {
            while (v[i] < pivot) ++i;
            while (v[j] > pivot) --j;
            if (i <= j) {
                const uint32_t t = v[i];
                v[i] = v[j];
                v[j] = t;
                ++i;
                --j;
            }
        }
        if (target <= j) hi = j;
        else if (target >= i) lo = i;
        else return v[target];
    }
    return v[target];
}

/*
 * knn.cu:136-231 restated: the k-th smallest key (by IEEE-754 bit pattern,
 * which orders non-negative floats numerically) is the k-distance squared;
 * the neighbour list holds every key strictly below it in ascending index
 * order, then the lowest-index keys equal to it until k entries are filled
 * (knn.cu:190-224). `scratch` holds n u32.
 */
void oracle_select_row(const float* keys, int n, int k, uint32_t* scratch, int* idx_out,
                       float* d2_out, float* kth_out) {
    for (int j = 0; j < n; ++j) scratch[j] = bits_of(keys[j]);
    const uint32_t kth = kth_smallest_u32(scratch, n, k);
    int below = 0;
    for (int j = 0; j < n; ++j) below += bits_of(keys[j]) < kth;
    int slot_lt = 0, slot_eq = below;
    for (int j = 0; j < n; ++j) {
        const uint32_t u = bits_of(keys[j]);
        if (u < kth) {
            idx_out[slot_lt] = j;
            d2_out[slot_lt] = keys[j];
            ++slot_lt;
        } else if (u == kth && slot_eq < k) {
            idx_out[slot_eq] = j;
            d2_out[slot_eq] = keys[j];
            ++slot_eq;
        }
    }
    *kth_out = float_of(kth);
}

/* kNN for rows [row0, row0+rows); outputs are row-major (rows x k). */
void oracle_knn_rows(const float* xs, int n, int d, int k, int row0, int rows, int* idx_out,
                     float* d2_out, float* kdist2_out) {
    float* keys = (float*)malloc(sizeof(float) * (size_t)n);
    uint32_t* scratch = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)n);
    for (int r = 0; r < rows; ++r) {
        const int i = row0 + r;
        oracle_key_row(xs, n, d, i, keys);
        oracle_select_row(keys, n, k, scratch, idx_out + (size_t)r * k, d2_out + (size_t)r * k,
                          kdist2_out + r);
    }
    free(keys);
    free(scratch);
}

/* ---------------------------------------------------------------------------
 * FP64 tie-break of near-tie rows (an addition with no reference code; it
 * settles rows on the (distance, index) order of the reference's FP64 test
 * oracle, tests/cpp/test_culof.cu:55-66). With K and K' the k-th and (k+1)-th
 * FP32 keys of a row and rho = (1 + g) / (1 - g), g = (d + 4) 2^-24: a row
 * with K' <= K rho is re-selected over the band {j : e32_j <= K rho} by
 * (e64, index), e64 the FP64 squared distance accumulated in ascending
 * feature order with separate multiply and add; other rows keep the FP32
 * result. Slot order follows the reference's rule on the keys that decided
 * the row. `tie_out[r]` is 1 for rows settled in FP64.
 * ------------------------------------------------------------------------- */

typedef struct {
    uint64_t e64;
    int idx;
    float e32;
} band_t;

static int band_cmp(const void* a, const void* b) {
    const band_t* x = (const band_t*)a;
    const band_t* y = (const band_t*)b;
    if (x->e64 != y->e64) return x->e64 < y->e64 ? -1 : 1;
    return (x->idx > y->idx) - (x->idx < y->idx);
}

static uint64_t word_of(float key, int idx) {
    uint32_t kb = bits_of(key);
    if (kb > 0x7F800000u) kb = 0x7F800000u;
    return ((uint64_t)kb << 32) | (uint32_t)idx;
}

double oracle_tie_band_ratio(int d) {
    const double g = (d + 4) * 5.9604644775390625e-8;
    return (1.0 + g) / (1.0 - g);
}

void oracle_knn_ties(const float* xs, int n, int d, int k, int row0, int rows, int* idx_out,
                     float* d2_out, float* kdist2_out, int* tie_out) {
    if (k + 1 > n - 1) {
        oracle_knn_rows(xs, n, d, k, row0, rows, idx_out, d2_out, kdist2_out);
        for (int r = 0; r < rows; ++r) tie_out[r] = 0;
        return;
    }
    const int k1 = k + 1;
    const double rho = oracle_tie_band_ratio(d);
    int* idx1 = (int*)malloc(sizeof(int) * (size_t)k1);
    float* d21 = (float*)malloc(sizeof(float) * (size_t)k1);
    band_t* band = (band_t*)malloc(sizeof(band_t) * (size_t)n);
    for (int r = 0; r < rows; ++r) {
        const int i = row0 + r;
        float kd1;
        oracle_knn_rows(xs, n, d, k1, i, 1, idx1, d21, &kd1);
        int tmax = 0;
        uint64_t wmax = word_of(d21[0], idx1[0]);
        for (int t = 1; t < k1; ++t) {
            const uint64_t w = word_of(d21[t], idx1[t]);
            if (w > wmax) {
                wmax = w;
                tmax = t;
            }
        }
        uint64_t wk = 0;
        for (int t = 0; t < k1; ++t) {
            if (t == tmax) continue;
            const uint64_t w = word_of(d21[t], idx1[t]);
            if (w > wk) wk = w;
        }
        const float K = float_of((uint32_t)(wk >> 32));
        const float K1 = float_of((uint32_t)(wmax >> 32));
        int* io = idx_out + (size_t)r * k;
        float* dout = d2_out + (size_t)r * k;
        if (isfinite(K) && (double)K1 <= (double)K * rho) {
            tie_out[r] = 1;
            const double bound = (double)K * rho;
            const float* xi = xs + (size_t)i * d;
            int m = 0;
            for (int j = 0; j < n; ++j) {
                if (j == i) continue;
                const float* xj = xs + (size_t)j * d;
                const float e = key_of(xi, xj, d);
                if (!((double)e <= bound)) continue;
                double s2 = 0.0;
                for (int t = 0; t < d; ++t) {
                    const double diff = (double)xi[t] - (double)xj[t];
                    s2 = s2 + diff * diff;
                }
                uint64_t u;
                memcpy(&u, &s2, sizeof u);
                band[m].e64 = u;
                band[m].idx = j;
                band[m].e32 = e;
                ++m;
            }
            qsort(band, (size_t)m, sizeof(band_t), band_cmp);
            const uint64_t kth = band[k - 1].e64;
            for (int t = 0; t < k; ++t) {
                int slot = t;
                if (band[t].e64 < kth) {
                    slot = 0;
                    for (int f = 0; f < k; ++f) slot += band[f].e64 < kth && band[f].idx < band[t].idx;
                }
                io[slot] = band[t].idx;
                dout[slot] = band[t].e32;
            }
            kdist2_out[r] = band[k - 1].e32;
        } else {
            tie_out[r] = 0;
            int slot = 0;
            for (int pass = 0; pass < 2; ++pass) {
                for (int t = 0; t < k1; ++t) {
                    if (t == tmax) continue;
                    if ((d21[t] < K) == (pass == 0)) {
                        io[slot] = idx1[t];
                        dout[slot] = d21[t];
                        ++slot;
                    }
                }
            }
            kdist2_out[r] = K;
        }
    }
    free(idx1);
    free(d21);
    free(band);
}

/* ---------------------------------------------------------------------------
 * Density and score: lof.cu:32-62. `mutation` plants a known mistake so the
 * tests can show their checks catch it: 1 = k-distance read from slot k-2's
 * key instead of the k-th key, 2 = reach distance without the max(),
 * 3 = inverted density ratio.
 * ------------------------------------------------------------------------- */

void oracle_lrd(const int* idx, const float* d2, const float* kdist2, int n, int k, int mutation,
                float* lrd) {
    for (int i = 0; i < n; ++i) {
        double sum = 0.0;
        for (int t = 0; t < k; ++t) {
            const int j = idx[(size_t)i * k + t];
            float reach2 = fmaxf(kdist2[j], d2[(size_t)i * k + t]);
            if (mutation == 2) reach2 = d2[(size_t)i * k + t];
            sum += sqrt((double)reach2);
        }
        lrd[i] = (float)(1.0 / (sum / k + 1e-10));
    }
}

void oracle_scores(const int* idx, const float* lrd, int n, int k, int mutation, float* scores) {
    for (int i = 0; i < n; ++i) {
        const double self = (double)lrd[i];
        double sum = 0.0;
        for (int t = 0; t < k; ++t) {
            const double other = (double)lrd[idx[(size_t)i * k + t]];
            sum += (mutation == 3) ? self / other : other / self;
        }
        scores[i] = (float)(sum / k);
    }
}

/*
 * The whole FP32 forward (pipeline.cu:131-207): centre, keys, select, LRD,
 * score. Optional outputs (idx, d2, kdist2, lrd) may be NULL; idx/d2 are
 * row-major (n x k).
 */
int oracle_lof(const float* x, int n, int d, int k, int normalize, int mutation, float* scores,
               int* idx_out, float* d2_out, float* kdist2_out, float* lrd_out) {
    if (!x || n < 2 || d < 1 || k < 1 || k >= n) return 1;
    float* xc = (float*)malloc(sizeof(float) * (size_t)n * d);
    float* norms = (float*)malloc(sizeof(float) * (size_t)n);
    int* idx = idx_out ? idx_out : (int*)malloc(sizeof(int) * (size_t)n * k);
    float* d2 = d2_out ? d2_out : (float*)malloc(sizeof(float) * (size_t)n * k);
    float* kd = kdist2_out ? kdist2_out : (float*)malloc(sizeof(float) * (size_t)n);
    float* lrd = lrd_out ? lrd_out : (float*)malloc(sizeof(float) * (size_t)n);

    oracle_center(x, n, d, normalize, xc, norms, NULL, NULL);
    /* key rows: the input itself, or the z-scored rows when normalising */
    oracle_knn_rows(normalize ? xc : x, n, d, k, 0, n, idx, d2, kd);
    if (mutation == 1 && k >= 2) {
        for (int i = 0; i < n; ++i) {
            /* slot order is ascending index, so take the (k-1)-th key by value */
            float best = -1.0f;
            for (int t = 0; t < k; ++t) {
                const float v = d2[(size_t)i * k + t];
                if (v < kd[i] && v > best) best = v;
            }
            if (best >= 0.0f) kd[i] = best;
        }
    }
    oracle_lrd(idx, d2, kd, n, k, mutation, lrd);
    oracle_scores(idx, lrd, n, k, mutation, scores);

    free(xc);
    free(norms);
    if (!idx_out) free(idx);
    if (!d2_out) free(d2);
    if (!kdist2_out) free(kd);
    if (!lrd_out) free(lrd);
    return 0;
}

/* ---------------------------------------------------------------------------
 * New points against a fitted set (no reference code: scikit-learn's novelty
 * mode, sklearn/neighbors/_lof.py score_samples, restated in the FP32/FP64 mix
 * of the forward above). The forward on x gives the fitted k-distances and
 * densities; query row r is keyed against every fitted key row (the raw rows,
 * or z-scored with the fitted mean and scale), nothing excluded, selected by
 * the same rule and slot order, then
 *   lrd_q = 1 / (mean_t sqrt(max(kd[j_t], e_t)) + 1e-10),
 *   lof_q = mean_t lrd[j_t] / lrd_q,
 * both accumulated in FP64 in slot order. idx_out / d2_out (m x k, row-major)
 * are optional.
 * ------------------------------------------------------------------------- */

int oracle_lof_novelty(const float* x, int n, int d, const float* q, int m, int k, int normalize,
                       float* scores_out, int* idx_out, float* d2_out) {
    if (!x || !q || n < 2 || d < 1 || k < 1 || k >= n || m < 0) return 1;
    float* xc = (float*)malloc(sizeof(float) * (size_t)n * d);
    float* norms = (float*)malloc(sizeof(float) * (size_t)n);
    double* mean = (double*)malloc(sizeof(double) * (size_t)d);
    double* inv_std = (double*)malloc(sizeof(double) * (size_t)d);
    float* fit_scores = (float*)malloc(sizeof(float) * (size_t)n);
    float* kd = (float*)malloc(sizeof(float) * (size_t)n);
    float* lrd = (float*)malloc(sizeof(float) * (size_t)n);
    float* qs = (float*)malloc(sizeof(float) * (size_t)(m > 0 ? m : 1) * d);
    float* keys = (float*)malloc(sizeof(float) * (size_t)n);
    uint32_t* scratch = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)n);
    int* idx = (int*)malloc(sizeof(int) * (size_t)k);
    float* d2 = (float*)malloc(sizeof(float) * (size_t)k);

    oracle_lof(x, n, d, k, normalize, 0, fit_scores, NULL, NULL, kd, lrd);
    oracle_center(x, n, d, normalize, xc, norms, mean, inv_std);
    for (int r = 0; r < m; ++r) {
        for (int t = 0; t < d; ++t) {
            const float v = q[(size_t)r * d + t];
            qs[(size_t)r * d + t] = normalize ? (float)(((double)v - mean[t]) * inv_std[t]) : v;
        }
    }
    const float* xs = normalize ? xc : x;
    for (int r = 0; r < m; ++r) {
        const float* qr = qs + (size_t)r * d;
        for (int j = 0; j < n; ++j) keys[j] = key_of(qr, xs + (size_t)j * d, d);
        float kth;
        oracle_select_row(keys, n, k, scratch, idx, d2, &kth);
        double reach = 0.0;
        for (int t = 0; t < k; ++t) reach += sqrt((double)fmaxf(kd[idx[t]], d2[t]));
        const float lrd_q = (float)(1.0 / (reach / k + 1e-10));
        double ratio = 0.0;
        for (int t = 0; t < k; ++t) ratio += (double)lrd[idx[t]] / (double)lrd_q;
        scores_out[r] = (float)(ratio / k);
        if (idx_out) memcpy(idx_out + (size_t)r * k, idx, sizeof(int) * (size_t)k);
        if (d2_out) memcpy(d2_out + (size_t)r * k, d2, sizeof(float) * (size_t)k);
    }
    free(xc); free(norms); free(mean); free(inv_std); free(fit_scores); free(kd); free(lrd);
    free(qs); free(keys); free(scratch); free(idx); free(d2);
    return 0;
}

/* ---------------------------------------------------------------------------
 * FP64 brute-force LOF: tests/cpp/test_culof.cu:39-80 restated. Distances in
 * double from the uncentred input, ties ordered by (distance, index), the
 * k-distance is the distance to the k-th neighbour in that order.
 * ------------------------------------------------------------------------- */

typedef struct {
    double dist;
    int idx;
} pair_t;

static int pair_less(const pair_t* a, const pair_t* b) {
    return a->dist != b->dist ? a->dist < b->dist : a->idx < b->idx;
}

/* Partition so that the k smallest pairs (by pair_less) occupy v[0..k). */
static void select_pairs(pair_t* v, int n, int k) {
    int lo = 0, hi = n - 1;
    const int target = k - 1;
    while (lo < hi) {
        const pair_t pivot = v[lo + (hi - lo) / 2];
        int i = lo, j = hi;
        while (i <= j) {
            while (pair_less(&v[i], &pivot)) ++i;
            while (pair_less(&pivot, &v[j])) --j;
            if (i <= j) {
                const pair_t t = v[i];
                v[i] = v[j];
                v[j] = t;
                ++i;
                --j;
            }
        }
        if (target <= j) hi = j;
        else if (target >= i) lo = i;
        else break;
    }
}

static int pair_cmp(const void* a, const void* b) {
    const pair_t* x = (const pair_t*)a;
    const pair_t* y = (const pair_t*)b;
    if (pair_less(x, y)) return -1;
    if (pair_less(y, x)) return 1;
    return 0;
}

int oracle_lof_fp64(const float* x, int n, int d, int k, int mutation, double* scores,
                    int* idx_out, double* kdist_out) {
    if (!x || n < 2 || d < 1 || k < 1 || k >= n) return 1;
    pair_t* row = (pair_t*)malloc(sizeof(pair_t) * (size_t)n);
    int* nbr = idx_out ? idx_out : (int*)malloc(sizeof(int) * (size_t)n * k);
    double* nd = (double*)malloc(sizeof(double) * (size_t)n * k);
    double* kdist = kdist_out ? kdist_out : (double*)malloc(sizeof(double) * (size_t)n);
    double* lrd = (double*)malloc(sizeof(double) * (size_t)n);

    for (int i = 0; i < n; ++i) {
        for (int j = 0; j < n; ++j) {
            double s = 0.0;
            for (int t = 0; t < d; ++t) {
                const double diff = (double)x[(size_t)i * d + t] - (double)x[(size_t)j * d + t];
                s += diff * diff;
            }
            row[j].dist = (j == i) ? INFINITY : sqrt(s);
            row[j].idx = j;
        }
        select_pairs(row, n, k);
        qsort(row, (size_t)k, sizeof(pair_t), pair_cmp);
        for (int t = 0; t < k; ++t) {
            nbr[(size_t)i * k + t] = row[t].idx;
            nd[(size_t)i * k + t] = row[t].dist;
        }
        kdist[i] = row[(mutation == 1 && k >= 2) ? k - 2 : k - 1].dist;
    }
    for (int i = 0; i < n; ++i) {
        double s = 0.0;
        for (int t = 0; t < k; ++t) {
            const int j = nbr[(size_t)i * k + t];
            const double dij = nd[(size_t)i * k + t];
            s += (mutation == 2) ? dij : (kdist[j] > dij ? kdist[j] : dij);
        }
        lrd[i] = 1.0 / (s / k + 1e-10);
    }
    for (int i = 0; i < n; ++i) {
        double s = 0.0;
        for (int t = 0; t < k; ++t) {
            const int j = nbr[(size_t)i * k + t];
            s += (mutation == 3) ? lrd[i] / lrd[j] : lrd[j] / lrd[i];
        }
        scores[i] = s / k;
    }
    free(row);
    free(nd);
    free(lrd);
    if (!idx_out) free(nbr);
    if (!kdist_out) free(kdist);
    return 0;
}
