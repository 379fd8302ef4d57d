/*
 * culof C ABI: the FFI boundary of libculof_amd.so.
 *
 * Plain pointers, sizes and status codes; no C++ types, no torch types. Every
 * entry point returns CULOF_OK (0) or a status code, and on failure leaves a
 * message for culof_last_error() (per calling thread). Nothing here throws.
 *
 * Host-pointer entry points are synchronous. Device-pointer entry points
 * (suffix _device and the stage functions) enqueue on `stream` (a hipStream_t,
 * NULL for the default stream) and return without synchronising, except where
 * a comment says otherwise.
 *
 * Layout conventions shared by every entry point:
 *   points   row-major (n_points x n_dims) float32
 *   nbr_idx  int32, transposed: neighbour slot t of point i at [t * n_points + i]
 *   nbr_d2   float32, same layout, squared distance of that neighbour
 *   kdist2   float32 (n_points), squared k-distance
 * Slot order inside a neighbour list is the reference's: every neighbour
 * strictly closer than the k-th key in ascending index, then the neighbours
 * tied with the k-th key in ascending index (src/cuda/knn.cu:190-224). Slot
 * k-1 therefore always holds the k-th neighbour in (distance, index) order.
 */

#ifndef CULOF_C_H
#define CULOF_C_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

enum culof_status {
    CULOF_OK = 0,
    CULOF_INVALID_ARGUMENT = 1, /* std::invalid_argument in the C++ API */
    CULOF_DEVICE_ERROR = 2,     /* CudaError in the C++ API */
    CULOF_NO_DEVICE = 3         /* CudaError("no ... device available") */
};

/* Precision forms of the distance screen. Selection is always by the exact
 * FP32 difference-form keys, so both forms return the same neighbours, the
 * same k-distances and the same scores; they differ in speed only.
 *
 * CULOF_FORM_FP64_TIES may be OR-ed into either form: rows whose k-th and
 * (k+1)-th FP32 keys are closer than FP32 resolution ((d + 4) 2^-24 relative,
 * exact ties included) are re-selected on FP64 distances with the
 * (distance, index) order of the reference's FP64 test oracle
 * (tests/cpp/test_culof.cu:55-66); other rows are unchanged. This changes
 * results on such rows by design: their neighbour sets become the FP64 ones.
 * No reference counterpart. */
enum culof_form {
    CULOF_FORM_PLAIN = 0,        /* FP32 Gram screen (MFMA or VALU) */
    CULOF_FORM_BF16_SCREEN = 1,  /* BF16 MFMA screen, wider error margin */
    CULOF_FORM_FP64_TIES = 256   /* flag: FP64 tie-break of near-tie rows */
};

/* Message of the last failed call on this thread ("" if none). */
const char* culof_last_error(void);

/* 1 if a usable GPU is present, else 0.
 * Replaces culof::cuda_available (reference include/culof.h:28,
 * src/cuda/pipeline.cu:115-118). */
int culof_device_available(void);

/* Writes a NUL-terminated description of the current device into buf.
 * Replaces culof::device_info (reference include/culof.h:31,
 * src/cuda/pipeline.cu:120-129). */
int culof_device_info(char* buf, size_t capacity);

/* LOF scores of host data, written to host `scores` (n_points floats).
 * Replaces culof::lof (reference include/culof.h:46-47,
 * src/cuda/pipeline.cu:131-207); same argument checks and semantics. */
int culof_lof(const float* points, int n_points, int n_dims, int k, int normalize,
              float* scores);

/* Device-resident forward. `points` and every output are device pointers.
 * Optional outputs (nbr_idx, nbr_d2, kdist2, lrd) may be NULL. The reference
 * has no device-pointer entry point; this is culof::lof without the two host
 * copies (src/cuda/pipeline.cu:159-160,203-205), for callers whose data already
 * lives in HBM (the torch binding). Never synchronises. */
int culof_lof_device(const float* points, int n_points, int n_dims, int k, int normalize,
                     int form, float* scores, int* nbr_idx, float* nbr_d2, float* kdist2,
                     float* lrd, void* stream);

/* Bytes of device scratch culof_lof_device_noalloc holds per device. */
#define CULOF_NOALLOC_SCRATCH_BYTES (256u * 1024u * 1024u)

/* culof_lof_device without any device allocation: same arguments, checks and
 * results (bit-identical), with every scratch buffer taken from a fixed
 * array of CULOF_NOALLOC_SCRATCH_BYTES that each device holds from the first
 * call on it until the process ends. No counterpart in the reference; it is
 * for callers that cannot allocate (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc). It calls no hipMalloc,
 * hipMallocAsync or hipFree. A call whose scratch would exceed the array
 * returns CULOF_INVALID_ARGUMENT before launching anything: at most
 * 4 (3 p + 4 k + 16) bytes per point with p the padded n_dims (4, 8, 16, 32,
 * 64, 128, or n_dims rounded up to 4), plus about 2 MiB, e.g. 550k points at
 * n_dims = 8, k = 20. Every call on a device shares that device's array: calls that overlap in time on one device
 * (on different streams or threads) race and must be serialised by the
 * caller. It takes the list sweep rather than the collect pipeline, so it is
 * slower than culof_lof_device on large inputs. Never synchronises. */
int culof_lof_device_noalloc(const float* points, int n_points, int n_dims, int k, int normalize,
                             int form, float* scores, int* nbr_idx, float* nbr_d2, float* kdist2,
                             float* lrd, void* stream);

/* New points scored against a fitted set. No counterpart in the reference
 * (its estimator has no novelty mode, README.md:29-34); this is scikit-learn's
 * LocalOutlierFactor(novelty=True).fit(points).score_samples(queries) with the
 * sign of culof_lof: scores_new[r] (n_queries floats) is the LOF of query row r
 * against the fitted rows -- its k nearest fitted rows (none excluded; exact
 * FP32 keys, the reference's tie rule), their k-distances and densities from
 * the forward on `points`. scores_fit (optional, n_points floats) receives
 * the fitted rows' own LOF; nbr_idx_new / nbr_d2_new (optional, k x n_queries,
 * slot t of query r at [t * n_queries + r]) receive each query's fitted
 * neighbours and squared-distance keys in the reference's slot order. All
 * pointers are device pointers; `form` selects the fitted forward's screen and
 * tie rule (the queries use the index rule). For callers that score held-out
 * data. Never synchronises. */
int culof_lof_novelty_device(const float* points, int n_points, int n_dims, const float* queries,
                             int n_queries, int k, int normalize, int form, float* scores_fit,
                             float* scores_new, int* nbr_idx_new, float* nbr_d2_new, void* stream);

/* Backward of the device-resident forward: grad_points (n_points x n_dims,
 * device) = d(sum_i grad_scores[i] * LOF_i) / d points, with the kNN sets and
 * the reach-distance max() branches fixed from a forward call on the same
 * points, k and normalize (pass that call's nbr_idx, nbr_d2 and kdist2).
 * Distances are recomputed from the points; sums run in FP64 in a fixed order
 * (reverse neighbour lists, no floating-point atomics), so repeated calls are
 * bit-identical. A zero distance contributes a zero subgradient. n_dims <= 256.
 * No reference counterpart: the reference is forward-only. */
int culof_lof_backward_device(const float* points, int n_points, int n_dims, int k,
                              int normalize, const int* nbr_idx, const float* nbr_d2,
                              const float* kdist2, const float* grad_scores, float* grad_points,
                              void* stream);

/* ---- stage entry points (device pointers), for tests and integrators ----
 * They replace the internal stage interface of the reference
 * (include/culof_internal.cuh:31-64). */

/* Centre (and with `normalize` z-score) each feature with FP64 accumulation
 * and write the centred float32 rows (n_points x n_dims) plus the
 * FP64-accumulated squared norms of those rows.
 * Replaces detail::standardize + detail::squared_norms
 * (include/culof_internal.cuh:31,34; src/cuda/preprocess.cu:31-98). */
int culof_stage_center(const float* points, int n_points, int n_dims, int normalize,
                       float* centred, float* norms, void* stream);

/* Exact kNN of query rows [row0, row0 + rows) over all n_points candidates of
 * `points` (centred, and z-scored with `normalize`, internally). Keys are
 * difference-form FP32 squared distances sum_t (a_t - b_t)^2 of the input rows
 * (of the z-scored rows with `normalize`), +inf on the diagonal; ties go to
 * the lower index. method 0 = the production path for `form` (a screened
 * sweep on the matrix cores where instantiated), 1 = per-row radix select
 * recomputing keys every pass, 2 = the VALU-screened sweep (form 0 only); all
 * return the same bits. Outputs use stride n_points (slot t of point i at
 * [t * n_points + i]). Replaces the SGEMM + detail::select_knn tile step
 * (src/cuda/pipeline.cu:175-195, include/culof_internal.cuh:52). */
int culof_stage_knn(const float* points, int n_points, int n_dims, int k, int normalize,
                    int form, int row0, int rows, int method, int* nbr_idx, float* nbr_d2,
                    float* kdist2, void* stream);

/* Exact radix select over stored key rows: keys is row-major (rows x n_cols)
 * float32 (non-negative or +inf; NaN is read as +inf). Outputs use stride
 * `rows` (slot t of row r at [t * rows + r]). This is the reference's selection
 * rule applied to a materialised tile (src/cuda/knn.cu:119-232). */
int culof_stage_select_rows(const float* keys, int rows, int n_cols, int k, int* nbr_idx,
                            float* nbr_d2, float* kdist2, void* stream);

/* lrd[i] = float(1 / (mean_t sqrt(double(max(kdist2[j], nbr_d2))) + 1e-10)).
 * Replaces detail::compute_lrd (src/cuda/lof.cu:32-46,66-71). */
int culof_stage_lrd(const int* nbr_idx, const float* nbr_d2, const float* kdist2, int n_points,
                    int k, float* lrd, void* stream);

/* scores[i] = float(sum_t double(lrd[j]) / double(lrd[i]) / k).
 * Replaces detail::compute_scores (src/cuda/lof.cu:49-62,73-78). */
int culof_stage_scores(const int* nbr_idx, const float* lrd, int n_points, int k, float* scores,
                       void* stream);

/* Rows per tile that keep n_points * 4 bytes per row within budget_bytes,
 * clamped to [1, n_points]. Same arithmetic as detail::choose_tile_rows
 * (src/cuda/pipeline.cu:33-38). */
int culof_choose_tile_rows(int n_points, size_t budget_bytes);

/* ---- kernel timing (no reference counterpart; used by the benchmark) ----
 * When enabled, every launch of the library's kernels is bracketed by HIP
 * events on the stream it runs on. culof_profile_read synchronises those
 * events and returns the summed duration and launch count of the kernels
 * whose name contains `name`. */
int culof_profile_enable(int on);
int culof_profile_reset(void);
int culof_profile_read(const char* name, double* total_ms, long* launches);
/* Work counter beside the timers: pairs of 32-row tiles (a query tile and a
 * candidate tile, 32 x 32 row pairs) the symmetric collect passes screened on
 * the current device while timing was enabled, since the last reset (the
 * pruned pass skips the rest). Synchronises the device. */
int culof_profile_tile_pairs(long long* tile_pairs);

#ifdef __cplusplus
}
#endif

#endif /* CULOF_C_H */
