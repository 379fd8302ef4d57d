// Internal stage interfaces of libculof_amd (HIP, gfx950).
//
// Counterpart of the reference's include/culof_internal.cuh:31-64, taking a
// hipStream_t. Every launcher checks its own launch and throws CudaError on a
// HIP failure; the C ABI layer (api.cpp) turns exceptions into status codes.

#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <functional>
#include <string>

#include "culof.h"

namespace culof {
namespace detail {

void check_hip(hipError_t err, const char* file, int line, const char* expr);

#define CULOF_CHECK(expr) ::culof::detail::check_hip((expr), __FILE__, __LINE__, #expr)

// Bit pattern of +inf: the key of a point against itself and of a NaN expansion.
constexpr uint32_t kInfBits = 0x7F800000u;

constexpr int kPrepWork = 256 + 3;  // doubles per feature of center_and_norms' work buffer
constexpr int kMaxBackwardDims = 256;

// Appends v to a device queue of `cap` entries. Every producer appends a row
// at most once per queue, and queues are sized for every row, so the cap is
// never reached; an append past it is dropped rather than written out of
// bounds, and consumers read min(*count, cap) entries.
__device__ __forceinline__ void queue_push(int* list, int* count, int cap, int v) {
    const int at = atomicAdd(count, 1);
    if (at < cap) list[at] = v;
}

// Zeroes `bytes` (a multiple of 4) of device memory on `stream` with a kernel:
// unlike hipMemsetAsync it records a plain kernel node under graph capture.
void zero_fill(void* dst, size_t bytes, hipStream_t stream);

// Precision forms of the distance screen (culof_form in culof_c.h).
constexpr int kFormPlain = 0;  // FP32 screen
constexpr int kFormBF16 = 1;   // BF16 MFMA screen
constexpr int kFormMask = 0xFF;
constexpr int kFormTiesFp64 = 0x100;  // flag: settle near-tie rows in FP64

// Screen form of a form word, or -1 when the word is not a valid one.
inline int screen_form(int form) {
    if (form & ~(kFormMask | kFormTiesFp64)) return -1;
    const int screen = form & kFormMask;
    return (screen == kFormPlain || screen == kFormBF16) ? screen : -1;
}

// Feature count padded for the fused sweep: the centred rows are stored with
// this stride, zero-filled, so a row loads as whole float4s.
int padded_dims(int n_dims);

// ---- kernel timing hooks (see culof_profile_* in culof_c.h) ----
// The device's tile-pair work counter while timing is on, else null.
unsigned long long* profile_tile_counter();
unsigned long long* profile_tile_counter_for(int device, bool create);
struct ProfileScope {
    ProfileScope(const char* name, hipStream_t stream);
    ~ProfileScope();
    ProfileScope(const ProfileScope&) = delete;
    ProfileScope& operator=(const ProfileScope&) = delete;

private:
    int slot_ = -1;
    hipStream_t stream_ = nullptr;
};

// ---- prep.hip: FP64 centring and norms (reference src/cuda/preprocess.cu) ----
// Writes `centred` with row stride `stride` (>= n_dims, padding zeroed) and
// FP64-accumulated squared norms. `work` must hold kPrepWork * n_dims doubles:
// partials (256 per feature), then mean, inv_std and scaled (1.0 / 0.0) per feature.
// Optional: `copy` receives the input rows at the same stride (zero padding);
// *guard (zeroed by the caller) is set when some norm exceeds FLT_MAX / 16.
void center_and_norms(const float* points, int n_points, int n_dims, bool scale, float* centred,
                      int stride, float* norms, double* work, float* copy, int* guard,
                      hipStream_t stream);
// The same transform, with the mean and inv_std center_and_norms left in
// `work`, applied to other rows (new points scored against a fitted set).
void apply_rows(const float* rows, int n_rows, int n_dims, const double* work, float* out,
                int stride, float* norms, float* copy, hipStream_t stream);

// Selection keys are difference-form FP32 squared distances
// e(i, j) = sum_t (a_t - b_t)^2 over the key rows (the input rows, or the
// z-scored rows when normalising), an FMA chain in ascending t; +inf on the
// diagonal and for NaN. Both kNN paths below select exactly by (e, index).

// ---- sweep.hip: fused Gram screen + exact top-k per query row ----
// Returns false when (n_dims, k) has no sweep instantiation; the caller then
// uses the radix path for every row.
bool sweep_supported(int n_dims, int k);
// Queries [row0, row0 + rows). `keyrows` and `centred` share the padded
// stride; `norms` are of the centred rows; *guard set means every row goes to
// the fallback list. Rows whose k-th key is +inf are appended to
// fallback_rows (fallback_cap entries) with fallback_count incremented; the
// radix path redoes them.
void sweep_knn(const float* keyrows, const float* centred, int stride, const float* norms,
               const int* guard, int n_points, int n_dims, int k, int row0, int rows,
               int* nbr_idx, float* nbr_d2, float* kdist2, int* fallback_rows,
               int* fallback_count, int fallback_cap, hipStream_t stream);
// The same sweep for rows of another set (qkeys / qcentred / qnorms, same
// stride and transform) against all n_points rows, nothing excluded; outputs
// with stride n_queries; rows the sweep cannot settle go to fallback_rows.
void sweep_query_knn(const float* qkeys, const float* qcentred, const float* qnorms, int n_queries,
                     const float* keyrows, const float* centred, int stride, const float* norms,
                     const int* guard, int n_points, int n_dims, int k, int* nbr_idx, float* nbr_d2,
                     float* kdist2, int* fallback_rows, int* fallback_count, int fallback_cap,
                     hipStream_t stream);

// ---- mfma.hip: matrix-core screened sweep (FP32 or BF16 screen) ----
// Row stride the form's screen expects (BF16 needs multiples of 16).
int mfma_padded_dims(int n_dims, int form);
// Largest k of the list kernels (the in-register list sweeps, the seeded list
// pass); the collect pipeline (candidate order, seeds, collect, select) keeps
// its lists in LDS and serves k up to kCollectKMax.
constexpr int kListKMax = 64;
constexpr int kCollectKMax = 256;
// (n_dims, k, form) has a matrix-core collect pipeline instantiation.
bool mfma_supported(int n_dims, int k, int form);
// ... and a matrix-core list sweep (runs without the candidate order).
bool mfma_list_supported(int n_dims, int k, int form);
int mfma_slices(int n_points, int rows);
// Device scratch the launch needs (screen operand copy, adjusted norms,
// per-slice partial lists).
size_t mfma_scratch_bytes(int n_points, int dp, int form, int rows, int k, bool collect);
struct CandidateOrder;
// ord (optional): stream candidates in that order; qlist (optional): the
// original indices of the `rows` queries (else row0 + local); tseed
// (optional, indexed by original query index): an upper bound of each
// query's k-th key from seed_thresholds. fallback_rows holds fallback_cap
// entries (every row of the call).
void mfma_knn(const float* keyrows, const float* centred, int stride, const float* norms,
              const int* guard, int n_points, int n_dims, int k, int form, int row0, int rows,
              const CandidateOrder* ord, const int* qlist, const float* tseed,
              void* scratch, int* nbr_idx, float* nbr_d2, float* kdist2, int* fallback_rows,
              int* fallback_count, int fallback_cap, hipStream_t stream);

// ---- radix.hip: per-row exact radix select (reference src/cuda/knn.cu) ----
// Recompute mode: keys rebuilt from `keyrows` (row stride `stride`) every
// pass. With `row_count` (device) null the query rows are [row0, row0 + rows);
// otherwise they are row_list[0 .. min(*row_count, rows)) (`rows` is then the
// list's capacity) and the launch is a fixed-size grid-stride loop, so no host
// read-back is needed.
void radix_knn(const float* keyrows, int stride, int n_points, int n_dims, int k, int row0,
               int rows, const int* row_list, const int* row_count, int* nbr_idx, float* nbr_d2,
               float* kdist2, hipStream_t stream);
// Query mode: rows of `qrows` (another row set, same stride) against all
// n_points key rows, nothing excluded; outputs with stride n_queries. All
// queries, or row_list[0 .. min(*row_count, n_queries)) (device) when
// row_count is given.
void radix_query_knn(const float* qrows, const float* keyrows, int stride, int n_points, int n_dims,
                     int k, int n_queries, const int* row_list, const int* row_count, int* nbr_idx,
                     float* nbr_d2, float* kdist2, hipStream_t stream);
// Stored keys for listed rows: exact keys of row_list[0 .. min(*row_count,
// cap)) against all n_points key rows, `batch` rows at a time into `keys`
// (batch x n_points floats, brute_batch_rows), each batch then selected by
// the stored radix select; outputs at the rows' original indices. The rows
// many collected candidates would flood (heavy rows, overflowed buffers) take
// it where no list kernel serves the shape.
constexpr size_t kBruteKeyBytes = 512ull << 20;  // key buffer of one batch, at most
int brute_batch_rows(int n_points, int cap);
void brute_knn_rows(const float* keyrows, int stride, int n_points, int k, const int* row_list,
                    const int* row_count, int cap, float* keys, int batch, int* nbr_idx, float* nbr_d2,
                    float* kdist2, hipStream_t stream);
// Stored mode: keys row-major (rows x n_cols); outputs with stride `rows`.
void radix_select_rows(const float* keys, int rows, int n_cols, int k, int* nbr_idx,
                       float* nbr_d2, float* kdist2, hipStream_t stream);

// ---- order.hip: candidate order and per-query seeds (performance only) ----
struct CandidateOrder {
    int* order = nullptr;     // position -> original index
    int* inv = nullptr;       // original index -> position
    float* xs_ord = nullptr;  // key rows in candidate order (stride dp)
    int dp = 0;
    const float* xs = nullptr;  // the key rows the order was built from (natural order)
    // work space left free once the order is built (n, n, 4096 and 64 ints)
    int* tmp_order = nullptr;
    int* tmp_counts = nullptr;
    int* tmp_hist = nullptr;
    int* tmp_small = nullptr;
};
bool order_supported(int dp);
size_t order_scratch_bytes(int n, int dp);
CandidateOrder build_order(const float* xs, int n, int dp, void* scratch, hipStream_t s);
// tseed[i] for every point i = an upper bound of its k-th key (the k-th among
// a window of candidates near it in the order). key_dp: the padded feature
// count of the keys (<= o.dp; the rest of each row is zero padding).
void seed_thresholds(const CandidateOrder& o, int n, int key_dp, int k, float* tseed, hipStream_t s);
// Moves the rows whose seeds are the largest (at most n / 100, by a histogram
// of the seeds) to the end of the order, each part keeping its order, and
// rebuilds inv and xs_ord. Changes no result: it tightens the tile bounds the
// symmetric pass prunes with.
void order_large_seeds_last(CandidateOrder& o, int n, const float* tseed, hipStream_t s);

// ---- ties.hip: FP64 tie-break over a (k+1)-neighbour result ----
// Relative band (1 + g') / (1 - g') of FP32 keys that FP64 may reorder.
double tie_band_ratio(int d);
// Near-tie rows settled through the matrix-core collect pass (mfma.hip): a
// queue of rows with their band bounds and the buffers of that pass.
struct TieBand {
    const float* centred;  // centred rows (stride = the key rows' stride)
    const float* norms;
    const int* guard;
    int form;              // screen form: kFormPlain or kFormBF16
    int cap;               // rows the queue holds (tie_band_capacity)
    int* rows;             // queued original row indices, cap ints
    int* count;            // queue length; rows past cap went to the brute-force queue
    double* bound;         // per queue slot: the band bound K rho
    float* tband;          // per original row (n floats): smallest float >= its bound
    void* scratch;         // tie_band_scratch_bytes
};
// Rows the band queue holds for a call over `rows` rows; 0 when the collect
// pass does not serve (n_dims, stride, form).
int tie_band_capacity(int n_points, int n_dims, int stride, int form, int rows);
size_t tie_band_scratch_bytes(int n_points, int n_dims, int form, int cap);
// Collect pass over the queued rows with their bounds as thresholds, then the
// FP64 band select; rows it cannot hold are appended to brute_rows (brute_cap
// entries).
void tie_band_select(const float* keyrows, int stride, int n_points, int k, const TieBand& band,
                     int* nbr_idx, float* nbr_d2, float* kdist2, int* brute_rows, int* brute_count,
                     int brute_cap, hipStream_t stream);

// idx1 / d21: the (k+1)-lists of rows [row0, row0 + rows) (stride n). Writes
// the k-lists of those rows; tie_rows needs `rows` ints, tie_count one int.
// With `band`, near-tie rows go through tie_band_select first and only the
// rows it cannot hold are scanned by brute force.
void tie_break_rows(const float* keyrows, int stride, int n, int d, int k, int row0, int rows,
                    const int* idx1, const float* d21, int* nbr_idx, float* nbr_d2, float* kdist2,
                    int* tie_rows, int* tie_count, const TieBand* band, hipStream_t s);

// ---- backward.hip: d(sum w LOF)/dX with kNN sets and max() branches fixed ----
size_t backward_scratch_bytes(int n, int d, int k);
// keyrows: n x d (stride d) rows the forward keyed on. inv_std / scaled (per
// feature, from center_and_norms' work buffer) are given when the forward
// normalised, null otherwise.
void lof_backward(const float* keyrows, int n, int d, int k, const int* nbr_idx, const float* nbr_d2,
                  const float* kdist2, const float* grad_scores, const double* inv_std,
                  const double* scaled, float* grad_points, void* scratch, hipStream_t s);

// ---- density.hip: reference src/cuda/lof.cu ----
void compute_lrd(const int* nbr_idx, const float* nbr_d2, const float* kdist2, int n_points,
                 int k, float* lrd, hipStream_t stream);
void compute_scores(const int* nbr_idx, const float* lrd, int n_points, int k, float* scores,
                    hipStream_t stream);
// LOF of new points: mean of lrd_fit over each query's fitted neighbours
// divided by the query's own lrd_query (scikit-learn's novelty score_samples).
void compute_query_scores(const int* nbr_idx, const float* lrd_fit, const float* lrd_query,
                          int n_queries, int k, float* scores, hipStream_t stream);

// Same arithmetic as the reference's choose_tile_rows (src/cuda/pipeline.cu:33-38).
int choose_tile_rows(int n_points, size_t budget_bytes);

// ---- api.hip: pieces the no-allocation entry point (noalloc.hip) shares ----
// Runs f, turning its exceptions into a culof_status and culof_last_error().
int run_guarded(const std::function<void()>& f);
// culof_lof_device with every scratch buffer taken from [region, region +
// bytes) and no allocator called; a call needing more is rejected before
// anything is launched.
void forward_in_region(const float* points, int n, int d, int k, bool normalize, int form,
                       float* scores, int* nbr_idx, float* nbr_d2, float* kdist2, float* lrd,
                       void* region, size_t bytes, hipStream_t s);

}  // namespace detail
}  // namespace culof
