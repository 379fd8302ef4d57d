// Local reachability density and LOF score: the two O(n k) stages.
//
// Term for term the reference's src/cuda/lof.cu:32-62 (and scikit-learn's
// LocalOutlierFactor): reach distances are taken on squared keys and rooted
// once in FP64 (sqrt is monotonic, so max(sqrt a, sqrt b) = sqrt max(a, b)),
// summed in FP64 in slot order, with scikit-learn's 1e-10 added to the mean
// before inverting; the score is the FP64 mean of lrd[j] / lrd[i] in slot
// order. One thread per point; the transposed neighbour layout makes the
// nbr_idx / nbr_d2 reads coalesced, and kdist2[j] / lrd[j] are gathers.

#include "internal.hpp"

namespace culof {
namespace detail {

namespace {

constexpr int kBlock = 256;
constexpr double kLrdEpsilon = 1e-10;  // scikit-learn's guard

__global__ void __launch_bounds__(kBlock)
lrd_kernel(const int* __restrict__ nbr_idx, const float* __restrict__ nbr_d2,
           const float* __restrict__ kdist2, int n, int k, float* __restrict__ lrd) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double sum = 0.0;
    for (int t = 0; t < k; ++t) {
        const size_t o = static_cast<size_t>(t) * n + i;
        const int j = nbr_idx[o];
        sum += sqrt(static_cast<double>(fmaxf(kdist2[j], nbr_d2[o])));
    }
    lrd[i] = static_cast<float>(1.0 / (sum / k + kLrdEpsilon));
}

// lrd: the neighbours' densities; lrd_self: row i's own (the same array for
// the fitted points, the queries' densities when scoring new points).
__global__ void __launch_bounds__(kBlock)
score_kernel(const int* __restrict__ nbr_idx, const float* __restrict__ lrd,
             const float* __restrict__ lrd_self, int n, int k, float* __restrict__ scores) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double self = static_cast<double>(lrd_self[i]);
    double sum = 0.0;
    for (int t = 0; t < k; ++t) {
        sum += static_cast<double>(lrd[nbr_idx[static_cast<size_t>(t) * n + i]]) / self;
    }
    scores[i] = static_cast<float>(sum / k);
}

}  // namespace

void compute_lrd(const int* nbr_idx, const float* nbr_d2, const float* kdist2, int n_points,
                 int k, float* lrd, hipStream_t stream) {
    ProfileScope prof("lrd", stream);
    lrd_kernel<<<(n_points + kBlock - 1) / kBlock, kBlock, 0, stream>>>(nbr_idx, nbr_d2, kdist2,
                                                                          n_points, k, lrd);
    CULOF_CHECK(hipGetLastError());
}

void compute_scores(const int* nbr_idx, const float* lrd, int n_points, int k, float* scores,
                    hipStream_t stream) {
    ProfileScope prof("score", stream);
    score_kernel<<<(n_points + kBlock - 1) / kBlock, kBlock, 0, stream>>>(nbr_idx, lrd, lrd, n_points,
                                                                            k, scores);
    CULOF_CHECK(hipGetLastError());
}

void compute_query_scores(const int* nbr_idx, const float* lrd_fit, const float* lrd_query,
                          int n_queries, int k, float* scores, hipStream_t stream) {
    ProfileScope prof("score", stream);
    if (n_queries <= 0) return;
    score_kernel<<<(n_queries + kBlock - 1) / kBlock, kBlock, 0, stream>>>(nbr_idx, lrd_fit, lrd_query,
                                                                             n_queries, k, scores);
    CULOF_CHECK(hipGetLastError());
}

}  // namespace detail
}  // namespace culof
