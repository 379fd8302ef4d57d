// FP64 tie-break for rows whose k / (k+1) boundary is below FP32 resolution.
//
// The FP32 key e32 of a pair (difference form, see sweep.hip) differs from the
// exact squared distance by at most a relative g = (d + 2) 2^-24. With K the
// k-th key of a row and K' its (k+1)-th, the FP32 neighbour set can differ from
// the one an FP64 computation would choose only if K' <= K (1 + g) / (1 - g)
// (an exact tie K' == K included): then some pair straddling the boundary
// might be ordered the other way by its true distance. Such a row is settled
// again in FP64: every candidate whose e32 lies at or below K rho
// (rho = (1 + g') / (1 - g'), g' = (d + 4) 2^-24 leaving room for the FP64
// rounding) is re-keyed with e64 = sum_t (a_t - b_t)^2 accumulated in FP64
// with separate multiply and add in ascending feature order, and the k
// smallest (e64, index) pairs form the list: the (distance, index) order of
// the reference's FP64 test oracle (tests/cpp/test_culof.cu:55-66). No pair
// outside the band can beat the k pairs of the FP32 set, so the band holds
// the whole FP64 set.
//
// The input is a kNN computed with k + 1 neighbours (slot order of the
// reference). Rows that are not near-ties keep the FP32 result: their k-list
// is the (k+1)-list minus its largest word, re-ordered into the reference's
// slot order. Near-tie rows are re-selected over the band: slots follow the
// same rule on FP64 keys (pairs below the k-th FP64 key in ascending index,
// then the tied ones in ascending index); nbr_d2 and kdist2 keep the FP32
// keys of the chosen pairs, kdist2 that of the k-th pair in FP64 order.

#include "internal.hpp"

#include <stdexcept>

namespace culof {
namespace detail {

namespace {

constexpr int kBlock = 256;
constexpr int kCap = 2048;          // band entries held in LDS between compactions
constexpr int kRound = 4 * kBlock;  // candidates examined per round
constexpr int kMaxTieStride = 1024; // query row staged in LDS (floats)

__device__ __forceinline__ uint64_t word_of(float key, int idx) {
    const uint32_t kb = __float_as_uint(key);
    return (static_cast<uint64_t>(kb > kInfBits ? kInfBits : kb) << 32) | static_cast<uint32_t>(idx);
}

// Smallest float >= b (b >= 0; +inf past FLT_MAX).
__device__ __forceinline__ float float_at_least(double b) {
    const float f = static_cast<float>(b);
    return static_cast<double>(f) < b ? __uint_as_float(__float_as_uint(f) + 1u) : f;
}

// Split the (k+1)-lists: emit rows that are not near-ties, queue the rest
// (to the band queue while it has room, else to the brute-force queue).
__global__ void __launch_bounds__(kBlock)
tie_split_kernel(const int* __restrict__ idx1, const float* __restrict__ d21, int n, int row0,
                 int rows, int k, double rho, int* __restrict__ nbr_idx,
                 float* __restrict__ nbr_d2, float* __restrict__ kdist2, int* __restrict__ tie_rows,
                 int* __restrict__ tie_count, TieBand band) {
    const int r = blockIdx.x * kBlock + threadIdx.x;
    if (r >= rows) return;
    const int i = row0 + r;
    const size_t ns = static_cast<size_t>(n);
    uint64_t wmax = 0;
    int tmax = 0;
    for (int t = 0; t <= k; ++t) {
        const uint64_t w = word_of(d21[t * ns + i], idx1[t * ns + i]);
        if (w > wmax || t == 0) {
            wmax = w;
            tmax = t;
        }
    }
    uint64_t wk = 0;
    for (int t = 0; t <= k; ++t) {
        if (t == tmax) continue;
        const uint64_t w = word_of(d21[t * ns + i], idx1[t * ns + i]);
        wk = w > wk ? w : wk;
    }
    const float K = __uint_as_float(static_cast<uint32_t>(wk >> 32));
    const float K1 = __uint_as_float(static_cast<uint32_t>(wmax >> 32));
    const double bound = static_cast<double>(K) * rho;
    if (isfinite(K) && static_cast<double>(K1) <= bound) {
        const int at = band.rows ? atomicAdd(band.count, 1) : band.cap;
        if (at < band.cap) {
            band.rows[at] = i;
            band.bound[at] = bound;
            band.tband[i] = float_at_least(bound);
        } else {
            queue_push(tie_rows, tie_count, rows, i);
        }
        return;
    }
    // Keys below K keep their ascending-index order; keys equal to K follow.
    int slot = 0;
    for (int pass = 0; pass < 2; ++pass) {
        for (int t = 0; t <= k; ++t) {
            if (t == tmax) continue;
            const float key = d21[t * ns + i];
            const bool below = key < K;
            if (below == (pass == 0)) {
                nbr_idx[slot * ns + i] = idx1[t * ns + i];
                nbr_d2[slot * ns + i] = key;
                ++slot;
            }
        }
    }
    kdist2[i] = K;
}

struct Band {
    uint64_t e64[kCap];   // FP64 key bits (non-negative doubles order as integers)
    int idx[kCap];
    float e32[kCap];
    int count;
};

__device__ __forceinline__ bool band_less(const Band& b, int x, int y) {
    return b.e64[x] != b.e64[y] ? b.e64[x] < b.e64[y] : b.idx[x] < b.idx[y];
}

// Bitonic sort of the whole buffer (entries past `count` hold +inf sentinels).
__device__ void band_sort(Band& b) {
    for (int size = 2; size <= kCap; size <<= 1) {
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            for (int p = threadIdx.x; p < kCap / 2; p += kBlock) {
                const int lo = 2 * p - (p & (stride - 1));
                const int hi = lo + stride;
                const bool up = (lo & size) == 0;
                if (band_less(b, hi, lo) == up) {
                    const uint64_t e = b.e64[lo];
                    b.e64[lo] = b.e64[hi];
                    b.e64[hi] = e;
                    const int x = b.idx[lo];
                    b.idx[lo] = b.idx[hi];
                    b.idx[hi] = x;
                    const float f = b.e32[lo];
                    b.e32[lo] = b.e32[hi];
                    b.e32[hi] = f;
                }
            }
            __syncthreads();
        }
    }
}

__device__ void band_clear_from(Band& b, int from) {
    for (int p = from + static_cast<int>(threadIdx.x); p < kCap; p += kBlock) {
        b.e64[p] = ~0ull;
        b.idx[p] = 0x7FFFFFFF;
        b.e32[p] = __uint_as_float(kInfBits);
    }
}

// One block per queued row (grid-stride over the queue): FP32 keys of every
// candidate, FP64 keys of those in the band, k smallest by (e64, index).
__global__ void __launch_bounds__(kBlock)
tie_resolve_kernel(const float* __restrict__ xs, int stride, int n, int d, int k, double rho,
                   const int* __restrict__ idx1, const float* __restrict__ d21,
                   const int* __restrict__ tie_rows, const int* __restrict__ tie_count, int tie_cap,
                   int* __restrict__ nbr_idx, float* __restrict__ nbr_d2,
                   float* __restrict__ kdist2) {
    __shared__ Band b;
    __shared__ double s_bound;
    __shared__ __attribute__((aligned(16))) float s_q[kMaxTieStride];
    const size_t ns = static_cast<size_t>(n);
    const int queued = min(*tie_count, tie_cap);
    for (int q = blockIdx.x; q < queued; q += gridDim.x) {
        const int i = tie_rows[q];
        if (threadIdx.x == 0) {
            // k-th smallest word of the (k+1)-list = the largest but one
            uint64_t w1 = 0, w2 = 0;
            for (int t = 0; t <= k; ++t) {
                const uint64_t w = word_of(d21[t * ns + i], idx1[t * ns + i]);
                if (w > w1) {
                    w2 = w1;
                    w1 = w;
                } else if (w > w2) {
                    w2 = w;
                }
            }
            s_bound = static_cast<double>(__uint_as_float(static_cast<uint32_t>(w2 >> 32))) * rho;
            b.count = 0;
        }
        band_clear_from(b, 0);
        __syncthreads();
        const double bound = s_bound;
        const float* xi = xs + static_cast<size_t>(i) * stride;
        for (int t = threadIdx.x; t < stride; t += kBlock) s_q[t] = xi[t];
        __syncthreads();
        const int d4 = (d + 3) / 4;  // rows are zero-padded to stride (a multiple of 4)
        for (int base = 0; base < n; base += kRound) {
#pragma unroll 1
            for (int u = 0; u < kRound / kBlock; ++u) {
                const int j = base + u * kBlock + static_cast<int>(threadIdx.x);
                if (j >= n || j == i) continue;
                const float* xj = xs + static_cast<size_t>(j) * stride;
                // FP32 key, features in ascending order (padding adds exact zeros)
                float e = 0.0f;
                const float4* q4 = reinterpret_cast<const float4*>(s_q);
                const float4* c4 = reinterpret_cast<const float4*>(xj);
                for (int v = 0; v < d4; ++v) {
                    const float4 a = q4[v];
                    const float4 b = c4[v];
                    float df = a.x - b.x;
                    e = __builtin_fmaf(df, df, e);
                    df = a.y - b.y;
                    e = __builtin_fmaf(df, df, e);
                    df = a.z - b.z;
                    e = __builtin_fmaf(df, df, e);
                    df = a.w - b.w;
                    e = __builtin_fmaf(df, df, e);
                }
                if (!(static_cast<double>(e) <= bound)) continue;  // NaN stays out
                double s = 0.0;
                for (int t = 0; t < d; ++t) {
                    const double df = __dsub_rn(static_cast<double>(xi[t]), static_cast<double>(xj[t]));
                    s = __dadd_rn(s, __dmul_rn(df, df));
                }
                const int slot = atomicAdd(&b.count, 1);
                b.e64[slot] = static_cast<uint64_t>(__double_as_longlong(s));
                b.idx[slot] = j;
                b.e32[slot] = e;
            }
            __syncthreads();
            if (b.count > kCap - kRound) {  // keep the k best, make room
                band_sort(b);
                const int keep = b.count < k ? b.count : k;
                __syncthreads();
                band_clear_from(b, keep);
                if (threadIdx.x == 0) b.count = keep;
                __syncthreads();
            }
        }
        band_sort(b);
        // Slot order on FP64 keys: below the k-th in ascending index, then the ties.
        const uint64_t kth = b.e64[k - 1];
        if (static_cast<int>(threadIdx.x) < k) {
            const int t = threadIdx.x;
            int slot = t;
            if (b.e64[t] < kth) {
                slot = 0;
                for (int f = 0; f < k; ++f) slot += (b.e64[f] < kth && b.idx[f] < b.idx[t]);
            }
            nbr_idx[slot * ns + i] = b.idx[t];
            nbr_d2[slot * ns + i] = b.e32[t];
            if (t == k - 1) kdist2[i] = b.e32[t];
        }
        __syncthreads();
    }
}

}  // namespace

double tie_band_ratio(int d) {
    const double g = (d + 4) * 5.9604644775390625e-8;  // (d + 4) 2^-24
    return (1.0 + g) / (1.0 - g);
}

void tie_break_rows(const float* keyrows, int stride, int n, int d, int k, int row0, int rows,
                    const int* idx1, const float* d21, int* nbr_idx, float* nbr_d2, float* kdist2,
                    int* tie_rows, int* tie_count, const TieBand* band, hipStream_t s) {
    if (rows <= 0) return;
    if (k > kBlock) throw std::invalid_argument("FP64 tie-break supports k <= 256");
    if (stride > kMaxTieStride || stride % 4 != 0)
        throw std::invalid_argument("FP64 tie-break supports n_dims <= 1024");
    const double rho = tie_band_ratio(d);
    ProfileScope prof("tie_break", s);
    zero_fill(tie_count, sizeof(int), s);
    TieBand queue{};
    if (band && band->cap > 0) {
        queue = *band;
        zero_fill(queue.count, sizeof(int), s);
    }
    tie_split_kernel<<<(rows + kBlock - 1) / kBlock, kBlock, 0, s>>>(
        idx1, d21, n, row0, rows, k, rho, nbr_idx, nbr_d2, kdist2, tie_rows, tie_count, queue);
    CULOF_CHECK(hipGetLastError());
    if (queue.rows) {
        tie_band_select(keyrows, stride, n, k, queue, nbr_idx, nbr_d2, kdist2, tie_rows, tie_count, rows, s);
    }
    const int grid = rows < 1024 ? rows : 1024;
    tie_resolve_kernel<<<grid, kBlock, 0, s>>>(keyrows, stride, n, d, k, rho, idx1, d21, tie_rows,
                                               tie_count, rows, nbr_idx, nbr_d2, kdist2);
    CULOF_CHECK(hipGetLastError());
}

}  // namespace detail
}  // namespace culof
