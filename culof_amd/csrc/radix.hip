// Per-row exact radix select on IEEE-754 FP32 key bit patterns.
//
// The reference's selection algorithm (src/cuda/knn.cu:119-232), rebuilt for
// wave64: for non-negative floats the unsigned order of the bit pattern is the
// numeric order, so three LDS-histogram passes over 11 + 11 + 10 bits pin the
// k-th smallest key exactly, and that key is the squared k-distance. A fourth
// pass gathers every key strictly below it in ascending index, then the
// lowest-index keys equal to it, with slots assigned by a block-wide prefix sum
// (no atomics, so the output is deterministic).
//
// Two key sources share the kernel:
//   * stored: rows of a materialised key tile (the reference's data flow);
//   * recompute: difference-form keys rebuilt from the key rows on every pass
//     with the same FP32 FMA chain as the fused sweep's admission test, so this
//     path is the exact fallback for rows the sweep hands back (k-th key +inf)
//     and the general path for k or d beyond the sweep's instantiations.
//
// The block scan is a wave64 shuffle scan plus one LDS step across the 4
// waves; the histogram bin search scans per-thread slices of 8 bins.

#include "internal.hpp"

#include <algorithm>
#include <stdexcept>

namespace culof {
namespace detail {

namespace {

constexpr int kBlock = 256;
constexpr int kBins = 2048;

// Exclusive prefix sum over the block; `total` receives the block sum.
__device__ int block_exclusive_scan(int v, int& total, int* wave_sums) {
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    int x = v;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const int y = __shfl_up(x, off, 64);
        if (lane >= off) x += y;
    }
    if (lane == 63) wave_sums[wave] = x;
    __syncthreads();
    int before = 0;
    total = 0;
#pragma unroll
    for (int w = 0; w < kBlock / 64; ++w) {
        const int s = wave_sums[w];
        before += (w < wave) ? s : 0;
        total += s;
    }
    __syncthreads();  // wave_sums is reused by the next scan
    return before + x - v;
}

struct RadixShared {
    int hist[kBins];
    int wave_sums[kBlock / 64];
    int found_bin;
    int found_below;
};

// Locate the bin holding the need-th smallest key and fold it into the prefix.
__device__ void refine(RadixShared& sh, int bins, int shift, int need, uint32_t& prefix,
                       int& below) {
    const int per = bins / kBlock;
    int local = 0;
    for (int b = threadIdx.x * per; b < (static_cast<int>(threadIdx.x) + 1) * per; ++b)
        local += sh.hist[b];
    int total = 0;
    const int excl = block_exclusive_scan(local, total, sh.wave_sums);
    // Exactly one slice holds the need-th key (1 <= need <= total).
    if (excl < need && excl + local >= need) {
        int acc = excl;
        int bin = threadIdx.x * per;
        for (int b = threadIdx.x * per; b < (static_cast<int>(threadIdx.x) + 1) * per; ++b) {
            if (acc + sh.hist[b] >= need) {
                bin = b;
                break;
            }
            acc += sh.hist[b];
        }
        sh.found_bin = bin;
        sh.found_below = acc;
    }
    __syncthreads();
    prefix |= static_cast<uint32_t>(sh.found_bin) << shift;
    below += sh.found_below;
    __syncthreads();
}

struct StoredKeys {
    const float* row;
    __device__ uint32_t operator()(int j) const {
        const float v = row[j];
        if (v != v) return kInfBits;                  // NaN reads as +inf
        return v > 0.0f ? __float_as_uint(v) : 0u;    // -0 and negatives read as +0
    }
};

struct RecomputedKeys {
    const float* q;   // query key row, in LDS
    const float* xs;  // key rows
    int stride;
    int d;
    int self;
    __device__ uint32_t operator()(int j) const {
        if (j == self) return kInfBits;
        const float* xj = xs + static_cast<size_t>(j) * stride;
        float e = 0.0f;
        for (int t = 0; t < d; ++t) {
            const float diff = q[t] - xj[t];
            e = __builtin_fmaf(diff, diff, e);
        }
        return (e != e) ? kInfBits : __float_as_uint(e);
    }
};

// f(key(j)) for every j in [0, n) of this thread's share, four keys loaded
// before any is used (histogram counts do not depend on the order).
template <class Keys, class F>
__device__ __forceinline__ void for_each_key(const Keys& key, int n, F&& f) {
    int j = threadIdx.x;
    for (; j + 3 * kBlock < n; j += 4 * kBlock) {
        const uint32_t u0 = key(j);
        const uint32_t u1 = key(j + kBlock);
        const uint32_t u2 = key(j + 2 * kBlock);
        const uint32_t u3 = key(j + 3 * kBlock);
        f(u0);
        f(u1);
        f(u2);
        f(u3);
    }
    for (; j < n; j += kBlock) f(key(j));
}

template <class Keys>
__device__ void select_row(const Keys& key, RadixShared& sh, int n, int k, int out_col,
                           int out_stride, int* __restrict__ nbr_idx, float* __restrict__ nbr_d2,
                           float* __restrict__ kdist2) {
    uint32_t prefix = 0;
    int below = 0;

    // pass 0: top 11 bits
    for (int b = threadIdx.x; b < kBins; b += kBlock) sh.hist[b] = 0;
    __syncthreads();
    for_each_key(key, n, [&](uint32_t u) { atomicAdd(&sh.hist[u >> 21], 1); });
    __syncthreads();
    refine(sh, kBins, 21, k, prefix, below);

    // pass 1: middle 11 bits of keys matching the top 11
    for (int b = threadIdx.x; b < kBins; b += kBlock) sh.hist[b] = 0;
    __syncthreads();
    for_each_key(key, n, [&](uint32_t u) {
        if ((u >> 21) == (prefix >> 21)) atomicAdd(&sh.hist[(u >> 10) & 0x7FFu], 1);
    });
    __syncthreads();
    refine(sh, kBins, 10, k - below, prefix, below);

    // pass 2: low 10 bits of keys matching the top 22
    for (int b = threadIdx.x; b < kBins; b += kBlock) sh.hist[b] = 0;
    __syncthreads();
    for_each_key(key, n, [&](uint32_t u) {
        if ((u >> 10) == (prefix >> 10)) atomicAdd(&sh.hist[u & 0x3FFu], 1);
    });
    __syncthreads();
    refine(sh, 1024, 0, k - below, prefix, below);

    const uint32_t kth = prefix;  // exactly the k-th smallest key
    if (threadIdx.x == 0) kdist2[out_col] = __uint_as_float(kth);

    // gather: all keys < kth by ascending index, then the lowest-index keys == kth
    int emitted_lt = 0;
    int emitted_eq = 0;
    const size_t os = static_cast<size_t>(out_stride);
    for (int base = 0; base < n; base += kBlock) {
        if (emitted_lt >= below && below + emitted_eq >= k) break;  // every slot filled
        const int j = base + static_cast<int>(threadIdx.x);
        const uint32_t u = (j < n) ? key(j) : 0xFFFFFFFFu;
        const int lt = (j < n && u < kth) ? 1 : 0;
        const int eq = (j < n && u == kth) ? 1 : 0;
        int total = 0;
        const int off = block_exclusive_scan(lt | (eq << 16), total, sh.wave_sums);
        if (lt) {
            const int slot = emitted_lt + (off & 0xFFFF);
            if (slot < k) {
                nbr_idx[slot * os + out_col] = j;
                nbr_d2[slot * os + out_col] = __uint_as_float(u);
            }
        }
        if (eq) {
            const int slot = below + emitted_eq + ((off >> 16) & 0xFFFF);
            if (slot < k) {
                nbr_idx[slot * os + out_col] = j;
                nbr_d2[slot * os + out_col] = __uint_as_float(u);
            }
        }
        emitted_lt += total & 0xFFFF;
        emitted_eq += (total >> 16) & 0xFFFF;
    }
}

__global__ void __launch_bounds__(kBlock)
radix_stored_kernel(const float* __restrict__ keys, int rows, int n_cols, int k,
                    int* __restrict__ nbr_idx, float* __restrict__ nbr_d2,
                    float* __restrict__ kdist2) {
    __shared__ RadixShared sh;
    const int r = blockIdx.x;
    if (r >= rows) return;
    const StoredKeys key{keys + static_cast<size_t>(r) * n_cols};
    select_row(key, sh, n_cols, k, r, rows, nbr_idx, nbr_d2, kdist2);
}

// Grid-stride over query rows so a launch sized for the worst case costs
// nothing when the row list is empty. With `count` null, rows are
// [row0, row0 + rows); otherwise row_list[0 .. *count).
__global__ void __launch_bounds__(kBlock)
radix_recompute_kernel(const float* __restrict__ xs, int stride, int d, int n, int k, int row0,
                       int rows,
                       const int* __restrict__ row_list, const int* __restrict__ count,
                       int* __restrict__ nbr_idx, float* __restrict__ nbr_d2,
                       float* __restrict__ kdist2) {
    extern __shared__ float s_q[];  // stride floats
    __shared__ RadixShared sh;
    const int total = count ? min(*count, rows) : rows;  // rows: the list's capacity
    for (int r = blockIdx.x; r < total; r += gridDim.x) {
        const int i = row_list ? row_list[r] : row0 + r;
        __syncthreads();  // s_q reuse across rows
        for (int t = threadIdx.x; t < stride; t += kBlock) s_q[t] = xs[static_cast<size_t>(i) * stride + t];
        __syncthreads();
        const RecomputedKeys key{s_q, xs, stride, d, i};
        select_row(key, sh, n, k, i, n, nbr_idx, nbr_d2, kdist2);
    }
}

// Queries from another row set: row r of `qrows` against every row of
// `xs` (both at row stride `stride`), nothing excluded. Outputs with stride
// n_queries. Grid-stride over the queries [0, n_queries), or over
// row_list[0 .. *count) when count is given.
__global__ void __launch_bounds__(kBlock)
radix_query_kernel(const float* __restrict__ qrows, const float* __restrict__ xs, int stride, int d,
                   int n, int k, int n_queries, const int* __restrict__ row_list,
                   const int* __restrict__ count, int* __restrict__ nbr_idx, float* __restrict__ nbr_d2,
                   float* __restrict__ kdist2) {
    extern __shared__ float s_q[];  // stride floats
    __shared__ RadixShared sh;
    const int total = count ? min(*count, n_queries) : n_queries;
    for (int i = blockIdx.x; i < total; i += gridDim.x) {
        const int r = row_list ? row_list[i] : i;
        __syncthreads();  // s_q reuse across rows
        for (int t = threadIdx.x; t < stride; t += kBlock) s_q[t] = qrows[static_cast<size_t>(r) * stride + t];
        __syncthreads();
        const RecomputedKeys key{s_q, xs, stride, d, -1};
        select_row(key, sh, n, k, r, n_queries, nbr_idx, nbr_d2, kdist2);
    }
}

// Exact keys of a batch of listed query rows against every key row, for the
// stored select below: keys[b * n + j] for list entries r0 + b of
// row_list[0 .. min(*count, cap)), b < batch. A block stages QB query rows in
// LDS and keys 256 candidates (one per thread) against all of them, the
// candidate row read once as float4s; the chain is the exact FMA chain of
// every other path (zero padding adds exact zeros), the diagonal and NaN +inf.
template <int QB>
__global__ void __launch_bounds__(kBlock)
brute_keys_kernel(const float* __restrict__ xs, int stride, int n, const int* __restrict__ row_list,
                  const int* __restrict__ count, int cap, int r0, int batch, float* __restrict__ keys) {
    extern __shared__ float s_rows[];  // QB * stride floats
    const int total = min(*count, cap);
    const int qb0 = r0 + static_cast<int>(blockIdx.y) * QB;
    const int end = min(total, r0 + batch);
    if (qb0 >= end) return;  // whole block
    const int nq = min(QB, end - qb0);
    for (int e = threadIdx.x; e < nq * stride; e += kBlock) {
        const int q = e / stride;
        s_rows[e] = xs[static_cast<size_t>(row_list[qb0 + q]) * stride + (e - q * stride)];
    }
    __syncthreads();
    const int j = blockIdx.x * kBlock + threadIdx.x;
    if (j >= n) return;
    float acc[QB];
#pragma unroll
    for (int q = 0; q < QB; ++q) acc[q] = 0.0f;
    const float4* c4 = reinterpret_cast<const float4*>(xs + static_cast<size_t>(j) * stride);
    const float4* q4 = reinterpret_cast<const float4*>(s_rows);
    const int f4 = stride / 4;
#pragma unroll 2
    for (int v = 0; v < f4; ++v) {
        const float4 c = c4[v];
#pragma unroll
        for (int q = 0; q < QB; ++q) {
            const float4 a = q4[q * f4 + v];
            float df = a.x - c.x;
            acc[q] = __builtin_fmaf(df, df, acc[q]);
            df = a.y - c.y;
            acc[q] = __builtin_fmaf(df, df, acc[q]);
            df = a.z - c.z;
            acc[q] = __builtin_fmaf(df, df, acc[q]);
            df = a.w - c.w;
            acc[q] = __builtin_fmaf(df, df, acc[q]);
        }
    }
#pragma unroll
    for (int q = 0; q < QB; ++q) {
        if (q < nq) {
            const int qi = row_list[qb0 + q];
            const float key = (j == qi || acc[q] != acc[q]) ? __uint_as_float(kInfBits) : acc[q];
            keys[static_cast<size_t>(qb0 - r0 + q) * n + j] = key;
        }
    }
}

// The stored select over a batch of brute_keys_kernel rows; outputs at the
// rows' original indices (stride n).
__global__ void __launch_bounds__(kBlock)
radix_list_stored_kernel(const float* __restrict__ keys, int n, int k, const int* __restrict__ row_list,
                         const int* __restrict__ count, int cap, int r0, int* __restrict__ nbr_idx,
                         float* __restrict__ nbr_d2, float* __restrict__ kdist2) {
    __shared__ RadixShared sh;
    const int r = r0 + static_cast<int>(blockIdx.x);
    if (r >= min(*count, cap)) return;
    const StoredKeys key{keys + static_cast<size_t>(blockIdx.x) * n};
    select_row(key, sh, n, k, row_list[r], n, nbr_idx, nbr_d2, kdist2);
}

}  // namespace

void radix_query_knn(const float* qrows, const float* keyrows, int stride, int n_points, int n_dims,
                     int k, int n_queries, const int* row_list, const int* row_count, int* nbr_idx,
                     float* nbr_d2, float* kdist2, hipStream_t stream) {
    ProfileScope prof("radix_knn", stream);
    if (n_queries <= 0) return;
    const int grid = std::min(n_queries, 2048);
    radix_query_kernel<<<grid, kBlock, sizeof(float) * stride, stream>>>(
        qrows, keyrows, stride, n_dims, n_points, k, n_queries, row_list, row_count, nbr_idx, nbr_d2,
        kdist2);
    CULOF_CHECK(hipGetLastError());
}

void radix_knn(const float* keyrows, int stride, int n_points, int n_dims, int k, int row0,
               int rows, const int* row_list, const int* row_count, int* nbr_idx, float* nbr_d2,
               float* kdist2, hipStream_t stream) {
    ProfileScope prof("radix_knn", stream);
    const int grid = row_count ? std::min(rows, 2048) : rows;
    if (grid <= 0) return;
    radix_recompute_kernel<<<grid, kBlock, sizeof(float) * stride, stream>>>(
        keyrows, stride, n_dims, n_points, k, row0, rows, row_list, row_count, nbr_idx, nbr_d2,
        kdist2);
    CULOF_CHECK(hipGetLastError());
}

int brute_batch_rows(int n_points, int cap) {
    const size_t per_row = static_cast<size_t>(std::max(n_points, 1)) * sizeof(float);
    const int rows = static_cast<int>(std::min<size_t>(kBruteKeyBytes / per_row, 1 << 20));
    return std::max(1, std::min(cap, std::max(rows, 64)));
}

void brute_knn_rows(const float* keyrows, int stride, int n_points, int k, const int* row_list,
                    const int* row_count, int cap, float* keys, int batch, int* nbr_idx, float* nbr_d2,
                    float* kdist2, hipStream_t stream) {
    ProfileScope prof("brute_rows", stream);
    if (cap <= 0) return;
    if (stride % 4 != 0) throw std::invalid_argument("brute_knn_rows: row stride must be whole float4s");
    const unsigned cgrid = static_cast<unsigned>((n_points + kBlock - 1) / kBlock);
    // Batches past the device count exit at once (no host read-back).
    for (int r0 = 0; r0 < cap; r0 += batch) {
        const int b = std::min(batch, cap - r0);
        if (stride <= 128) {
            constexpr int QB = 32;
            brute_keys_kernel<QB><<<dim3(cgrid, (b + QB - 1) / QB), kBlock, sizeof(float) * QB * stride, stream>>>(
                keyrows, stride, n_points, row_list, row_count, cap, r0, b, keys);
        } else {
            constexpr int QB = 16;
            brute_keys_kernel<QB><<<dim3(cgrid, (b + QB - 1) / QB), kBlock, sizeof(float) * QB * stride, stream>>>(
                keyrows, stride, n_points, row_list, row_count, cap, r0, b, keys);
        }
        CULOF_CHECK(hipGetLastError());
        radix_list_stored_kernel<<<b, kBlock, 0, stream>>>(keys, n_points, k, row_list, row_count, cap, r0,
                                                          nbr_idx, nbr_d2, kdist2);
        CULOF_CHECK(hipGetLastError());
    }
}

void radix_select_rows(const float* keys, int rows, int n_cols, int k, int* nbr_idx,
                       float* nbr_d2, float* kdist2, hipStream_t stream) {
    ProfileScope prof("radix_select_rows", stream);
    if (rows <= 0) return;
    radix_stored_kernel<<<rows, kBlock, 0, stream>>>(keys, rows, n_cols, k, nbr_idx, nbr_d2,
                                                     kdist2);
    CULOF_CHECK(hipGetLastError());
}

}  // namespace detail
}  // namespace culof
