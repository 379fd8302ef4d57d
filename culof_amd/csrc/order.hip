// Candidate order and per-query seeds for the matrix-core sweep.
//
// Neither changes any result: selection is by exact keys with the original
// index as the tie-breaker, whatever order the candidates stream in. They
// change how much work the sweep does. In arbitrary order each query's list
// improves about k ln(n / k) times before it settles, and every improvement
// is an exact re-key plus a list insertion. Here:
//
//   * the points are ordered by two levels of nearest-pivot cells
//     (64 pivots, then 64 sub-pivots inside each level-1 cell), so points
//     that are close in the order are mostly close in space;
//   * each query gets a seed T_q: the k-th smallest exact key among a window
//     of W candidates around its place in that order (itself excluded; the
//     queries of one block share the window).
//     Any k candidates bound the k-th neighbour from above, so T_q >= the
//     true k-distance; the sweep starts every list full of placeholder words
//     (T_q, 0xFFFFFFFF), which admit exactly the candidates with key <= T_q.
//
// With a tight seed only candidates near the final k-th key ever pass the
// screen, so the sweep's cost approaches that of the screen itself.

#include "internal.hpp"

#include <algorithm>
#include <stdexcept>
#include <utility>

namespace culof {
namespace detail {

namespace {

constexpr int kPiv = 64;     // pivots per level
constexpr int kBlock = 256;
constexpr int kKeys = kPiv * kPiv;

// Orders a wave's LDS writes before its later reads (and reads before later
// writes) without a block barrier.
__device__ __forceinline__ void seed_wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Index of the nearest of `count` pivot rows (LDS or global) to row x, by
// difference-form FP32 squared distance; ties to the lower pivot.
template <int DP>
__device__ int nearest_pivot(const float (&x)[DP], const float* piv, int count) {
    float best = __uint_as_float(kInfBits);
    int arg = 0;
    for (int t = 0; t < count; ++t) {
        const float4* p4 = reinterpret_cast<const float4*>(piv + static_cast<size_t>(t) * DP);
        float e = 0.0f;
#pragma unroll
        for (int v = 0; v < DP / 4; ++v) {
            const float4 q = p4[v];
            float df = x[4 * v] - q.x;
            e = __builtin_fmaf(df, df, e);
            df = x[4 * v + 1] - q.y;
            e = __builtin_fmaf(df, df, e);
            df = x[4 * v + 2] - q.z;
            e = __builtin_fmaf(df, df, e);
            df = x[4 * v + 3] - q.w;
            e = __builtin_fmaf(df, df, e);
        }
        if (e < best) {
            best = e;
            arg = t;
        }
    }
    return arg;
}

template <int DP>
__device__ void load_row(const float* __restrict__ row, float (&x)[DP]) {
    const float4* r4 = reinterpret_cast<const float4*>(row);
#pragma unroll
    for (int v = 0; v < DP / 4; ++v) {
        const float4 q = r4[v];
        x[4 * v] = q.x;
        x[4 * v + 1] = q.y;
        x[4 * v + 2] = q.z;
        x[4 * v + 3] = q.w;
    }
}

// Level-1 pivots: rows floor(t n / 64) of the key rows. The cells use the
// first DP features of rows of STRIDE floats (wider rows are ordered on
// their first 128 features: the order only steers work).
template <int DP, int STRIDE = DP>
__global__ void __launch_bounds__(kBlock)
level1_kernel(const float* __restrict__ xs, int n, int* __restrict__ cell,
              int* __restrict__ hist) {
    __shared__ __attribute__((aligned(16))) float piv[kPiv * DP];
    for (int c = threadIdx.x; c < kPiv * DP; c += kBlock) {
        const int t = c / DP;
        const long long row = static_cast<long long>(t) * n / kPiv;
        piv[c] = xs[row * STRIDE + (c - t * DP)];
    }
    __shared__ int s_hist[kPiv];
    if (threadIdx.x < kPiv) s_hist[threadIdx.x] = 0;
    __syncthreads();
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i < n) {
        float x[DP];
        load_row<DP>(xs + static_cast<size_t>(i) * STRIDE, x);
        const int c = nearest_pivot<DP>(x, piv, n < kPiv ? n : kPiv);
        cell[i] = c;
        atomicAdd(&s_hist[c], 1);
    }
    __syncthreads();
    // One global add per bin per block: the 64 bins are hot.
    if (threadIdx.x < kPiv && s_hist[threadIdx.x]) atomicAdd(&hist[threadIdx.x], s_hist[threadIdx.x]);
}

// Exclusive scan of `bins` counts (bins <= 4096) in one block; also resets
// the cursors used by the scatter.
__global__ void __launch_bounds__(kBlock)
scan_bins_kernel(const int* __restrict__ hist, int bins, int* __restrict__ offset,
                 int* __restrict__ cursor) {
    __shared__ int part[kBlock];
    const int per = (bins + kBlock - 1) / kBlock;
    const int b0 = threadIdx.x * per;
    int s = 0;
    for (int b = b0; b < b0 + per && b < bins; ++b) s += hist[b];
    part[threadIdx.x] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        int run = 0;
        for (int t = 0; t < kBlock; ++t) {
            const int v = part[t];
            part[t] = run;
            run += v;
        }
    }
    __syncthreads();
    int run = part[threadIdx.x];
    for (int b = b0; b < b0 + per && b < bins; ++b) {
        offset[b] = run;
        cursor[b] = 0;
        run += hist[b];
    }
}

__global__ void __launch_bounds__(kBlock)
scatter_kernel(const int* __restrict__ key, int n, const int* __restrict__ offset,
               int* __restrict__ cursor, int* __restrict__ order) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const int b = key[i];
    order[offset[b] + atomicAdd(&cursor[b], 1)] = i;
}

// Scatter for few hot bins (<= kPiv): ranks within the block from LDS
// atomics, then one global reservation per bin per block.
__global__ void __launch_bounds__(kBlock)
scatter_few_kernel(const int* __restrict__ key, int n, const int* __restrict__ offset,
                   int* __restrict__ cursor, int* __restrict__ order) {
    __shared__ int s_cnt[kPiv];
    __shared__ int s_base[kPiv];
    if (threadIdx.x < kPiv) s_cnt[threadIdx.x] = 0;
    __syncthreads();
    const int i = blockIdx.x * kBlock + threadIdx.x;
    const int b = i < n ? key[i] : 0;
    const int rank = i < n ? atomicAdd(&s_cnt[b], 1) : 0;
    __syncthreads();
    if (threadIdx.x < kPiv && s_cnt[threadIdx.x])
        s_base[threadIdx.x] = offset[threadIdx.x] + atomicAdd(&cursor[threadIdx.x], s_cnt[threadIdx.x]);
    __syncthreads();
    if (i < n) order[s_base[b] + rank] = i;
}

// Level-2 pivots of every level-1 cell: members at floor(t m / 64) of the
// cell's range in the level-1 order (fewer when the cell is smaller).
__global__ void __launch_bounds__(kBlock)
level2_pivots_kernel(const float* __restrict__ xs, int dp, int stride, const int* __restrict__ order1,
                     const int* __restrict__ offset1, const int* __restrict__ hist1,
                     float* __restrict__ piv2) {
    const int c = blockIdx.x;
    const int m = hist1[c];
    for (int e = threadIdx.x; e < kPiv * dp; e += kBlock) {
        const int t = e / dp;
        float v = 0.0f;
        if (t < m) {
            const int member = order1[offset1[c] + static_cast<int>(static_cast<long long>(t) * m / kPiv)];
            v = xs[static_cast<size_t>(member) * stride + (e - t * dp)];
        }
        piv2[static_cast<size_t>(c) * kPiv * dp + e] = v;
    }
}

// Final key of each point: level-1 cell * 64 + nearest level-2 pivot of its cell.
template <int DP, int STRIDE = DP>
__global__ void __launch_bounds__(kBlock)
level2_kernel(const float* __restrict__ xs, int n, const int* __restrict__ order1,
              const int* __restrict__ cell1, const int* __restrict__ hist1,
              const float* __restrict__ piv2, int* __restrict__ key, int* __restrict__ hist) {
    const int p = blockIdx.x * kBlock + threadIdx.x;
    if (p >= n) return;
    const int i = order1[p];
    const int c = cell1[i];
    const int m = hist1[c];
    float x[DP];
    load_row<DP>(xs + static_cast<size_t>(i) * STRIDE, x);
    const int s = nearest_pivot<DP>(x, piv2 + static_cast<size_t>(c) * kPiv * DP, m < kPiv ? m : kPiv);
    const int kk = c * kPiv + s;
    key[i] = kk;
    atomicAdd(&hist[kk], 1);
}

__global__ void __launch_bounds__(kBlock)
inverse_kernel(const int* __restrict__ order, int n, int* __restrict__ inv) {
    const int p = blockIdx.x * kBlock + threadIdx.x;
    if (p < n) inv[order[p]] = p;
}

// Rows of `src` (stride dp) gathered into the candidate order.
__global__ void __launch_bounds__(kBlock)
gather_rows_kernel(const float* __restrict__ src, int dp, const int* __restrict__ order, int n,
                   float* __restrict__ dst) {
    const size_t c4 = static_cast<size_t>(blockIdx.x) * kBlock + threadIdx.x;
    const int per = dp / 4;
    if (c4 >= static_cast<size_t>(n) * per) return;
    const size_t p = c4 / per;
    const int v = static_cast<int>(c4 - p * per);
    reinterpret_cast<float4*>(dst)[c4] =
        reinterpret_cast<const float4*>(src)[static_cast<size_t>(order[p]) * per + v];
}

// Seeds for the points at sorted positions [p0, p0 + 4 QPW) (one block, QPW
// queries per wave): exact keys against the W candidates of a window centred
// on the block in the candidate order, streamed through LDS in chunks of
// 64 RPL rows (lane l keys rows l, l + 64, ... of the chunk against the
// wave's QPW queries; the next piece is fetched into registers meanwhile),
// then each query's k-th smallest key by a bitwise search on the key bits.
// DP: the features keyed (a multiple of 4); STRIDE: the row stride of xs_ord
// (>= DP; the features past DP are zero padding, which adds exact zeros).
// Rows wider than 128 features reach LDS in pieces of 128 features, the keys
// carried across the pieces in ascending feature order (the same FMA chain).
template <int DP, int STRIDE, int W, int QPW, int RPL>
__global__ void __launch_bounds__(kBlock)
seed_kernel(const float* __restrict__ xs_ord, const int* __restrict__ order, int n, int k,
            float* __restrict__ tseed) {
    constexpr int kQ = 4 * QPW;                // queries per block
    constexpr int kRows = 64 * RPL;            // candidate rows per chunk
    constexpr int kChunks = W / kRows;
    constexpr int FC = DP < 128 ? DP : 128;    // features per LDS piece
    constexpr int kPieces = DP / FC;           // pieces per chunk
    constexpr int kPitch = FC + 4;             // LDS row pitch (floats)
    constexpr int kF4 = kRows * FC / 4;        // float4s per piece
    constexpr int kLoads = (kF4 + kBlock - 1) / kBlock;
    static_assert(DP % FC == 0, "whole pieces");
    // Wide rows: the wave's queries are read by scalar loads (their rows are
    // wave-uniform), so the LDS pipe serves only the candidate rows; from
    // LDS every step would read QPW query broadcasts besides one candidate.
    constexpr bool kScalarQ = DP >= 32;
    __shared__ __attribute__((aligned(16))) float s_q[kScalarQ ? 1 : kQ][kScalarQ ? 1 : DP];
    __shared__ __attribute__((aligned(16))) float s_c[kRows * kPitch];
    __shared__ uint32_t s_few[4][64];  // a wave's keys <= M, compacted
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int p0 = blockIdx.x * kQ;
    const int w = n < W ? n : W;
    int c0 = p0 + kQ / 2 - w / 2;
    c0 = c0 < 0 ? 0 : (c0 > n - w ? n - w : c0);
    if constexpr (!kScalarQ) {
        for (int e = threadIdx.x; e < kQ * DP / 4; e += kBlock) {
            const int t = e / (DP / 4);
            const int v = e - t * (DP / 4);
            const int pos = p0 + t < n ? p0 + t : n - 1;
            reinterpret_cast<float4*>(&s_q[t][0])[v] =
                reinterpret_cast<const float4*>(xs_ord + static_cast<size_t>(pos) * STRIDE)[v];
        }
    }
    const float* qrow[QPW];
#pragma unroll
    for (int t = 0; t < QPW; ++t) {
        const int pos = p0 + wave * QPW + t < n ? p0 + wave * QPW + t : n - 1;
        qrow[t] = xs_ord + static_cast<size_t>(__builtin_amdgcn_readfirstlane(pos)) * STRIDE;
    }
    float4 nxt[kLoads];
    // piece `pc` = chunk pc / kPieces, features [FC (pc % kPieces), + FC)
    auto fetch = [&](int pc) {
        const int ch = pc / kPieces;
        const int f0 = (pc - ch * kPieces) * FC;
#pragma unroll
        for (int u = 0; u < kLoads; ++u) {
            const int e = threadIdx.x + kBlock * u;
            const int r = e / (FC / 4);
            const int c = ch * kRows + r;
            nxt[u] = (e < kF4 && c < w)
                         ? reinterpret_cast<const float4*>(xs_ord + static_cast<size_t>(c0 + c) * STRIDE + f0)[e - r * (FC / 4)]
                         : make_float4(0.f, 0.f, 0.f, 0.f);
        }
    };
    fetch(0);
    // The keys of each query, W / 64 per lane, kept in a shift register so the
    // chunk loop can stay rolled: every chunk shifts RPL old keys out of the
    // front and its RPL new keys in at the back.
    uint32_t key[QPW][W / 64];
#pragma unroll 1
    for (int ch = 0; ch < kChunks; ++ch) {
        float e[RPL][QPW];
#pragma unroll
        for (int rr = 0; rr < RPL; ++rr) {
#pragma unroll
            for (int t = 0; t < QPW; ++t) e[rr][t] = 0.0f;
        }
#pragma unroll 1
        for (int pc = 0; pc < kPieces; ++pc) {
            __syncthreads();
#pragma unroll
            for (int u = 0; u < kLoads; ++u) {
                const int e4 = threadIdx.x + kBlock * u;
                const int r = e4 / (FC / 4);
                if (e4 < kF4) reinterpret_cast<float4*>(s_c + r * kPitch)[e4 - r * (FC / 4)] = nxt[u];
            }
            __syncthreads();
            if (ch * kPieces + pc + 1 < kChunks * kPieces) fetch(ch * kPieces + pc + 1);
            const int q4 = pc * (FC / 4);  // the piece's first float4 of a query row
#pragma unroll
            for (int rr = 0; rr < RPL; ++rr) {
                const int r = lane + 64 * rr;
                const float4* row = reinterpret_cast<const float4*>(s_c + r * kPitch);
#pragma unroll 4
                for (int v = 0; v < FC / 4; ++v) {
                    const float4 cv = row[v];
#pragma unroll
                    for (int t = 0; t < QPW; ++t) {
                        float4 qv;
                        if constexpr (kScalarQ) {
                            qv = reinterpret_cast<const float4*>(qrow[t])[q4 + v];
                        } else {
                            qv = reinterpret_cast<const float4*>(&s_q[wave * QPW + t][0])[q4 + v];
                        }
                        float df = qv.x - cv.x;
                        e[rr][t] = __builtin_fmaf(df, df, e[rr][t]);
                        df = qv.y - cv.y;
                        e[rr][t] = __builtin_fmaf(df, df, e[rr][t]);
                        df = qv.z - cv.z;
                        e[rr][t] = __builtin_fmaf(df, df, e[rr][t]);
                        df = qv.w - cv.w;
                        e[rr][t] = __builtin_fmaf(df, df, e[rr][t]);
                    }
                }
            }
        }
        uint32_t fresh[QPW][RPL];
#pragma unroll
        for (int rr = 0; rr < RPL; ++rr) {
            const int c = ch * kRows + lane + 64 * rr;
#pragma unroll
            for (int t = 0; t < QPW; ++t) {
                const bool self = c0 + c == p0 + wave * QPW + t;
                fresh[t][rr] = (c >= w || self || e[rr][t] != e[rr][t]) ? kInfBits : __float_as_uint(e[rr][t]);
            }
        }
#pragma unroll
        for (int t = 0; t < QPW; ++t) {
#pragma unroll
            for (int u = 0; u + RPL < W / 64; ++u) key[t][u] = key[t][u + RPL];
#pragma unroll
            for (int rr = 0; rr < RPL; ++rr) key[t][W / 64 - RPL + rr] = fresh[t][rr];
        }
    }
    // Largest ans with #(keys < ans) < k is the k-th smallest key K. Keys
    // are non-negative (bit 31 clear), and the seed only has to bound K from
    // above, so the search stops kSeedLowBits short: the prefix P found has
    // P <= K < P + 2^kSeedLowBits, and P with its low bits set is returned
    // (at most 2^-13 relative above K; a third fewer search steps).
    //
    // The search runs over the few keys that can matter: M, the k-th
    // smallest of the 64 lane minima, bounds K (the k lanes whose minima are
    // at most M hold k keys at most M), so K is the k-th smallest of the keys
    // <= M. Every probe of the search gets the same answer from those keys
    // as from all of them (a probe at most M + 1 counts the same keys below
    // it; one above M finds at least k either way), so the result is the
    // same bits. When at most 64 keys are <= M they are compacted one per
    // lane through LDS and searched with one compare per step instead of
    // W / 64; otherwise (many equal keys) the search runs over all keys.
    constexpr int kSeedLowBits = 10;
#pragma unroll
    for (int t = 0; t < QPW; ++t) {
        const int pos = p0 + wave * QPW + t;
        uint32_t ans = 0;
        bool searched = false;
        if (k <= 64) {
            uint32_t m = key[t][0];
#pragma unroll
            for (int u = 1; u < W / 64; ++u) m = min(m, key[t][u]);
            uint32_t M = 0;  // the k-th smallest lane minimum
            for (int b = 30; b >= 0; --b) {
                const uint32_t probe = M | (1u << b);
                if (__popcll(__ballot(m < probe)) < k) M = probe;
            }
            int c = 0;
#pragma unroll
            for (int u = 0; u < W / 64; ++u) c += key[t][u] <= M ? 1 : 0;
            int incl = c;  // inclusive prefix over the lanes
#pragma unroll
            for (int off = 1; off < 64; off <<= 1) {
                const int y = __shfl_up(incl, off, 64);
                if (lane >= off) incl += y;
            }
            const int total = __shfl(incl, 63, 64);
            if (total <= 64) {
                int at = incl - c;
#pragma unroll
                for (int u = 0; u < W / 64; ++u)
                    if (key[t][u] <= M) s_few[wave][at++] = key[t][u];
                seed_wave_sync();
                const uint32_t v = lane < total ? s_few[wave][lane] : 0xFFFFFFFFu;
                seed_wave_sync();  // the next query's writes come after every lane's read
                for (int b = 30; b >= kSeedLowBits; --b) {
                    const uint32_t probe = ans | (1u << b);
                    if (__popcll(__ballot(v < probe)) < k) ans = probe;
                }
                searched = true;
            }
        }
        if (!searched) {
            for (int b = 30; b >= kSeedLowBits; --b) {
                const uint32_t probe = ans | (1u << b);
                int below = 0;
#pragma unroll
                for (int u = 0; u < W / 64; ++u) below += __popcll(__ballot(key[t][u] < probe));
                if (below < k) ans = probe;
            }
        }
        ans |= (1u << kSeedLowBits) - 1u;
        if (lane == 0 && pos < n) tseed[order[pos]] = __uint_as_float(ans > kInfBits ? kInfBits : ans);
    }
}

// ---- rows with the largest seeds last ----
//
// A seed far above its neighbours' (an outlier, whose k-th neighbour is far
// away) widens the bounding box and the threshold of its tile, and the
// symmetric pass can then rule out few tile pairs by their boxes. Such rows
// are moved to the end of the candidate order, each part keeping its order:
// at most about n / 100 of them, those whose seed falls in the top bins of a
// histogram over the seeds' bit patterns (1/16 octave bins).
constexpr int kSeedBins = 4096;  // bits >> 19 of a non-negative float
constexpr int kSplitSpan = 1024;  // positions per block of the partition

// Seeds of one data set crowd into a few bins: counts go to a histogram in
// LDS first (kSplitSpan seeds per block), then its nonzero bins to memory.
__global__ void __launch_bounds__(kBlock)
seed_hist_kernel(const float* __restrict__ tseed, int n, int* __restrict__ hist) {
    __shared__ int s_hist[kSeedBins];
    for (int b = threadIdx.x; b < kSeedBins; b += kBlock) s_hist[b] = 0;
    __syncthreads();
    for (int e = 0; e < kSplitSpan / kBlock; ++e) {
        const int i = blockIdx.x * kSplitSpan + e * kBlock + threadIdx.x;
        if (i < n) atomicAdd(&s_hist[__float_as_uint(tseed[i]) >> 19], 1);
    }
    __syncthreads();
    for (int b = threadIdx.x; b < kSeedBins; b += kBlock)
        if (s_hist[b]) atomicAdd(&hist[b], s_hist[b]);
}

// small[0] = first bin of the large seeds, small[1] = their count: the
// histogram is staged in LDS, then one thread walks it down from the top.
__global__ void __launch_bounds__(kBlock)
seed_split_bin_kernel(const int* __restrict__ hist, int target, int* __restrict__ small) {
    constexpr int kPer = kSeedBins / kBlock;  // bins per thread, consecutive
    __shared__ int s_hist[kSeedBins];
    __shared__ int s_part[kBlock];
    int part = 0;
    for (int e = 0; e < kPer; ++e) {
        const int b = threadIdx.x * kPer + e;
        s_hist[b] = hist[b];
        part += s_hist[b];
    }
    s_part[threadIdx.x] = part;
    __syncthreads();
    if (threadIdx.x != 0) return;
    // whole groups of kPer bins from the top while they fit, then single bins
    int group = kBlock;
    int count = 0;
    while (group > 0 && count + s_part[group - 1] <= target) count += s_part[--group];
    int bin = group * kPer;
    while (bin > 0 && count + s_hist[bin - 1] <= target) count += s_hist[--bin];
    small[0] = bin;
    small[1] = count;
}

__device__ __forceinline__ bool large_seed(const float* __restrict__ tseed, const int* __restrict__ order,
                                           int p, int bin) {
    return static_cast<int>(__float_as_uint(tseed[order[p]]) >> 19) >= bin;
}

// counts[b] = positions of block b (kSplitSpan of them) that stay in front.
__global__ void __launch_bounds__(kBlock)
split_count_kernel(const float* __restrict__ tseed, const int* __restrict__ order, int n,
                   const int* __restrict__ small, int* __restrict__ counts) {
    __shared__ int s_sum[kBlock / 64];
    const int bin = small[0];
    int c = 0;
    for (int e = 0; e < kSplitSpan / kBlock; ++e) {
        const int p = blockIdx.x * kSplitSpan + e * kBlock + threadIdx.x;
        c += p < n && !large_seed(tseed, order, p, bin);
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) c += __shfl_xor(c, off, 64);
    if ((threadIdx.x & 63) == 0) s_sum[threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x == 0) {
        int t = 0;
        for (int w = 0; w < kBlock / 64; ++w) t += s_sum[w];
        counts[blockIdx.x] = t;
    }
}

// Stable partition: the positions of block b that stay go to offsets[b] + their
// rank among them; the others to (n - n_large) + their rank among those.
__global__ void __launch_bounds__(kBlock)
split_scatter_kernel(const float* __restrict__ tseed, const int* __restrict__ order, int n,
                     const int* __restrict__ small, const int* __restrict__ offsets,
                     int* __restrict__ out) {
    __shared__ int s_wave[kBlock / 64];
    const int bin = small[0];
    const int front_total = n - small[1];
    constexpr int kPer = kSplitSpan / kBlock;  // consecutive positions per thread
    const int p0 = blockIdx.x * kSplitSpan + threadIdx.x * kPer;
    bool stay[kPer];
    int mine = 0;
#pragma unroll
    for (int e = 0; e < kPer; ++e) {
        const int p = p0 + e;
        stay[e] = p < n && !large_seed(tseed, order, p, bin);
        mine += stay[e];
    }
    // exclusive scan of `mine` over the block
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    int x = mine;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const int y = __shfl_up(x, off, 64);
        if (lane >= off) x += y;
    }
    if (lane == 63) s_wave[wave] = x;
    __syncthreads();
    int before = offsets[blockIdx.x];
    for (int w = 0; w < wave; ++w) before += s_wave[w];
    before += x - mine;  // kept positions before p0 in the whole order
#pragma unroll
    for (int e = 0; e < kPer; ++e) {
        const int p = p0 + e;
        if (p >= n) break;
        if (stay[e]) {
            out[before++] = order[p];
        } else {
            out[front_total + (p - before)] = order[p];
        }
    }
}

template <int DP, int STRIDE = DP>
void launch_order(const float* xs, int n, int* cell1, int* hist1, int* offset1, int* cursor1,
                  int* order1, float* piv2, int* key, int* hist, int* offset, int* cursor,
                  int* order, int* inv, hipStream_t s) {
    const unsigned grid = static_cast<unsigned>((n + kBlock - 1) / kBlock);
    zero_fill(hist1, sizeof(int) * kPiv, s);
    zero_fill(hist, sizeof(int) * kKeys, s);
    level1_kernel<DP, STRIDE><<<grid, kBlock, 0, s>>>(xs, n, cell1, hist1);
    scan_bins_kernel<<<1, kBlock, 0, s>>>(hist1, kPiv, offset1, cursor1);
    scatter_few_kernel<<<grid, kBlock, 0, s>>>(cell1, n, offset1, cursor1, order1);
    level2_pivots_kernel<<<kPiv, kBlock, 0, s>>>(xs, DP, STRIDE, order1, offset1, hist1, piv2);
    level2_kernel<DP, STRIDE><<<grid, kBlock, 0, s>>>(xs, n, order1, cell1, hist1, piv2, key, hist);
    scan_bins_kernel<<<1, kBlock, 0, s>>>(hist, kKeys, offset, cursor);
    scatter_kernel<<<grid, kBlock, 0, s>>>(key, n, offset, cursor, order);
    inverse_kernel<<<grid, kBlock, 0, s>>>(order, n, inv);
}

template <int DP, int STRIDE = DP>
void launch_seed(const float* xs_ord, const int* order, int n, int k, float* tseed, hipStream_t s) {
    constexpr int RPL = DP <= 8 ? 8 : (DP <= 32 ? 2 : 1);  // about 16 KiB of rows per chunk
    if (k > kListKMax) {
        // Large k: the k-th key of a 2048-candidate window (a window of 1024
        // leaves seeds several times the true k-th rank, and the candidate
        // sets grow with them).
        constexpr int W = 2048;
        constexpr int QPW = 2;
        const unsigned grid = static_cast<unsigned>((n + 4 * QPW - 1) / (4 * QPW));
        seed_kernel<DP, STRIDE, W, QPW, RPL><<<grid, kBlock, 0, s>>>(xs_ord, order, n, k, tseed);
        return;
    }
    // Narrow rows afford a wider window (tighter seeds) for the same work.
    constexpr int W = (DP <= 8 || DP > 32) ? 1024 : 2048;
    constexpr int QPW = DP <= 32 ? 2 : 4;
    const unsigned grid = static_cast<unsigned>((n + 4 * QPW - 1) / (4 * QPW));
    seed_kernel<DP, STRIDE, W, QPW, RPL><<<grid, kBlock, 0, s>>>(xs_ord, order, n, k, tseed);
}

}  // namespace

size_t order_scratch_bytes(int n, int dp) {
    const size_t nn = static_cast<size_t>(n);
    size_t b = 0;
    b += 6 * nn * sizeof(int);                            // cell1, order1, key, order, inv, spare
    b += static_cast<size_t>(kPiv) * kPiv * dp * sizeof(float);  // level-2 pivots
    b += (3 * kPiv + 3 * kKeys) * sizeof(int);            // histograms, offsets, cursors
    b += nn * dp * sizeof(float);                         // key rows in order
    b += 16 * 256;
    return b;
}

bool order_supported(int dp) {
    return dp == 4 || dp == 8 || dp == 16 || dp == 32 || dp == 64 || dp == 128 || dp == 256 || dp == 512;
}

CandidateOrder build_order(const float* xs, int n, int dp, void* scratch, hipStream_t s) {
    if (!order_supported(dp)) throw std::invalid_argument("build_order: unsupported row stride");
    unsigned char* p = static_cast<unsigned char*>(scratch);
    auto take = [&](size_t bytes) {
        void* r = p;
        p += (bytes + 255) / 256 * 256;
        return r;
    };
    const size_t nn = static_cast<size_t>(n);
    int* cell1 = static_cast<int*>(take(nn * sizeof(int)));
    int* order1 = static_cast<int*>(take(nn * sizeof(int)));
    int* key = static_cast<int*>(take(nn * sizeof(int)));
    CandidateOrder o;
    o.order = static_cast<int*>(take(nn * sizeof(int)));
    o.inv = static_cast<int*>(take(nn * sizeof(int)));
    float* piv2 = static_cast<float*>(take(static_cast<size_t>(kPiv) * kPiv * dp * sizeof(float)));
    int* hist1 = static_cast<int*>(take(kPiv * sizeof(int)));
    int* offset1 = static_cast<int*>(take(kPiv * sizeof(int)));
    int* cursor1 = static_cast<int*>(take(kPiv * sizeof(int)));
    int* hist = static_cast<int*>(take(kKeys * sizeof(int)));
    int* offset = static_cast<int*>(take(kKeys * sizeof(int)));
    int* cursor = static_cast<int*>(take(kKeys * sizeof(int)));
    o.xs_ord = static_cast<float*>(take(nn * dp * sizeof(float)));
    o.dp = dp;
    o.xs = xs;
    // free once the order is built: work space of order_large_seeds_last
    o.tmp_order = order1;
    o.tmp_counts = key;
    o.tmp_hist = hist;
    o.tmp_small = offset1;
    ProfileScope prof("order", s);
    switch (dp) {
        case 4: launch_order<4>(xs, n, cell1, hist1, offset1, cursor1, order1, piv2, key, hist, offset, cursor, o.order, o.inv, s); break;
        case 8: launch_order<8>(xs, n, cell1, hist1, offset1, cursor1, order1, piv2, key, hist, offset, cursor, o.order, o.inv, s); break;
        case 16: launch_order<16>(xs, n, cell1, hist1, offset1, cursor1, order1, piv2, key, hist, offset, cursor, o.order, o.inv, s); break;
        case 32: launch_order<32>(xs, n, cell1, hist1, offset1, cursor1, order1, piv2, key, hist, offset, cursor, o.order, o.inv, s); break;
        case 64: launch_order<64>(xs, n, cell1, hist1, offset1, cursor1, order1, piv2, key, hist, offset, cursor, o.order, o.inv, s); break;
        case 128: launch_order<128>(xs, n, cell1, hist1, offset1, cursor1, order1, piv2, key, hist, offset, cursor, o.order, o.inv, s); break;
        case 256: launch_order<128, 256>(xs, n, cell1, hist1, offset1, cursor1, order1, piv2, key, hist, offset, cursor, o.order, o.inv, s); break;
        default: launch_order<128, 512>(xs, n, cell1, hist1, offset1, cursor1, order1, piv2, key, hist, offset, cursor, o.order, o.inv, s); break;
    }
    const size_t cells4 = nn * (dp / 4);
    gather_rows_kernel<<<static_cast<unsigned>((cells4 + kBlock - 1) / kBlock), kBlock, 0, s>>>(
        xs, dp, o.order, n, o.xs_ord);
    CULOF_CHECK(hipGetLastError());
    return o;
}

void order_large_seeds_last(CandidateOrder& o, int n, const float* tseed, hipStream_t s) {
    ProfileScope prof("order_split", s);
    const int target = n / 100;
    if (target < 1) return;
    const unsigned grid = static_cast<unsigned>((n + kBlock - 1) / kBlock);
    const int blocks = (n + kSplitSpan - 1) / kSplitSpan;
    zero_fill(o.tmp_hist, sizeof(int) * kSeedBins, s);
    seed_hist_kernel<<<blocks, kBlock, 0, s>>>(tseed, n, o.tmp_hist);
    seed_split_bin_kernel<<<1, kBlock, 0, s>>>(o.tmp_hist, target, o.tmp_small);
    int* counts = o.tmp_counts;
    int* offsets = o.tmp_counts + blocks;
    split_count_kernel<<<blocks, kBlock, 0, s>>>(tseed, o.order, n, o.tmp_small, counts);
    scan_bins_kernel<<<1, kBlock, 0, s>>>(counts, blocks, offsets, o.tmp_counts + 2 * blocks);
    split_scatter_kernel<<<blocks, kBlock, 0, s>>>(tseed, o.order, n, o.tmp_small, offsets, o.tmp_order);
    CULOF_CHECK(hipGetLastError());
    std::swap(o.order, o.tmp_order);
    inverse_kernel<<<grid, kBlock, 0, s>>>(o.order, n, o.inv);
    const size_t cells4 = static_cast<size_t>(n) * (o.dp / 4);
    gather_rows_kernel<<<static_cast<unsigned>((cells4 + kBlock - 1) / kBlock), kBlock, 0, s>>>(
        o.xs, o.dp, o.order, n, o.xs_ord);
    CULOF_CHECK(hipGetLastError());
}

void seed_thresholds(const CandidateOrder& o, int n, int key_dp, int k, float* tseed, hipStream_t s) {
    ProfileScope prof("seed", s);
    if (o.dp == 16 && key_dp < 16) {  // narrow rows padded for the BF16 screen: key the real width
        if (key_dp == 4) {
            launch_seed<4, 16>(o.xs_ord, o.order, n, k, tseed, s);
        } else {
            launch_seed<8, 16>(o.xs_ord, o.order, n, k, tseed, s);
        }
        CULOF_CHECK(hipGetLastError());
        return;
    }
    switch (o.dp) {
        case 4: launch_seed<4>(o.xs_ord, o.order, n, k, tseed, s); break;
        case 8: launch_seed<8>(o.xs_ord, o.order, n, k, tseed, s); break;
        case 16: launch_seed<16>(o.xs_ord, o.order, n, k, tseed, s); break;
        case 32: launch_seed<32>(o.xs_ord, o.order, n, k, tseed, s); break;
        case 64: launch_seed<64>(o.xs_ord, o.order, n, k, tseed, s); break;
        case 128: launch_seed<128>(o.xs_ord, o.order, n, k, tseed, s); break;
        case 256: launch_seed<256>(o.xs_ord, o.order, n, k, tseed, s); break;
        default: launch_seed<512>(o.xs_ord, o.order, n, k, tseed, s); break;
    }
    CULOF_CHECK(hipGetLastError());
}

}  // namespace detail
}  // namespace culof
