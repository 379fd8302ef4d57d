// Fused distance screen + exact k-nearest-neighbour sweep (the production
// selection for d <= 128, k <= 64).
//
// What it computes is the reference's selection (src/cuda/knn.cu:119-232 over
// the distance tile of src/cuda/pipeline.cu:175-195): for each query row i,
// the k smallest keys, ties at the k-th key broken toward the lower index, the
// k-th key itself as the squared k-distance, and the neighbours written in the
// reference's slot order. The key is the difference-form FP32 squared distance
// e(i, j) = sum_t (a_t - b_t)^2 over the key rows (+inf on the diagonal).
//
// How it computes it is MI355X-first: no distance tile exists, so the stage
// reads the points and writes O(n k) results and nothing else from HBM.
//
//   * A lane owns a query row; its rows live in VGPRs. A wave holds 64
//     queries; the 4 waves of a block share the same 64 queries and split the
//     candidate chunks between them.
//   * Candidates stream through LDS in chunks (one private region per wave,
//     register double-buffered) and are read back as wave-wide broadcasts.
//   * Screen: the reference's Gram expansion fma(-2, x_i.x_j, |x_i|^2 + |x_j|^2)
//     over the centred rows, d FMAs + 1 add + 1 FMA + 1 compare per pair. Its
//     error against e is at most c (|x_i|^2 + |x_j|^2) + c e with
//     c = (d + 16) 2^-24 (derivation below), so a candidate can enter the
//     top-k only if  raw - c |x_j|^2 <= T (1 + c) + c |x_i|^2,  T the lane's
//     current k-th key. The c |x_j|^2 term is folded into the staged norm.
//   * Admission: a candidate that passes the screen gets its exact key e and,
//     if (e, j) beats the lane's k-th entry, goes into a small per-lane LDS
//     ring. The ring is merged into the lane's list only when some lane's ring
//     is nearly full, so the wave pays for the divergent insert rarely.
//   * The list is exact: sorted by the 64-bit word (e bits << 32 | j), which
//     is the (key, index) order of the reference's tie rule.
//   * Every wave publishes its k-th key in LDS; each wave's list holds k keys
//     at or below it, so the smallest published key bounds the final k-th key
//     and all waves screen against it.
//   * Chunks are visited starting at the block's own rows (data stored in
//     locality order tightens the threshold from the first chunk).
//   * Wave 0 merges the other lists and writes the neighbours in slot order.
//
// Screen bound. With u = 2^-24, N = exact squared norms of the centred rows,
// D = exact squared distance of the key rows: the FMA chain of x_i.x_j errs by
// at most d u (N_i + N_j) / 2 (1 + O(u)); the norms, the add and the final FMA
// each add at most u of their magnitudes; centring the key rows moves D by at
// most u (2 D + N_i + N_j); and e differs from D by at most (d + 2) u e. Summed
// with slack for the rounding of the threshold arithmetic itself this stays
// below (d + 16) u (N_i + N_j + T). The CPU tests check the inequality on
// adversarial data (far-off clusters, duplicates, cancellation).

#include "internal.hpp"

#include <stdexcept>

namespace culof {
namespace detail {

namespace {

constexpr uint64_t kSentinel = ~0ull;  // above every real (key, index) word
constexpr int kWaves = 4;              // waves per block, all on the same 64 queries
constexpr int kGroup = 8;              // candidates screened per branch
constexpr int kRing = 16;              // ring entries per lane (> kGroup)

__device__ __forceinline__ uint64_t pack(uint32_t key, uint32_t idx) {
    return (static_cast<uint64_t>(key) << 32) | idx;
}

__device__ __forceinline__ float key_as_float(uint64_t w) {
    const uint32_t k = static_cast<uint32_t>(w >> 32);
    return __uint_as_float(k > kInfBits ? kInfBits : k);
}

// Orders this wave's LDS writes before its later LDS reads by other lanes.
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <int DP>
struct SweepShape {
    // Candidates per chunk: 64 for narrow rows, fewer for wide ones so the
    // per-wave staging regions stay small.
    static constexpr int kChunk = DP <= 16 ? 64 : (DP <= 32 ? 32 : (DP <= 64 ? 16 : 8));
    static constexpr int kVec = kChunk * DP / 4 / 64;  // float4 loads per lane per row set
    static_assert(kVec >= 1, "chunk too small for one float4 per lane");
    // Query key rows stay in VGPRs when narrow; wide ones are re-read from
    // global memory on the (rare) admission path.
    static constexpr bool kKeyRowInRegs = DP <= 16;
};

// The list is kept sorted in descending order with its k valid entries in
// slots [0, k) and a low dummy (0) in slots [k, KMAX), so the current k-th
// entry -- the admission threshold -- is always a[0]. Every register index is
// a compile-time constant: a runtime a[k - 1] would push the list to scratch.
template <int KMAX>
__device__ __forceinline__ void list_init(uint64_t (&a)[KMAX], int k) {
#pragma unroll
    for (int t = 0; t < KMAX; ++t) a[t] = (t < k) ? kSentinel : 0ull;
}

// Replace the largest valid entry by x (x < a[0]) and restore the order with
// one compare-exchange pass; the dummies stop x at the end of the valid run.
template <int KMAX>
__device__ __forceinline__ void list_insert(uint64_t (&a)[KMAX], uint64_t x) {
    a[0] = x;
#pragma unroll
    for (int t = 0; t + 1 < KMAX; ++t) {
        const uint64_t hi = a[t] < a[t + 1] ? a[t + 1] : a[t];
        const uint64_t lo = a[t] < a[t + 1] ? a[t] : a[t + 1];
        a[t] = hi;
        a[t + 1] = lo;
    }
}

template <int DP, int KMAX>
__global__ void __launch_bounds__(64 * kWaves)
sweep_kernel(const float* __restrict__ xs, const float* __restrict__ xc,
             const float* __restrict__ norms, const int* __restrict__ guard, int n, int k,
             float screen_c, int row0, int rows, int* __restrict__ nbr_idx,
             float* __restrict__ nbr_d2, float* __restrict__ kdist2, int* __restrict__ fb_rows,
             int* __restrict__ fb_count, int fb_cap, const float* __restrict__ qxs, const float* __restrict__ qxc,
             const float* __restrict__ qnorms, int same_set, int out_stride) {
    using Shape = SweepShape<DP>;
    constexpr int CH = Shape::kChunk;
    constexpr int NV = Shape::kVec;
    constexpr int V4 = DP / 4;  // float4s per row

    // LDS carve-up. The sweep-time regions and the merge region alias.
    constexpr size_t kRowsBytes = sizeof(float) * kWaves * CH * DP;  // one row set
    constexpr size_t kNormBytes = sizeof(float) * kWaves * CH;
    constexpr size_t kRingBytes = sizeof(uint64_t) * kWaves * kRing * 64;
    constexpr size_t kSweepBytes = 2 * kRowsBytes + kNormBytes + kRingBytes;
    constexpr size_t kMergeBytes = sizeof(uint64_t) * KMAX * 64;
    constexpr size_t kUnion = kSweepBytes > kMergeBytes ? kSweepBytes : kMergeBytes;
    __shared__ __attribute__((aligned(16))) unsigned char smem[kUnion];
    __shared__ uint32_t s_tkey[kWaves][64];

    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    float4* s_cc = reinterpret_cast<float4*>(smem) + static_cast<size_t>(wave) * CH * V4;
    float4* s_cs = reinterpret_cast<float4*>(smem + kRowsBytes) + static_cast<size_t>(wave) * CH * V4;
    float* s_cm = reinterpret_cast<float*>(smem + 2 * kRowsBytes) + wave * CH;
    uint64_t* s_ring = reinterpret_cast<uint64_t*>(smem + 2 * kRowsBytes + kNormBytes) + wave * kRing * 64;
    uint64_t* s_merge = reinterpret_cast<uint64_t*>(smem);

    const int local = blockIdx.x * 64 + lane;
    const bool valid = local < rows;
    const int qi = row0 + (valid ? local : 0);

    if (*guard) {
        // Some norm is too large for the screen to stay finite: the exact
        // radix path takes every row.
        if (wave == 0 && valid) queue_push(fb_rows, fb_count, fb_cap, qi);
        return;
    }

    // Query rows in registers.
    float qc[DP];
    float qs[Shape::kKeyRowInRegs ? DP : 1];
    {
        const float4* src = reinterpret_cast<const float4*>(qxc + static_cast<size_t>(qi) * DP);
#pragma unroll
        for (int v = 0; v < V4; ++v) {
            const float4 t = src[v];
            qc[4 * v + 0] = t.x;
            qc[4 * v + 1] = t.y;
            qc[4 * v + 2] = t.z;
            qc[4 * v + 3] = t.w;
        }
        if constexpr (Shape::kKeyRowInRegs) {
            const float4* srs = reinterpret_cast<const float4*>(qxs + static_cast<size_t>(qi) * DP);
#pragma unroll
            for (int v = 0; v < V4; ++v) {
                const float4 t = srs[v];
                qs[4 * v + 0] = t.x;
                qs[4 * v + 1] = t.y;
                qs[4 * v + 2] = t.z;
                qs[4 * v + 3] = t.w;
            }
        }
    }
    const float nq = qnorms[qi];
    const float cq = screen_c * nq;
    const float up = 1.0f + 4.0f * 5.9604645e-8f;  // rounds the screen threshold upward

    uint64_t a[KMAX];
    list_init<KMAX>(a, k);
    // Screen threshold; invalid lanes never pass (nothing is below -inf).
    float Tq = valid ? __uint_as_float(kInfBits) : -__uint_as_float(kInfBits);
    int cnt = 0;
    s_tkey[wave][lane] = kInfBits;
    __syncthreads();  // every bound initialised before any wave reads one

    const int n_chunks = (n + CH - 1) / CH;
    // Rotate the start so neighbouring blocks read different chunks first (a
    // block of the candidate set starts at its own rows). Query rows of
    // another set may outnumber the candidates: the start is taken modulo the
    // chunk count so every block still visits each chunk exactly once.
    const int c0 = ((row0 + static_cast<int>(blockIdx.x) * 64) / CH) % n_chunks;

    // Register prefetch of a chunk's rows and adjusted norms.
    float4 pre_c[NV];
    float4 pre_s[NV];
    float pre_m = 0.0f;
    auto load_chunk = [&](int m) {
        int chunk = c0 + m;
        if (chunk >= n_chunks) chunk -= n_chunks;
        const int base = chunk * CH;
#pragma unroll
        for (int u = 0; u < NV; ++u) {
            const int v = lane + 64 * u;
            const int c = v / V4;
            const int part = v - c * V4;
            const int j = base + c;
            const size_t o = static_cast<size_t>(j) * V4 + part;
            pre_c[u] = (j < n) ? reinterpret_cast<const float4*>(xc)[o] : make_float4(0.f, 0.f, 0.f, 0.f);
            pre_s[u] = (j < n) ? reinterpret_cast<const float4*>(xs)[o] : make_float4(0.f, 0.f, 0.f, 0.f);
        }
        if (lane < CH) {
            const int j = base + lane;
            // |x_j|^2 (1 - c): the candidate's share of the screen margin.
            // Past the end: NaN, which fails every compare.
            pre_m = (j < n) ? norms[j] * (1.0f - screen_c) : __uint_as_float(0x7FC00000u);
        }
        return base;
    };

    auto merge_ring = [&]() {
        for (int r = 0; r < kRing; ++r) {
            if (!__any(r < cnt)) break;
            if (r < cnt) {
                const uint64_t p = s_ring[r * 64 + lane];
                if (p < a[0]) list_insert<KMAX>(a, p);
            }
        }
        cnt = 0;
        const uint32_t tk = static_cast<uint32_t>(a[0] >> 32);
        s_tkey[wave][lane] = tk > kInfBits ? kInfBits : tk;
    };

    auto refresh_bound = [&]() {
        if (!valid) return;
        float t = key_as_float(a[0]);
#pragma unroll
        for (int w = 0; w < kWaves; ++w) t = fminf(t, __uint_as_float(s_tkey[w][lane]));
        // T (1 + c) + c |x_i|^2, rounded up; +inf stays +inf.
        Tq = __builtin_fmaf(t, 1.0f + screen_c, cq) * up;
    };

    int m = wave;
    int base = (m < n_chunks) ? load_chunk(m) : 0;
    for (; m < n_chunks; m += kWaves) {
        // Stage the prefetched chunk in this wave's LDS regions.
#pragma unroll
        for (int u = 0; u < NV; ++u) {
            s_cc[lane + 64 * u] = pre_c[u];
            s_cs[lane + 64 * u] = pre_s[u];
        }
        if (lane < CH) s_cm[lane] = pre_m;
        wave_lds_sync();
        const int cur_base = base;
        if (m + kWaves < n_chunks) base = load_chunk(m + kWaves);  // next chunk in flight
        refresh_bound();

        constexpr int G = CH < kGroup ? CH : kGroup;
        for (int c = 0; c < CH; c += G) {
            if (__any(cnt > kRing - G)) {
                merge_ring();
                refresh_bound();
            }
            // Screen G candidates branch-free: G independent FMA chains whose
            // LDS reads can all be in flight together.
            float raw[G];
#pragma unroll
            for (int g = 0; g < G; ++g) {
                const float4* cv = s_cc + (c + g) * V4;
                float acc = 0.0f;
#pragma unroll
                for (int v = 0; v < V4; ++v) {
                    const float4 t = cv[v];
                    acc = __builtin_fmaf(qc[4 * v + 0], t.x, acc);
                    acc = __builtin_fmaf(qc[4 * v + 1], t.y, acc);
                    acc = __builtin_fmaf(qc[4 * v + 2], t.z, acc);
                    acc = __builtin_fmaf(qc[4 * v + 3], t.w, acc);
                }
                raw[g] = __builtin_fmaf(-2.0f, acc, nq + s_cm[c + g]);
            }
            float lo = raw[0];
#pragma unroll
            for (int g = 1; g < G; ++g) lo = fminf(lo, raw[g]);  // NaN (past the end) drops out
            if (lo <= Tq) {
#pragma unroll
                for (int g = 0; g < G; ++g) {
                    if (raw[g] <= Tq) {
                        // Exact key: difference form over the key rows.
                        const int j = cur_base + c + g;
                        const float4* sv = s_cs + (c + g) * V4;
                        const float4* qv = reinterpret_cast<const float4*>(qxs + static_cast<size_t>(qi) * DP);
                        float e = 0.0f;
#pragma unroll
                        for (int v = 0; v < V4; ++v) {
                            const float4 t = sv[v];
                            float4 q4;
                            if constexpr (Shape::kKeyRowInRegs) {
                                q4 = make_float4(qs[4 * v + 0], qs[4 * v + 1], qs[4 * v + 2], qs[4 * v + 3]);
                            } else {
                                q4 = qv[v];
                            }
                            float df = q4.x - t.x;
                            e = __builtin_fmaf(df, df, e);
                            df = q4.y - t.y;
                            e = __builtin_fmaf(df, df, e);
                            df = q4.z - t.z;
                            e = __builtin_fmaf(df, df, e);
                            df = q4.w - t.w;
                            e = __builtin_fmaf(df, df, e);
                        }
                        const uint32_t kb = ((same_set && j == qi) || e != e) ? kInfBits : __float_as_uint(e);
                        const uint64_t p = pack(kb, static_cast<uint32_t>(j));
                        if (p < a[0]) {
                            s_ring[cnt * 64 + lane] = p;
                            ++cnt;
                        }
                    }
                }
            }
        }
        wave_lds_sync();
    }
    merge_ring();

    // ---- merge the W lists into wave 0 ----
    __syncthreads();  // every wave done with its sweep regions
    for (int w = 1; w < kWaves; ++w) {
        if (wave == w) {
#pragma unroll
            for (int t = 0; t < KMAX; ++t) s_merge[t * 64 + lane] = a[t];
        }
        __syncthreads();
        if (wave == 0) {
            // Smallest first (the lists are descending), so the scan can stop
            // as soon as no lane has an entry below its threshold.
            for (int t = k - 1; t >= 0; --t) {
                const uint64_t e = s_merge[t * 64 + lane];
                if (!__any(e < a[0])) break;
                if (e < a[0]) list_insert<KMAX>(a, e);
            }
        }
        __syncthreads();
    }
    if (wave != 0) return;

    // ---- emit in the reference's slot order ----
    // Stage the valid entries ascending: slot t <- a[k - 1 - t].
#pragma unroll
    for (int t = 0; t < KMAX; ++t) {
        if (t < k) s_merge[(k - 1 - t) * 64 + lane] = a[t];
    }
    wave_lds_sync();
    if (!valid) return;
    const uint32_t kth = static_cast<uint32_t>(a[0] >> 32);
    if (kth >= kInfBits) {
        // Fewer than k finite keys: NaN keys (read as +inf, like the
        // reference's) may then belong to the set, and the screen never admits
        // them. The exact radix path redoes the row.
        queue_push(fb_rows, fb_count, fb_cap, qi);
        return;
    }
    kdist2[qi] = __uint_as_float(kth);
    // Entries are sorted by (key, index): the ones below the k-th key form a
    // prefix [0, below); the tied ones follow already in ascending index, so
    // their slot is their position. The prefix is re-ranked by index.
    int below = 0;
    for (int t = 0; t < k; ++t) below += static_cast<uint32_t>(s_merge[t * 64 + lane] >> 32) < kth;
    const size_t ns = static_cast<size_t>(out_stride);
    for (int e = 0; e < k; ++e) {
        const uint64_t w = s_merge[e * 64 + lane];
        const uint32_t idx = static_cast<uint32_t>(w);
        int slot = e;
        if (e < below) {
            slot = 0;
            for (int f = 0; f < below; ++f) slot += static_cast<uint32_t>(s_merge[f * 64 + lane]) < idx;
        }
        nbr_idx[static_cast<size_t>(slot) * ns + qi] = static_cast<int>(idx);
        nbr_d2[static_cast<size_t>(slot) * ns + qi] = __uint_as_float(static_cast<uint32_t>(w >> 32));
    }
}

struct SweepArgs {
    const float* xs;
    const float* xc;
    const float* norms;
    const int* guard;
    int n;
    int k;
    float screen_c;
    int row0;
    int rows;
    int* nbr_idx;
    float* nbr_d2;
    float* kdist2;
    int* fb_rows;
    int* fb_count;
    int fb_cap;  // entries fb_rows holds
    // Query rows: the candidates themselves (same_set, the point itself
    // excluded, outputs at stride n) or another row set (outputs at stride rows).
    const float* qxs;
    const float* qxc;
    const float* qnorms;
    int same_set;
    int out_stride;
};

template <int DP, int KMAX>
void launch_sweep(const SweepArgs& a, hipStream_t stream) {
    const int grid = (a.rows + 63) / 64;
    sweep_kernel<DP, KMAX><<<grid, 64 * kWaves, 0, stream>>>(
        a.xs, a.xc, a.norms, a.guard, a.n, a.k, a.screen_c, a.row0, a.rows, a.nbr_idx, a.nbr_d2,
        a.kdist2, a.fb_rows, a.fb_count, a.fb_cap, a.qxs, a.qxc, a.qnorms, a.same_set, a.out_stride);
}

template <int DP>
bool dispatch_k(const SweepArgs* args, int k, hipStream_t stream) {
    if (k <= 8) {
        if (args) launch_sweep<DP, 8>(*args, stream);
        return true;
    }
    if (k <= 16) {
        if (args) launch_sweep<DP, 16>(*args, stream);
        return true;
    }
    if (k <= 32) {
        if (args) launch_sweep<DP, 32>(*args, stream);
        return true;
    }
    if constexpr (DP <= 64) {
        if (k <= 64) {
            if (args) launch_sweep<DP, 64>(*args, stream);
            return true;
        }
    }
    return false;
}

// With args null, only reports whether an instantiation exists.
bool dispatch(int dp, int k, const SweepArgs* args, hipStream_t stream) {
    switch (dp) {
        case 4: return dispatch_k<4>(args, k, stream);
        case 8: return dispatch_k<8>(args, k, stream);
        case 16: return dispatch_k<16>(args, k, stream);
        case 32: return dispatch_k<32>(args, k, stream);
        case 64: return dispatch_k<64>(args, k, stream);
        case 128: return dispatch_k<128>(args, k, stream);
        default: return false;
    }
}

}  // namespace

int padded_dims(int n_dims) {
    // Powers of two up to 512: the widths the matrix-core pipeline is built for.
    for (int p : {4, 8, 16, 32, 64, 128, 256, 512}) {
        if (n_dims <= p) return p;
    }
    return (n_dims + 3) / 4 * 4;
}

bool sweep_supported(int n_dims, int k) { return dispatch(padded_dims(n_dims), k, nullptr, nullptr); }

void sweep_knn(const float* keyrows, const float* centred, int stride, const float* norms,
               const int* guard, int n_points, int n_dims, int k, int row0, int rows,
               int* nbr_idx, float* nbr_d2, float* kdist2, int* fallback_rows,
               int* fallback_count, int fallback_cap, hipStream_t stream) {
    if (stride != padded_dims(n_dims)) throw std::invalid_argument("sweep_knn: row stride mismatch");
    if (rows <= 0) return;
    ProfileScope prof("sweep_knn", stream);
    const SweepArgs args{keyrows, centred, norms, guard, n_points, k,
                         static_cast<float>(n_dims + 16) * 5.9604645e-8f,  // (d + 16) 2^-24
                         row0, rows, nbr_idx, nbr_d2, kdist2, fallback_rows, fallback_count,
                         fallback_cap, keyrows, centred, norms, 1, n_points};
    if (!dispatch(stride, k, &args, stream)) throw std::invalid_argument("sweep_knn: unsupported (n_dims, k)");
    CULOF_CHECK(hipGetLastError());
}

void sweep_query_knn(const float* qkeys, const float* qcentred, const float* qnorms, int n_queries,
                     const float* keyrows, const float* centred, int stride, const float* norms,
                     const int* guard, int n_points, int n_dims, int k, int* nbr_idx, float* nbr_d2,
                     float* kdist2, int* fallback_rows, int* fallback_count, int fallback_cap,
                     hipStream_t stream) {
    if (stride != padded_dims(n_dims)) throw std::invalid_argument("sweep_query_knn: row stride mismatch");
    if (n_queries <= 0) return;
    ProfileScope prof("sweep_knn", stream);
    const SweepArgs args{keyrows, centred, norms, guard, n_points, k,
                         static_cast<float>(n_dims + 16) * 5.9604645e-8f,  // (d + 16) 2^-24
                         0, n_queries, nbr_idx, nbr_d2, kdist2, fallback_rows, fallback_count,
                         fallback_cap, qkeys, qcentred, qnorms, 0, n_queries};
    if (!dispatch(stride, k, &args, stream)) throw std::invalid_argument("sweep_query_knn: unsupported (n_dims, k)");
    CULOF_CHECK(hipGetLastError());
}

}  // namespace detail
}  // namespace culof
