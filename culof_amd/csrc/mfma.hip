// MFMA-screened exact kNN: the production selection for every d <= 128 and
// k <= 64, in two precision forms of the screen.
//
// Same result as sweep.hip and radix.hip -- the k smallest difference-form
// FP32 keys e(i, j) = sum_t (a_t - b_t)^2 per query, ties toward the lower
// index, slot order of the reference (src/cuda/knn.cu:190-224) -- with the
// distance screen on the matrix cores:
//
//   * Per 32 x 32 tile the screen is one chain of MFMAs, C = m + A B with
//     A = candidate rows (32 x DP), B = -2 x query rows (DP x 32) and the
//     accumulator pre-loaded with m_i = |x_i|^2 (1 - c). Form FP32 uses
//     v_mfma_f32_32x32x2_f32 on the centred rows; form BF16 uses
//     v_mfma_f32_32x32x16_bf16 on split hi / lo copies (products exact,
//     FP32 accumulation). The query's own |x_q|^2 is folded into the
//     threshold, so the VALU does one min per pair and one compare per tile.
//     A wave owns 32 queries (lanes l and l + 32 serve query l % 32); the 4
//     waves of a block stream the same candidate tiles through LDS.
//   * A pair passes when acc <= T (1 + c) + c |x_q|^2 - |x_q|^2, c the form's
//     relative error bound (below), so a candidate with exact key <= T always
//     passes. Three ways use it:
//     - mfma_sym_collect_kernel (full runs of 32,768 rows or more): T is a
//       seed (an upper bound of the k-th key), each unordered pair of tiles is
//       computed once and tested against both rows' thresholds, and passing
//       candidates are appended to per-row buffers;
//     - mfma_collect_kernel (other runs, the second pass, near-tie rows): the
//       same against every candidate tile of the query's slice;
//     - mfma_sweep_kernel (CULOF_ORDER=0, and the list pass for rows whose
//       candidate sets outgrow the buffers): T is the lane's current k-th key,
//       passing candidates are queued in a per-lane LDS ring and inserted into
//       a sorted register list by exact key, slices merged by
//       merge_slices_kernel.
//     The collect kernels leave the exact keys to the select kernels
//     (collect_select_kernel, big_select_kernel).
//
// Screen bound. Let u = 2^-24, N the exact squared norms of the centred rows.
// The accumulated value differs from m_i - 2 x_q.x_i by at most 2 g (N_q + N_i),
// g = (DP + 8) u covering any FP32 accumulation order of the MFMA chain and the
// product roundings. The BF16 split form adds 3.1 2^-16 (N_q + N_i) for the
// terms it drops (see ScreenBF16) and sums 3 DP products. With the norm,
// centring and exact-key roundings (about 3.2 u N and (d + 5) u T) this gives
//   FP32: c = (2 DP + 32) u,   BF16 split: c = 3.1 2^-16 + (12 DP + 64) u,
// and the threshold arithmetic adds a 4 u (T + |x_q|^2) margin for its own
// rounding. A true top-k candidate therefore always passes; the exact keys
// decide, so every form selects the same neighbours.

#include "internal.hpp"

#include <algorithm>
#include <cstdlib>
#include <stdexcept>
#include <type_traits>

namespace culof {
namespace detail {

namespace {

typedef float floatx16 __attribute__((ext_vector_type(16)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

constexpr uint64_t kSentinel = ~0ull;
constexpr int kTilesPerBlock = 4;  // query tiles (waves) per block
constexpr int kRing = 32;          // queued screen passes per lane
constexpr float kU = 5.9604645e-8f;  // 2^-24
constexpr int kWideMin = 256;    // padded widths from here on take the K-chunked collect pass
constexpr int kWideChunk = 64;   // features per K chunk of that pass

__device__ __forceinline__ uint64_t pack(uint32_t key, uint32_t idx) {
    return (static_cast<uint64_t>(key) << 32) | idx;
}

__device__ __forceinline__ float key_as_float(uint64_t w) {
    const uint32_t k = static_cast<uint32_t>(w >> 32);
    return __uint_as_float(k > kInfBits ? kInfBits : k);
}

__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Sorted register list, descending, threshold at a[0] (see sweep.hip).
// Lanes below this one whose bit is set in a wave-wide mask.
__device__ __forceinline__ int lanes_below(uint64_t mask) {
    return __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(mask >> 32),
                                     __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(mask), 0u));
}

template <int KMAX>
__device__ __forceinline__ void list_init(uint64_t (&a)[KMAX], int k) {
#pragma unroll
    for (int t = 0; t < KMAX; ++t) a[t] = (t < k) ? kSentinel : 0ull;
}

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

// ---- screen operand formats ----

// FP32: row i stored as [x_0, x_2, ..., x_{DP-2}, x_1, x_3, ..., x_{DP-1}], so
// the lane serving row i, k-half h, reads DP/2 contiguous floats for its
// DP/2 MFMAs (instruction m consumes k = 2m + h).
struct ScreenF32 {
    template <int DP>
    struct Frag {
        float v[DP / 2];
    };
    using Elem = float;
    static constexpr int kRowMul = 1;  // stored elements per feature
    // This lane's half of a row: DP / 2 contiguous floats.
    template <int DP>
    __device__ static void load_a(const Elem* half_row, Frag<DP>& f) {
        if constexpr (DP % 8 == 0) {
            const float4* p = reinterpret_cast<const float4*>(half_row);
#pragma unroll
            for (int v = 0; v < DP / 8; ++v) {
                const float4 t = p[v];
                f.v[4 * v + 0] = t.x;
                f.v[4 * v + 1] = t.y;
                f.v[4 * v + 2] = t.z;
                f.v[4 * v + 3] = t.w;
            }
        } else {  // DP == 4: two floats
            const float2 t = reinterpret_cast<const float2*>(half_row)[0];
            f.v[0] = t.x;
            f.v[1] = t.y;
        }
    }
    __device__ static constexpr int half_offset(int dp, int h) { return h * (dp / 2); }
    // B fragment of the query: -2 x_q[2m + h] (natural-order centred row).
    template <int DP>
    __device__ static void load_b(const float* xc_row, int h, Frag<DP>& f) {
#pragma unroll
        for (int m = 0; m < DP / 2; ++m) f.v[m] = -2.0f * xc_row[2 * m + h];
    }
    template <int DP>
    __device__ static floatx16 run(const Frag<DP>& a, const Frag<DP>& b, floatx16 acc) {
#pragma unroll
        for (int m = 0; m < DP / 2; ++m) acc = __builtin_amdgcn_mfma_f32_32x32x2f32(a.v[m], b.v[m], acc, 0, 0, 0);
        return acc;
    }
    // The same products for two query fragments, the A values read from LDS
    // step by step (one A fragment never sits whole in registers).
    template <int DP>
    __device__ static void run2(const Elem* half_row, const Frag<DP>& b0, const Frag<DP>& b1, floatx16& acc0,
                                floatx16& acc1) {
        static_assert(DP % 8 == 0, "float4 steps");
        const float4* p = reinterpret_cast<const float4*>(half_row);
#pragma unroll
        for (int v = 0; v < DP / 8; ++v) {
            const float4 t = p[v];
            const float a[4] = {t.x, t.y, t.z, t.w};
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                acc0 = __builtin_amdgcn_mfma_f32_32x32x2f32(a[i], b0.v[4 * v + i], acc0, 0, 0, 0);
                acc1 = __builtin_amdgcn_mfma_f32_32x32x2f32(a[i], b1.v[4 * v + i], acc1, 0, 0, 0);
            }
        }
    }
    __host__ __device__ static float slack(int dp) { return (2.0f * dp + 32.0f) * kU; }
};

// BF16 split: every operand x is carried as x_hi = bf16(x) and
// x_lo = bf16(x - x_hi) (the residual is exact in FP32), and each K step runs
// three MFMAs, a_hi b_hi + a_hi b_lo + a_lo b_hi, accumulating in FP32. The
// products of bf16 values are exact; the dropped a_lo b_lo term and the two
// second roundings leave about 3 2^-16 |a_t b_t| per feature, against 2^-7
// for plain bf16 products, at three BF16 MFMAs per step (5.3x the FP32
// matrix rate for the same K). Row i of the operand holds DP hi values then
// DP lo values, natural feature order; instruction m consumes k in
// [16m, 16m + 16), the lane of k-half h the 8 values at 16m + 8h.
struct ScreenBF16 {
    template <int DP>
    struct Frag {
        bf16x8 hi[DP / 16];
        bf16x8 lo[DP / 16];
    };
    using Elem = __bf16;
    static constexpr int kRowMul = 2;  // stored elements per feature
    template <int DP>
    __device__ static void load_a(const Elem* half_row, Frag<DP>& f) {
        const bf16x8* p = reinterpret_cast<const bf16x8*>(half_row);
#pragma unroll
        for (int m = 0; m < DP / 16; ++m) {
            f.hi[m] = p[2 * m];
            f.lo[m] = p[DP / 8 + 2 * m];
        }
    }
    __device__ static constexpr int half_offset(int, int h) { return 8 * h; }
    template <int DP>
    __device__ static void load_b(const float* xc_row, int h, Frag<DP>& f) {
#pragma unroll
        for (int m = 0; m < DP / 16; ++m) {
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                const float v = -2.0f * xc_row[16 * m + 8 * h + e];
                const __bf16 hi = static_cast<__bf16>(v);
                f.hi[m][e] = hi;
                f.lo[m][e] = static_cast<__bf16>(v - static_cast<float>(hi));
            }
        }
    }
    template <int DP>
    __device__ static floatx16 run(const Frag<DP>& a, const Frag<DP>& b, floatx16 acc) {
#pragma unroll
        for (int m = 0; m < DP / 16; ++m) {
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a.lo[m], b.hi[m], acc, 0, 0, 0);
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a.hi[m], b.lo[m], acc, 0, 0, 0);
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a.hi[m], b.hi[m], acc, 0, 0, 0);
        }
        return acc;
    }
    // The same products for two query fragments (each in the order of run),
    // the A values read from LDS step by step.
    template <int DP>
    __device__ static void run2(const Elem* half_row, const Frag<DP>& b0, const Frag<DP>& b1, floatx16& acc0,
                                floatx16& acc1) {
        const bf16x8* p = reinterpret_cast<const bf16x8*>(half_row);
#pragma unroll
        for (int m = 0; m < DP / 16; ++m) {
            const bf16x8 ahi = p[2 * m];
            const bf16x8 alo = p[DP / 8 + 2 * m];
            acc0 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(alo, b0.hi[m], acc0, 0, 0, 0);
            acc1 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(alo, b1.hi[m], acc1, 0, 0, 0);
            acc0 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ahi, b0.lo[m], acc0, 0, 0, 0);
            acc1 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ahi, b1.lo[m], acc1, 0, 0, 0);
            acc0 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ahi, b0.hi[m], acc0, 0, 0, 0);
            acc1 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ahi, b1.hi[m], acc1, 0, 0, 0);
        }
    }
    // The two phases of run2 for a screen in two levels: the hi x hi
    // products alone (a coarse value within coarse_slack of the full one),
    // then the lo x hi and hi x lo corrections, for the tiles that need them.
    // hi(a) hi(b) + lo(a) hi(b) + hi(a) lo(b) are the products of run2; only
    // their order of accumulation differs, which the slack below does not
    // depend on.
    template <int DP>
    __device__ static void run2_hi(const Elem* half_row, const Frag<DP>& b0, const Frag<DP>& b1, floatx16& acc0,
                                   floatx16& acc1) {
        const bf16x8* p = reinterpret_cast<const bf16x8*>(half_row);
#pragma unroll
        for (int m = 0; m < DP / 16; ++m) {
            const bf16x8 ahi = p[2 * m];
            acc0 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ahi, b0.hi[m], acc0, 0, 0, 0);
            acc1 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ahi, b1.hi[m], acc1, 0, 0, 0);
        }
    }
    template <int DP>
    __device__ static void run2_lo(const Elem* half_row, const Frag<DP>& b0, const Frag<DP>& b1, floatx16& acc0,
                                   floatx16& acc1) {
        const bf16x8* p = reinterpret_cast<const bf16x8*>(half_row);
#pragma unroll
        for (int m = 0; m < DP / 16; ++m) {
            const bf16x8 ahi = p[2 * m];
            const bf16x8 alo = p[DP / 8 + 2 * m];
            acc0 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(alo, b0.hi[m], acc0, 0, 0, 0);
            acc1 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(alo, b1.hi[m], acc1, 0, 0, 0);
            acc0 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ahi, b0.lo[m], acc0, 0, 0, 0);
            acc1 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ahi, b1.lo[m], acc1, 0, 0, 0);
        }
    }
    // Coarse slack: |full - coarse| <= coarse_slack(DP) (N_q + N_c). The
    // corrections are products with a bf16 residual, |lo(x)| <= 2^-8 |x|, so
    // together at most 2^-7 (1 + 2^-8) sum |a_t b_t| <= 2^-7 (1 + 2^-8)
    // (N_q + N_c) (b = -2 x_q, so sum |a_t b_t| <= N_c + N_q); their 2 DP
    // additions to an accumulator of magnitude at most 3.1 (N_q + N_c) add
    // 2 DP x 2u x 3.1 (counted at 2 ulp per addition, as above). The factor
    // 1.01 and 16u leave room for the roundings of the widened thresholds.
    __host__ __device__ static constexpr float coarse_slack(int dp) {
        return 1.01f * (0.0078430176f + 12.4f * dp * kU) + 16.0f * kU;
    }
    // 3.1 2^-16 for the split products, FP32 accumulation of 3 DP terms whose
    // magnitudes sum to at most 2 (N_q + N_i) (counted at 2 ulp per addition,
    // in case the matrix core truncates), and the norm, centring and
    // threshold roundings.
    __host__ __device__ static float slack(int dp) {
        return 4.7302246e-5f + (12.0f * dp + 64.0f) * kU;
    }
};

// LDS chunk geometry: candidate tiles per chunk (up to 8, about 16 KiB of
// screen operand per buffer) and the row pitch in float4s. Rows of 64 bytes
// or more get one float4 of padding so the 32 rows a wave reads in one
// ds_read_b128 fall in different banks.
template <class Op, int DP>
struct ChunkShape {
    static constexpr int kRowBytes = DP * Op::kRowMul * static_cast<int>(sizeof(typename Op::Elem));
    static constexpr int kRowF4 = kRowBytes / 16 > 0 ? kRowBytes / 16 : 1;
    static constexpr int kPitchF4 = kRowF4 >= 4 ? kRowF4 + 1 : kRowF4;
    static constexpr int kTileBytes = 32 * kPitchF4 * 16;
    static constexpr int kFit = 16384 / kTileBytes;
    static constexpr int kTiles = kFit >= 8 ? 8 : kFit >= 4 ? 4 : kFit >= 2 ? 2 : 1;  // a power of two
    // A operand plus the tile norms (4 bytes per row)
    static constexpr size_t kBytes = static_cast<size_t>(kTiles) * (kTileBytes + 32 * 4);
    // the symmetric pass also stages the rows' column thresholds (4 bytes per
    // row) and their per-(tile, half) maxima (16 bytes per tile, 8 used)
    static constexpr size_t kSymBytes = static_cast<size_t>(kTiles) * (kTileBytes + 32 * 4 * 2 + 16);
    // Streamed by LDS DMA (two query tiles per wave): two tiles per chunk.
    static constexpr int kDmaTiles = 2;
    static constexpr size_t kDmaSymBytes = static_cast<size_t>(kDmaTiles) * (kTileBytes + 32 * 4 * 2 + 16);
};

// Exact key of a query row against key row xs_row: the same FMA chain as the
// oracle and the other sweeps. Q is either a register array (indexed with
// constants only) or a pointer to the query's key row in global memory.
template <int DP, class Q>
__device__ __forceinline__ float exact_key(const Q& q, const float* __restrict__ xs_row) {
    const float4* c4 = reinterpret_cast<const float4*>(xs_row);
    float e = 0.0f;
#pragma unroll
    for (int v = 0; v < DP / 4; ++v) {
        const float4 t = c4[v];
        float df = q[4 * v + 0] - t.x;
        e = __builtin_fmaf(df, df, e);
        df = q[4 * v + 1] - t.y;
        e = __builtin_fmaf(df, df, e);
        df = q[4 * v + 2] - t.z;
        e = __builtin_fmaf(df, df, e);
        df = q[4 * v + 3] - t.w;
        e = __builtin_fmaf(df, df, e);
    }
    return e;
}

// exact_key with the query row in LDS (16-byte aligned), read as float4s.
// Rows wider than 128 features go in unrolled pieces of 64 (the chain stays
// in ascending feature order), so a key never holds a whole wide row in
// registers.
template <int DP>
__device__ __forceinline__ float exact_key_q4(const float* q, const float* __restrict__ xs_row) {
    constexpr int kPiece = DP <= 128 ? DP : 64;
    float e = 0.0f;
#pragma unroll 1
    for (int f0 = 0; f0 < DP; f0 += kPiece) {
        const float4* q4 = reinterpret_cast<const float4*>(q + f0);
        const float4* c4 = reinterpret_cast<const float4*>(xs_row + f0);
#pragma unroll
        for (int v = 0; v < kPiece / 4; ++v) {
            const float4 a = q4[v];
            const float4 t = c4[v];
            float df = a.x - t.x;
            e = __builtin_fmaf(df, df, e);
            df = a.y - t.y;
            e = __builtin_fmaf(df, df, e);
            df = a.z - t.z;
            e = __builtin_fmaf(df, df, e);
            df = a.w - t.w;
            e = __builtin_fmaf(df, df, e);
        }
    }
    return e;
}

// Writes one query's list (ascending in `stage`, k entries at stride 64) in the
// reference's slot order: keys below the k-th ascending by index, then the
// tied ones already in ascending index. Returns false for a +inf k-th key.
__device__ bool emit_list(const uint64_t* stage, int k, int n, int qi, int* __restrict__ nbr_idx,
                          float* __restrict__ nbr_d2, float* __restrict__ kdist2) {
    const uint32_t kth = static_cast<uint32_t>(stage[(k - 1) * 64] >> 32);
    if (kth >= kInfBits) return false;
    kdist2[qi] = __uint_as_float(kth);
    int below = 0;
    for (int t = 0; t < k; ++t) below += static_cast<uint32_t>(stage[t * 64] >> 32) < kth;
    const size_t ns = static_cast<size_t>(n);
    for (int e = 0; e < k; ++e) {
        const uint64_t w = stage[e * 64];
        const uint32_t idx = static_cast<uint32_t>(w);
        int slot = e;
        if (e < below) {
            slot = 0;
            for (int f = 0; f < below; ++f) slot += static_cast<uint32_t>(stage[f * 64]) < idx;
        }
        nbr_idx[static_cast<size_t>(slot) * ns + qi] = static_cast<int>(idx);
        nbr_d2[static_cast<size_t>(slot) * ns + qi] = __uint_as_float(static_cast<uint32_t>(w >> 32));
    }
    return true;
}

struct MfmaArgs {
    const void* xa;      // screen operand rows (form layout), candidate order
    const float* xc;     // centred rows, natural order, stride DP
    const float* xs;     // key rows, natural order, stride DP (queries)
    const float* xs_cand;  // key rows in candidate order (== xs without an order)
    const int* cand;     // candidate position -> original index (null: identity)
    const int* qlist;    // local query -> original index (null: row0 + local)
    const float* tseed;  // per original query: upper bound of its k-th key (optional)
    const float* madj;   // |x_i|^2 (1 - c), candidate order
    const float* norms;  // |x_i|^2, natural order
    const int* guard;
    int n;
    int k;
    float c;
    int row0;
    int rows;
    int slices;
    uint64_t* partial;   // [slices][rows][k] when slices > 1
    int collect;         // 1: collect kernels (needs tseed), 0: list kernels
    const int* row_count;  // optional device count bounding rows
    uint32_t* cbuf;      // collect mode: [rows][slices][cap] candidate positions
    int* ccount;         // collect mode: [rows][slices] entries (> cap: overflowed)
    int cap;
    int* ov_rows;        // collect mode: rows whose slice buffer overflowed (second pass)
    int* ov_count;
    int ov_cap;
    int* big_rows;       // collect mode: locals with more than kSelectMax entries
    int* big_count;
    int* list_rows;      // collect mode: rows for the seeded list pass
    int* list_count;
    int list_cap;
    int* nbr_idx;
    float* nbr_d2;
    float* kdist2;
    int* fb_rows;
    int* fb_count;
    int fb_cap;          // entries fb_rows holds
    const int* heavy;    // symmetric pass: 1 for rows (by position) queued for the list pass up front
    const int* qpos;     // original index -> candidate position (null: identity)
    // First-pass wave select of a run in candidate order: each row's list goes
    // to stage[local * (2k + 2)] (k indices, k keys, the k-th key, 1 = emitted)
    // and unstage_kernel writes the transposed outputs; null: written directly.
    int* stage;
};

// Words per staged record: k indices, k keys, the k-th key and the emitted
// flag, padded to whole 16-byte groups.
__host__ __device__ constexpr int stage_words(int k) { return (2 * k + 2 + 3) & ~3; }

// Streams candidate tiles [t_begin, t_end) in chunks of CT tiles through a
// double-buffered LDS stage shared by the block's 4 waves: while the waves
// screen chunk c out of one buffer, every thread holds its share of chunk
// c + 1 in registers and writes it to the other buffer afterwards; one
// barrier per chunk. Tile v is read from v, or v - wrap when v >= wrap
// (t_end <= 2 wrap). For each tile, visit(acc, v) gets the accumulator
// C = m + A B of this lane (16 candidate rows, one query); with kSym it is
// visit(acc, v, wmax, wrow): also this half's maximum column threshold over
// the tile (wmax[tile].x for half 0, .y for half 1) and the tile's staged
// column thresholds (8 float4s: rows 4i..4i+3). The operand arrays are padded
// to whole tiles (zero rows, NaN norms and thresholds); tiles past t_end are
// never visited. `stage` holds 2 * ChunkShape::kBytes (kSymBytes with kSym).
//
// With two query tiles per wave (QT = 2, the BF16 screen) at 128 features
// or more the screen runs in two levels: coarse(acc, v, wmax, nmax, t) gets
// each query tile's accumulator after the hi x hi products alone, with this
// half's maximum column threshold and maximum adjusted norm over the tile
// (nmax, staged like wmax), and returns whether the tile can hold a pair
// that passes; only then are the correction products added and the tile
// visited. Other shapes never call coarse.
template <class Op, int DP, bool kSym, int QT = 1, class Visit, class Coarse>
__device__ __forceinline__ void stream_range(const MfmaArgs& p, const typename Op::template Frag<DP>* bq,
                                             int j, int h, unsigned char* stage, int t_begin, int t_end,
                                             int wrap, const float* __restrict__ wcol,
                                             const float2* __restrict__ wmax, Visit&& visit,
                                             const float2* __restrict__ nmax, Coarse&& coarse,
                                             const int* __restrict__ clist = nullptr, int nlist = 0) {
    static_assert(QT == 1 || (QT == 2 && kSym), "two query tiles per wave: symmetric pass only");
    using Frag = typename Op::template Frag<DP>;
    using Elem = typename Op::Elem;
    using Shape = ChunkShape<Op, DP>;
    // Two query tiles per wave: the A tiles go global -> LDS by DMA, two per
    // chunk (no staging registers).
    constexpr bool kDma = QT == 2;
    // The two-level screen pays where it skips most of the corrections: at
    // 128 features boxes hardly separate tiles, so the pruning keeps far pairs
    // that the coarse test drops; at 64 most kept pairs are near ones.
    constexpr bool kTwoLevel = QT == 2 && DP >= 128;
    constexpr int CT = kDma ? Shape::kDmaTiles : Shape::kTiles;
    constexpr int kRowsPerChunk = 32 * CT;
    constexpr int kRowF4 = Shape::kRowF4;     // float4s per global row (DP * elem >= 16 bytes)
    constexpr int kPitchF4 = Shape::kPitchF4; // float4s per LDS row
    constexpr int kTileF4 = 32 * kRowF4;      // float4s per A tile in global memory
    constexpr int kChunkF4 = CT * kTileF4;
    constexpr int kLdsTileF4 = 32 * kPitchF4;
    constexpr int kLdsChunkF4 = CT * kLdsTileF4;
    constexpr int kLoadsA = ((kDma ? kLdsChunkF4 : kChunkF4) + 255) / 256;  // per thread
    constexpr int kChunkM4 = kRowsPerChunk / 4;       // float4s of norms per chunk
    static_assert(DP * Op::kRowMul * sizeof(Elem) % 16 == 0, "screen rows must be whole float4s");
    static_assert(2 * kChunkM4 + CT <= 256, "one staging thread per float4");
    float4* s_a = reinterpret_cast<float4*>(stage);
    float4* s_m = s_a + 2 * kLdsChunkF4;
    float4* s_w = s_m + 2 * kChunkM4;              // kSym: column thresholds
    float2* s_x = reinterpret_cast<float2*>(s_w + 2 * kChunkM4);  // kSym: per-tile maxima
    float2* s_nx = s_x + 2 * CT;                   // QT = 2: per-tile maximum adjusted norms

    // Chunks of CT tiles from t_begin, or (clist) only the listed chunk
    // starts, nlist of them (the pruned symmetric pass).
    const int n_chunks = clist ? nlist : (t_end - t_begin + CT - 1) / CT;
    auto chunk_start = [&](int ch) { return clist ? clist[ch] : t_begin + ch * CT; };
    const float4* ga = static_cast<const float4*>(p.xa);
    const float4* gm = reinterpret_cast<const float4*>(p.madj);
    const float4* gw = reinterpret_cast<const float4*>(wcol);
    const float nan = __uint_as_float(0x7FC00000u);
    auto phys = [&](int v) { return v >= wrap ? v - wrap : v; };

    float4 ra[kDma ? 1 : kLoadsA];
    // Separate registers for norms and column thresholds: one register
    // loaded by both would make the wave wait for the whole fetch right after
    // issuing it (a load may not overwrite a register another load still owns).
    float4 rm = make_float4(nan, nan, nan, nan);
    float4 rw = make_float4(nan, nan, nan, nan);
    float2 rx = make_float2(nan, nan);
    // threads [0, kChunkM4) stage norms, [kChunkM4, 2 kChunkM4) column
    // thresholds, [2 kChunkM4, 2 kChunkM4 + CT) the per-tile maxima
    const int tid = threadIdx.x;
    // dbuf: the LDS buffer the DMA writes (kDma only)
    auto fetch = [&](int t0, int dbuf) {
        if constexpr (kDma) {
            // Instruction u of wave w fills LDS float4s [256 u + 64 w, +64),
            // one per lane, in the padded layout: slot idx is float4
            // idx % kPitchF4 of chunk row idx / kPitchF4. Lanes on a pad slot,
            // past the chunk or on a tile past t_end write nothing: those
            // slots are never read, or only by tiles that are never visited.
            float4* dst = s_a + dbuf * kLdsChunkF4 + (tid & ~63);
#pragma unroll
            for (int u = 0; u < kLoadsA; ++u) {
                const int idx = tid + 256 * u;
                const int row = idx / kPitchF4;
                const int f = idx - row * kPitchF4;
                const int vt = t0 + row / 32;
                if (idx < kLdsChunkF4 && f < kRowF4 && vt < t_end)
                    __builtin_amdgcn_global_load_lds(
                        ga + static_cast<size_t>(phys(vt)) * kTileF4 + (row % 32) * kRowF4 + f, dst + 256 * u, 16,
                        0, 0);
            }
            // Wave 0 also moves the chunk's norms, column thresholds (lanes
            // [0, kChunkM4), one float4 each) and per-tile maxima (lanes
            // [0, 2 CT), one float each).
            if (tid < 64) {
                const int vt = t0 + tid / 8;
                if (tid < kChunkM4 && vt < t_end) {
                    const size_t at = static_cast<size_t>(phys(vt)) * 8 + tid % 8;
                    __builtin_amdgcn_global_load_lds(gm + at, s_m + dbuf * kChunkM4, 16, 0, 0);
                    __builtin_amdgcn_global_load_lds(gw + at, s_w + dbuf * kChunkM4, 16, 0, 0);
                }
                const int vx = t0 + tid / 2;
                if (tid < 2 * CT && vx < t_end) {
                    __builtin_amdgcn_global_load_lds(reinterpret_cast<const float*>(wmax + phys(vx)) + (tid & 1),
                                                     s_x + dbuf * CT, 4, 0, 0);
                    if constexpr (kTwoLevel)
                        __builtin_amdgcn_global_load_lds(reinterpret_cast<const float*>(nmax + phys(vx)) + (tid & 1),
                                                         s_nx + dbuf * CT, 4, 0, 0);
                }
            }
        } else {
#pragma unroll
            for (int u = 0; u < kLoadsA; ++u) {
                const int idx = tid + 256 * u;
                const int vt = t0 + idx / kTileF4;
                const bool live = idx < kChunkF4 && vt < t_end;
                ra[u] = live ? ga[static_cast<size_t>(phys(vt)) * kTileF4 + idx % kTileF4]
                             : make_float4(0.f, 0.f, 0.f, 0.f);
            }
        }
        if constexpr (kDma) return;
        if (tid < kChunkM4) {
            const int vt = t0 + tid / 8;
            rm = vt < t_end ? gm[static_cast<size_t>(phys(vt)) * 8 + tid % 8] : make_float4(nan, nan, nan, nan);
        }
        if constexpr (kSym) {
            if (tid >= kChunkM4 && tid < 2 * kChunkM4) {
                const int e = tid - kChunkM4;
                const int vt = t0 + e / 8;
                rw = vt < t_end ? gw[static_cast<size_t>(phys(vt)) * 8 + e % 8] : make_float4(nan, nan, nan, nan);
            }
            if (tid >= 2 * kChunkM4 && tid < 2 * kChunkM4 + CT) {
                const int vt = t0 + tid - 2 * kChunkM4;
                rx = vt < t_end ? wmax[phys(vt)] : make_float2(nan, nan);
            }
        }
    };
    auto stage_chunk = [&](int buf) {
        if constexpr (kDma) return;  // the DMA wrote the buffer
#pragma unroll
        for (int u = 0; u < kLoadsA; ++u) {
            const int idx = tid + 256 * u;
            const int row = idx / kRowF4;
            if (idx < kChunkF4) s_a[buf * kLdsChunkF4 + row * kPitchF4 + (idx - row * kRowF4)] = ra[u];
        }
        if (tid < kChunkM4) s_m[buf * kChunkM4 + tid] = rm;
        if constexpr (kSym) {
            if (tid >= kChunkM4 && tid < 2 * kChunkM4) s_w[buf * kChunkM4 + tid - kChunkM4] = rw;
            if (tid >= 2 * kChunkM4 && tid < 2 * kChunkM4 + CT) s_x[buf * CT + tid - 2 * kChunkM4] = rx;
        }
    };
    // Accumulator initialised with the 16 adjusted norms of this lane's rows.
    auto init_acc = [&](const float4* sm, int tt) {
        floatx16 acc;
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            const float4 t = sm[tt * 8 + 2 * b + h];
            acc[4 * b + 0] = t.x;
            acc[4 * b + 1] = t.y;
            acc[4 * b + 2] = t.z;
            acc[4 * b + 3] = t.w;
        }
        return acc;
    };
    // kSym: this half's column maxima of the chunk's tiles (the two-level
    // screen reads them from LDS where used, to spare registers)
    float wm[kSym && !kTwoLevel ? CT : 1];
    auto wmax_at = [&](int buf, int tt) {
        if constexpr (kTwoLevel) {
            const float2 x = s_x[buf * CT + tt];
            return h ? x.y : x.x;
        } else {
            return wm[tt];
        }
    };
    auto call = [&](const floatx16& acc, int buf, int tt, int v, int t) {
        if constexpr (kSym) {
            visit(acc, v, wmax_at(buf, tt), s_w + buf * kChunkM4 + tt * 8, t);
        } else {
            visit(acc, v);
        }
    };

    // Every DMA of a chunk has landed once its issuing wave waits for its
    // memory counter; the barrier after that publishes the buffer.
    auto dma_wait = [&]() {
        if constexpr (kDma) __builtin_amdgcn_s_waitcnt(0);
    };
    if (n_chunks > 0) {
        fetch(chunk_start(0), 0);
        stage_chunk(0);
    }
    dma_wait();
    __syncthreads();
    for (int ch = 0; ch < n_chunks; ++ch) {
        const int buf = ch & 1;
        const int t0 = chunk_start(ch);
        if (ch + 1 < n_chunks) fetch(chunk_start(ch + 1), buf ^ 1);  // next chunk in flight
        const float4* sa = s_a + buf * kLdsChunkF4;
        const float4* sm = s_m + buf * kChunkM4;
        if constexpr (kSym && !kTwoLevel) {
#pragma unroll
            for (int tt = 0; tt < CT; ++tt) {
                const float2 x = s_x[buf * CT + tt];
                wm[tt] = h ? x.y : x.x;
            }
        }
        if constexpr (kTwoLevel) {
            // One tile per step against two query fragments: two independent
            // MFMA chains per A value read; the hi x hi products first, the
            // corrections (for both query tiles) only where a coarse test
            // passes.
#pragma unroll
            for (int tt = 0; tt < CT; ++tt) {
                floatx16 acc0 = init_acc(sm, tt);
                floatx16 acc1 = acc0;
                const Elem* arow = reinterpret_cast<const Elem*>(sa + (tt * 32 + j) * kPitchF4) +
                                   Op::half_offset(DP, h);
                Op::template run2_hi<DP>(arow, bq[0], bq[1], acc0, acc1);
                const int v = t0 + tt;
                if (v < t_end) {
                    const float2 nx = s_nx[buf * CT + tt];
                    const float nmh = h ? nx.y : nx.x;
                    const float wmh = wmax_at(buf, tt);
                    if (coarse(acc0, v, wmh, nmh, 0) || coarse(acc1, v, wmh, nmh, 1)) {
                        Op::template run2_lo<DP>(arow, bq[0], bq[1], acc0, acc1);
                        call(acc0, buf, tt, v, 0);
                        call(acc1, buf, tt, v, 1);
                    }
                }
            }
        } else if constexpr (QT == 2) {
            // One tile per step against two query fragments: two independent
            // MFMA chains per A value read.
#pragma unroll
            for (int tt = 0; tt < CT; ++tt) {
                floatx16 acc0 = init_acc(sm, tt);
                floatx16 acc1 = acc0;
                Op::template run2<DP>(reinterpret_cast<const Elem*>(sa + (tt * 32 + j) * kPitchF4) +
                                          Op::half_offset(DP, h),
                                      bq[0], bq[1], acc0, acc1);
                if (t0 + tt < t_end) {
                    call(acc0, buf, tt, t0 + tt, 0);
                    call(acc1, buf, tt, t0 + tt, 1);
                }
            }
        } else {
            // Two tiles per step: two independent MFMA chains hide each
            // other's result latency.
#pragma unroll
            for (int tt = 0; tt < CT; tt += 2) {
                Frag f0, f1;
                Op::template load_a<DP>(reinterpret_cast<const Elem*>(sa + (tt * 32 + j) * kPitchF4) +
                                            Op::half_offset(DP, h), f0);
                floatx16 acc0 = init_acc(sm, tt);
                const bool two = tt + 1 < CT;
                floatx16 acc1;
                if (two) {
                    Op::template load_a<DP>(reinterpret_cast<const Elem*>(sa + ((tt + 1) * 32 + j) * kPitchF4) +
                                                Op::half_offset(DP, h), f1);
                    acc1 = init_acc(sm, tt + 1);
                }
                acc0 = Op::template run<DP>(f0, bq[0], acc0);
                if (two) acc1 = Op::template run<DP>(f1, bq[0], acc1);
                if (t0 + tt < t_end) call(acc0, buf, tt, t0 + tt, 0);
                if (two && t0 + tt + 1 < t_end) call(acc1, buf, tt + 1, t0 + tt + 1, 0);
            }
        }
        if (ch + 1 < n_chunks) stage_chunk(buf ^ 1);
        dma_wait();
        __syncthreads();
    }
}

// The block's candidate slice (blockIdx.y of p.slices) of all candidate tiles.
template <class Op, int DP, class Visit>
__device__ __forceinline__ void stream_tiles(const MfmaArgs& p, const typename Op::template Frag<DP>& bq,
                                             int j, int h, unsigned char* stage, Visit&& visit) {
    const int n_tiles = (p.n + 31) / 32;
    const int t_begin = static_cast<int>((static_cast<long long>(n_tiles) * blockIdx.y) / p.slices);
    const int t_end = static_cast<int>((static_cast<long long>(n_tiles) * (blockIdx.y + 1)) / p.slices);
    stream_range<Op, DP, false>(p, &bq, j, h, stage, t_begin, t_end, n_tiles, nullptr, nullptr, visit, nullptr,
                                [](const floatx16&, int, float, float, int) { return true; });
}

// Wide lists need the whole register file: let those instantiations run at
// two waves per SIMD rather than spill the list to scratch.
template <class Op, int DP, int KMAX>
__global__ void __launch_bounds__(64 * kTilesPerBlock)
__attribute__((amdgpu_waves_per_eu(1, KMAX >= 64 ? 2 : 8)))
mfma_sweep_kernel(MfmaArgs p) {
    using Frag = typename Op::template Frag<DP>;
    constexpr bool kQRegs = DP <= 16;  // query key row in VGPRs

    // LDS: rings + two candidate chunks during the sweep; afterwards one wave
    // at a time stages its lists in the same bytes.
    constexpr size_t kRingBytes = sizeof(uint32_t) * kTilesPerBlock * kRing * 64;
    constexpr size_t kChunkBytes = ChunkShape<Op, DP>::kBytes;
    constexpr size_t kSweepBytes = kRingBytes + 2 * kChunkBytes;
    constexpr size_t kStageBytes = sizeof(uint64_t) * KMAX * 64;
    __shared__ __attribute__((aligned(16))) unsigned char smem[kSweepBytes > kStageBytes ? kSweepBytes : kStageBytes];

    const int rows = p.row_count ? min(p.rows, *p.row_count) : p.rows;
    if (static_cast<int>(blockIdx.x) * kTilesPerBlock * 32 >= rows) return;  // whole block
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int j = lane & 31;   // query within the tile
    const int h = lane >> 5;   // k-half / candidate-row half
    const int local = (blockIdx.x * kTilesPerBlock + wave) * 32 + j;
    const bool valid = local < rows;
    const int qi = p.qlist ? p.qlist[valid ? local : 0] : p.row0 + (valid ? local : 0);
    const int n = p.n;

    if (*p.guard) {
        if (valid && h == 0 && blockIdx.y == 0) queue_push(p.fb_rows, p.fb_count, p.fb_cap, qi);
        return;
    }

    // Query operands.
    Frag bq;
    Op::template load_b<DP>(p.xc + static_cast<size_t>(qi) * DP, h, bq);
    float qs[kQRegs ? DP : 1];
    if constexpr (kQRegs) {
#pragma unroll
        for (int t = 0; t < DP; ++t) qs[t] = p.xs[static_cast<size_t>(qi) * DP + t];
    }
    const float* q_global = p.xs + static_cast<size_t>(qi) * DP;
    const float nq = p.norms[qi];
    const float c = p.c;

    uint64_t a[KMAX];
    list_init<KMAX>(a, p.k);
    if (p.tseed) {
        // Placeholders (T_q, 0xFFFFFFFF) admit exactly the keys <= T_q; at
        // least k real candidates have such keys, so none survives the merge.
        const uint64_t fake = pack(__float_as_uint(p.tseed[qi]), 0xFFFFFFFFu);
#pragma unroll
        for (int t = 0; t < KMAX; ++t) a[t] = (t < p.k) ? fake : 0ull;
    }
    // Each of the two lanes serving a query keeps k exact keys over its half of
    // the candidates, so the smaller of the two k-th keys bounds the final one:
    // both screen against it (lane l and l ^ 32 swap theirs).
    auto threshold = [&]() {
        float t = key_as_float(a[0]);
        t = fminf(t, __shfl_xor(t, 32, 64));
        if (!valid) return -__uint_as_float(kInfBits);
        return (__builtin_fmaf(t, 1.0f + c, c * nq) - nq) + 4.0f * kU * (t + nq);
    };
    float Tn = threshold();
    int cnt = 0;
    uint32_t* ring = reinterpret_cast<uint32_t*>(smem) + wave * kRing * 64;

    // Forced inline: an out-of-line call would put the list and the
    // arguments it captures in scratch memory.
    auto flush = [&]() __attribute__((always_inline)) {
        // Exact keys of the queued candidates, four at a time so their row
        // loads overlap, then sorted inserts.
        for (int r = 0; r < kRing; r += 4) {
            if (!__any(r < cnt)) break;
            float e[4];
            uint32_t id[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const bool live = r + u < cnt;
                id[u] = live ? ring[(r + u) * 64 + lane] : 0u;
                const float* row = p.xs_cand + static_cast<size_t>(id[u]) * DP;
                if constexpr (kQRegs) {
                    e[u] = live ? exact_key<DP>(qs, row) : 0.0f;
                } else {
                    e[u] = live ? exact_key<DP>(q_global, row) : 0.0f;
                }
            }
            // One copy of the insert network, fed through a shift register so
            // no register array is indexed at run time.
            uint64_t w0 = kSentinel, w1 = kSentinel, w2 = kSentinel, w3 = kSentinel;
            {
                uint64_t w[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const bool live = r + u < cnt;
                    const uint32_t orig = (p.cand && live) ? static_cast<uint32_t>(p.cand[id[u]]) : id[u];
                    const uint32_t kb = (static_cast<int>(orig) == qi || e[u] != e[u]) ? kInfBits
                                                                                       : __float_as_uint(e[u]);
                    w[u] = live ? pack(kb, orig) : kSentinel;
                }
                w0 = w[0];
                w1 = w[1];
                w2 = w[2];
                w3 = w[3];
            }
#pragma unroll 1
            for (int u = 0; u < 4; ++u) {
                if (w0 < a[0]) list_insert<KMAX>(a, w0);
                w0 = w1;
                w1 = w2;
                w2 = w3;
            }
        }
        cnt = 0;
        Tn = threshold();
    };

    // Screen one tile's accumulator against the threshold; queue passes. The
    // common case (no lane passes) costs a min tree and one compare; when some
    // lane passes, each register is tested with one wave-wide compare whose
    // mask is checked on the scalar unit, so registers nobody passes cost no
    // divergent branch.
    auto screen = [&](const floatx16& acc, int tile) {
        float lo = acc[0];
#pragma unroll
        for (int r = 1; r < 16; ++r) lo = __builtin_fminf(lo, acc[r]);  // NaN rows drop out
        if (__any(lo <= Tn)) {
            if (__any(cnt > kRing - 16)) flush();
            const uint32_t c0 = static_cast<uint32_t>(tile * 32 + 4 * h);
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const bool pass = acc[r] <= Tn;
                if (__ballot(pass)) {
                    if (pass) {
                        ring[cnt * 64 + lane] = c0 + static_cast<uint32_t>(8 * (r / 4) + (r % 4));
                        ++cnt;
                    }
                }
            }
        }
    };

    stream_tiles<Op, DP>(p, bq, j, h, smem + kRingBytes, screen);
    flush();

    // ---- fold the two half-lists of each query (lane j + 32 into lane j) and
    // write the result; the waves take turns on the shared staging bytes ----
    __syncthreads();  // every ring drained
    uint64_t* stage = reinterpret_cast<uint64_t*>(smem);
    for (int w = 0; w < kTilesPerBlock; ++w) {
        if (wave == w) {
            if (h == 1) {
#pragma unroll
                for (int t = 0; t < KMAX; ++t) stage[t * 64 + lane] = a[t];
            }
            wave_lds_sync();
            if (h == 0) {
                for (int t = p.k - 1; t >= 0; --t) {
                    const uint64_t e = stage[t * 64 + lane + 32];
                    if (e < a[0]) list_insert<KMAX>(a, e);
                }
            }
            wave_lds_sync();
            if (h == 0 && valid) {
                if (p.slices > 1) {
                    // Partial list of this slice, ascending.
                    uint64_t* dst = p.partial + (static_cast<size_t>(blockIdx.y) * p.rows + local) * p.k;
#pragma unroll
                    for (int t = 0; t < KMAX; ++t) {
                        if (t < p.k) dst[p.k - 1 - t] = a[t];
                    }
                } else {
#pragma unroll
                    for (int t = 0; t < KMAX; ++t) {
                        if (t < p.k) stage[(p.k - 1 - t) * 64 + lane] = a[t];
                    }
                    if (!emit_list(stage + lane, p.k, n, qi, p.nbr_idx, p.nbr_d2, p.kdist2)) {
                        queue_push(p.fb_rows, p.fb_count, p.fb_cap, qi);
                    }
                }
            }
        }
        __syncthreads();
    }
}

// ---- collect mode: fixed per-query threshold, no list in the sweep ----
//
// With a seed T_q (an upper bound of the query's k-th key) the screen
// threshold never needs to move: every candidate whose exact key can be
// <= T_q passes, and the k nearest are among them. The sweep then only
// appends the positions of passing candidates to a per-(query, slice) buffer
// (the two lanes serving a query share it: in each tile the k-half 0 lane's
// passes go first); the select kernels key them exactly and choose. Without
// the register list the sweep runs at several waves per SIMD and spends its
// VALU on the screen alone.
template <class Op, int DP>
__global__ void __launch_bounds__(64 * kTilesPerBlock)
mfma_collect_kernel(MfmaArgs p) {
    using Frag = typename Op::template Frag<DP>;
    __shared__ __attribute__((aligned(16))) unsigned char smem[2 * ChunkShape<Op, DP>::kBytes];

    const int rows = p.row_count ? min(p.rows, *p.row_count) : p.rows;
    if (static_cast<int>(blockIdx.x) * kTilesPerBlock * 32 >= rows) return;  // whole block
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int j = lane & 31;
    const int h = lane >> 5;
    const int local = (blockIdx.x * kTilesPerBlock + wave) * 32 + j;
    const bool valid = local < rows;
    const int qi = p.qlist ? p.qlist[valid ? local : 0] : p.row0 + (valid ? local : 0);

    if (*p.guard) {
        if (valid && h == 0 && blockIdx.y == 0) queue_push(p.fb_rows, p.fb_count, p.fb_cap, qi);
        return;
    }

    Frag bq;
    Op::template load_b<DP>(p.xc + static_cast<size_t>(qi) * DP, h, bq);
    const float nq = p.norms[qi];
    const float c = p.c;
    const float t = p.tseed[qi];
    // Same screen inequality as the list sweep, at the fixed threshold T_q.
    const float Tn = valid ? (__builtin_fmaf(t, 1.0f + c, c * nq) - nq) + 4.0f * kU * (t + nq)
                           : -__uint_as_float(kInfBits);
    const size_t slot = static_cast<size_t>(valid ? local : 0) * p.slices + blockIdx.y;
    uint32_t* buf = p.cbuf + slot * p.cap;
    const int cap = p.cap;
    int cnt = 0;  // the same in both lanes of a query

    auto collect = [&](const floatx16& acc, int tile) {
        float lo = acc[0];
#pragma unroll
        for (int r = 1; r < 16; ++r) lo = __builtin_fminf(lo, acc[r]);  // NaN rows drop out
        if (__any(lo <= Tn)) {
            uint32_t mask = 0;
#pragma unroll
            for (int r = 0; r < 16; ++r) mask |= (acc[r] <= Tn ? 1u : 0u) << r;
            const int mine = __popc(mask);
            const int other = __shfl_xor(mine, 32, 64);
            int pos = cnt + (h ? other : 0);
            const uint32_t c0 = static_cast<uint32_t>(tile * 32 + 4 * h);
            while (mask) {
                const int r = __ffs(mask) - 1;
                mask &= mask - 1;
                if (pos < cap) buf[pos] = c0 + static_cast<uint32_t>(8 * (r / 4) + (r % 4));
                ++pos;
            }
            cnt += mine + other;
        }
    };
    stream_tiles<Op, DP>(p, bq, j, h, smem, collect);
    if (valid && h == 0) p.ccount[slot] = cnt;
}

// ---- wide rows: the collect pass with the contraction in K chunks ----
//
// A query fragment of more than 128 features does not fit in registers beside
// the candidate fragments, so for DP >= kWideMin the contraction runs in
// chunks of kWideChunk features: each step stages one chunk of kWideTiles
// candidate tiles in LDS (double-buffered through registers, one barrier per
// step) and adds its products to those tiles' accumulators; the query chunk
// comes from the operand rows themselves (position qpos[qi]), unscaled, so
// the chain accumulates D = -m/2 + A x_q. Every intermediate of that chain is
// exactly -1/2 times the one of C = m + A (-2 x_q) (scaling by a power of two
// commutes with rounding), so acc = -2 D is the value the narrow pass would
// screen, tested against the same threshold.
constexpr int kWideTiles = 2;

template <class Op, int DP>
__global__ void __launch_bounds__(64 * kTilesPerBlock)
mfma_wide_collect_kernel(MfmaArgs p) {
    constexpr int KC = kWideChunk;
    constexpr int NK = DP / KC;
    constexpr int CT = kWideTiles;
    static_assert(DP % KC == 0 && NK >= 2, "wide rows are whole chunks");
    using Frag = typename Op::template Frag<KC>;
    using Elem = typename Op::Elem;
    constexpr int kRowF4 = KC * Op::kRowMul * static_cast<int>(sizeof(Elem)) / 16;  // float4s per chunk row
    constexpr int kPitchF4 = kRowF4 + 1;    // LDS row pitch: the 32 rows of a read fall in different banks
    constexpr int kSubF4 = 32 * kRowF4;     // one (tile, chunk) block in global memory
    constexpr int kLdsTileF4 = 32 * kPitchF4;
    constexpr int kStepF4 = CT * kSubF4;
    constexpr int kLoads = (kStepF4 + 255) / 256;  // per thread
    __shared__ float4 s_a[2][CT * kLdsTileF4];
    __shared__ float4 s_m[2][CT * 8];

    const int rows = p.row_count ? min(p.rows, *p.row_count) : p.rows;
    if (static_cast<int>(blockIdx.x) * kTilesPerBlock * 32 >= rows) return;  // whole block
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int j = lane & 31;
    const int h = lane >> 5;
    const int local = (blockIdx.x * kTilesPerBlock + wave) * 32 + j;
    const bool valid = local < rows;
    const int qi = p.qlist ? p.qlist[valid ? local : 0] : p.row0 + (valid ? local : 0);

    if (*p.guard) {
        if (valid && h == 0 && blockIdx.y == 0) queue_push(p.fb_rows, p.fb_count, p.fb_cap, qi);
        return;
    }

    const float4* ga = static_cast<const float4*>(p.xa);
    const int qp = p.qpos ? p.qpos[qi] : qi;
    // chunk 0 of the query's operand row; chunk kc is kSubF4 further on
    const float4* qrow = ga + static_cast<size_t>(qp / 32) * NK * kSubF4 + (qp % 32) * kRowF4;
    const float nq = p.norms[qi];
    const float c = p.c;
    const float t = p.tseed[qi];
    const float Tn = valid ? (__builtin_fmaf(t, 1.0f + c, c * nq) - nq) + 4.0f * kU * (t + nq)
                           : -__uint_as_float(kInfBits);
    const size_t slot = static_cast<size_t>(valid ? local : 0) * p.slices + blockIdx.y;
    uint32_t* buf = p.cbuf + slot * p.cap;
    const int cap = p.cap;
    int cnt = 0;  // the same in both lanes of a query

    auto collect = [&](const floatx16& acc, int tile) {
        float lo = acc[0];
#pragma unroll
        for (int r = 1; r < 16; ++r) lo = __builtin_fminf(lo, acc[r]);  // NaN rows drop out
        if (__any(lo <= Tn)) {
            uint32_t mask = 0;
#pragma unroll
            for (int r = 0; r < 16; ++r) mask |= (acc[r] <= Tn ? 1u : 0u) << r;
            const int mine = __popc(mask);
            const int other = __shfl_xor(mine, 32, 64);
            int pos = cnt + (h ? other : 0);
            const uint32_t c0 = static_cast<uint32_t>(tile * 32 + 4 * h);
            while (mask) {
                const int r = __ffs(mask) - 1;
                mask &= mask - 1;
                if (pos < cap) buf[pos] = c0 + static_cast<uint32_t>(8 * (r / 4) + (r % 4));
                ++pos;
            }
            cnt += mine + other;
        }
    };

    const int n_tiles = (p.n + 31) / 32;
    const int t_begin = static_cast<int>((static_cast<long long>(n_tiles) * blockIdx.y) / p.slices);
    const int t_end = static_cast<int>((static_cast<long long>(n_tiles) * (blockIdx.y + 1)) / p.slices);
    const int steps = (t_end - t_begin + CT - 1) / CT * NK;
    const float4* gm = reinterpret_cast<const float4*>(p.madj);
    const float nan = __uint_as_float(0x7FC00000u);
    float4 ra[kLoads];
    float4 rm = make_float4(nan, nan, nan, nan);
    // Step s: tiles t_begin + CT (s / NK) + [0, CT), chunk s % NK; the tiles'
    // norms travel with chunk 0 (threads [0, 8 CT), one float4 each).
    auto fetch = [&](int st) {
        const int g = st / NK;
        const int kc = st - g * NK;
        const int v0 = t_begin + g * CT;
#pragma unroll
        for (int u = 0; u < kLoads; ++u) {
            const int idx = tid + 256 * u;
            const int tt = idx / kSubF4;
            const int vt = v0 + tt;
            ra[u] = (idx < kStepF4 && vt < t_end)
                        ? ga[(static_cast<size_t>(vt) * NK + kc) * kSubF4 + (idx - tt * kSubF4)]
                        : make_float4(0.f, 0.f, 0.f, 0.f);
        }
        if (kc == 0 && tid < CT * 8) {
            const int vt = v0 + tid / 8;
            rm = vt < t_end ? gm[static_cast<size_t>(vt) * 8 + tid % 8] : make_float4(nan, nan, nan, nan);
        }
    };
    auto stage = [&](int st, int b) {
#pragma unroll
        for (int u = 0; u < kLoads; ++u) {
            const int idx = tid + 256 * u;
            if (idx < kStepF4) {
                const int tt = idx / kSubF4;
                const int r = idx - tt * kSubF4;
                const int row = r / kRowF4;
                s_a[b][tt * kLdsTileF4 + row * kPitchF4 + (r - row * kRowF4)] = ra[u];
            }
        }
        if (st % NK == 0 && tid < CT * 8) s_m[b][tid] = rm;
    };

    floatx16 acc[CT];
    if (steps > 0) {
        fetch(0);
        stage(0, 0);
    }
    __syncthreads();
    for (int st = 0; st < steps; ++st) {
        const int b = st & 1;
        const int g = st / NK;
        const int kc = st - g * NK;
        if (st + 1 < steps) fetch(st + 1);  // next step in flight
        if (kc == 0) {
#pragma unroll
            for (int tt = 0; tt < CT; ++tt) {
#pragma unroll
                for (int bb = 0; bb < 4; ++bb) {
                    const float4 m4 = s_m[b][tt * 8 + 2 * bb + h];
                    acc[tt][4 * bb + 0] = -0.5f * m4.x;
                    acc[tt][4 * bb + 1] = -0.5f * m4.y;
                    acc[tt][4 * bb + 2] = -0.5f * m4.z;
                    acc[tt][4 * bb + 3] = -0.5f * m4.w;
                }
            }
        }
        Frag bq;
        Op::template load_a<KC>(reinterpret_cast<const Elem*>(qrow + static_cast<size_t>(kc) * kSubF4) +
                                    Op::half_offset(KC, h),
                                bq);
#pragma unroll
        for (int tt = 0; tt < CT; ++tt) {
            Frag fa;
            Op::template load_a<KC>(reinterpret_cast<const Elem*>(&s_a[b][tt * kLdsTileF4 + j * kPitchF4]) +
                                        Op::half_offset(KC, h),
                                    fa);
            acc[tt] = Op::template run<KC>(fa, bq, acc[tt]);
        }
        if (kc == NK - 1) {
#pragma unroll
            for (int tt = 0; tt < CT; ++tt) {
                const int v = t_begin + g * CT + tt;
                if (v < t_end) {
                    floatx16 a2;
#pragma unroll
                    for (int r = 0; r < 16; ++r) a2[r] = -2.0f * acc[tt][r];
                    collect(a2, v);
                }
            }
        }
        if (st + 1 < steps) stage(st + 1, b ^ 1);
        __syncthreads();
    }
    if (valid && h == 0) p.ccount[slot] = cnt;
}

// ---- symmetric collect: a full run, queries and candidates in one order ----
//
// One screen value serves both rows of a pair. With acc = m_c - 2 x_q . x_c
// (m_c = |x_c|^2 (1 - c), the accumulator's initial value), the row test
// acc <= Tn_q collects c for q. With the roles swapped the pass would compare
// acc' = m_q - 2 x_q . x_c against Tn_c; acc - m_c + m_q is that value up to
// the same rounding bound the slack c already covers (the bound is symmetric
// in the two rows), so the column test acc - W_c <= L_q collects q for c, with
// W_c = Tn_c + m_c + 8u (|Tn_c| + 2 |x_c|^2 + T_c) rounded up (computed in
// FP64 once per row) and L_q = 4u |x_q|^2 - m_q: the margins cover the
// rounding of Tn_c and of the two extra operations (|acc| <= 2|x_c|^2 +
// |x_q|^2 + c S). So each unordered pair of tiles is computed once: block B
// (query rows [128 B, 128 B + 128), positions = local rows) takes candidate
// super-tiles B + o (mod NB) for o in [0, H_B], H_B = (NB - 1) / 2, plus
// o = NB / 2 when NB is even and B < NB / 2 - every unordered pair of
// distinct super-tiles exactly once; blockIdx.y splits the offsets into S
// slices. o = 0, its own super-tile, runs the row test only (that covers
// every ordered pair in it).
//
// A row has 2S buffer slices. Row entries go to slice y (only this wave
// writes it; the count stays in registers, as in the plain collect pass).
// Column entries are kept per candidate row as (c, mask of this wave's 32
// queries) words, appended with no atomics to a chunk of a shared pool that
// this wave owns (one atomic per kChunkWords words takes a new chunk; its
// header records the wave's query tile and slice); scatter_pairs_kernel
// expands them into slice S + y of row c afterwards (one atomic per word).
// Words that find the pool exhausted flag their rows, which then count as
// overflowed. The per-(tile, half) maximum of W gives the coarse column test
// min(acc) - Wmax <= L_q (rounding is monotone, so it misses nothing), two
// instructions per tile.
constexpr int kChunkWords = 128;

// QT query tiles per wave (2 for the wide BF16 screen: each A value read from
// LDS then feeds two MFMA chains, halving the operand reads per pair); a
// block covers 4 QT query tiles and the ring is over blocks of 128 QT rows.
template <class Op, int DP>
constexpr int sym_query_tiles() {
    return (std::is_same_v<Op, ScreenBF16> && DP >= 64) ? 2 : 1;
}

// Candidate tiles per LDS chunk of the symmetric pass (stream_range's CT).
template <class Op, int DP>
constexpr int sym_chunk_tiles() {
    return sym_query_tiles<Op, DP>() == 2 ? ChunkShape<Op, DP>::kDmaTiles : ChunkShape<Op, DP>::kTiles;
}

template <class Op, int DP>
__global__ void __launch_bounds__(64 * kTilesPerBlock)
__attribute__((amdgpu_waves_per_eu(DP >= 128 || sym_query_tiles<Op, DP>() == 2 ? 2 : 4, 8)))
mfma_sym_collect_kernel(MfmaArgs p, const float* __restrict__ wcol, const float2* __restrict__ wmax,
                        const float2* __restrict__ nmax, uint64_t* __restrict__ pool, int* __restrict__ chunk_region,
                        int* __restrict__ chunk_count, int* __restrict__ next_chunk, int max_chunks,
                        int* __restrict__ col_over, const int* __restrict__ live_chunks,
                        const int* __restrict__ live_count, int live_stride, int n_slices) {
    using Frag = typename Op::template Frag<DP>;
    constexpr int QT = sym_query_tiles<Op, DP>();
    constexpr int kBlockTiles = kTilesPerBlock * QT;
    __shared__ __attribute__((aligned(16)))
    unsigned char smem[2 * (QT == 2 ? ChunkShape<Op, DP>::kDmaSymBytes : ChunkShape<Op, DP>::kSymBytes)];

    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int j = lane & 31;
    const int h = lane >> 5;
    // A 1-D grid of nb x ns blocks, dispatched from the last block of
    // positions backwards (the rows with the largest seeds sit there and
    // keep the most chunks after pruning: started first, they do not trail),
    // the slices of a block together.
    const int ns = n_slices;
    const int nb = static_cast<int>(gridDim.x) / ns;
    const int B = nb - 1 - static_cast<int>(blockIdx.x) / ns;
    const int ys = static_cast<int>(blockIdx.x) % ns;  // this block's slice
    const int qtile0 = (B * kTilesPerBlock + wave) * QT;

    int q[QT];
    bool valid[QT];
    int qi[QT];
#pragma unroll
    for (int t = 0; t < QT; ++t) {
        q[t] = (qtile0 + t) * 32 + j;  // position == local row
        valid[t] = q[t] < p.n;
        qi[t] = p.qlist[valid[t] ? q[t] : 0];
    }

    if (*p.guard) {
#pragma unroll
        for (int t = 0; t < QT; ++t)
            if (valid[t] && h == 0 && ys == 0) queue_push(p.fb_rows, p.fb_count, p.fb_cap, qi[t]);
        return;
    }

    Frag bq[QT];
    float Tn[QT], Lq[QT];
    // Two-level screen (QT = 2): the coarse value is within g (N_q + N_c) of
    // the full one, so a tile can hold a passing pair only where
    // min(coarse) <= Tn + g (N_q + N_c) (row test) or
    // min(coarse) - Wmax <= L_q + g (N_q + N_c) plus the rounding of the
    // column difference (column test); N_c <= nmax / (1 - c) and the 1.01 in
    // g covers that factor. gq holds the per-query part of the margin (with
    // 8u of |Tn| and |L_q| for the roundings of the widened thresholds; a
    // row that tests nothing, Tn = -inf, adds no |Tn|, which would make its
    // column margin infinite).
    float gq[QT];
    constexpr float g = QT == 2 && DP >= 128 ? ScreenBF16::coarse_slack(DP) : 0.0f;
    const float c = p.c;
#pragma unroll
    for (int t = 0; t < QT; ++t) {
        Op::template load_b<DP>(p.xc + static_cast<size_t>(qi[t]) * DP, h, bq[t]);
        const float nq = p.norms[qi[t]];
        const float tq = p.tseed[qi[t]];
        // A heavy row collects nothing for itself (the list pass serves it)
        // but still serves as a candidate of the other rows.
        const bool tests = valid[t] && !p.heavy[q[t]];
        Tn[t] = tests ? (__builtin_fmaf(tq, 1.0f + c, c * nq) - nq) + 4.0f * kU * (tq + nq)
                      : -__uint_as_float(kInfBits);
        Lq[t] = valid[t] ? 4.0f * kU * nq - p.madj[q[t]] : __uint_as_float(0x7FC00000u);
        gq[t] = g * nq + 8.0f * kU * ((tests ? __builtin_fabsf(Tn[t]) : 0.0f) + __builtin_fabsf(Lq[t]));
    }
    const int h_b = (nb - 1) / 2 + ((nb % 2 == 0 && B < nb / 2) ? 1 : 0);
    const long long span = h_b + 1;  // offsets 0..h_b
    const int o0 = static_cast<int>(span * ys / ns);
    const int o1 = static_cast<int>(span * (ys + 1) / ns);
    const int wrap = kBlockTiles * nb;
    const int v_col = kBlockTiles * (B + 1);  // tiles of other blocks from here
    const int cap = p.cap;
    const int nslot = 2 * ns;
    uint32_t* qbuf[QT];
    int cnt[QT];      // row entries (the same in both lanes of a query)
    int chunk[QT];    // this wave's pool chunk per query tile (the same in every lane)
    int fill[QT];     // words used in it
#pragma unroll
    for (int t = 0; t < QT; ++t) {
        qbuf[t] = p.cbuf + (static_cast<size_t>(valid[t] ? q[t] : 0) * nslot + ys) * cap;
        cnt[t] = 0;
        chunk[t] = -1;
        fill[t] = kChunkWords;
    }
    auto room_for_two = [&](int t) {  // called by the whole wave
        if (fill[t] + 2 <= kChunkWords) return;
        if (lane == 0 && chunk[t] >= 0 && chunk[t] < max_chunks) chunk_count[chunk[t]] = fill[t];
        int id = 0;
        if (lane == 0) id = atomicAdd(next_chunk, 1);
        chunk[t] = __shfl(id, 0, 64);
        fill[t] = 0;
        if (lane == 0 && chunk[t] < max_chunks) chunk_region[chunk[t]] = (qtile0 + t) * ns + ys;
    };

    auto visit = [&](const floatx16& acc, int v, float wm, const float4* wrow, int t) {
        float lo = acc[0];
#pragma unroll
        for (int r = 1; r < 16; ++r) lo = __builtin_fminf(lo, acc[r]);  // NaN rows drop out
        const bool other_tile = v >= v_col;  // the same in the whole wave
        const bool rowp = lo <= Tn[t];
        const bool colp = other_tile && lo - wm <= Lq[t];
        if (!__any(rowp || colp)) return;
        const int pt = v >= wrap ? v - wrap : v;
        const uint32_t c0 = static_cast<uint32_t>(pt * 32 + 4 * h);
        if (__any(rowp)) {
            uint32_t mask = 0;
#pragma unroll
            for (int r = 0; r < 16; ++r) mask |= (acc[r] <= Tn[t] ? 1u : 0u) << r;
            const int mine = __popc(mask);
            const int other = __shfl_xor(mine, 32, 64);
            int pos = cnt[t] + (h ? other : 0);
            while (mask) {
                const int r = __ffs(mask) - 1;
                mask &= mask - 1;
                if (pos < cap) qbuf[t][pos] = c0 + static_cast<uint32_t>(8 * (r / 4) + (r % 4));
                ++pos;
            }
            cnt[t] += mine + other;
        }
        if (other_tile && __any(colp)) {
#pragma unroll
            for (int b = 0; b < 4; ++b) {
                const float4 w = wrow[2 * b + h];
                const float wv[4] = {w.x, w.y, w.z, w.w};
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const uint64_t bal = __ballot(acc[4 * b + i] - wv[i] <= Lq[t]);
                    if (bal == 0) continue;
                    room_for_two(t);
                    // lane 0 writes half 0's word (row 8b + i), lane 32 half 1's (row 8b + 4 + i)
                    const uint32_t lo32 = static_cast<uint32_t>(bal);
                    const uint32_t hi32 = static_cast<uint32_t>(bal >> 32);
                    const uint32_t m = h ? hi32 : lo32;
                    if (j == 0 && m != 0) {
                        const int pos = fill[t] + (h && lo32 != 0 ? 1 : 0);
                        const uint32_t crow = c0 + static_cast<uint32_t>(8 * b + i);
                        if (chunk[t] < max_chunks) {
                            pool[static_cast<size_t>(chunk[t]) * kChunkWords + pos] =
                                (static_cast<uint64_t>(crow) << 32) | m;
                        } else {
                            atomicOr(col_over + crow, 1);
                        }
                    }
                    fill[t] += (lo32 != 0) + (hi32 != 0);
                }
            }
        }
    };
    // Coarse test of the two-level screen (QT = 2), on the hi x hi value.
    auto coarse = [&](const floatx16& acc, int v, float wm, float nm, int t) {
        float lo = acc[0];
#pragma unroll
        for (int r = 1; r < 16; ++r) lo = __builtin_fminf(lo, acc[r]);  // NaN rows drop out
        const float m = __builtin_fmaf(g, nm, gq[t]);
        const bool rowp = lo <= Tn[t] + m;
        const bool colp = v >= v_col && lo - wm <= __builtin_fmaf(8.0f * kU, __builtin_fabsf(wm), Lq[t] + m);
        return __any(rowp || colp);
    };
    // Only the chunks sym_live_chunks_kernel kept (when it ran).
    const int region = B * ns + ys;
    const int* clist = live_chunks ? live_chunks + static_cast<size_t>(region) * live_stride : nullptr;
    const int nlist = live_chunks ? live_count[region] : 0;
    stream_range<Op, DP, true, QT>(p, bq, j, h, smem, kBlockTiles * (B + o0), kBlockTiles * (B + o1), wrap,
                                   wcol, wmax, visit, nmax, coarse, clist, nlist);
#pragma unroll
    for (int t = 0; t < QT; ++t) {
        if (valid[t] && h == 0) p.ccount[static_cast<size_t>(q[t]) * nslot + ys] = cnt[t];
        if (lane == 0 && chunk[t] >= 0 && chunk[t] < max_chunks) chunk_count[chunk[t]] = fill[t];
    }
}

// ---- pruning of the symmetric pass by tile bounds ----
//
// A pair of rows matters to the pass only when its exact key can be at most
// the query's seed (row test) or the candidate's (column test); rows flagged
// heavy test nothing. Per tile of 32 positions: the bounding box of its key
// rows and the largest seed among its rows that test. A chunk of candidate
// tiles whose box lies farther from the block's query box than both the
// block's and the chunk's largest seeds (the lower bound, rounded, scaled
// down by (2 DP + 16) 2^-23 to cover its own rounding and the exact keys')
// holds no such pair and is skipped. Results do not change. Boxes are
// stored feature-major: the lower corners [DP][tiles], then the upper
// corners [DP][tiles] (groups likewise), so the box tests, one tile per
// thread, read consecutive words.
template <int DP>
__global__ void __launch_bounds__(256)
tile_bounds_kernel(const float* __restrict__ xs_ord, const int* __restrict__ order,
                   const float* __restrict__ tseed, const int* __restrict__ heavy, int n, int tiles,
                   float* __restrict__ box, float* __restrict__ tmax) {
    const int e = blockIdx.x * 256 + threadIdx.x;  // (tile, feature)
    if (e >= tiles * DP) return;
    const int t = e / DP;
    const int f = e - t * DP;
    float lo = __uint_as_float(kInfBits);
    float hi = -__uint_as_float(kInfBits);
    float tm = -__uint_as_float(kInfBits);
    for (int r = 0; r < 32; ++r) {
        const int pos = t * 32 + r;
        if (pos >= n) break;
        const float v = xs_ord[static_cast<size_t>(pos) * DP + f];
        lo = __builtin_fminf(lo, v);  // NaN coordinates drop out (their keys are +inf)
        hi = __builtin_fmaxf(hi, v);
        if (f == 0 && !heavy[pos]) tm = __builtin_fmaxf(tm, tseed[order[pos]]);
    }
    box[static_cast<size_t>(f) * tiles + t] = lo;
    box[static_cast<size_t>(DP + f) * tiles + t] = hi;
    if (f == 0) tmax[t] = tm;
}

// Bounds of groups of kPruneGroup consecutive tiles (positions), for the
// first test of sym_live_chunks_kernel.
constexpr int kPruneGroup = 8;

template <int DP>
__global__ void __launch_bounds__(256)
group_bounds_kernel(const float* __restrict__ box, const float* __restrict__ tmax, int tiles,
                    float* __restrict__ gbox, float* __restrict__ gtmax) {
    const int groups = (tiles + kPruneGroup - 1) / kPruneGroup;
    const int e = blockIdx.x * 256 + threadIdx.x;  // (group, feature)
    if (e >= groups * DP) return;
    const int g = e / DP;
    const int f = e - g * DP;
    float lo = __uint_as_float(kInfBits);
    float hi = -__uint_as_float(kInfBits);
    float tm = -__uint_as_float(kInfBits);
    for (int t = g * kPruneGroup; t < min(tiles, (g + 1) * kPruneGroup); ++t) {
        lo = __builtin_fminf(lo, box[static_cast<size_t>(f) * tiles + t]);
        hi = __builtin_fmaxf(hi, box[static_cast<size_t>(DP + f) * tiles + t]);
        if (f == 0) tm = __builtin_fmaxf(tm, tmax[t]);
    }
    gbox[static_cast<size_t>(f) * groups + g] = lo;
    gbox[static_cast<size_t>(DP + f) * groups + g] = hi;
    if (f == 0) gtmax[g] = tm;
}

// Squared gap between two boxes: the first feature-major in global memory
// (feature f of its lower corner at lo[f * stride], of its upper corner at
// hi[f * stride]; the threads of a wave read consecutive boxes, so each load
// is coalesced), the second in LDS (DP floats each).
template <int DP>
__device__ __forceinline__ float box_gap2(const float* __restrict__ lo, const float* __restrict__ hi,
                                          size_t stride, const float* blo, const float* bhi) {
    float lb = 0.0f;
#pragma unroll 8
    for (int f = 0; f < DP; ++f) {
        const float g = __builtin_fmaxf(0.0f, __builtin_fmaxf(lo[f * stride] - bhi[f], blo[f] - hi[f * stride]));
        lb = __builtin_fmaf(g, g, lb);
    }
    return lb;
}

// For block B and slice y of the symmetric pass: the starts of its candidate
// chunks (CT tiles from kBlockTiles (B + o0)) that may hold a pair, in
// ascending order, and their count. A chunk stays when one of its tiles
// does: the tile's box within reach of the block's box for the larger of the
// block's and the tile's largest seeds (one thread per tile, 256 tiles per
// round). The block's own tiles (lower bound 0) always stay.
template <int DP, int CT, int kBlockTiles>
__global__ void __launch_bounds__(256)
sym_live_chunks_kernel(const float* __restrict__ box, const float* __restrict__ tmax,
                       const float* __restrict__ gbox, const float* __restrict__ gtmax, int nb, int ns,
                       int stride, int* __restrict__ live, int* __restrict__ live_count,
                       unsigned long long* __restrict__ visited) {
    static_assert(256 % CT == 0, "whole chunks per round");
    __shared__ __attribute__((aligned(16))) float s_lo[DP];
    __shared__ __attribute__((aligned(16))) float s_hi[DP];
    __shared__ float s_tq;
    __shared__ int s_ok[256];
    __shared__ int s_wave[4];
    __shared__ int s_tiles;  // candidate tiles of the kept chunks
    const int B = blockIdx.x;
    const int y = blockIdx.y;
    const int tid = threadIdx.x;
    const int h_b = (nb - 1) / 2 + ((nb % 2 == 0 && B < nb / 2) ? 1 : 0);
    const long long span = h_b + 1;
    const int o0 = static_cast<int>(span * y / ns);
    const int o1 = static_cast<int>(span * (y + 1) / ns);
    const int wrap = kBlockTiles * nb;  // tiles (the box arrays' stride)
    const int n_groups = (wrap + kPruneGroup - 1) / kPruneGroup;
    const int t_begin = kBlockTiles * (B + o0);
    const int t_end = kBlockTiles * (B + o1);
    const int nc = (t_end - t_begin + CT - 1) / CT;
    for (int f = tid; f < DP; f += 256) {
        float lo = __uint_as_float(kInfBits);
        float hi = -__uint_as_float(kInfBits);
        for (int t = kBlockTiles * B; t < kBlockTiles * (B + 1); ++t) {
            lo = __builtin_fminf(lo, box[static_cast<size_t>(f) * wrap + t]);
            hi = __builtin_fmaxf(hi, box[static_cast<size_t>(DP + f) * wrap + t]);
        }
        s_lo[f] = lo;
        s_hi[f] = hi;
    }
    if (tid == 0) {
        float tq = -__uint_as_float(kInfBits);
        for (int t = kBlockTiles * B; t < kBlockTiles * (B + 1); ++t) tq = __builtin_fmaxf(tq, tmax[t]);
        s_tq = tq;
        s_tiles = 0;
    }
    __syncthreads();
    const float shrink = 1.0f - (2.0f * DP + 16.0f) * 2.0f * kU;
    int* out = live + static_cast<size_t>(B * ns + y) * stride;
    constexpr int kChunksPerRound = 256 / CT;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    int total = 0;
    for (int c0 = 0; c0 < nc; c0 += kChunksPerRound) {
        const int v = t_begin + c0 * CT + tid;
        int ok = 0;
        if (v < t_end) {
            const size_t pt = static_cast<size_t>(v >= wrap ? v - wrap : v);
            // the tile's group first (its box holds the tile's, its seed
            // bound is at least the tile's): most far tiles fall away there,
            // and the group's bounds are the same loads for its 8 threads
            const size_t pg = pt / kPruneGroup;
            const float lbg = box_gap2<DP>(gbox + pg, gbox + static_cast<size_t>(DP) * n_groups + pg, n_groups,
                                           s_lo, s_hi);
            if (!(lbg * shrink > __builtin_fmaxf(s_tq, gtmax[pg]))) {
                const float lb = box_gap2<DP>(box + pt, box + static_cast<size_t>(DP) * wrap + pt, wrap, s_lo, s_hi);
                const float thr = __builtin_fmaxf(s_tq, tmax[pt]);
                ok = !(lb * shrink > thr);  // NaN keeps
            }
        }
        s_ok[tid] = ok;
        __syncthreads();
        const int c = c0 + tid;
        bool keep = false;
        if (tid < kChunksPerRound && c < nc) {
#pragma unroll
            for (int e = 0; e < CT; ++e) keep |= s_ok[tid * CT + e] != 0;
        }
        // compact in ascending chunk order
        const uint64_t bal = __ballot(keep);
        if (lane == 0) s_wave[wave] = __popcll(bal);
        __syncthreads();
        int before = total;
        for (int w = 0; w < wave; ++w) before += s_wave[w];
        if (keep) {
            out[before + lanes_below(bal)] = t_begin + c * CT;
            if (visited) atomicAdd(&s_tiles, min(CT, t_end - (t_begin + c * CT)));
        }
        total += s_wave[0] + s_wave[1] + s_wave[2] + s_wave[3];
        __syncthreads();
    }
    if (tid == 0) live_count[B * ns + y] = total;
    if (visited && tid == 0) atomicAdd(visited, static_cast<unsigned long long>(s_tiles) * kBlockTiles);
}

// The work counter for an unpruned symmetric pass: every tile pair it visits.
__global__ void add_counter_kernel(unsigned long long* __restrict__ c, unsigned long long v) {
    if (threadIdx.x == 0 && blockIdx.x == 0) atomicAdd(c, v);
}

// Column words of each pool chunk (written by query tile t, slice y) into
// slice S + y of their rows: one atomic per word reserves its popcount, then
// the queries t * 32 + bit are written. Rows whose words found the pool
// exhausted are marked overflowed (their first column slice is set past cap).
__global__ void __launch_bounds__(kChunkWords)
scatter_pairs_kernel(const uint64_t* __restrict__ pool, const int* __restrict__ chunk_region,
                     const int* __restrict__ chunk_count, const int* __restrict__ next_chunk,
                     int max_chunks, int ns, int cap, const int* __restrict__ guard,
                     uint32_t* __restrict__ cbuf, int* __restrict__ ccount) {
    if (*guard) return;
    const int chunks = min(*next_chunk, max_chunks);
    for (int ch = blockIdx.x; ch < chunks; ch += gridDim.x) {
        const int count = chunk_count[ch];
        const int r = chunk_region[ch];
        const uint64_t* src = pool + static_cast<size_t>(ch) * kChunkWords;
        const uint32_t q0 = static_cast<uint32_t>(r / ns) * 32;
        const int y = r % ns;
        for (int e = threadIdx.x; e < count; e += kChunkWords) {  // one thread per word
            const uint64_t w = src[e];
            uint32_t m = static_cast<uint32_t>(w);
            const size_t slot = static_cast<size_t>(w >> 32) * (2 * ns) + ns + y;
            int pos = atomicAdd(ccount + slot, __popc(m));
            uint32_t* dst = cbuf + slot * cap;
            while (m) {
                const int bit = __ffs(m) - 1;
                m &= m - 1;
                if (pos < cap) dst[pos] = q0 + static_cast<uint32_t>(bit);
                ++pos;
            }
        }
    }
}

__global__ void __launch_bounds__(256)
mark_overflow_kernel(const int* __restrict__ col_over, int n, int ns, int cap, int* __restrict__ ccount) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n && col_over[i]) ccount[static_cast<size_t>(i) * (2 * ns) + ns] = cap + 1;
}

// Rows whose candidate set would be huge (outliers in sparse space, whose
// seed threshold covers much of the data) skip the collect passes: the
// number of m sample rows (every (n/m)-th position, so spread over all cells;
// m = 8192 / d within [64, 512])
// within the row's seed T estimates its candidate count; above kHeavyCount
// the row is queued for the seeded list pass while the queue has room, and
// flagged. In the symmetric pass a flagged row collects nothing and is never
// a column row, so such rows cannot flood the column buffers. The estimate
// only steers work; results do not depend on it.
constexpr int kHeavyCount = 65536;
constexpr int kHeavyChunk = 64;  // sample rows staged in LDS at a time

template <int DP>
__global__ void __launch_bounds__(256)
heavy_rows_kernel(const float* __restrict__ xs_ord, const int* __restrict__ order,
                  const float* __restrict__ tseed, const int* __restrict__ guard, int n, int m,
                  int* __restrict__ heavy, int* __restrict__ list_rows, int* __restrict__ list_count,
                  int list_cap) {
    __shared__ __attribute__((aligned(16))) float s_smp[kHeavyChunk * DP];
    const int q = blockIdx.x * 256 + threadIdx.x;
    const bool valid = q < n;
    if (*guard) {
        // The symmetric pass hands every row to the radix path: queue none
        // here, or those rows would reach the fallback queue twice.
        if (valid) heavy[q] = 0;
        return;
    }
    float xq[DP];
    const float* row = xs_ord + static_cast<size_t>(valid ? q : 0) * DP;
#pragma unroll
    for (int t = 0; t < DP; ++t) xq[t] = row[t];
    const float T = tseed[order[valid ? q : 0]];
    int hits = 0;
    for (int s0 = 0; s0 < m; s0 += kHeavyChunk) {
        __syncthreads();
        for (int e = threadIdx.x; e < kHeavyChunk * DP; e += 256) {
            const int sj = s0 + e / DP;
            const size_t pos = static_cast<size_t>(static_cast<long long>(sj) * n / m);
            s_smp[e] = sj < m ? xs_ord[pos * DP + e % DP] : __uint_as_float(0x7FC00000u);
        }
        __syncthreads();
#pragma unroll 2
        for (int r = 0; r < kHeavyChunk; ++r) {
            float e = 0.0f;
#pragma unroll
            for (int t = 0; t < DP; ++t) {
                const float df = xq[t] - s_smp[r * DP + t];
                e = __builtin_fmaf(df, df, e);
            }
            hits += e <= T;
        }
    }
    if (!valid) return;
    int flag = 0;
    if (static_cast<long long>(hits) * n > static_cast<long long>(kHeavyCount) * m) {
        const int at = atomicAdd(list_count, 1);
        if (at < list_cap) {
            list_rows[at] = order[q];
            flag = 1;
        }
    }
    heavy[q] = flag;
}

// Column thresholds of the symmetric pass, per candidate position (NaN past n):
// W = Tn + m + 8u (|Tn| + 2 |x|^2 + T), in FP64, rounded up to a float (-inf
// for a heavy row); and per (tile, half) the maximum over the half's 16 rows
// (8b + 4h + i).
__global__ void __launch_bounds__(256)
column_thresholds_kernel(const float* __restrict__ tseed, const float* __restrict__ norms,
                         const float* __restrict__ madj, const int* __restrict__ cand,
                         const int* __restrict__ heavy, int n, int padded_rows, float c,
                         float* __restrict__ wcol) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= padded_rows) return;
    if (i >= n) {
        wcol[i] = __uint_as_float(0x7FC00000u);
        return;
    }
    if (heavy[i]) {
        wcol[i] = -__uint_as_float(kInfBits);
        return;
    }
    const int o = cand[i];
    const double t = tseed[o];
    const double nc = norms[o];
    const double u = 5.9604644775390625e-8;
    const double cd = c;
    const double tn = t * (1.0 + cd) + cd * nc - nc + 4.0 * u * (t + nc);
    const double w = tn + static_cast<double>(madj[i]) + 8.0 * u * (fabs(tn) + 2.0 * nc + t);
    const float f = static_cast<float>(w);
    wcol[i] = static_cast<double>(f) < w ? __uint_as_float(__float_as_uint(f) + 1u) : f;
}

// Per (tile, half): the largest column threshold and the largest adjusted
// norm of the half's 16 rows (rows 8 b + 4 h + i, the rows one lane of the
// screen holds).
__global__ void __launch_bounds__(256)
column_maxima_kernel(const float* __restrict__ wcol, const float* __restrict__ madj, int tiles,
                     float2* __restrict__ wmax, float2* __restrict__ nmax) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= 2 * tiles) return;
    const int tile = e >> 1;
    const int h = e & 1;
    float m = __uint_as_float(0x7FC00000u);
    float a = __uint_as_float(0x7FC00000u);
    for (int b = 0; b < 4; ++b) {
        for (int i = 0; i < 4; ++i) {
            m = __builtin_fmaxf(m, wcol[tile * 32 + 8 * b + 4 * h + i]);
            a = __builtin_fmaxf(a, madj[tile * 32 + 8 * b + 4 * h + i]);
        }
    }
    reinterpret_cast<float*>(wmax)[e] = m;
    reinterpret_cast<float*>(nmax)[e] = a;
}

// Collected entries per query the wave select holds; larger sets go to the
// block select, which streams any number through LDS.
constexpr int kSelectMax = 1024;
// The wave select's own capacity: narrow rows (short candidate lists) take
// 512, which halves its LDS and doubles the waves per CU; their rare longer
// lists go to the block select.
template <int DP>
constexpr int wave_select_max() {
    return DP <= 32 ? 512 : kSelectMax;
}
constexpr int kBigChunk = 2048;  // words sorted per step of the block select


// (key, original index) word of candidate position `pos` for query qi.
template <int DP>
__device__ __forceinline__ uint64_t collected_word(const MfmaArgs& p, const float* s_q, uint32_t pos, int qi) {
    const uint32_t orig = p.cand ? static_cast<uint32_t>(p.cand[pos]) : pos;
    const float key = exact_key_q4<DP>(s_q, p.xs_cand + static_cast<size_t>(pos) * DP);
    const uint32_t kb = (static_cast<int>(orig) == qi || key != key) ? kInfBits : __float_as_uint(key);
    return pack(kb, orig);
}

// The k smallest words of a row, in any order in sel[0, k) (k <= 64): written
// in the reference's slot order by the calling wave (all 64 lanes).
__device__ void emit_selected(const MfmaArgs& p, const uint64_t* sel, int qi, int lane, int local = -1) {
    // Called by a whole wave. Lane f holds word f; the k-th word is their
    // maximum. Sorting the lanes by (group, index) -- group 0 for keys below
    // the k-th, 1 for keys equal to it, 2 for lanes past k -- puts every word
    // at its slot: below-k-th keys in ascending index, then the tied ones.
    const int k = p.k;
    const bool valid = lane < k;
    const uint64_t me = valid ? sel[lane] : 0;
    uint64_t kw = me;
#pragma unroll
    for (int m = 1; m < 64; m <<= 1) {
        const uint32_t hi = static_cast<uint32_t>(__shfl_xor(static_cast<int>(kw >> 32), m, 64));
        const uint32_t lo = static_cast<uint32_t>(__shfl_xor(static_cast<int>(kw), m, 64));
        const uint64_t o = (static_cast<uint64_t>(hi) << 32) | lo;
        kw = o > kw ? o : kw;
    }
    const uint32_t kth = static_cast<uint32_t>(kw >> 32);
    uint32_t key = static_cast<uint32_t>(me >> 32);
    const uint32_t grp = !valid ? 2u : (key < kth ? 0u : 1u);
    uint64_t s = (static_cast<uint64_t>(grp) << 32) | static_cast<uint32_t>(me);
#pragma unroll
    for (int size = 2; size <= 64; size <<= 1) {
#pragma unroll
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            const uint32_t ohi = static_cast<uint32_t>(__shfl_xor(static_cast<int>(s >> 32), stride, 64));
            const uint32_t olo = static_cast<uint32_t>(__shfl_xor(static_cast<int>(s), stride, 64));
            const uint32_t okey = static_cast<uint32_t>(__shfl_xor(static_cast<int>(key), stride, 64));
            const uint64_t os = (static_cast<uint64_t>(ohi) << 32) | olo;
            const bool up = (lane & size) == 0;
            const bool lower = (lane & stride) == 0;
            const bool swap = (lower == up) ? (os < s) : (os > s);
            if (swap) {
                s = os;
                key = okey;
            }
        }
    }
    if (local >= 0) {  // staged: one contiguous record per row
        int* rec = p.stage + static_cast<size_t>(local) * stage_words(k);
        if (valid) {
            rec[lane] = static_cast<int>(static_cast<uint32_t>(s));
            rec[k + lane] = static_cast<int>(key);
        }
        if (lane == 0) {
            rec[2 * k] = static_cast<int>(kth);
            rec[2 * k + 1] = 1;
        }
        return;
    }
    if (valid) {
        const size_t ns = static_cast<size_t>(p.n);
        p.nbr_idx[static_cast<size_t>(lane) * ns + qi] = static_cast<int>(static_cast<uint32_t>(s));
        p.nbr_d2[static_cast<size_t>(lane) * ns + qi] = __uint_as_float(key);
        if (lane == 0) p.kdist2[qi] = __uint_as_float(kth);
    }
}

// k smallest of the `total` words in s_words (WPL words per lane), emitted.
template <int WPL>
__device__ void select_words(const MfmaArgs& p, const uint64_t* s_words, int total, int qi, int lane,
                             uint64_t* s_sel, int local = -1) {
    uint64_t w[WPL];
#pragma unroll
    for (int u = 0; u < WPL; ++u) {
        const int e = lane + 64 * u;
        w[u] = e < total ? s_words[e] : kSentinel;
    }
    // The k-th smallest word: the k-th key by a bitwise search on the key
    // bits, then, only when that key is tied, the right index among the tied.
    // (Keys are at most +inf, bit 31 clear; at least k of them are real.)
    uint32_t kh = 0;
    for (int b = 30; b >= 0; --b) {
        const uint32_t probe = kh | (1u << b);
        int below = 0;
#pragma unroll
        for (int u = 0; u < WPL; ++u) below += __popcll(__ballot(static_cast<uint32_t>(w[u] >> 32) < probe));
        if (below < p.k) kh = probe;
    }
    int below = 0, tied = 0;
#pragma unroll
    for (int u = 0; u < WPL; ++u) {
        const uint32_t hi = static_cast<uint32_t>(w[u] >> 32);
        below += __popcll(__ballot(hi < kh));
        tied += __popcll(__ballot(hi == kh));
    }
    const int need = p.k - below;  // 1 <= need <= tied
    uint32_t lo = 0;
    if (need < tied) {
        for (int b = 31; b >= 0; --b) {
            const uint32_t probe = lo | (1u << b);
            int cnt = 0;
#pragma unroll
            for (int u = 0; u < WPL; ++u)
                cnt += __popcll(__ballot(static_cast<uint32_t>(w[u] >> 32) == kh && static_cast<uint32_t>(w[u]) < probe));
            if (cnt < need) lo = probe;
        }
    } else {
        // every tied word is taken: the k-th is the largest of them
#pragma unroll
        for (int u = 0; u < WPL; ++u)
            if (static_cast<uint32_t>(w[u] >> 32) == kh) lo = max(lo, static_cast<uint32_t>(w[u]));
#pragma unroll
        for (int m = 1; m < 64; m <<= 1) lo = max(lo, static_cast<uint32_t>(__shfl_xor(static_cast<int>(lo), m, 64)));
    }
    const uint64_t ans = (static_cast<uint64_t>(kh) << 32) | lo;
    if (kh >= kInfBits) {
        if (lane == 0) queue_push(p.fb_rows, p.fb_count, p.fb_cap, qi);
        return;
    }
    int off = 0;
#pragma unroll
    for (int u = 0; u < WPL; ++u) {
        const bool m = w[u] <= ans;
        const uint64_t mask = __ballot(m);
        if (m) s_sel[off + lanes_below(mask)] = w[u];
        off += __popcll(mask);
    }
    wave_lds_sync();
    emit_selected(p, s_sel, qi, lane, local);
}

// emit_selected for 64 < k <= 64 KW: the same slot order, the bitonic network
// over 64 KW words (strides below 64 between lanes by shuffles, the others
// between registers of one lane).
template <int KW>
__device__ __forceinline__ void emit_selected_wide(const MfmaArgs& p, const uint64_t* sel, int qi, int lane,
                                                   int local) {
    // Word e = 64 u + lane sits in register u of its lane; the k-th word is
    // their maximum. Sorting the words by (group, index) -- group 0 for keys
    // below the k-th, 1 for keys equal to it, 2 past k -- puts every word at
    // its slot: below-k-th keys in ascending index, then the tied ones. The
    // bitonic network runs over 64 KW words: strides below 64 between lanes
    // (shuffles), the others between registers of one lane.
    constexpr int N = 64 * KW;
    const int k = p.k;
    uint64_t me[KW];
    uint64_t kw = 0;
#pragma unroll
    for (int u = 0; u < KW; ++u) {
        const int e = 64 * u + lane;
        me[u] = e < k ? sel[e] : 0;
        kw = me[u] > kw ? me[u] : kw;
    }
#pragma unroll
    for (int m = 1; m < 64; m <<= 1) {
        const uint32_t hi = static_cast<uint32_t>(__shfl_xor(static_cast<int>(kw >> 32), m, 64));
        const uint32_t lo = static_cast<uint32_t>(__shfl_xor(static_cast<int>(kw), m, 64));
        const uint64_t o = (static_cast<uint64_t>(hi) << 32) | lo;
        kw = o > kw ? o : kw;
    }
    const uint32_t kth = static_cast<uint32_t>(kw >> 32);
    uint64_t s[KW];
    uint32_t key[KW];
#pragma unroll
    for (int u = 0; u < KW; ++u) {
        key[u] = static_cast<uint32_t>(me[u] >> 32);
        const uint32_t grp = 64 * u + lane >= k ? 2u : (key[u] < kth ? 0u : 1u);
        s[u] = (static_cast<uint64_t>(grp) << 32) | static_cast<uint32_t>(me[u]);
    }
#pragma unroll
    for (int size = 2; size <= N; size <<= 1) {
#pragma unroll
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            if (stride >= 64) {
                const int su = stride / 64;
#pragma unroll
                for (int u = 0; u < KW; ++u) {
                    if (u & su) continue;
                    const int u2 = u | su;
                    const bool up = ((64 * u + lane) & size) == 0;
                    if (up ? s[u2] < s[u] : s[u2] > s[u]) {
                        const uint64_t ts = s[u];
                        s[u] = s[u2];
                        s[u2] = ts;
                        const uint32_t tk = key[u];
                        key[u] = key[u2];
                        key[u2] = tk;
                    }
                }
            } else {
#pragma unroll
                for (int u = 0; u < KW; ++u) {
                    const uint32_t ohi = static_cast<uint32_t>(__shfl_xor(static_cast<int>(s[u] >> 32), stride, 64));
                    const uint32_t olo = static_cast<uint32_t>(__shfl_xor(static_cast<int>(s[u]), stride, 64));
                    const uint32_t okey = static_cast<uint32_t>(__shfl_xor(static_cast<int>(key[u]), stride, 64));
                    const uint64_t os = (static_cast<uint64_t>(ohi) << 32) | olo;
                    const bool up = ((64 * u + lane) & size) == 0;
                    const bool lower = (lane & stride) == 0;
                    const bool swap = (lower == up) ? (os < s[u]) : (os > s[u]);
                    if (swap) {
                        s[u] = os;
                        key[u] = okey;
                    }
                }
            }
        }
    }
    if (local >= 0) {  // staged record (see emit_selected)
        int* rec = p.stage + static_cast<size_t>(local) * stage_words(k);
#pragma unroll
        for (int u = 0; u < KW; ++u) {
            const int e = 64 * u + lane;
            if (e < k) {
                rec[e] = static_cast<int>(static_cast<uint32_t>(s[u]));
                rec[k + e] = static_cast<int>(key[u]);
            }
        }
        if (lane == 0) {
            rec[2 * k] = static_cast<int>(kth);
            rec[2 * k + 1] = 1;
        }
        return;
    }
    const size_t ns = static_cast<size_t>(p.n);
#pragma unroll
    for (int u = 0; u < KW; ++u) {
        const int e = 64 * u + lane;
        if (e < k) {
            p.nbr_idx[static_cast<size_t>(e) * ns + qi] = static_cast<int>(static_cast<uint32_t>(s[u]));
            p.nbr_d2[static_cast<size_t>(e) * ns + qi] = __uint_as_float(key[u]);
        }
    }
    if (lane == 0) p.kdist2[qi] = __uint_as_float(kth);
}

// The emit for 64 < k <= kCollectKMax (the selects' WIDE instantiations).
__device__ __forceinline__ void emit_wide(const MfmaArgs& p, const uint64_t* sel, int qi, int lane,
                                          int local = -1) {
    if (p.k <= 128) {
        emit_selected_wide<2>(p, sel, qi, lane, local);
    } else {
        emit_selected_wide<4>(p, sel, qi, lane, local);
    }
}

// select_words for k > 64: the selected words are compacted into
// s_words[0, k) (read whole by then) for the wide emit.
template <int WPL>
__device__ __forceinline__ void select_words_wide(const MfmaArgs& p, uint64_t* s_words, int total, int qi,
                                                  int lane, int local) {
    uint64_t w[WPL];
#pragma unroll
    for (int u = 0; u < WPL; ++u) {
        const int e = lane + 64 * u;
        w[u] = e < total ? s_words[e] : kSentinel;
    }
    // The k-th smallest word: the k-th key by a bitwise search on the key
    // bits, then, only when that key is tied, the right index among the tied.
    // (Keys are at most +inf, bit 31 clear; at least k of them are real.)
    uint32_t kh = 0;
    for (int b = 30; b >= 0; --b) {
        const uint32_t probe = kh | (1u << b);
        int below = 0;
#pragma unroll
        for (int u = 0; u < WPL; ++u) below += __popcll(__ballot(static_cast<uint32_t>(w[u] >> 32) < probe));
        if (below < p.k) kh = probe;
    }
    int below = 0, tied = 0;
#pragma unroll
    for (int u = 0; u < WPL; ++u) {
        const uint32_t hi = static_cast<uint32_t>(w[u] >> 32);
        below += __popcll(__ballot(hi < kh));
        tied += __popcll(__ballot(hi == kh));
    }
    const int need = p.k - below;  // 1 <= need <= tied
    uint32_t lo = 0;
    if (need < tied) {
        for (int b = 31; b >= 0; --b) {
            const uint32_t probe = lo | (1u << b);
            int cnt = 0;
#pragma unroll
            for (int u = 0; u < WPL; ++u)
                cnt += __popcll(__ballot(static_cast<uint32_t>(w[u] >> 32) == kh && static_cast<uint32_t>(w[u]) < probe));
            if (cnt < need) lo = probe;
        }
    } else {
        // every tied word is taken: the k-th is the largest of them
#pragma unroll
        for (int u = 0; u < WPL; ++u)
            if (static_cast<uint32_t>(w[u] >> 32) == kh) lo = max(lo, static_cast<uint32_t>(w[u]));
#pragma unroll
        for (int m = 1; m < 64; m <<= 1) lo = max(lo, static_cast<uint32_t>(__shfl_xor(static_cast<int>(lo), m, 64)));
    }
    const uint64_t ans = (static_cast<uint64_t>(kh) << 32) | lo;
    if (kh >= kInfBits) {
        if (lane == 0) queue_push(p.fb_rows, p.fb_count, p.fb_cap, qi);
        return;
    }
    wave_lds_sync();  // every lane holds its words: the buffer takes the selection
    int off = 0;
#pragma unroll
    for (int u = 0; u < WPL; ++u) {
        const bool m = w[u] <= ans;
        const uint64_t mask = __ballot(m);
        if (m) s_words[off + lanes_below(mask)] = w[u];
        off += __popcll(mask);
    }
    wave_lds_sync();
    emit_wide(p, s_words, qi, lane, local);
}

// The wave select for k > 64 (the WIDE instantiation of the select kernel).
__device__ __forceinline__ void wide_select(const MfmaArgs& p, uint64_t* s_words, int total, int qi, int lane,
                                            int local) {
    if (total <= 128) {
        select_words_wide<2>(p, s_words, total, qi, lane, local);
    } else if (total <= 256) {
        select_words_wide<4>(p, s_words, total, qi, lane, local);
    } else if (total <= 512) {
        select_words_wide<8>(p, s_words, total, qi, lane, local);
    } else {
        select_words_wide<16>(p, s_words, total, qi, lane, local);
    }
}

// The list pass's partial lists, a wave per row: the row's slices x k words
// (each slice's list ascending, padded with sentinels; every candidate in one
// slice) read coalesced into LDS, and the k smallest emitted in the
// reference's slot order by the wave select's own search (the same words and
// the same order as a serial merge of the lists). slices x k <= 64 WPL.
template <int WPL>
__global__ void __launch_bounds__(256)
merge_slices_wave_kernel(MfmaArgs p) {
    __shared__ uint64_t s_words[4][64 * WPL];
    __shared__ uint64_t s_sel[4][64];
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int local = blockIdx.x * 4 + wave;
    const int rows = p.row_count ? min(p.rows, *p.row_count) : p.rows;
    if (local >= rows) return;  // whole waves; no block barrier below
    const int qi = p.qlist ? p.qlist[local] : p.row0 + local;
    const int k = p.k;
    const int total = p.slices * k;
    for (int e = lane; e < total; e += 64) {
        const int s = e / k;
        s_words[wave][e] = p.partial[(static_cast<size_t>(s) * p.rows + local) * k + (e - s * k)];
    }
    wave_lds_sync();
    select_words<WPL>(p, s_words[wave], total, qi, lane, s_sel[wave]);
}

__device__ __forceinline__ int query_of(const MfmaArgs& p, int local) {
    return p.qlist ? p.qlist[local] : p.row0 + local;
}

// One wave per query of the collect pass: sets of up to wave_select_max<DP>() entries
// are keyed and selected here; larger ones are queued for the block select
// (big_rows) or, when a slice buffer overflowed, for a second collect pass
// with larger buffers (ov_rows).
// WIDE: the instantiation for 64 < k <= kCollectKMax (the k <= 64 one keeps
// its registers and occupancy free of the wider emit).
template <int DP, bool WIDE>
__global__ void __launch_bounds__(256)
collect_select_kernel(MfmaArgs p) {
    __shared__ __attribute__((aligned(16))) float s_q[4][DP];
    constexpr int kWaveMax = wave_select_max<DP>();
    __shared__ uint64_t s_words[4][kWaveMax];
    __shared__ uint64_t s_sel[4][64];
    __shared__ int s_pre[4][33];
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int local = blockIdx.x * 4 + wave;
    if (local >= p.rows) return;  // whole waves; no block barrier below
    // staged: the record says "not emitted" until the emit overwrites it
    if (p.stage && lane == 0) p.stage[static_cast<size_t>(local) * stage_words(p.k) + 2 * p.k + 1] = 0;
    const int stage_row = p.stage ? local : -1;
    if (*p.guard) return;         // the sweep queued every row for the radix path
    if (p.heavy && p.heavy[local]) return;  // already queued for the list pass
    const int qi = query_of(p, local);
    const int nb = p.slices;      // <= 32
    const size_t base = static_cast<size_t>(local) * nb;
    const int mine = lane < nb ? p.ccount[base + lane] : 0;
    if (lane < nb) s_pre[wave][lane + 1] = mine;
    for (int t = lane; t < DP; t += 64) s_q[wave][t] = p.xs[static_cast<size_t>(qi) * DP + t];
    wave_lds_sync();
    if (lane == 0) {
        int run = 0;
        s_pre[wave][0] = 0;
        for (int b = 0; b < nb; ++b) {
            run += s_pre[wave][b + 1];
            s_pre[wave][b + 1] = run;
        }
    }
    wave_lds_sync();
    const int total = s_pre[wave][nb];
    if (__any(mine > p.cap)) {
        if (lane == 0) {
            const int at = atomicAdd(p.ov_count, 1);
            const int at2 = at < p.ov_cap ? 0 : atomicAdd(p.list_count, 1);
            if (at < p.ov_cap) {
                p.ov_rows[at] = qi;
            } else if (at2 < p.list_cap) {
                p.list_rows[at2] = qi;
            } else {
                queue_push(p.fb_rows, p.fb_count, p.fb_cap, qi);
            }
        }
        return;
    }
    if (total > kWaveMax) {
        if (lane == 0) queue_push(p.big_rows, p.big_count, p.rows, local);
        return;
    }
    if (total < p.k) {  // cannot happen with a valid seed; stay exact regardless
        if (lane == 0) queue_push(p.fb_rows, p.fb_count, p.fb_cap, qi);
        return;
    }
    // Exact keys straight from the slice buffers, every lane busy: entry e's
    // slice is the last b with s_pre[b] <= e (a binary search; empty slices
    // share their prefix with the next one).
#pragma unroll 1
    for (int e = lane; e < total; e += 64) {
        int lo = 0, hi = nb;
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (s_pre[wave][mid] <= e) {
                lo = mid;
            } else {
                hi = mid;
            }
        }
        const uint32_t pos = p.cbuf[(base + lo) * p.cap + (e - s_pre[wave][lo])];
        s_words[wave][e] = collected_word<DP>(p, s_q[wave], pos, qi);
    }
    wave_lds_sync();
    if constexpr (WIDE) {
        wide_select(p, s_words[wave], total, qi, lane, stage_row);
    } else if (total <= 128) {
        select_words<2>(p, s_words[wave], total, qi, lane, s_sel[wave], stage_row);
    } else if (total <= 256) {
        select_words<4>(p, s_words[wave], total, qi, lane, s_sel[wave], stage_row);
    } else if (kWaveMax <= 512 || total <= 512) {
        select_words<8>(p, s_words[wave], total, qi, lane, s_sel[wave], stage_row);
    } else if constexpr (kWaveMax > 512) {
        select_words<16>(p, s_words[wave], total, qi, lane, s_sel[wave], stage_row);
    }
}

// The staged records of the first-pass wave select into the reference's
// transposed outputs: one thread per original row i (record of position
// qpos[i]), so each slot's stores are consecutive words. Rows the select did
// not emit (flag 0) belong to the later passes and are left alone.
// The record is read as 16-byte groups, eight at a time, so a wave's 64
// records are consumed while their cache lines are still resident.
__global__ void __launch_bounds__(256)
unstage_kernel(const int* __restrict__ stage, const int* __restrict__ qpos, int n, int k,
               int* __restrict__ nbr_idx, float* __restrict__ nbr_d2, float* __restrict__ kdist2) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const int words = stage_words(k);
    const int4* rec = reinterpret_cast<const int4*>(stage + static_cast<size_t>(qpos[i]) * words);
    const int flag_at = 2 * k + 1;
    if (reinterpret_cast<const int*>(rec)[flag_at] == 0) return;
    const size_t ns = static_cast<size_t>(n);
    for (int g0 = 0; g0 < words / 4; g0 += 8) {
        int4 v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) v[u] = g0 + u < words / 4 ? rec[g0 + u] : make_int4(0, 0, 0, 0);
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int w[4] = {v[u].x, v[u].y, v[u].z, v[u].w};
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                const int e = 4 * (g0 + u) + c;
                if (e < k) {
                    nbr_idx[static_cast<size_t>(e) * ns + i] = w[c];
                } else if (e < 2 * k) {
                    nbr_d2[static_cast<size_t>(e - k) * ns + i] = __int_as_float(w[c]);
                } else if (e == 2 * k) {
                    kdist2[i] = __int_as_float(w[c]);
                }
            }
        }
    }
}

// One block per queued row (grid-stride). The row's collected entries are
// keyed in rounds that fill the free part of an LDS buffer of kBigChunk words;
// once the buffer has been sorted, only words below the running k-th word
// (the k-th smallest seen so far; the words are distinct) enter it, so later
// rounds mostly key and drop. The buffer is sorted (bitonic, over the next
// power of two it fills) when less than kBigMinRoom words are free or the
// entries run out, keeping the k smallest. Any number of entries.
// rows_list[r] is the row's local index in the collect pass it came from
// (null: r itself); that pass's buffers are p.cbuf / p.ccount / p.cap.
constexpr int kBigThreads = 512;
constexpr int kBigMinRoom = 512;

template <int DP, bool WIDE>
__global__ void __launch_bounds__(kBigThreads)
big_select_kernel(MfmaArgs p, const int* __restrict__ rows_list, const int* __restrict__ rows_count) {
    __shared__ __attribute__((aligned(16))) float s_q[DP];
    __shared__ uint64_t s_w[kBigChunk];
    __shared__ int s_pre[33];
    __shared__ int s_fill;
    if (*p.guard) return;
    const int count = min(*rows_count, p.rows);
    const int nb = p.slices;
    const int tid = threadIdx.x;
    for (int r = blockIdx.x; r < count; r += gridDim.x) {
        const int local = rows_list ? rows_list[r] : r;
        const int qi = query_of(p, local);
        const size_t base = static_cast<size_t>(local) * nb;
        __syncthreads();
        if (tid < nb) s_pre[tid + 1] = p.ccount[base + tid];
        for (int t = tid; t < DP; t += kBigThreads) s_q[t] = p.xs[static_cast<size_t>(qi) * DP + t];
        __syncthreads();
        if (tid == 0) {
            int run = 0;
            s_pre[0] = 0;
            bool over = false;
            for (int b = 0; b < nb; ++b) {
                over |= s_pre[b + 1] > p.cap;
                run += s_pre[b + 1];
                s_pre[b + 1] = run;
            }
            if (over) s_pre[0] = -1;
        }
        __syncthreads();
        if (s_pre[0] < 0 || s_pre[nb] < p.k) {
            // A slice buffer overflowed: the seeded list sweep redoes the row
            // when a queue is given (radix when it is full or absent).
            if (tid == 0) {
                const int at = (s_pre[0] < 0 && p.ov_rows) ? atomicAdd(p.ov_count, 1) : p.ov_cap;
                if (at < p.ov_cap) {
                    p.ov_rows[at] = qi;
                } else {
                    queue_push(p.fb_rows, p.fb_count, p.fb_cap, qi);
                }
            }
            continue;
        }
        const int total = s_pre[nb];
        int used = 0;               // words at s_w[0, used)
        bool sorted = false;        // s_w[0, used) is the k smallest so far, used == k
        uint64_t thr = kSentinel;   // after a sort: the k-th smallest word so far
        for (int e0 = 0; e0 < total || !sorted;) {
            if (e0 < total) {
                const int take = min(kBigChunk - used, total - e0);
                if (tid == 0) s_fill = used;
                __syncthreads();
                for (int i = tid; i < take; i += kBigThreads) {
                    const int e = e0 + i;
                    // entry e's slice: the last b with s_pre[b] <= e
                    int lo = 0, hi = nb;
                    while (hi - lo > 1) {
                        const int mid = (lo + hi) >> 1;
                        if (s_pre[mid] <= e) {
                            lo = mid;
                        } else {
                            hi = mid;
                        }
                    }
                    const uint32_t pos = p.cbuf[(base + lo) * p.cap + (e - s_pre[lo])];
                    const uint64_t w = collected_word<DP>(p, s_q, pos, qi);
                    if (w < thr) s_w[atomicAdd(&s_fill, 1)] = w;
                }
                __syncthreads();
                const int grown = s_fill;
                sorted = sorted && grown == used;
                used = grown;
                e0 += take;
                if (e0 < total && kBigChunk - used >= kBigMinRoom) continue;
                if (sorted) continue;
            }
            // Sort s_w[0, used) (padded to a power of two) and keep the k smallest.
            int m = 64;
            while (m < used) m <<= 1;
            for (int i = used + tid; i < m; i += kBigThreads) s_w[i] = kSentinel;
            __syncthreads();
            for (int size = 2; size <= m; size <<= 1) {
                for (int stride = size >> 1; stride > 0; stride >>= 1) {
                    for (int q = tid; q < m / 2; q += kBigThreads) {
                        const int lo = 2 * q - (q & (stride - 1));
                        const int hi = lo + stride;
                        const bool up = (lo & size) == 0;
                        const uint64_t a = s_w[lo], b2 = s_w[hi];
                        if ((b2 < a) == up) {
                            s_w[lo] = b2;
                            s_w[hi] = a;
                        }
                    }
                    __syncthreads();
                }
            }
            used = p.k;
            thr = s_w[p.k - 1];
            sorted = true;
        }
        if (tid < 64) {
            if (static_cast<uint32_t>(s_w[p.k - 1] >> 32) >= kInfBits) {
                if (tid == 0) queue_push(p.fb_rows, p.fb_count, p.fb_cap, qi);
            } else {
                if constexpr (WIDE) {
                    emit_wide(p, s_w, qi, tid);
                } else {
                    emit_selected(p, s_w, qi, tid);
                }
            }
        }
    }
}

// ---- FP64 band select over a collect pass (near-tie rows) ----
//
// A near-tie row (its k-th and (k+1)-th FP32 keys closer than FP32 can tell
// apart) is settled on FP64 keys over its band {j : e32(i, j) <= bound}; the
// collect pass ran with a threshold of at least `bound`, so the band is among
// the collected entries. The block keys them in FP32, keeps the band, re-keys
// it in FP64 (separate multiply and add, ascending feature order) and emits
// the k smallest (e64, index) pairs in the reference's slot order: below the
// k-th FP64 key in ascending index, then the tied ones in ascending index.
// nbr_d2 / kdist2 keep the FP32 keys of the chosen pairs. Rows whose slice
// buffer overflowed or whose collected set exceeds kBandMax go to the
// brute-force queue (the whole candidate range, same rule).
constexpr int kBandMax = 2048;  // collected entries one block holds

template <int DP>
__global__ void __launch_bounds__(256)
band_select_kernel(MfmaArgs p, const double* __restrict__ bound_of, int* __restrict__ brute_rows,
                   int* __restrict__ brute_count) {
    __shared__ __attribute__((aligned(16))) float s_q[DP];
    __shared__ uint64_t s_e64[kBandMax];  // non-negative doubles order as integers
    __shared__ int s_idx[kBandMax];
    __shared__ float s_e32[kBandMax];
    __shared__ int s_rank[kBandMax];
    __shared__ int s_pre[33];
    __shared__ int s_m;
    __shared__ uint64_t s_kth;
    __shared__ float s_kd;
    if (*p.guard) return;  // the collect pass queued every row as brute force
    const int count = min(*p.row_count, p.rows);
    const int nb = p.slices;
    const size_t ns = static_cast<size_t>(p.n);
    for (int r = blockIdx.x; r < count; r += gridDim.x) {
        const int qi = p.qlist[r];
        const double bound = bound_of[r];
        const size_t base = static_cast<size_t>(r) * nb;
        __syncthreads();
        if (threadIdx.x < nb) s_pre[threadIdx.x + 1] = p.ccount[base + threadIdx.x];
        for (int t = threadIdx.x; t < DP; t += 256) s_q[t] = p.xs[static_cast<size_t>(qi) * DP + t];
        if (threadIdx.x == 0) s_m = 0;
        __syncthreads();
        if (threadIdx.x == 0) {
            int run = 0;
            bool over = false;
            s_pre[0] = 0;
            for (int b = 0; b < nb; ++b) {
                over |= s_pre[b + 1] > p.cap;
                run += s_pre[b + 1];
                s_pre[b + 1] = run;
            }
            if (over || run > kBandMax) s_pre[0] = -1;
        }
        __syncthreads();
        if (s_pre[0] < 0) {
            if (threadIdx.x == 0) queue_push(brute_rows, brute_count, p.fb_cap, qi);
            continue;
        }
        const int total = s_pre[nb];
        for (int e = threadIdx.x; e < total; e += 256) {
            int b = 0;
            while (e >= s_pre[b + 1]) ++b;
            const uint32_t pos = p.cbuf[(base + b) * p.cap + (e - s_pre[b])];
            const int orig = p.cand ? p.cand[pos] : static_cast<int>(pos);
            if (orig == qi) continue;
            const float* row = p.xs_cand + static_cast<size_t>(pos) * DP;
            const float e32 = exact_key_q4<DP>(s_q, row);
            if (!(static_cast<double>(e32) <= bound)) continue;  // NaN stays out
            double e64 = 0.0;
#pragma unroll 16
            for (int t = 0; t < DP; ++t) {  // zero padding adds exact zeros
                const double df = __dsub_rn(static_cast<double>(s_q[t]), static_cast<double>(row[t]));
                e64 = __dadd_rn(e64, __dmul_rn(df, df));
            }
            const int at = atomicAdd(&s_m, 1);
            s_e64[at] = static_cast<uint64_t>(__double_as_longlong(e64));
            s_idx[at] = orig;
            s_e32[at] = e32;
        }
        __syncthreads();
        const int m = s_m;
        if (m < p.k) {  // cannot happen (the FP32 k-list lies in the band); stay exact regardless
            if (threadIdx.x == 0) queue_push(brute_rows, brute_count, p.fb_cap, qi);
            continue;
        }
        // Rank of each entry in (e64, index) order; the k-th is the one of rank k - 1.
        for (int t = threadIdx.x; t < m; t += 256) {
            const uint64_t et = s_e64[t];
            const int it = s_idx[t];
            int rank = 0;
            for (int u = 0; u < m; ++u) {
                const uint64_t eu = s_e64[u];
                rank += eu < et || (eu == et && s_idx[u] < it);
            }
            s_rank[t] = rank;
            if (rank == p.k - 1) {
                s_kth = et;
                s_kd = s_e32[t];
            }
        }
        __syncthreads();
        const uint64_t kth = s_kth;
        for (int t = threadIdx.x; t < m; t += 256) {
            const int rank = s_rank[t];
            if (rank >= p.k) continue;
            int slot = rank;  // ties of the k-th sit at ranks [below, k) in index order
            if (s_e64[t] < kth) {
                slot = 0;
                const int it = s_idx[t];
                for (int u = 0; u < m; ++u) slot += s_e64[u] < kth && s_idx[u] < it;
            }
            p.nbr_idx[static_cast<size_t>(slot) * ns + qi] = s_idx[t];
            p.nbr_d2[static_cast<size_t>(slot) * ns + qi] = s_e32[t];
        }
        if (threadIdx.x == 0) p.kdist2[qi] = s_kd;
    }
}

// One thread per query: fold the slices' partial lists, then emit.
template <int KMAX>
__global__ void __launch_bounds__(64)
merge_slices_kernel(const uint64_t* __restrict__ partial, int slices, int rows, int row0,
                    const int* __restrict__ qlist, const int* __restrict__ row_count, int n, int k, int* __restrict__ nbr_idx,
                    float* __restrict__ nbr_d2, float* __restrict__ kdist2,
                    int* __restrict__ fb_rows, int* __restrict__ fb_count, int fb_cap) {
    __shared__ uint64_t stage[KMAX][64];
    const int lane = threadIdx.x;
    const int local = blockIdx.x * 64 + lane;
    if (local >= (row_count ? min(rows, *row_count) : rows)) return;
    const int qi = qlist ? qlist[local] : row0 + local;
    uint64_t a[KMAX];
    list_init<KMAX>(a, k);
    for (int s = 0; s < slices; ++s) {
        const uint64_t* src = partial + (static_cast<size_t>(s) * rows + local) * k;
        for (int t = 0; t < k; ++t) {
            const uint64_t e = src[t];
            if (e >= a[0]) break;  // ascending: the rest cannot enter
            list_insert<KMAX>(a, e);
        }
    }
#pragma unroll
    for (int t = 0; t < KMAX; ++t) {
        if (t < k) stage[k - 1 - t][lane] = a[t];
    }
    if (!emit_list(&stage[0][lane], k, n, qi, nbr_idx, nbr_d2, kdist2)) {
        queue_push(fb_rows, fb_count, fb_cap, qi);
    }
}

// ---- operand preparation ----
//
// Rows of DP <= 128 features are stored whole in the form layout. Wider rows
// (the K-chunked collect pass) are stored in chunks of kWideChunk features:
// block (tile, chunk) holds the 32 rows of a candidate tile, each as the form
// layout of that chunk's features, so one chunk of a tile is contiguous.
// Element (row i, feature f) of a padded width dp in the operand layout, in
// elements (floats for FP32; bf16 values for the split form, hi then lo).
__device__ __forceinline__ size_t operand_at(size_t i, int f, int dp, bool bf16, bool lo) {
    const bool wide = dp >= kWideMin;
    const int w = wide ? kWideChunk : dp;  // features per stored (sub)row
    const int kc = wide ? f / kWideChunk : 0;
    const int fl = f - kc * w;
    const size_t sub = wide ? ((i / 32) * (dp / kWideChunk) + kc) * 32 + i % 32 : i;
    if (bf16) return sub * 2 * w + (lo ? w : 0) + fl;
    return sub * w + ((fl & 1) ? w / 2 + fl / 2 : fl / 2);
}

// Screen operands for rows [0, padded_rows): rows past n are zero.
// Screen operands for candidate positions [0, padded_rows) (position p holds
// row cand[p], or row p without an order); rows past n are zero.
__global__ void __launch_bounds__(256)
screen_rows_f32_kernel(const float* __restrict__ xc, const int* __restrict__ cand, int n,
                       int padded_rows, int dp, float* __restrict__ xa) {
    const size_t cell = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (cell >= static_cast<size_t>(padded_rows) * dp) return;
    const size_t i = cell / dp;
    const int f = static_cast<int>(cell % dp);
    const size_t at = operand_at(i, f, dp, false, false);
    if (i >= static_cast<size_t>(n)) {
        xa[at] = 0.0f;
        return;
    }
    const size_t row = cand ? static_cast<size_t>(cand[i]) : i;
    xa[at] = xc[row * dp + f];
}

// Split BF16 operands: row p = [bf16(x) for each feature, bf16(x - bf16(x)) ...].
__global__ void __launch_bounds__(256)
screen_rows_bf16_kernel(const float* __restrict__ xc, const int* __restrict__ cand, int n,
                        int dp, size_t padded_cells, __bf16* __restrict__ xb) {
    const size_t cell = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (cell >= padded_cells) return;
    const size_t i = cell / dp;
    const int f = static_cast<int>(cell - i * dp);
    const size_t at_hi = operand_at(i, f, dp, true, false);
    const size_t at_lo = operand_at(i, f, dp, true, true);
    if (i >= static_cast<size_t>(n)) {
        xb[at_hi] = static_cast<__bf16>(0.0f);
        xb[at_lo] = static_cast<__bf16>(0.0f);
        return;
    }
    const size_t src = cand ? static_cast<size_t>(cand[i]) : i;
    const float v = xc[src * dp + f];
    const __bf16 hi = static_cast<__bf16>(v);
    xb[at_hi] = hi;
    xb[at_lo] = static_cast<__bf16>(v - static_cast<float>(hi));
}

// |x|^2 (1 - c) per candidate position in [0, padded_rows); NaN past n,
// which fails every compare.
__global__ void __launch_bounds__(256)
adjusted_norms_kernel(const float* __restrict__ norms, const int* __restrict__ cand, int n,
                      int padded_rows, float c, float* __restrict__ madj) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= padded_rows) return;
    madj[i] = i < n ? norms[cand ? cand[i] : i] * (1.0f - c) : __uint_as_float(0x7FC00000u);
}

template <class Op, int DP, int KMAX>
void launch_mfma(const MfmaArgs& a, hipStream_t stream) {
    const int qtiles = (a.rows + 31) / 32;
    // An odd grid width (blocks past the rows return at once): workgroups
    // are dealt to the 8 XCDs in turn, so with a width that is a multiple of
    // 8 every XCD would run all slices of a fixed set of query blocks (the
    // heavy rows' list pass at n = 500k, d = 64: 8.5 ms against 5.0 ms).
    int gx = (qtiles + kTilesPerBlock - 1) / kTilesPerBlock;
    if (a.slices > 1 && gx % 2 == 0) ++gx;
    const dim3 grid(gx, a.slices);
    {
        ProfileScope prof("mfma_sweep", stream);
        mfma_sweep_kernel<Op, DP, KMAX><<<grid, 64 * kTilesPerBlock, 0, stream>>>(a);
    }
    if (a.slices > 1) {
        ProfileScope prof("merge_slices", stream);
        const int words = a.slices * a.k;
        const unsigned wgrid = static_cast<unsigned>((a.rows + 3) / 4);
        MfmaArgs m = a;
        m.stage = nullptr;  // the merged lists go straight to the outputs
        if (words <= 128) {
            merge_slices_wave_kernel<2><<<wgrid, 256, 0, stream>>>(m);
        } else if (words <= 256) {
            merge_slices_wave_kernel<4><<<wgrid, 256, 0, stream>>>(m);
        } else if (words <= 512) {
            merge_slices_wave_kernel<8><<<wgrid, 256, 0, stream>>>(m);
        } else if (words <= 1024) {
            merge_slices_wave_kernel<16><<<wgrid, 256, 0, stream>>>(m);
        } else {
            merge_slices_kernel<KMAX><<<(a.rows + 63) / 64, 64, 0, stream>>>(
                a.partial, a.slices, a.rows, a.row0, a.qlist, a.row_count, a.n, a.k, a.nbr_idx, a.nbr_d2,
                a.kdist2, a.fb_rows, a.fb_count, a.fb_cap);
        }
    }
}

template <class Op, int DP>
void launch_collect(const MfmaArgs& a, hipStream_t stream) {
    const int qtiles = (a.rows + 31) / 32;
    const dim3 grid((qtiles + kTilesPerBlock - 1) / kTilesPerBlock, a.slices);
    const unsigned big_grid = static_cast<unsigned>(std::min(a.rows, 1024));
    {
        ProfileScope prof(a.big_rows ? "mfma_collect" : "second_collect", stream);
        if constexpr (DP >= kWideMin) {
            mfma_wide_collect_kernel<Op, DP><<<grid, 64 * kTilesPerBlock, 0, stream>>>(a);
        } else {
            mfma_collect_kernel<Op, DP><<<grid, 64 * kTilesPerBlock, 0, stream>>>(a);
        }
    }
    const bool wide = a.k > kListKMax;
    if (a.big_rows) {
        // First pass: a wave per row, large sets to the block select.
        {
            ProfileScope prof("collect_select", stream);
            if (wide) {
                collect_select_kernel<DP, true><<<(a.rows + 3) / 4, 256, 0, stream>>>(a);
            } else {
                collect_select_kernel<DP, false><<<(a.rows + 3) / 4, 256, 0, stream>>>(a);
            }
            if (a.stage) {
                unstage_kernel<<<(a.n + 255) / 256, 256, 0, stream>>>(a.stage, a.qpos, a.n, a.k, a.nbr_idx,
                                                                      a.nbr_d2, a.kdist2);
            }
        }
        ProfileScope prof("big_select", stream);
        if (wide) {
            big_select_kernel<DP, true><<<big_grid, kBigThreads, 0, stream>>>(a, a.big_rows, a.big_count);
        } else {
            big_select_kernel<DP, false><<<big_grid, kBigThreads, 0, stream>>>(a, a.big_rows, a.big_count);
        }
    } else {
        // Second pass over overflowed rows: all large.
        ProfileScope prof("second_select", stream);
        if (wide) {
            big_select_kernel<DP, true><<<big_grid, kBigThreads, 0, stream>>>(a, nullptr, a.row_count);
        } else {
            big_select_kernel<DP, false><<<big_grid, kBigThreads, 0, stream>>>(a, nullptr, a.row_count);
        }
    }
}

// Buffers of the symmetric first pass besides the row slices.
struct SymBuffers {
    float* wcol;
    float2* wmax;
    float2* nmax;        // per (tile, half): largest adjusted norm (the two-level screen's margin)
    uint64_t* pool;      // column words, in chunks of kChunkWords
    int* chunk_region;   // per chunk: query tile * S + slice of the wave that owns it
    int* chunk_count;    // per chunk: words written
    int* next_chunk;     // chunks handed out
    int max_chunks;
    int* col_over;
    int slices;  // S: offset slices (the rows have 2S buffer slices)
    // pruning: per-tile boxes and largest seeds, per (block, slice) the live chunks
    float* box = nullptr;
    float* tmax = nullptr;
    float* gbox = nullptr;   // per group of kPruneGroup tiles
    float* gtmax = nullptr;
    int* live = nullptr;
    int* live_count = nullptr;
    int live_stride = 0;
};

// Symmetric first pass of a full run, the column pairs into their rows, then
// the selects.
template <class Op, int DP>
void launch_sym(const MfmaArgs& a, const SymBuffers& sb, hipStream_t stream) {
    constexpr int kBlockTiles = kTilesPerBlock * sym_query_tiles<Op, DP>();
    constexpr int rows_per_block = 32 * kBlockTiles;
    const int nb = (a.n + rows_per_block - 1) / rows_per_block;
    if (sb.live) {
        ProfileScope prof("sym_prune", stream);
        const int tiles = kBlockTiles * nb;
        tile_bounds_kernel<DP><<<(tiles * DP + 255) / 256, 256, 0, stream>>>(
            a.xs_cand, a.cand, a.tseed, a.heavy, a.n, tiles, sb.box, sb.tmax);
        const int groups = (tiles + kPruneGroup - 1) / kPruneGroup;
        group_bounds_kernel<DP><<<(groups * DP + 255) / 256, 256, 0, stream>>>(sb.box, sb.tmax, tiles, sb.gbox,
                                                                              sb.gtmax);
        sym_live_chunks_kernel<DP, sym_chunk_tiles<Op, DP>(), kBlockTiles>
            <<<dim3(nb, sb.slices), 256, 0, stream>>>(sb.box, sb.tmax, sb.gbox, sb.gtmax, nb, sb.slices,
                                                      sb.live_stride, sb.live, sb.live_count,
                                                      profile_tile_counter());
        CULOF_CHECK(hipGetLastError());
    } else if (unsigned long long* counter = profile_tile_counter()) {
        unsigned long long pairs = 0;
        for (int B = 0; B < nb; ++B) {
            const int h_b = (nb - 1) / 2 + ((nb % 2 == 0 && B < nb / 2) ? 1 : 0);
            pairs += static_cast<unsigned long long>(h_b + 1) * kBlockTiles * kBlockTiles;
        }
        add_counter_kernel<<<1, 64, 0, stream>>>(counter, pairs);
    }
    {
        ProfileScope prof("mfma_collect", stream);
        mfma_sym_collect_kernel<Op, DP><<<nb * sb.slices, 64 * kTilesPerBlock, 0, stream>>>(
            a, sb.wcol, sb.wmax, sb.nmax, sb.pool, sb.chunk_region, sb.chunk_count, sb.next_chunk, sb.max_chunks,
            sb.col_over, sb.live, sb.live_count, sb.live_stride, sb.slices);
    }
    {
        ProfileScope prof("scatter_pairs", stream);
        scatter_pairs_kernel<<<8192, kChunkWords, 0, stream>>>(sb.pool, sb.chunk_region, sb.chunk_count,
                                                      sb.next_chunk, sb.max_chunks, sb.slices, a.cap,
                                                      a.guard, a.cbuf, a.ccount);
        mark_overflow_kernel<<<(a.n + 255) / 256, 256, 0, stream>>>(sb.col_over, a.n, sb.slices, a.cap, a.ccount);
    }
    const bool wide = a.k > kListKMax;
    {
        ProfileScope prof("collect_select", stream);
        if (wide) {
            collect_select_kernel<DP, true><<<(a.rows + 3) / 4, 256, 0, stream>>>(a);
        } else {
            collect_select_kernel<DP, false><<<(a.rows + 3) / 4, 256, 0, stream>>>(a);
        }
        if (a.stage) {
            unstage_kernel<<<(a.n + 255) / 256, 256, 0, stream>>>(a.stage, a.qpos, a.n, a.k, a.nbr_idx, a.nbr_d2,
                                                                  a.kdist2);
        }
    }
    ProfileScope prof("big_select", stream);
    if (wide) {
        big_select_kernel<DP, true><<<std::min(a.rows, 1024), kBigThreads, 0, stream>>>(a, a.big_rows, a.big_count);
    } else {
        big_select_kernel<DP, false><<<std::min(a.rows, 1024), kBigThreads, 0, stream>>>(a, a.big_rows, a.big_count);
    }
}

template <class Op, int DP>
bool dispatch_k(const MfmaArgs* a, int k, hipStream_t stream) {
    if (a && a->collect) {
        if (k > kCollectKMax) return false;
        launch_collect<Op, DP>(*a, stream);
        return true;
    }
    if constexpr (DP >= kWideMin) {
        return false;  // no list sweep for wide rows
    } else {
        if (k <= 16) {
            if (a) launch_mfma<Op, DP, 16>(*a, stream);
            return true;
        }
        if (k <= 32) {
            if (a) launch_mfma<Op, DP, 32>(*a, stream);
            return true;
        }
        if (k <= 48) {  // e.g. k + 1 = 33 of the FP64 tie-break at k = 32: two waves per SIMD
            if (a) launch_mfma<Op, DP, 48>(*a, stream);
            return true;
        }
        if (k <= 64) {
            if (a) launch_mfma<Op, DP, 64>(*a, stream);
            return true;
        }
        return false;
    }
}

// Calls f(std::integral_constant<int, DP>) for the padded widths Op is built for.
template <class Op, class F>
bool for_dp(int dp, F&& f) {
    using std::integral_constant;
    if constexpr (std::is_same_v<Op, ScreenF32>) {
        switch (dp) {
            case 4: return f(integral_constant<int, 4>{});
            case 8: return f(integral_constant<int, 8>{});
            default: break;
        }
    }
    switch (dp) {
        case 16: return f(integral_constant<int, 16>{});
        case 32: return f(integral_constant<int, 32>{});
        case 64: return f(integral_constant<int, 64>{});
        case 128: return f(integral_constant<int, 128>{});
        case 256: return f(integral_constant<int, 256>{});
        case 512: return f(integral_constant<int, 512>{});
        default: return false;
    }
}

template <class Op>
bool dispatch(int dp, int k, const MfmaArgs* a, hipStream_t stream) {
    return for_dp<Op>(dp, [&](auto D) { return dispatch_k<Op, decltype(D)::value>(a, k, stream); });
}

// The band pass: collect over the queued near-tie rows, then the FP64 select.
template <class Op, int DP>
void launch_band(const MfmaArgs& a, const double* bound, int* brute_rows, int* brute_count,
                 hipStream_t stream) {
    const int qtiles = (a.rows + 31) / 32;
    const dim3 grid((qtiles + kTilesPerBlock - 1) / kTilesPerBlock, a.slices);
    {
        ProfileScope prof("tie_collect", stream);
        if constexpr (DP >= kWideMin) {
            mfma_wide_collect_kernel<Op, DP><<<grid, 64 * kTilesPerBlock, 0, stream>>>(a);
        } else {
            mfma_collect_kernel<Op, DP><<<grid, 64 * kTilesPerBlock, 0, stream>>>(a);
        }
    }
    ProfileScope prof("tie_band", stream);
    band_select_kernel<DP><<<std::min(a.rows, 1024), 256, 0, stream>>>(a, bound, brute_rows, brute_count);
}

}  // namespace

int mfma_padded_dims(int n_dims, int form) {
    const int dp = padded_dims(n_dims);
    return (form == kFormBF16 && dp < 16) ? 16 : dp;
}

bool mfma_supported(int n_dims, int k, int form) {
    const int dp = mfma_padded_dims(n_dims, form);
    auto any = [](auto) { return true; };
    return k <= kCollectKMax && (form == kFormBF16 ? for_dp<ScreenBF16>(dp, any) : for_dp<ScreenF32>(dp, any));
}

bool mfma_list_supported(int n_dims, int k, int form) {
    const int dp = mfma_padded_dims(n_dims, form);
    return form == kFormBF16 ? dispatch<ScreenBF16>(dp, k, nullptr, nullptr)
                             : dispatch<ScreenF32>(dp, k, nullptr, nullptr);
}

int mfma_slices(int n_points, int rows) {
    // Each slice restarts every query's list, so split the candidates only as
    // far as needed for about 8 query tiles per SIMD (1024 SIMDs), and never
    // below ~96 candidate tiles per slice.
    const long long waves = (rows + 31) / 32;
    long long s = (8192 + waves - 1) / waves;
    const long long max_by_len = std::max(1, ((n_points + 31) / 32) / 96);
    s = std::min<long long>(std::min<long long>(s, 8), max_by_len);
    return static_cast<int>(std::max<long long>(s, 1));
}

// First-pass buffer capacity per (row, slice): kSelectMax entries up to
// k = 64; beyond, 16 k (at most 4096), since a seed's candidate set grows with
// k (it is the k-th key among a fixed window) and a full slice sends the row
// to the slower passes.
int first_cap(int k) {
    return k <= kListKMax ? kSelectMax : std::min(4096, (16 * k + 255) / 256 * 256);
}

// Collect buffers: first_cap(k) positions per (query, slice) in the first pass;
// rows whose slice buffer overflows (at most kSecondRows of them) are
// collected again with kSecondCap per slice, then go to the block select.
// The second pass serves few rows, so it splits the candidates into many
// short slices to keep its latency low. Rows that overflow it too (candidate
// sets of many thousands, e.g. outliers in high dimension, whose distances
// concentrate) go to the seeded list sweep over as many slices (kListRows of
// them), which keeps no candidate buffer at all.
// Pool chunks of the symmetric pass's column words: 256 words (two chunks)
// per row, at least 8192 chunks.
int sym_chunks(int n_points) { return std::max(2 * n_points, 8192); }

// Operand rows are padded to whole blocks of query tiles (the symmetric pass
// pairs blocks; the other kernels ignore the extra rows).
size_t padded_positions(int n_points) {
    const size_t b = 2 * 32 * kTilesPerBlock;  // blocks of up to two query tiles per wave
    return (static_cast<size_t>(n_points) + b - 1) / b * b;
}

// Screen operand rows (form layout) and adjusted norms, in candidate order
// (cand: position -> original index; null = natural order).
void prepare_screen(const float* centred, const float* norms, const int* cand, int n_points, int dp,
                    bool bf16, float c, void* xa, float* madj, hipStream_t stream) {
    ProfileScope prof("screen_prep", stream);
    const int padded_rows = static_cast<int>(padded_positions(n_points));
    const size_t padded_cells = static_cast<size_t>(padded_rows) * dp;
    const unsigned grid = static_cast<unsigned>((padded_cells + 255) / 256);
    if (bf16) {
        screen_rows_bf16_kernel<<<grid, 256, 0, stream>>>(centred, cand, n_points, dp, padded_cells,
                                                            static_cast<__bf16*>(xa));
    } else {
        screen_rows_f32_kernel<<<grid, 256, 0, stream>>>(centred, cand, n_points, padded_rows, dp,
                                                           static_cast<float*>(xa));
    }
    CULOF_CHECK(hipGetLastError());
    adjusted_norms_kernel<<<(padded_rows + 255) / 256, 256, 0, stream>>>(norms, cand, n_points,
                                                                          padded_rows, c, madj);
    CULOF_CHECK(hipGetLastError());
}

constexpr int kSecondRows = 1024;
constexpr int kSecondCap = 4096;
constexpr int kListRows = 8192;
int second_slices(int n_points) {
    const int tiles = (n_points + 31) / 32;
    return std::max(1, std::min(32, tiles / 8));
}

// Symmetric first pass (a full run with the candidate order): 2S slices of
// first_cap(k) entries per row (S row slices, S column slices); a row with a full
// slice goes to the second pass.
// CULOF_SYM=0 turns it off (the same results; for measurement).
constexpr int kSymMinRows = 32768;
bool sym_enabled() {
    const char* env = std::getenv("CULOF_SYM");
    return !(env && env[0] == '0');
}
// Offsets of the half ring split so the pass has about 16 waves per SIMD (a
// short tail when the last round of waves is partly filled), with at least 8
// super-tiles per slice and at most 16 slices (the selects take 32 buffer
// slices per row).
// Query tiles per wave of the symmetric pass (sym_query_tiles on the host).
int sym_qt(int dp, int form) { return form == kFormBF16 && dp >= 64 ? 2 : 1; }

int sym_slices(int n_points, int qt) {
    const int rows_per_block = 32 * kTilesPerBlock * qt;
    const int nb = (n_points + rows_per_block - 1) / rows_per_block;
    const int by_waves = (16384 + kTilesPerBlock * nb - 1) / (kTilesPerBlock * nb);
    const int by_len = std::max(1, (nb / 2 + 1) / 8);
    const int cap = std::min(by_len, 16);
    int s = std::max(1, std::min(by_waves, cap));
    // An odd count. Workgroups are dealt to the 8 XCDs in turn and a block's
    // slices are consecutive workgroups, so with an even count every XCD
    // would run the same one or few slices of all blocks, their candidate
    // streams crowding the same memory at once (measured on MI355X: the BF16
    // pass at n = 100k, d = 128 takes 3.2 ms with 9, 11, 13 or 15 slices and
    // 3.4-5.7 ms with 6, 8, 10, 12 or 16; at n = 500k, d = 64, FP32, 13 ms
    // with 3 slices against 18 ms with 2).
    if (s % 2 == 0) s = s + 1 <= cap ? s + 1 : s - 1;
    return s;
}

// CULOF_PRUNE=0 turns the symmetric pass's tile pruning off (the same
// results; for measurement).
bool prune_enabled() {
    const char* env = std::getenv("CULOF_PRUNE");
    return !(env && env[0] == '0');
}

// Sizes of the pruning buffers of a symmetric pass over n_points rows with ns
// slices: tiles (boxes and largest seeds) and live-chunk entries per
// (block, slice).
struct PruneShape {
    int nb = 0;
    int tiles = 0;
    int stride = 0;
};

PruneShape prune_shape(int n_points, int dp, int form, int ns) {
    PruneShape ps;
    const int qt = sym_qt(dp, form);
    const int block_tiles = kTilesPerBlock * qt;
    int ct = 1;
    auto chunk = [&](auto D) {
        constexpr int DPv = decltype(D)::value;
        if constexpr (DPv < kWideMin) {
            ct = form == kFormBF16 ? sym_chunk_tiles<ScreenBF16, DPv>() : sym_chunk_tiles<ScreenF32, DPv>();
        }
        return true;
    };
    if (form == kFormBF16) {
        for_dp<ScreenBF16>(dp, chunk);
    } else {
        for_dp<ScreenF32>(dp, chunk);
    }
    ps.nb = (n_points + 32 * block_tiles - 1) / (32 * block_tiles);
    ps.tiles = block_tiles * ps.nb;
    const int span = (ps.nb - 1) / 2 + 2;
    const int per = (span + ns - 1) / ns;
    ps.stride = (block_tiles * per + ct - 1) / ct + 1;
    return ps;
}

size_t prune_bytes(int n_points, int dp, int form, int ns) {
    const PruneShape ps = prune_shape(n_points, dp, form, ns);
    const size_t groups = (static_cast<size_t>(ps.tiles) + kPruneGroup - 1) / kPruneGroup;
    return (static_cast<size_t>(ps.tiles) + groups) * (2 * dp + 1) * sizeof(float) +
           static_cast<size_t>(ps.nb) * ns * (ps.stride + 1) * sizeof(int) + 6 * 256;
}

size_t mfma_scratch_bytes(int n_points, int dp, int form, int rows, int k, bool collect) {
    const size_t padded = padded_positions(n_points);
    const size_t rows_bytes = form == kFormBF16 ? 2 * sizeof(uint16_t) : sizeof(float);
    const size_t operand = padded * dp * rows_bytes;
    const size_t madj = sizeof(float) * padded;
    const int s = mfma_slices(n_points, rows);
    size_t lists = 0;
    if (collect) {
        const size_t slots = static_cast<size_t>(rows) * s;
        const size_t second = static_cast<size_t>(std::min(rows, kSecondRows)) * second_slices(n_points);
        const size_t ss = static_cast<size_t>(sym_slices(n_points, sym_qt(dp, form)));
        const size_t sym = static_cast<size_t>(rows) * (2 * ss * (first_cap(k) + 1) + 2) * sizeof(uint32_t) +
                           padded * sizeof(float) + 2 * (padded / 32 * sizeof(float2) + 256) +
                           static_cast<size_t>(sym_chunks(n_points)) * (kChunkWords * sizeof(uint64_t) + 2 * sizeof(int)) +
                           prune_bytes(n_points, dp, form, static_cast<int>(ss)) + 2048;
        // (the symmetric pass serves full runs only)
        const bool sym_run = rows == n_points && n_points >= kSymMinRows;
        const size_t first = std::max(slots * (first_cap(k) + 1) * sizeof(uint32_t), sym_run ? sym : 0) + 256;
        lists = first + static_cast<size_t>(rows) * 2 * sizeof(int) + 4 * sizeof(int) +
                second * kSecondCap * sizeof(uint32_t) + second * sizeof(int) +
                static_cast<size_t>(std::min(rows, kListRows)) * (sizeof(int) +
                    (k <= kListKMax && dp < kWideMin ? sizeof(uint64_t) * second_slices(n_points) * k : 0)) + 16 * 256;
        // the wave select's staged records (a full run in candidate order)
        lists += static_cast<size_t>(rows) * stage_words(k) * sizeof(int) + 256;
        if (k > kListKMax || dp >= kWideMin) {  // the list rows' stored keys
            lists += static_cast<size_t>(brute_batch_rows(n_points, std::min(rows, kListRows))) * n_points *
                     sizeof(float) + 256;
        }
    } else if (s > 1) {
        lists = sizeof(uint64_t) * s * static_cast<size_t>(rows) * k;
    }
    return operand + 256 + madj + 256 + lists + 256;
}

void mfma_knn(const float* keyrows, const float* centred, int stride, const float* norms,
              const int* guard, int n_points, int n_dims, int k, int form, int row0, int rows,
              const CandidateOrder* ord, const int* qlist, const float* tseed,
              void* scratch, int* nbr_idx, float* nbr_d2, float* kdist2, int* fallback_rows,
              int* fallback_count, int fallback_cap, hipStream_t stream) {
    const int dp = mfma_padded_dims(n_dims, form);
    if (stride != dp) throw std::invalid_argument("mfma_knn: row stride mismatch");
    if (rows <= 0) return;
    if (ord && ord->dp != dp) throw std::invalid_argument("mfma_knn: order built for another stride");
    const bool bf16 = form == kFormBF16;
    const float c = bf16 ? ScreenBF16::slack(dp) : ScreenF32::slack(dp);
    const int padded_rows = static_cast<int>(padded_positions(n_points));
    const size_t padded_cells = static_cast<size_t>(padded_rows) * dp;
    unsigned char* base = static_cast<unsigned char*>(scratch);
    auto take = [&](size_t bytes) {
        void* r = base;
        base += (bytes + 255) / 256 * 256;
        return r;
    };
    void* xa = take(padded_cells * 4);  // f32, or bf16 hi + lo
    float* madj = static_cast<float*>(take(sizeof(float) * padded_rows));
    const int slices = mfma_slices(n_points, rows);
    if (slices > 32) throw std::invalid_argument("mfma_knn: too many slices");
    const bool collect = tseed != nullptr;
    MfmaArgs args{xa, centred, keyrows, ord ? ord->xs_ord : keyrows, ord ? ord->order : nullptr,
                  qlist, tseed, madj, norms, guard, n_points, k, c, row0, rows,
                  slices, nullptr, collect ? 1 : 0, nullptr, nullptr, nullptr, kSelectMax,
                  nullptr, nullptr, 0, nullptr, nullptr, nullptr, nullptr, 0, nbr_idx, nbr_d2,
                  kdist2, fallback_rows, fallback_count, fallback_cap, nullptr, ord ? ord->inv : nullptr};
    MfmaArgs second = args;
    MfmaArgs list = args;
    float* brute_keys = nullptr;  // stored keys of the list rows where no list kernel serves
    int brute_batch = 0;
    // A full run in candidate order computes each pair of tiles once.
    // (Below kSymMinRows the extra launches of the symmetric pass cost more
    // than the halved matrix work saves.)
    // (Wide rows take the K-chunked pass, which has no symmetric form.)
    const bool sym = collect && ord && qlist == ord->order && row0 == 0 && rows == n_points &&
                     n_points >= kSymMinRows && dp < kWideMin && sym_enabled();
    SymBuffers sb{};
    int* heavy = nullptr;
    if (collect) {
        if (sym) {
            sb.slices = sym_slices(n_points, sym_qt(dp, form));
            args.slices = 2 * sb.slices;
            args.cap = first_cap(k);
            const size_t slots = static_cast<size_t>(rows) * args.slices;
            args.cbuf = static_cast<uint32_t*>(take(slots * args.cap * sizeof(uint32_t)));
            args.ccount = static_cast<int*>(take(slots * sizeof(int)));
            heavy = static_cast<int*>(take(static_cast<size_t>(rows) * sizeof(int)));
            args.heavy = heavy;
            sb.wcol = static_cast<float*>(take(sizeof(float) * padded_rows));
            sb.wmax = static_cast<float2*>(take(sizeof(float2) * (padded_rows / 32)));
            sb.nmax = static_cast<float2*>(take(sizeof(float2) * (padded_rows / 32)));
            sb.max_chunks = sym_chunks(n_points);
            const size_t chunks = static_cast<size_t>(sb.max_chunks);
            sb.pool = static_cast<uint64_t*>(take(chunks * kChunkWords * sizeof(uint64_t)));
            sb.chunk_region = static_cast<int*>(take(chunks * sizeof(int)));
            sb.chunk_count = static_cast<int*>(take(chunks * sizeof(int)));
            sb.next_chunk = static_cast<int*>(take(sizeof(int)));
            sb.col_over = static_cast<int*>(take(static_cast<size_t>(rows) * sizeof(int)));
            if (prune_enabled()) {
                const PruneShape ps = prune_shape(n_points, dp, form, sb.slices);
                sb.box = static_cast<float*>(take(static_cast<size_t>(ps.tiles) * 2 * dp * sizeof(float)));
                sb.tmax = static_cast<float*>(take(static_cast<size_t>(ps.tiles) * sizeof(float)));
                const size_t groups = (static_cast<size_t>(ps.tiles) + kPruneGroup - 1) / kPruneGroup;
                sb.gbox = static_cast<float*>(take(groups * 2 * dp * sizeof(float)));
                sb.gtmax = static_cast<float*>(take(groups * sizeof(float)));
                sb.live_stride = ps.stride;
                sb.live = static_cast<int*>(take(static_cast<size_t>(ps.nb) * sb.slices * ps.stride * sizeof(int)));
                sb.live_count = static_cast<int*>(take(static_cast<size_t>(ps.nb) * sb.slices * sizeof(int)));
            }
            zero_fill(args.ccount, slots * sizeof(int), stream);
            zero_fill(sb.next_chunk, sizeof(int), stream);
            zero_fill(sb.col_over, static_cast<size_t>(rows) * sizeof(int), stream);
        } else {
            const size_t slots = static_cast<size_t>(rows) * slices;
            args.cap = first_cap(k);
            args.cbuf = static_cast<uint32_t*>(take(slots * args.cap * sizeof(uint32_t)));
            args.ccount = static_cast<int*>(take(slots * sizeof(int)));
        }
        args.big_rows = static_cast<int*>(take(static_cast<size_t>(rows) * sizeof(int)));
        // A full run in candidate order stages its first-pass lists by
        // position and writes the transposed outputs in original order.
        if (ord && qlist == ord->order && row0 == 0 && rows == n_points) {
            args.stage = static_cast<int*>(take(static_cast<size_t>(rows) * stage_words(k) * sizeof(int)));
        }
        args.ov_cap = std::min(rows, kSecondRows);
        args.ov_rows = static_cast<int*>(take(static_cast<size_t>(args.ov_cap) * sizeof(int)));
        int* counters = static_cast<int*>(take(4 * sizeof(int)));
        args.big_count = counters;
        args.ov_count = counters + 1;
        zero_fill(counters, 4 * sizeof(int), stream);
        // The second pass: the overflowed rows, with large buffers.
        second = args;
        second.qlist = args.ov_rows;
        second.row_count = args.ov_count;
        second.row0 = 0;
        second.rows = args.ov_cap;
        second.cap = kSecondCap;
        second.slices = second_slices(n_points);
        const size_t slots2 = static_cast<size_t>(args.ov_cap) * second.slices;
        second.cbuf = static_cast<uint32_t*>(take(slots2 * kSecondCap * sizeof(uint32_t)));
        second.ccount = static_cast<int*>(take(slots2 * sizeof(int)));
        second.big_rows = nullptr;
        second.big_count = nullptr;
        second.heavy = nullptr;
        second.stage = nullptr;
        // Its overflow queue feeds the list pass.
        second.ov_cap = std::min(rows, kListRows);
        second.ov_rows = static_cast<int*>(take(static_cast<size_t>(second.ov_cap) * sizeof(int)));
        second.ov_count = counters + 2;
        list = args;
        list.collect = 0;
        list.qlist = second.ov_rows;
        list.row_count = second.ov_count;
        list.row0 = 0;
        list.rows = second.ov_cap;
        list.slices = second.slices;
        list.partial = list.slices > 1 && k <= kListKMax && dp < kWideMin ? static_cast<uint64_t*>(take(
                           sizeof(uint64_t) * list.slices * static_cast<size_t>(list.rows) * k))
                                       : nullptr;
        list.cbuf = nullptr;
        list.ccount = nullptr;
        list.heavy = nullptr;
        list.stage = nullptr;
        if (k > kListKMax || dp >= kWideMin) {
            brute_batch = brute_batch_rows(n_points, list.rows);
            brute_keys = static_cast<float*>(take(static_cast<size_t>(brute_batch) * n_points * sizeof(float)));
        }
        // First-pass rows that find the second pass's queue full go there too.
        args.list_rows = second.ov_rows;
        args.list_count = second.ov_count;
        args.list_cap = second.ov_cap;
    } else if (slices > 1) {
        args.partial = static_cast<uint64_t*>(take(sizeof(uint64_t) * slices * static_cast<size_t>(rows) * k));
    }

    prepare_screen(centred, norms, ord ? ord->order : nullptr, n_points, dp, bf16, c, xa, madj, stream);
    auto run = [&](const MfmaArgs& a) {
        const bool ok = bf16 ? dispatch<ScreenBF16>(dp, k, &a, stream)
                             : dispatch<ScreenF32>(dp, k, &a, stream);
        if (!ok) throw std::invalid_argument("mfma_knn: unsupported (n_dims, k, form)");
        CULOF_CHECK(hipGetLastError());
    };
    if (sym) {
        {
            // Rows with huge candidate sets go to the list pass up front.
            ProfileScope prof("heavy_rows", stream);
            // ~8k sample features per row (64..512 rows): the estimate only
            // has to tell rows near kHeavyCount candidates from rows near n.
            const int m = std::min(n_points, std::clamp(8192 / dp, 64, 512));
            auto go = [&](auto D) {
                if constexpr (decltype(D)::value < kWideMin) {
                    heavy_rows_kernel<decltype(D)::value><<<(n_points + 255) / 256, 256, 0, stream>>>(
                        ord->xs_ord, ord->order, tseed, guard, n_points, m, heavy, second.ov_rows,
                        second.ov_count, second.ov_cap);
                }
                return true;
            };
            for_dp<ScreenF32>(dp, go);
            CULOF_CHECK(hipGetLastError());
        }
        column_thresholds_kernel<<<(padded_rows + 255) / 256, 256, 0, stream>>>(
            tseed, norms, madj, ord->order, heavy, n_points, padded_rows, c, sb.wcol);
        column_maxima_kernel<<<(2 * (padded_rows / 32) + 255) / 256, 256, 0, stream>>>(
            sb.wcol, madj, padded_rows / 32, sb.wmax, sb.nmax);
        CULOF_CHECK(hipGetLastError());
        auto go_bf16 = [&](auto D) {
            if constexpr (decltype(D)::value < kWideMin) launch_sym<ScreenBF16, decltype(D)::value>(args, sb, stream);
            return decltype(D)::value < kWideMin;
        };
        auto go_f32 = [&](auto D) {
            if constexpr (decltype(D)::value < kWideMin) launch_sym<ScreenF32, decltype(D)::value>(args, sb, stream);
            return decltype(D)::value < kWideMin;
        };
        const bool ok = k <= kCollectKMax && (bf16 ? for_dp<ScreenBF16>(dp, go_bf16) : for_dp<ScreenF32>(dp, go_f32));
        if (!ok) throw std::invalid_argument("mfma_knn: unsupported (n_dims, k, form)");
        CULOF_CHECK(hipGetLastError());
    } else {
        run(args);
    }
    if (collect) {
        run(second);
        if (k <= kListKMax && dp < kWideMin) {
            run(list);
        } else {
            // No list kernel holds this k or width: the rows queued for the
            // list pass (heavy rows, second-pass overflow) get their exact
            // keys stored and the radix select over them.
            brute_knn_rows(keyrows, stride, n_points, k, list.qlist, list.row_count, list.rows, brute_keys,
                           brute_batch, nbr_idx, nbr_d2, kdist2, stream);
        }
    }
}

// Band pass for near-tie rows: up to kBandRows queued rows, collected in
// natural candidate order over second_slices(n) slices of kBandCap entries.
// A band is a few entries beyond k, so the buffers are small; rows that
// overflow them take the brute-force scan.
constexpr int kBandRows = 32768;
constexpr int kBandCap = 128;

int tie_band_capacity(int n_points, int n_dims, int stride, int form, int rows) {
    const int dp = mfma_padded_dims(n_dims, form);
    const bool ok = form == kFormBF16 ? for_dp<ScreenBF16>(dp, [](auto) { return true; })
                                      : for_dp<ScreenF32>(dp, [](auto) { return true; });
    return ok && dp == stride && n_points > 0 ? std::min(rows, kBandRows) : 0;
}

size_t tie_band_scratch_bytes(int n_points, int n_dims, int form, int cap) {
    const int dp = mfma_padded_dims(n_dims, form);
    const size_t padded = padded_positions(n_points);
    const size_t slots = static_cast<size_t>(cap) * second_slices(n_points);
    return padded * dp * 4 + 256 + padded * sizeof(float) + 256 + slots * kBandCap * sizeof(uint32_t) +
           256 + slots * sizeof(int) + 256;
}

void tie_band_select(const float* keyrows, int stride, int n_points, int k, const TieBand& band,
                     int* nbr_idx, float* nbr_d2, float* kdist2, int* brute_rows, int* brute_count,
                     int brute_cap, hipStream_t stream) {
    const int dp = stride;
    const bool bf16 = band.form == kFormBF16;
    const float c = bf16 ? ScreenBF16::slack(dp) : ScreenF32::slack(dp);
    const int padded_rows = static_cast<int>(padded_positions(n_points));
    unsigned char* base = static_cast<unsigned char*>(band.scratch);
    auto take = [&](size_t bytes) {
        void* r = base;
        base += (bytes + 255) / 256 * 256;
        return r;
    };
    void* xa = take(static_cast<size_t>(padded_rows) * dp * 4);
    float* madj = static_cast<float*>(take(sizeof(float) * padded_rows));
    const int slices = second_slices(n_points);
    const size_t slots = static_cast<size_t>(band.cap) * slices;
    uint32_t* cbuf = static_cast<uint32_t*>(take(slots * kBandCap * sizeof(uint32_t)));
    int* ccount = static_cast<int*>(take(slots * sizeof(int)));
    prepare_screen(band.centred, band.norms, nullptr, n_points, dp, bf16, c, xa, madj, stream);
    MfmaArgs a{xa, band.centred, keyrows, keyrows, nullptr, band.rows, band.tband, madj, band.norms,
               band.guard, n_points, k, c, 0, band.cap, slices, nullptr, 1, band.count, cbuf, ccount,
               kBandCap, nullptr, nullptr, 0, nullptr, nullptr, nullptr, nullptr, 0, nbr_idx, nbr_d2,
               kdist2, brute_rows, brute_count, brute_cap};
    auto go_bf16 = [&](auto D) {
        launch_band<ScreenBF16, decltype(D)::value>(a, band.bound, brute_rows, brute_count, stream);
        return true;
    };
    auto go_f32 = [&](auto D) {
        launch_band<ScreenF32, decltype(D)::value>(a, band.bound, brute_rows, brute_count, stream);
        return true;
    };
    const bool ok = bf16 ? for_dp<ScreenBF16>(dp, go_bf16) : for_dp<ScreenF32>(dp, go_f32);
    if (!ok) throw std::invalid_argument("tie band: unsupported (n_dims, form)");
    CULOF_CHECK(hipGetLastError());
}

}  // namespace detail
}  // namespace culof
