// Backward pass: d(sum_i w_i LOF_i) / dX with the kNN sets and the
// reach-distance max() branches fixed from the forward.
//
// With N(i) the k neighbours of i (slot order of the forward), e_ij their
// squared distances, kd2_j the squared k-distance of j and nk(j) = N(j)[k-1]
// its k-th neighbour:
//   reach_ij = max(kd_j, d_ij)            (d = sqrt e, kd = sqrt kd2)
//   m_i      = (1/k) sum_{j in N(i)} reach_ij,   lrd_i = 1 / (m_i + 1e-10)
//   LOF_i    = (m_i + 1e-10) S_i,               S_i = (1/k) sum_{j in N(i)} lrd_j
// (src/cuda/lof.cu:32-62). Reverse-mode:
//   g_lrd_j = sum_{i : j in N(i)} w_i (m_i + eps) / k
//   g_m_i   = w_i S_i - g_lrd_i lrd_i^2
//   reach_ij takes branch "kd" when kd2_j > e_ij, else branch "d":
//     branch d  adds g_m_i / k to dL/dd_ij
//     branch kd adds g_m_i / k to g_kd_j, which goes to dL/dd(j, nk(j))
//   dd_ab / dx_a = (x_a - x_b) / d_ab, and 0 when d_ab = 0.
// When i is j's k-th neighbour both branches are the same function of the
// points, so the branch choice at equality does not matter there.
//
// Every sum runs in FP64 in a fixed order. Contributions owned by another
// point come through reverse neighbour lists (CSR): counted with integer
// atomics (each entry keeping the place it took), filled by those places, then
// each segment sorted by (slot, query) linear index, so
// the result is bit-identical from run to run with no floating-point atomics.
// The distances are recomputed from the key rows in FP64 (n k pairs) instead
// of storing anything n x n. With normalisation the gradient is carried back
// through the z-score (population standard deviation, constant features
// centred only), as for batch normalisation.

#include "internal.hpp"

#include <stdexcept>
#include <type_traits>

namespace culof {
namespace detail {

namespace {

constexpr int kBlock = 256;
constexpr double kEps = 1e-10;

// m_i and lrd_i in FP64 from the forward's neighbour lists.
__global__ void __launch_bounds__(kBlock)
point_m_kernel(const int* __restrict__ nbr, const float* __restrict__ d2,
               const float* __restrict__ kd2, int n, int k, double* __restrict__ m,
               double* __restrict__ lrd) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double s = 0.0;
    for (int t = 0; t < k; ++t) {
        const size_t o = static_cast<size_t>(t) * n + i;
        s += sqrt(static_cast<double>(fmaxf(kd2[nbr[o]], d2[o])));
    }
    const double mi = s / k;
    m[i] = mi;
    lrd[i] = 1.0 / (mi + kEps);
}

// S_i = mean lrd over N(i); a_i = w_i (m_i + eps) / k, what i sends to g_lrd_j.
__global__ void __launch_bounds__(kBlock)
point_s_kernel(const int* __restrict__ nbr, const double* __restrict__ lrd,
               const double* __restrict__ m, const float* __restrict__ w, int n, int k,
               double* __restrict__ S, double* __restrict__ a) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double s = 0.0;
    for (int t = 0; t < k; ++t) s += lrd[nbr[static_cast<size_t>(t) * n + i]];
    S[i] = s / k;
    a[i] = static_cast<double>(w[i]) * (m[i] + kEps) / k;
}

// ---- reverse lists ----

// count[p] = entries whose neighbour is p; each entry keeps the place it took
// in p's list (in whatever order the atomics ran: the lists are sorted later).
__global__ void __launch_bounds__(kBlock)
rev_count_kernel(const int* __restrict__ nbr, size_t nk, int* __restrict__ count, int* __restrict__ place) {
    const size_t e = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e < nk) place[e] = atomicAdd(&count[nbr[e]], 1);
}

// Exclusive scan, three steps: per-block scan + block totals, scan of the
// totals in one block, add back.
__device__ int block_scan_excl(int v, int& total) {
    __shared__ int wave_sums[kBlock / 64];
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
        before += (w < wave) ? wave_sums[w] : 0;
        total += wave_sums[w];
    }
    __syncthreads();
    return before + x - v;
}

__global__ void __launch_bounds__(kBlock)
scan_blocks_kernel(const int* __restrict__ in, int n, int* __restrict__ out,
                   int* __restrict__ block_totals) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    int total = 0;
    const int ex = block_scan_excl(i < n ? in[i] : 0, total);
    if (i < n) out[i] = ex;
    if (threadIdx.x == 0) block_totals[blockIdx.x] = total;
}

__global__ void __launch_bounds__(kBlock)
scan_totals_kernel(int* __restrict__ totals, int blocks) {
    int carry = 0;
    for (int base = 0; base < blocks; base += kBlock) {
        const int i = base + threadIdx.x;
        int total = 0;
        const int ex = block_scan_excl(i < blocks ? totals[i] : 0, total);
        if (i < blocks) totals[i] = carry + ex;
        carry += total;
    }
}

__global__ void __launch_bounds__(kBlock)
scan_add_kernel(int* __restrict__ out, int n, const int* __restrict__ totals) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i < n) out[i] += totals[blockIdx.x];
}

__global__ void __launch_bounds__(kBlock)
rev_fill_kernel(const int* __restrict__ nbr, size_t nk, const int* __restrict__ offset,
                const int* __restrict__ place, int* __restrict__ rev) {
    const size_t e = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e >= nk) return;
    rev[offset[nbr[e]] + place[e]] = static_cast<int>(e);  // e = slot * n + query
}

// Heapsort each segment ascending, one thread per point: O(L log L) even for
// hub points that appear in many lists.
__device__ void sift_down(int* a, int start, int end) {
    int root = start;
    while (2 * root + 1 <= end) {
        int child = 2 * root + 1;
        if (child + 1 <= end && a[child] < a[child + 1]) ++child;
        if (a[root] < a[child]) {
            const int t = a[root];
            a[root] = a[child];
            a[child] = t;
            root = child;
        } else {
            return;
        }
    }
}

__device__ void heap_sort(int* a, int len) {
    if (len < 2) return;
    for (int start = (len - 2) / 2; start >= 0; --start) sift_down(a, start, len - 1);
    for (int end = len - 1; end > 0; --end) {
        const int t = a[0];
        a[0] = a[end];
        a[end] = t;
        sift_down(a, 0, end - 1);
    }
}

// The kBlock consecutive points of a block own one contiguous CSR range, sorted
// in windows of whole lists that fit in LDS (kSortLds entries, 128 KB: the
// whole range while lists average up to 128 entries, so any k up to about 100
// takes one window). A window is copied to LDS and every entry finds its place
// in its own list by counting the smaller entries there (the entries of a list
// are distinct); the work is spread over the block's lanes by entry, so a hub
// point that sits in hundreds of lists costs L^2 / kSortThreads steps per lane
// instead of a serial sort of L entries by one lane. A single list longer than
// kSortLds is heap-sorted in place in global memory by one lane.
constexpr int kSortLds = 32768;
constexpr int kSortThreads = 1024;

__global__ void __launch_bounds__(kSortThreads)
rev_sort_kernel(int* __restrict__ rev, const int* __restrict__ offset,
                const int* __restrict__ count, int n) {
    __shared__ int s_seg[kSortLds];
    __shared__ int s_off[kBlock + 1];  // list starts relative to the range, then its length
    const int p0 = blockIdx.x * kBlock;
    const int pts = min(kBlock, n - p0);
    const int begin = offset[p0];
    const int span = offset[p0 + pts - 1] + count[p0 + pts - 1] - begin;
    if (static_cast<int>(threadIdx.x) < pts) s_off[threadIdx.x] = offset[p0 + threadIdx.x] - begin;
    if (threadIdx.x == 0) s_off[pts] = span;
    __syncthreads();
    // windows [a, b) of whole lists; every lane walks the same windows
    for (int a = 0; a < pts;) {
        const int w0 = s_off[a];
        if (s_off[a + 1] - w0 > kSortLds) {  // one list longer than LDS holds
            if (threadIdx.x == 0) heap_sort(rev + begin + w0, s_off[a + 1] - w0);
            ++a;
            continue;
        }
        int lo = a + 1, hi = pts;  // the last b in [a + 1, pts] with s_off[b] - w0 <= kSortLds
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (s_off[mid] - w0 <= kSortLds) lo = mid;
            else hi = mid - 1;
        }
        const int b = lo;
        const int len = s_off[b] - w0;
        for (int e = threadIdx.x; e < len; e += kSortThreads) s_seg[e] = rev[begin + w0 + e];
        __syncthreads();
        for (int q = threadIdx.x; q < len; q += kSortThreads) {
            // the list holding entry w0 + q: the last start <= it (empty lists
            // share a start with the next one, whose start is then also <= it)
            int l = a, h = b;  // s_off[l] <= w0 + q < s_off[h]
            while (h - l > 1) {
                const int mid = (l + h) >> 1;
                if (s_off[mid] <= w0 + q) l = mid;
                else h = mid;
            }
            const int seg_lo = s_off[l] - w0;
            const int seg_hi = s_off[l + 1] - w0;
            const int v = s_seg[q];
            int rank = 0;
#pragma unroll 4
            for (int r = seg_lo; r < seg_hi; ++r) rank += s_seg[r] < v ? 1 : 0;
            rev[begin + w0 + seg_lo + rank] = v;
        }
        __syncthreads();  // the next window reuses s_seg
        a = b;
    }
}

// Reverse lists are walked kBatch entries at a time: the loads of a batch are
// all issued before the first of its terms is added, so a long list costs one
// memory latency per batch rather than per entry, while the terms are still
// added one by one in list order. Past the end of a list a batch reads the
// list's last entry again (the loop only runs for non-empty lists) and drops
// its term. The sums of g_m and g_kd take 8 entries per batch; grad_x takes
// 2, as with its branch-free terms the waves per SIMD matter more than the
// loads in flight per wave (50k x 32: 8 entries, 154 VGPRs, 0.51 ms for the
// backward; 4, 94 VGPRs, 0.46 ms; 2, 0.42 ms; 1, 0.43 ms).
constexpr int kSumBatch = 8;
constexpr int kBatch = 2;

// A point in many lists would keep its lane or group busy long after the
// rest, so every pass over the reverse lists cuts them into chunks of kChunk
// entries: work item t < n takes the first chunk of point t's list, work item
// n + x the extra chunk x (chunk_start[p] + c - 1 is the c-th chunk of p for
// c >= 1), whose partial result a finishing pass adds to the point's in chunk
// order.
constexpr int kChunk = 64;

// extra[p] = chunks of p's reverse list after the first
__global__ void __launch_bounds__(kBlock)
chunk_count_kernel(const int* __restrict__ count, int n, int* __restrict__ extra) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < n) extra[p] = count[p] > 0 ? (count[p] - 1) / kChunk : 0;
}

// owner[chunk_start[p] + t] = p for t < extra[p]
__global__ void __launch_bounds__(kBlock)
chunk_owner_kernel(const int* __restrict__ extra, const int* __restrict__ chunk_start, int n,
                   int* __restrict__ owner) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    for (int t = 0; t < extra[p]; ++t) owner[chunk_start[p] + t] = p;
}

// Point and chunk of work item t; false past the last extra chunk.
__device__ bool chunk_of(int t, int n, const int* __restrict__ extra,
                         const int* __restrict__ chunk_start, const int* __restrict__ owner,
                         int& p, int& chunk) {
    if (t < n) {
        p = t;
        chunk = 0;
        return true;
    }
    const int x = t - n;
    if (x >= chunk_start[n - 1] + extra[n - 1]) return false;
    p = owner[x];
    chunk = x - chunk_start[p] + 1;
    return true;
}

// Work items: n points plus at most n k / kChunk extra chunks.
size_t chunk_items(size_t n, int k) { return n + n * k / kChunk; }

// sum_{i : p in N(i)} a_i over one chunk of p's list; the first chunk's sum
// goes to gm[p], an extra chunk's to part[x].
__global__ void __launch_bounds__(kBlock)
grad_m_chunk_kernel(const int* __restrict__ rev, const int* __restrict__ offset,
                    const int* __restrict__ count, const double* __restrict__ a,
                    const int* __restrict__ extra, const int* __restrict__ chunk_start,
                    const int* __restrict__ owner, int n, double* __restrict__ gm,
                    double* __restrict__ part) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    int p, chunk;
    if (!chunk_of(t, n, extra, chunk_start, owner, p, chunk)) return;
    double sum = 0.0;
    const int* seg = rev + offset[p];
    const int lo = chunk * kChunk;
    const int hi = min(lo + kChunk, count[p]);
    for (int r0 = lo; r0 < hi; r0 += kSumBatch) {
        int e[kSumBatch];
        double v[kSumBatch];
#pragma unroll
        for (int u = 0; u < kSumBatch; ++u) e[u] = seg[min(r0 + u, hi - 1)];
#pragma unroll
        for (int u = 0; u < kSumBatch; ++u) v[u] = a[e[u] % n];
#pragma unroll
        for (int u = 0; u < kSumBatch; ++u)
            if (r0 + u < hi) sum += v[u];
    }
    if (chunk == 0) gm[p] = sum;
    else part[t - n] = sum;
}

// g_m_p = w_p S_p - lrd_p^2 g_lrd_p, g_lrd_p = the chunk sums in chunk order
__global__ void __launch_bounds__(kBlock)
grad_m_finish_kernel(const int* __restrict__ extra, const int* __restrict__ chunk_start,
                     const double* __restrict__ part, const double* __restrict__ S,
                     const double* __restrict__ lrd, const float* __restrict__ w, int n,
                     double* __restrict__ gm) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    double g_lrd = gm[p];
    for (int t = 0; t < extra[p]; ++t) g_lrd += part[chunk_start[p] + t];
    gm[p] = static_cast<double>(w[p]) * S[p] - g_lrd * lrd[p] * lrd[p];
}

// g_kd_p = sum over i with p in N(i) whose reach took the kd branch of g_m_i / k,
// one chunk of p's list: the first chunk's sum to gkd[p], an extra one's to part[x]
__global__ void __launch_bounds__(kBlock)
grad_kd_chunk_kernel(const int* __restrict__ rev, const int* __restrict__ offset,
                     const int* __restrict__ count, const float* __restrict__ d2,
                     const float* __restrict__ kd2, const double* __restrict__ gm,
                     const int* __restrict__ extra, const int* __restrict__ chunk_start,
                     const int* __restrict__ owner, int n, int k, double* __restrict__ gkd,
                     double* __restrict__ part) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    int p, chunk;
    if (!chunk_of(t, n, extra, chunk_start, owner, p, chunk)) return;
    double sum = 0.0;
    const float kd2p = kd2[p];
    const int* seg = rev + offset[p];
    const int lo = chunk * kChunk;
    const int hi = min(lo + kChunk, count[p]);
    for (int r0 = lo; r0 < hi; r0 += kSumBatch) {
        int e[kSumBatch];
        float de[kSumBatch];
        double g[kSumBatch];
#pragma unroll
        for (int u = 0; u < kSumBatch; ++u) e[u] = seg[min(r0 + u, hi - 1)];
#pragma unroll
        for (int u = 0; u < kSumBatch; ++u) {
            de[u] = d2[e[u]];
            g[u] = gm[e[u] % n];
        }
#pragma unroll
        for (int u = 0; u < kSumBatch; ++u)
            if (r0 + u < hi && kd2p > de[u]) sum += g[u] / k;
    }
    if (chunk == 0) gkd[p] = sum;
    else part[t - n] = sum;
}

__global__ void __launch_bounds__(kBlock)
grad_kd_finish_kernel(const int* __restrict__ extra, const int* __restrict__ chunk_start,
                      const double* __restrict__ part, int n, double* __restrict__ gkd) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    double s = gkd[p];
    for (int t = 0; t < extra[p]; ++t) s += part[chunk_start[p] + t];
    gkd[p] = s;
}

// One 64-bit value moved between lanes by a DPP control (two 32-bit moves).
template <int Ctrl>
__device__ double dpp_f64(double v) {
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), Ctrl, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), Ctrl, 0xF, 0xF, false);
    return __hiloint2double(hi, lo);
}

// Sum over the G lanes of an aligned group, the same bits in every lane. A
// butterfly: at each level every lane adds the total of the other half of its
// block (after the level below, both halves hold their own totals in every
// lane), so each lane forms A + B or B + A, the same value. Within a row of 16
// lanes the partner comes through DPP (quad permutes, then the half-row and
// row mirrors), above it through a lane permute.
template <int G>
__device__ double group_sum(double x) {
    if constexpr (G >= 2) x += dpp_f64<0xB1>(x);   // quad_perm [1,0,3,2]
    if constexpr (G >= 4) x += dpp_f64<0x4E>(x);   // quad_perm [2,3,0,1]
    if constexpr (G >= 8) x += dpp_f64<0x141>(x);  // row_half_mirror
    if constexpr (G >= 16) x += dpp_f64<0x140>(x); // row_mirror
    if constexpr (G >= 32) x += __shfl_xor(x, 16, G);
    if constexpr (G >= 64) x += __shfl_xor(x, 32, G);
    return x;
}

// Each reverse entry e = slot * n + i (p = nbr[e] is the point whose list it
// sits in) replaced by its query i, with the coefficient of its term in p's
// gradient beside it: g_m_i / k when reach_ip took branch d, plus g_kd_i when
// p is i's k-th neighbour.
__global__ void __launch_bounds__(kBlock)
rev_coef_kernel(int* __restrict__ rev, size_t nk, const int* __restrict__ nbr,
                const float* __restrict__ d2, const float* __restrict__ kd2,
                const double* __restrict__ gm, const double* __restrict__ gkd, int n, int k,
                double* __restrict__ coef) {
    const size_t r = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (r >= nk) return;
    const int e = rev[r];
    const int i = e % n;
    const int slot = e / n;
    double c = 0.0;
    if (!(kd2[nbr[e]] > d2[e])) c += gm[i] / k;  // branch d of reach_ip
    if (slot == k - 1) c += gkd[i];             // p is the k-th neighbour of i
    rev[r] = i;
    coef[r] = c;
}

// A group of G lanes owns one point; lane l of the group holds features
// l, l + G, ... (F of them), so rows are read coalesced and the FP64 row of
// the gradient stays in registers.
template <int G, int F>
struct GroupRow {
    int lane;  // lane within the group
    int d;

    __device__ void load(const float* __restrict__ row, double (&v)[F]) const {
#pragma unroll
        for (int u = 0; u < F; ++u) {
            const int f = lane + G * u;
            v[u] = (f < d) ? static_cast<double>(row[f]) : 0.0;
        }
    }

    // A row of another point as FP32 values, features past d as zeros.
    __device__ void fetch(const float* __restrict__ row, float (&v)[F]) const {
#pragma unroll
        for (int u = 0; u < F; ++u) {
            const int f = lane + G * u;
            v[u] = (f < d) ? row[f] : 0.0f;
        }
    }

    // acc += c_t (x_a - x_t) / |x_a - x_t| for t = 0 .. N-1, in that order; a
    // term with c_t = 0, or whose points coincide, leaves acc as it is. No
    // branches, so the N group sums overlap.
    template <int N>
    __device__ void add_terms(const double (&xa)[F], const float (&xb)[N][F], const double (&c)[N],
                              double (&acc)[F]) const {
        double df[N][F];
        double ss[N];
#pragma unroll
        for (int t = 0; t < N; ++t) {
            ss[t] = 0.0;
#pragma unroll
            for (int u = 0; u < F; ++u) {
                df[t][u] = xa[u] - static_cast<double>(xb[t][u]);
                ss[t] = fma(df[t][u], df[t][u], ss[t]);
            }
        }
#pragma unroll
        for (int t = 0; t < N; ++t) ss[t] = group_sum<G>(ss[t]);
#pragma unroll
        for (int t = 0; t < N; ++t) {
            const bool live = c[t] != 0.0 && ss[t] != 0.0;
            const double sc = c[t] / sqrt(ss[t]);
#pragma unroll
            for (int u = 0; u < F; ++u) acc[u] = live ? fma(sc, df[t][u], acc[u]) : acc[u];
        }
    }
};

// Every term a point owns or receives (own list, own k-distance, reverse
// list) accumulated into its FP64 gradient row; O(n k d) in total. Terms are
// fetched kBatch at a time and added in list order. Group p (p < n) takes
// point p's own terms and the first chunk of its reverse list, group n + x
// extra chunk x, whose partial row chunk_finish_kernel adds to the point's.
template <int G, int F>
__global__ void __launch_bounds__(kBlock)
grad_x_kernel(const float* __restrict__ xs, int d, const int* __restrict__ nbr,
              const float* __restrict__ d2, const float* __restrict__ kd2,
              const int* __restrict__ src, const double* __restrict__ coef,
              const int* __restrict__ offset, const int* __restrict__ count,
              const double* __restrict__ gm, const double* __restrict__ gkd,
              const int* __restrict__ extra, const int* __restrict__ chunk_start,
              const int* __restrict__ owner, int n, int k, double* __restrict__ gx,
              double* __restrict__ part) {
    const int g = (blockIdx.x * kBlock + threadIdx.x) / G;  // whole groups take one branch
    int p, chunk;
    if (!chunk_of(g, n, extra, chunk_start, owner, p, chunk)) return;
    const GroupRow<G, F> row{static_cast<int>(threadIdx.x % G), d};
    double xp[F], acc[F];
    row.load(xs + static_cast<size_t>(p) * d, xp);
#pragma unroll
    for (int u = 0; u < F; ++u) acc[u] = 0.0;
    float xb[kBatch][F];
    double c[kBatch];
    // Every load of a batch is unconditional (past the end of a list it reads
    // the list's last entry again) and the terms are picked with selects, so
    // no load of a batch waits on another's value through a branch.
    if (chunk == 0) {
        // own list: branch-d terms, then the own k-distance
        const double own = gm[p] / k;
        for (int t0 = 0; t0 < k; t0 += kBatch) {
            int j[kBatch];
            float dt[kBatch];
#pragma unroll
            for (int u = 0; u < kBatch; ++u) {
                const size_t o = static_cast<size_t>(min(t0 + u, k - 1)) * n + p;
                j[u] = nbr[o];
                dt[u] = d2[o];
            }
#pragma unroll
            for (int u = 0; u < kBatch; ++u) {
                const float kj = kd2[j[u]];
                c[u] = (t0 + u < k && !(kj > dt[u])) ? own : 0.0;
                row.fetch(xs + static_cast<size_t>(j[u]) * d, xb[u]);
            }
            row.template add_terms<kBatch>(xp, xb, c, acc);
        }
        float xk[1][F];
        const double ck[1] = {gkd[p]};
        row.fetch(xs + static_cast<size_t>(nbr[static_cast<size_t>(k - 1) * n + p]) * d, xk[0]);
        row.template add_terms<1>(xp, xk, ck, acc);
    }
    // reverse list: p as neighbour j of query i, coefficients precomputed
    const int* sq = src + offset[p];
    const double* cq = coef + offset[p];
    const int r_lo = chunk * kChunk;
    const int r_hi = min(r_lo + kChunk, count[p]);
    for (int r0 = r_lo; r0 < r_hi; r0 += kBatch) {
        int i[kBatch];
        double ci[kBatch];
#pragma unroll
        for (int u = 0; u < kBatch; ++u) {
            const int r = min(r0 + u, r_hi - 1);
            i[u] = sq[r];
            ci[u] = cq[r];
        }
#pragma unroll
        for (int u = 0; u < kBatch; ++u) {
            row.fetch(xs + static_cast<size_t>(i[u]) * d, xb[u]);
            c[u] = r0 + u < r_hi ? ci[u] : 0.0;
        }
        row.template add_terms<kBatch>(xp, xb, c, acc);
    }
    double* out = chunk == 0 ? gx + static_cast<size_t>(p) * d : part + static_cast<size_t>(g - n) * d;
#pragma unroll
    for (int u = 0; u < F; ++u) {
        const int f = row.lane + G * u;
        if (f < d) out[f] = acc[u];
    }
}

// gx[p] += the partial rows of p's extra chunks, in chunk order
__global__ void __launch_bounds__(kBlock)
chunk_finish_kernel(const int* __restrict__ extra, const int* __restrict__ chunk_start,
                    const double* __restrict__ part, int n, int d, double* __restrict__ gx) {
    const size_t cell = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (cell >= static_cast<size_t>(n) * d) return;
    const int p = static_cast<int>(cell / d);
    const int f = static_cast<int>(cell % d);
    const int ex = extra[p];
    if (ex == 0) return;
    double v = gx[cell];
    for (int t = 0; t < ex; ++t) v += part[static_cast<size_t>(chunk_start[p] + t) * d + f];
    gx[cell] = v;
}

template <int G, int F>
void launch_grad_x(const float* xs, int d, const int* nbr, const float* d2, const float* kd2,
                   const int* src, const double* coef, const int* offset, const int* count,
                   const double* gm, const double* gkd, const int* extra, const int* chunk_start,
                   const int* owner, int n, int k, double* gx, double* part, hipStream_t s) {
    const size_t lanes = chunk_items(static_cast<size_t>(n), k) * G;
    grad_x_kernel<G, F><<<static_cast<unsigned>((lanes + kBlock - 1) / kBlock), kBlock, 0, s>>>(
        xs, d, nbr, d2, kd2, src, coef, offset, count, gm, gkd, extra, chunk_start, owner, n, k, gx, part);
}

// Per-feature sums for the z-score chain: sum_i g_it and sum_i g_it xs_it.
__global__ void __launch_bounds__(kBlock)
zscore_sums_kernel(const double* __restrict__ g, const float* __restrict__ xs, int n, int d,
                   double* __restrict__ sums) {
    const int f = blockIdx.x;
    __shared__ double s1[kBlock];
    __shared__ double s2[kBlock];
    double a = 0.0, b = 0.0;
    for (int i = threadIdx.x; i < n; i += kBlock) {
        const double gv = g[static_cast<size_t>(i) * d + f];
        a += gv;
        b += gv * static_cast<double>(xs[static_cast<size_t>(i) * d + f]);
    }
    s1[threadIdx.x] = a;
    s2[threadIdx.x] = b;
    __syncthreads();
    for (int off = kBlock / 2; off > 0; off >>= 1) {
        if (static_cast<int>(threadIdx.x) < off) {
            s1[threadIdx.x] += s1[threadIdx.x + off];
            s2[threadIdx.x] += s2[threadIdx.x + off];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        sums[2 * f] = s1[0];
        sums[2 * f + 1] = s2[0];
    }
}

// dL/dx = inv_std (g - mean g - xs mean(g xs)) for scaled features; for a
// constant feature (scale fixed at 1) only the centring term remains.
__global__ void __launch_bounds__(kBlock)
zscore_apply_kernel(const double* __restrict__ g, const float* __restrict__ xs,
                    const double* __restrict__ sums, const double* __restrict__ inv_std,
                    const double* __restrict__ scaled, int n, int d, float* __restrict__ out) {
    const size_t c = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (c >= static_cast<size_t>(n) * d) return;
    const int f = static_cast<int>(c % d);
    const double mg = sums[2 * f] / n;
    double v = g[c] - mg;
    if (scaled[f] != 0.0) v -= static_cast<double>(xs[c]) * (sums[2 * f + 1] / n);
    out[c] = static_cast<float>(v * inv_std[f]);
}

__global__ void __launch_bounds__(kBlock)
to_float_kernel(const double* __restrict__ g, size_t cells, float* __restrict__ out) {
    const size_t c = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (c < cells) out[c] = static_cast<float>(g[c]);
}

unsigned grid_for(size_t items, int block = kBlock) {
    return static_cast<unsigned>((items + block - 1) / block);
}

}  // namespace

size_t backward_scratch_bytes(int n, int d, int k) {
    const size_t nn = static_cast<size_t>(n);
    const size_t blocks = (nn + kBlock - 1) / kBlock;
    size_t b = 0;
    b += 6 * nn * sizeof(double);               // m, lrd, S, a, gm, gkd
    b += nn * d * sizeof(double);               // gradient in FP64
    b += 2 * static_cast<size_t>(d) * sizeof(double);  // z-score sums
    b += (3 * nn + blocks + 8) * sizeof(int);   // count, offset, extra chunks, block totals
    b += nn * k * sizeof(int);                  // reverse lists (then their queries)
    b += nn * k * sizeof(double);               // coefficients of the reverse terms
    const size_t chunks = nn * k / kChunk + 1;
    b += nn * sizeof(int);                      // chunk starts
    b += chunks * sizeof(int);                  // chunk owners
    b += chunks * d * sizeof(double);           // partial rows of the extra chunks
    b += 20 * 256;                              // alignment of the 17 pieces
    return b;
}

void lof_backward(const float* keyrows, int n, int d, int k, const int* nbr_idx, const float* nbr_d2,
                  const float* kdist2, const float* grad_scores, const double* inv_std,
                  const double* scaled, float* grad_points, void* scratch, hipStream_t s) {
    if (d > kMaxBackwardDims) throw std::invalid_argument("backward supports n_dims <= 256");
    const size_t nn = static_cast<size_t>(n);
    const size_t nk = nn * k;
    const size_t blocks = (nn + kBlock - 1) / kBlock;
    unsigned char* p = static_cast<unsigned char*>(scratch);
    auto take = [&](size_t bytes) {
        void* r = p;
        p += (bytes + 255) / 256 * 256;
        return r;
    };
    double* m = static_cast<double*>(take(nn * sizeof(double)));
    double* lrd = static_cast<double*>(take(nn * sizeof(double)));
    double* S = static_cast<double*>(take(nn * sizeof(double)));
    double* a = static_cast<double*>(take(nn * sizeof(double)));
    double* gm = static_cast<double*>(take(nn * sizeof(double)));
    double* gkd = static_cast<double*>(take(nn * sizeof(double)));
    double* gx = static_cast<double*>(take(nn * d * sizeof(double)));
    double* zsums = static_cast<double*>(take(2 * d * sizeof(double)));
    int* count = static_cast<int*>(take(nn * sizeof(int)));
    int* offset = static_cast<int*>(take(nn * sizeof(int)));
    int* extra = static_cast<int*>(take(nn * sizeof(int)));  // chunks per list after the first
    int* totals = static_cast<int*>(take((blocks + 8) * sizeof(int)));
    int* rev = static_cast<int*>(take(nk * sizeof(int)));
    double* coef = static_cast<double*>(take(nk * sizeof(double)));
    const size_t chunks = nk / kChunk + 1;
    int* chunk_start = static_cast<int*>(take(nn * sizeof(int)));
    int* owner = static_cast<int*>(take(chunks * sizeof(int)));
    double* part = static_cast<double*>(take(chunks * d * sizeof(double)));

    ProfileScope prof("lof_backward", s);
    point_m_kernel<<<grid_for(nn), kBlock, 0, s>>>(nbr_idx, nbr_d2, kdist2, n, k, m, lrd);
    point_s_kernel<<<grid_for(nn), kBlock, 0, s>>>(nbr_idx, lrd, m, grad_scores, n, k, S, a);
    int* place = reinterpret_cast<int*>(coef);  // n k ints, free until rev_coef_kernel
    zero_fill(count, nn * sizeof(int), s);
    rev_count_kernel<<<grid_for(nk), kBlock, 0, s>>>(nbr_idx, nk, count, place);
    scan_blocks_kernel<<<static_cast<unsigned>(blocks), kBlock, 0, s>>>(count, n, offset, totals);
    scan_totals_kernel<<<1, kBlock, 0, s>>>(totals, static_cast<int>(blocks));
    scan_add_kernel<<<static_cast<unsigned>(blocks), kBlock, 0, s>>>(offset, n, totals);
    rev_fill_kernel<<<grid_for(nk), kBlock, 0, s>>>(nbr_idx, nk, offset, place, rev);
    rev_sort_kernel<<<grid_for(nn), kSortThreads, 0, s>>>(rev, offset, count, n);
    chunk_count_kernel<<<grid_for(nn), kBlock, 0, s>>>(count, n, extra);
    scan_blocks_kernel<<<static_cast<unsigned>(blocks), kBlock, 0, s>>>(extra, n, chunk_start, totals);
    scan_totals_kernel<<<1, kBlock, 0, s>>>(totals, static_cast<int>(blocks));
    scan_add_kernel<<<static_cast<unsigned>(blocks), kBlock, 0, s>>>(chunk_start, n, totals);
    chunk_owner_kernel<<<grid_for(nn), kBlock, 0, s>>>(extra, chunk_start, n, owner);
    const unsigned items = grid_for(chunk_items(nn, k));
    grad_m_chunk_kernel<<<items, kBlock, 0, s>>>(rev, offset, count, a, extra, chunk_start, owner, n, gm, part);
    grad_m_finish_kernel<<<grid_for(nn), kBlock, 0, s>>>(extra, chunk_start, part, S, lrd, grad_scores, n, gm);
    grad_kd_chunk_kernel<<<items, kBlock, 0, s>>>(rev, offset, count, nbr_d2, kdist2, gm, extra, chunk_start,
                                                  owner, n, k, gkd, part);
    grad_kd_finish_kernel<<<grid_for(nn), kBlock, 0, s>>>(extra, chunk_start, part, n, gkd);
    rev_coef_kernel<<<grid_for(nk), kBlock, 0, s>>>(rev, nk, nbr_idx, nbr_d2, kdist2, gm, gkd, n, k, coef);
    // G lanes per point with F <= 4 features each: few lanes repeat the per-term
    // work (group sum, square root, division) while rows are still read in
    // G-wide runs.
    auto grad_x = [&](auto g, auto f) {
        launch_grad_x<decltype(g)::value, decltype(f)::value>(keyrows, d, nbr_idx, nbr_d2, kdist2, rev, coef,
                                                              offset, count, gm, gkd, extra, chunk_start,
                                                              owner, n, k, gx, part, s);
    };
    using I = std::integral_constant<int, 1>;
    using II = std::integral_constant<int, 2>;
    using IV = std::integral_constant<int, 4>;
    if (d <= 4) grad_x(std::integral_constant<int, 4>{}, I{});
    else if (d <= 8) grad_x(std::integral_constant<int, 8>{}, I{});
    else if (d <= 16) grad_x(std::integral_constant<int, 8>{}, II{});
    else if (d <= 32) grad_x(std::integral_constant<int, 8>{}, IV{});
    else if (d <= 64) grad_x(std::integral_constant<int, 16>{}, IV{});
    else if (d <= 128) grad_x(std::integral_constant<int, 32>{}, IV{});
    else grad_x(std::integral_constant<int, 64>{}, IV{});
    static_assert(kMaxBackwardDims <= 64 * 4, "the widest grad_x shape covers 256 features");
    chunk_finish_kernel<<<grid_for(nn * d), kBlock, 0, s>>>(extra, chunk_start, part, n, d, gx);
    if (inv_std) {
        zscore_sums_kernel<<<d, kBlock, 0, s>>>(gx, keyrows, n, d, zsums);
        zscore_apply_kernel<<<grid_for(nn * d), kBlock, 0, s>>>(gx, keyrows, zsums, inv_std, scaled, n, d, grad_points);
    } else {
        to_float_kernel<<<grid_for(nn * d), kBlock, 0, s>>>(gx, nn * d, grad_points);
    }
    CULOF_CHECK(hipGetLastError());
}

}  // namespace detail
}  // namespace culof
