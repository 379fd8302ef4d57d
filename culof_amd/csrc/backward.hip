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
// atomics, filled, then each segment sorted by (slot, query) linear index, so
// the result is bit-identical from run to run with no floating-point atomics.
// The distances are recomputed from the key rows in FP64 (n k pairs) instead
// of storing anything n x n. With normalisation the gradient is carried back
// through the z-score (population standard deviation, constant features
// centred only), as for batch normalisation.

#include "internal.hpp"

#include <stdexcept>

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

__global__ void __launch_bounds__(kBlock)
rev_count_kernel(const int* __restrict__ nbr, size_t nk, int* __restrict__ count) {
    const size_t e = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e < nk) atomicAdd(&count[nbr[e]], 1);
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
                int* __restrict__ cursor, int* __restrict__ rev) {
    const size_t e = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e >= nk) return;
    const int p = nbr[e];
    rev[offset[p] + atomicAdd(&cursor[p], 1)] = static_cast<int>(e);  // e = slot * n + query
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

// The kBlock consecutive points of a block own one contiguous CSR range; when
// it fits (kSortLds entries) it is sorted in LDS, each thread its own segment
// (insertion sort for short ones), else each segment in place in global memory.
constexpr int kSortLds = 8192;

__global__ void __launch_bounds__(kBlock)
rev_sort_kernel(int* __restrict__ rev, const int* __restrict__ offset,
                const int* __restrict__ count, int n) {
    __shared__ int s_seg[kSortLds];
    const int p0 = blockIdx.x * kBlock;
    const int p_last = min(p0 + kBlock, n) - 1;
    const int begin = offset[p0];
    const int span = offset[p_last] + count[p_last] - begin;
    const int p = p0 + threadIdx.x;
    if (span > kSortLds) {
        if (p < n) heap_sort(rev + offset[p], count[p]);
        return;
    }
    for (int e = threadIdx.x; e < span; e += kBlock) s_seg[e] = rev[begin + e];
    __syncthreads();
    if (p < n) {
        int* a = s_seg + (offset[p] - begin);
        const int len = count[p];
        if (len <= 32) {
            for (int i = 1; i < len; ++i) {
                const int v = a[i];
                int j = i - 1;
                while (j >= 0 && a[j] > v) {
                    a[j + 1] = a[j];
                    --j;
                }
                a[j + 1] = v;
            }
        } else {
            heap_sort(a, len);
        }
    }
    __syncthreads();
    for (int e = threadIdx.x; e < span; e += kBlock) rev[begin + e] = s_seg[e];
}

// g_m_p = w_p S_p - lrd_p^2 sum_{i : p in N(i)} a_i
__global__ void __launch_bounds__(kBlock)
grad_m_kernel(const int* __restrict__ rev, const int* __restrict__ offset,
              const int* __restrict__ count, const double* __restrict__ a,
              const double* __restrict__ S, const double* __restrict__ lrd,
              const float* __restrict__ w, int n, double* __restrict__ gm) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    double g_lrd = 0.0;
    const int* seg = rev + offset[p];
    for (int r = 0; r < count[p]; ++r) g_lrd += a[seg[r] % n];
    gm[p] = static_cast<double>(w[p]) * S[p] - g_lrd * lrd[p] * lrd[p];
}

// g_kd_p = sum over i with p in N(i) whose reach took the kd branch of g_m_i / k
__global__ void __launch_bounds__(kBlock)
grad_kd_kernel(const int* __restrict__ rev, const int* __restrict__ offset,
               const int* __restrict__ count, const float* __restrict__ d2,
               const float* __restrict__ kd2, const double* __restrict__ gm, int n, int k,
               double* __restrict__ gkd) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    double s = 0.0;
    const int* seg = rev + offset[p];
    for (int r = 0; r < count[p]; ++r) {
        const int e = seg[r];
        if (kd2[p] > d2[e]) s += gm[e % n] / k;
    }
    gkd[p] = s;
}

// A group of G lanes owns one point; lane l of the group holds features
// l, l + G, ... (F of them), so rows are read coalesced and the FP64 row of
// the gradient stays in registers. Squared distances are reduced across the
// group with an xor butterfly, which leaves the same bits in every lane.
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

    // acc += c (x_a - x_b) / |x_a - x_b|; nothing when the points coincide.
    __device__ void add_unit(const double (&xa)[F], const float* __restrict__ xb_row, double c,
                             double (&acc)[F]) const {
        if (c == 0.0) return;  // uniform across the group
        double df[F];
        double ss = 0.0;
#pragma unroll
        for (int u = 0; u < F; ++u) {
            const int f = lane + G * u;
            df[u] = xa[u] - ((f < d) ? static_cast<double>(xb_row[f]) : 0.0);
            ss = fma(df[u], df[u], ss);
        }
#pragma unroll
        for (int m = 1; m < G; m <<= 1) ss += __shfl_xor(ss, m, G);
        if (ss == 0.0) return;
        const double sc = c / sqrt(ss);
#pragma unroll
        for (int u = 0; u < F; ++u) acc[u] = fma(sc, df[u], acc[u]);
    }
};

// Every term a point owns or receives (own list, own k-distance, reverse
// list) accumulated into its FP64 gradient row; O(n k d) in total.
template <int G, int F>
__global__ void __launch_bounds__(kBlock)
grad_x_kernel(const float* __restrict__ xs, int d, const int* __restrict__ nbr,
              const float* __restrict__ d2, const float* __restrict__ kd2,
              const int* __restrict__ rev, const int* __restrict__ offset,
              const int* __restrict__ count, const double* __restrict__ gm,
              const double* __restrict__ gkd, int n, int k, double* __restrict__ gx) {
    const int p = (blockIdx.x * kBlock + threadIdx.x) / G;
    if (p >= n) return;  // whole groups leave together
    const GroupRow<G, F> row{static_cast<int>(threadIdx.x % G), d};
    double xp[F], acc[F];
    row.load(xs + static_cast<size_t>(p) * d, xp);
#pragma unroll
    for (int u = 0; u < F; ++u) acc[u] = 0.0;
    // own list: branch-d terms, then the own k-distance
    const double own = gm[p] / k;
    for (int t = 0; t < k; ++t) {
        const size_t o = static_cast<size_t>(t) * n + p;
        const int j = nbr[o];
        if (!(kd2[j] > d2[o])) row.add_unit(xp, xs + static_cast<size_t>(j) * d, own, acc);
    }
    row.add_unit(xp, xs + static_cast<size_t>(nbr[static_cast<size_t>(k - 1) * n + p]) * d, gkd[p],
                 acc);
    // reverse list: p as neighbour j of query i
    const int* seg = rev + offset[p];
    const int len = count[p];
    for (int r = 0; r < len; ++r) {
        const int e = seg[r];
        const int i = e % n;
        const int slot = e / n;
        double c = 0.0;
        if (!(kd2[p] > d2[e])) c += gm[i] / k;  // branch d of reach_ip
        if (slot == k - 1) c += gkd[i];         // p is the k-th neighbour of i
        row.add_unit(xp, xs + static_cast<size_t>(i) * d, c, acc);
    }
#pragma unroll
    for (int u = 0; u < F; ++u) {
        const int f = row.lane + G * u;
        if (f < d) gx[static_cast<size_t>(p) * d + f] = acc[u];
    }
}

template <int G, int F>
void launch_grad_x(const float* xs, int d, const int* nbr, const float* d2, const float* kd2,
                   const int* rev, const int* offset, const int* count, const double* gm,
                   const double* gkd, int n, int k, double* gx, hipStream_t s) {
    const size_t lanes = static_cast<size_t>(n) * G;
    grad_x_kernel<G, F><<<static_cast<unsigned>((lanes + kBlock - 1) / kBlock), kBlock, 0, s>>>(
        xs, d, nbr, d2, kd2, rev, offset, count, gm, gkd, n, k, gx);
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
    b += 7 * nn * sizeof(double);               // m, lrd, S, a, gm, gkd + slack
    b += nn * d * sizeof(double);               // gradient in FP64
    b += (3 * nn + blocks + 8) * sizeof(int);   // count, offset, cursor, block totals
    b += nn * k * sizeof(int);                  // reverse lists
    b += 16 * 256;                              // alignment
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
    int* cursor = static_cast<int*>(take(nn * sizeof(int)));
    int* totals = static_cast<int*>(take((blocks + 8) * sizeof(int)));
    int* rev = static_cast<int*>(take(nk * sizeof(int)));

    ProfileScope prof("lof_backward", s);
    point_m_kernel<<<grid_for(nn), kBlock, 0, s>>>(nbr_idx, nbr_d2, kdist2, n, k, m, lrd);
    point_s_kernel<<<grid_for(nn), kBlock, 0, s>>>(nbr_idx, lrd, m, grad_scores, n, k, S, a);
    zero_fill(count, nn * sizeof(int), s);
    zero_fill(cursor, nn * sizeof(int), s);
    rev_count_kernel<<<grid_for(nk), kBlock, 0, s>>>(nbr_idx, nk, count);
    scan_blocks_kernel<<<static_cast<unsigned>(blocks), kBlock, 0, s>>>(count, n, offset, totals);
    scan_totals_kernel<<<1, kBlock, 0, s>>>(totals, static_cast<int>(blocks));
    scan_add_kernel<<<static_cast<unsigned>(blocks), kBlock, 0, s>>>(offset, n, totals);
    rev_fill_kernel<<<grid_for(nk), kBlock, 0, s>>>(nbr_idx, nk, offset, cursor, rev);
    rev_sort_kernel<<<grid_for(nn), kBlock, 0, s>>>(rev, offset, count, n);
    grad_m_kernel<<<grid_for(nn), kBlock, 0, s>>>(rev, offset, count, a, S, lrd, grad_scores, n, gm);
    grad_kd_kernel<<<grid_for(nn), kBlock, 0, s>>>(rev, offset, count, nbr_d2, kdist2, gm, n, k, gkd);
    if (d <= 4) {
        launch_grad_x<4, 1>(keyrows, d, nbr_idx, nbr_d2, kdist2, rev, offset, count, gm, gkd, n, k, gx, s);
    } else if (d <= 8) {
        launch_grad_x<8, 1>(keyrows, d, nbr_idx, nbr_d2, kdist2, rev, offset, count, gm, gkd, n, k, gx, s);
    } else if (d <= 16) {
        launch_grad_x<16, 1>(keyrows, d, nbr_idx, nbr_d2, kdist2, rev, offset, count, gm, gkd, n, k, gx, s);
    } else if (d <= 32) {
        launch_grad_x<32, 1>(keyrows, d, nbr_idx, nbr_d2, kdist2, rev, offset, count, gm, gkd, n, k, gx, s);
    } else if (d <= 64) {
        launch_grad_x<64, 1>(keyrows, d, nbr_idx, nbr_d2, kdist2, rev, offset, count, gm, gkd, n, k, gx, s);
    } else if (d <= 128) {
        launch_grad_x<64, 2>(keyrows, d, nbr_idx, nbr_d2, kdist2, rev, offset, count, gm, gkd, n, k, gx, s);
    } else {
        launch_grad_x<64, kMaxBackwardDims / 64>(keyrows, d, nbr_idx, nbr_d2, kdist2, rev, offset, count, gm,
                                                 gkd, n, k, gx, s);
    }
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
