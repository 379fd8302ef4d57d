// FP64 centring (optionally z-scoring) and FP64-accumulated row norms.
//
// Semantics follow the reference's src/cuda/preprocess.cu:31-84: centring is
// unconditional, scaling uses the population standard deviation with a 1e-12
// guard, and the norms are taken over the centred float32 rows.
//
// The FP64 summation order is the reference's too, so the means come out bit
// for bit the same: per feature, 256 partial sums where partial p holds rows
// p, p + 256, p + 512, ... in ascending order, then a halving tree over the
// partials. Here the 256 x d partials are independent threads (coalesced:
// consecutive threads read consecutive features of a row) instead of one block
// per feature reading a strided column, and the tree runs in a second tiny
// kernel. The apply pass writes the padded row layout the sweep kernel reads.

#include "internal.hpp"

#include <algorithm>
#include <stdexcept>

namespace culof {
namespace detail {

namespace {

constexpr int kLanes = 256;  // partial sums per feature (the reference's block size)

// Partial p of feature f. With `mean` set, sums squared deviations instead.
__global__ void __launch_bounds__(256)
column_partials_kernel(const float* __restrict__ x, int n, int d, const double* __restrict__ mean,
                       double* __restrict__ part) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= kLanes * d) return;
    const int f = t % d;
    const int p = t / d;
    const bool dev = mean != nullptr;
    const double m = dev ? mean[f] : 0.0;
    double s = 0.0;
    int i = p;
    // Eight independent loads in flight, added strictly in row order.
    for (; i + 7 * kLanes < n; i += 8 * kLanes) {
        float v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) v[u] = x[static_cast<size_t>(i + u * kLanes) * d + f];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            if (dev) {
                const double dv = static_cast<double>(v[u]) - m;
                s = fma(dv, dv, s);
            } else {
                s += static_cast<double>(v[u]);
            }
        }
    }
    for (; i < n; i += kLanes) {
        const float v = x[static_cast<size_t>(i) * d + f];
        if (dev) {
            const double dv = static_cast<double>(v) - m;
            s = fma(dv, dv, s);
        } else {
            s += static_cast<double>(v);
        }
    }
    part[static_cast<size_t>(p) * d + f] = s;
}

// One block per feature: halving tree over the 256 partials.
// pass 0 writes mean = sum / n; pass 1 writes inv_std from the variance and
// flags whether the feature was scaled (1.0) or left centred (0.0).
__global__ void __launch_bounds__(256)
column_finalize_kernel(const double* __restrict__ part, int n, int d, int pass,
                       double* __restrict__ mean, double* __restrict__ inv_std,
                       double* __restrict__ scaled) {
    const int f = blockIdx.x;
    __shared__ double s[kLanes];
    s[threadIdx.x] = part[static_cast<size_t>(threadIdx.x) * d + f];
    __syncthreads();
    for (int off = kLanes / 2; off > 0; off >>= 1) {
        if (static_cast<int>(threadIdx.x) < off) s[threadIdx.x] += s[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        if (pass == 0) {
            mean[f] = s[0] / n;
            inv_std[f] = 1.0;
            scaled[f] = 0.0;
        } else {
            const double sd = sqrt(s[0] / n);
            // A constant feature stays centred at zero rather than divided by noise.
            inv_std[f] = (sd > 1e-12) ? 1.0 / sd : 1.0;
            scaled[f] = (sd > 1e-12) ? 1.0 : 0.0;
        }
    }
}

// One thread per row: centred float32 row (zero-padded to `stride`), its
// norm, and optionally a zero-padded copy of the input row (the key rows when
// not normalising). Raises *guard when a norm is too large for the Gram
// screen to stay finite (above FLT_MAX / 16, or not finite). The block's 256
// rows pass through LDS 16 features at a time, so the loads and stores of
// consecutive threads are consecutive features of a row; each thread still
// sums its row's squares in ascending feature order.
constexpr int kApplyRows = 256;
constexpr int kApplyChunk = 16;

__global__ void __launch_bounds__(kApplyRows)
center_apply_kernel(const float* __restrict__ x, int n, int d, const double* __restrict__ mean,
                    const double* __restrict__ inv_std, float* __restrict__ out, int stride,
                    float* __restrict__ norms, float* __restrict__ copy, int* __restrict__ guard) {
    __shared__ float s_v[kApplyRows][kApplyChunk + 1];
    const int tid = threadIdx.x;
    const int r0 = blockIdx.x * kApplyRows;
    double s = 0.0;
    for (int f0 = 0; f0 < stride; f0 += kApplyChunk) {
        // load: element e of the chunk is row e / kApplyChunk, feature e % kApplyChunk
#pragma unroll
        for (int u = 0; u < kApplyChunk; ++u) {
            const int e = tid + kApplyRows * u;
            const int r = e / kApplyChunk;
            const int f = f0 + (e - r * kApplyChunk);
            const int i = r0 + r;
            float xv = 0.0f;
            if (i < n && f < d) xv = x[static_cast<size_t>(i) * d + f];
            if (copy && i < n && f < stride) copy[static_cast<size_t>(i) * stride + f] = xv;
            s_v[r][e - r * kApplyChunk] = xv;
        }
        __syncthreads();
        // this thread's row, features in ascending order
        for (int c = 0; c < kApplyChunk; ++c) {
            const int f = f0 + c;
            float v = 0.0f;
            if (f < d) {
                v = static_cast<float>((static_cast<double>(s_v[tid][c]) - mean[f]) * inv_std[f]);
                s += static_cast<double>(v) * static_cast<double>(v);  // exact product
            }
            s_v[tid][c] = v;
        }
        __syncthreads();
#pragma unroll
        for (int u = 0; u < kApplyChunk; ++u) {
            const int e = tid + kApplyRows * u;
            const int r = e / kApplyChunk;
            const int f = f0 + (e - r * kApplyChunk);
            const int i = r0 + r;
            if (i < n && f < stride) out[static_cast<size_t>(i) * stride + f] = s_v[r][e - r * kApplyChunk];
        }
        __syncthreads();
    }
    const int i = r0 + tid;
    if (i >= n) return;
    const float nf = static_cast<float>(s);
    norms[i] = nf;
    if (guard && !(nf <= 2.1267647e37f)) atomicOr(guard, 1);  // FLT_MAX / 16, NaN
}

// Zero fill, 16 bytes per store where the region allows.
__global__ void __launch_bounds__(256)
zero_words_kernel(uint32_t* __restrict__ dst, size_t words) {
    const size_t stride = static_cast<size_t>(gridDim.x) * 256;
    size_t i = static_cast<size_t>(blockIdx.x) * 256 + threadIdx.x;
    if ((reinterpret_cast<uintptr_t>(dst) & 15) == 0) {
        uint4* d4 = reinterpret_cast<uint4*>(dst);
        const size_t quads = words / 4;
        for (size_t q = i; q < quads; q += stride) d4[q] = make_uint4(0u, 0u, 0u, 0u);
        for (size_t w = 4 * quads + i; w < words; w += stride) dst[w] = 0u;
        return;
    }
    for (; i < words; i += stride) dst[i] = 0u;
}

}  // namespace

void zero_fill(void* dst, size_t bytes, hipStream_t stream) {
    if (bytes == 0) return;
    if (bytes % 4 != 0) throw std::invalid_argument("zero_fill: size must be whole 32-bit words");
    const size_t words = bytes / 4;
    const unsigned grid = static_cast<unsigned>(std::min<size_t>((words / 4 + 255) / 256 + 1, 2048));
    zero_words_kernel<<<grid, 256, 0, stream>>>(static_cast<uint32_t*>(dst), words);
    CULOF_CHECK(hipGetLastError());
}

void center_and_norms(const float* points, int n_points, int n_dims, bool scale, float* centred,
                      int stride, float* norms, double* work, float* copy, int* guard,
                      hipStream_t stream) {
    double* part = work;                                   // 256 x d
    double* mean = work + static_cast<size_t>(kLanes) * n_dims;
    double* inv_std = mean + n_dims;
    double* scaled = inv_std + n_dims;
    const int threads = kLanes * n_dims;
    const int grid = (threads + 255) / 256;
    {
        ProfileScope prof("center_partials", stream);
        column_partials_kernel<<<grid, 256, 0, stream>>>(points, n_points, n_dims, nullptr, part);
        CULOF_CHECK(hipGetLastError());
    }
    column_finalize_kernel<<<n_dims, kLanes, 0, stream>>>(part, n_points, n_dims, 0, mean, inv_std,
                                                          scaled);
    CULOF_CHECK(hipGetLastError());
    if (scale) {
        column_partials_kernel<<<grid, 256, 0, stream>>>(points, n_points, n_dims, mean, part);
        CULOF_CHECK(hipGetLastError());
        column_finalize_kernel<<<n_dims, kLanes, 0, stream>>>(part, n_points, n_dims, 1, mean,
                                                              inv_std, scaled);
        CULOF_CHECK(hipGetLastError());
    }
    {
        ProfileScope prof("center_apply", stream);
        center_apply_kernel<<<(n_points + kApplyRows - 1) / kApplyRows, kApplyRows, 0, stream>>>(
            points, n_points, n_dims, mean, inv_std, centred, stride, norms, copy, guard);
        CULOF_CHECK(hipGetLastError());
    }
}

void apply_rows(const float* rows, int n_rows, int n_dims, const double* work, float* out,
                int stride, float* norms, float* copy, hipStream_t stream) {
    const double* mean = work + static_cast<size_t>(kLanes) * n_dims;
    const double* inv_std = mean + n_dims;
    if (n_rows <= 0) return;
    center_apply_kernel<<<(n_rows + kApplyRows - 1) / kApplyRows, kApplyRows, 0, stream>>>(
        rows, n_rows, n_dims, mean, inv_std, out, stride, norms, copy, nullptr);
    CULOF_CHECK(hipGetLastError());
}

}  // namespace detail
}  // namespace culof
