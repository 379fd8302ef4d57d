// Host orchestration, the C ABI (include/culof_c.h) and the C++ API
// (include/culof.h) of libculof_amd.
//
// The forward pass is the reference's culof::lof (src/cuda/pipeline.cu:131-207)
// re-planned for one MI355X:
//   centre + norms (prep.hip) -> screened exact kNN sweep for every query row
//   (mfma.hip on the matrix cores, or sweep.hip on the VALU; no distance tile
//   exists, so memory is O(n k)) -> exact radix redo of any row a sweep hands
//   back (radix.hip; a fixed grid-stride launch, no host read-back) -> LRD ->
//   score (density.hip).
// Scratch is stream-ordered (hipMallocAsync on the call's stream), so calls on
// different streams or threads never share memory and nothing synchronises
// the device mid-pipeline. Query rows go to the kNN kernels in tiles sized to
// the device memory left free (all rows at once unless memory is short);
// CULOF_TILE_ROWS forces a tile height. Results are bit-identical for any
// tile height because every query row is computed independently.

#include "internal.hpp"

#include "culof.h"
#include "culof_c.h"

#include <algorithm>
#include <atomic>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <mutex>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

namespace culof {
namespace detail {

void check_hip(hipError_t err, const char* file, int line, const char* expr) {
    if (err == hipSuccess) return;
    std::ostringstream os;
    os << "HIP error at " << file << ":" << line << " (" << expr << "): " << hipGetErrorString(err);
    throw CudaError(os.str());
}

int choose_tile_rows(int n_points, size_t budget_bytes) {
    const size_t row_bytes = static_cast<size_t>(n_points) * sizeof(float);
    size_t rows = budget_bytes / std::max<size_t>(row_bytes, 1);
    rows = std::clamp<size_t>(rows, 1, static_cast<size_t>(std::max(n_points, 1)));
    return static_cast<int>(rows);
}

// ---------------------------------------------------------------------------
// Kernel timing registry (culof_profile_*).
// ---------------------------------------------------------------------------

namespace {

struct ProfRecord {
    std::string name;
    hipEvent_t start = nullptr;
    hipEvent_t stop = nullptr;
};

std::atomic<bool> g_prof_on{false};
std::mutex g_prof_mu;
std::vector<ProfRecord> g_prof;

}  // namespace

// The tile-pair work counter of culof_profile_tile_pairs: one device word
// per device, allocated when timing is first enabled on it.
unsigned long long* profile_tile_counter_for(int device, bool create) {
    static std::mutex mu;
    static std::vector<std::pair<int, unsigned long long*>> counters;
    std::lock_guard<std::mutex> lock(mu);
    for (auto& c : counters)
        if (c.first == device) return c.second;
    if (!create) return nullptr;
    unsigned long long* p = nullptr;
    CULOF_CHECK(hipMalloc(&p, sizeof(unsigned long long)));
    CULOF_CHECK(hipMemset(p, 0, sizeof(unsigned long long)));
    counters.emplace_back(device, p);
    return p;
}

unsigned long long* profile_tile_counter() {
    if (!g_prof_on.load(std::memory_order_relaxed)) return nullptr;
    int device = 0;
    if (hipGetDevice(&device) != hipSuccess) return nullptr;
    return profile_tile_counter_for(device, false);
}

ProfileScope::ProfileScope(const char* name, hipStream_t stream) : stream_(stream) {
    if (!g_prof_on.load(std::memory_order_relaxed)) return;
    ProfRecord rec;
    rec.name = name;
    CULOF_CHECK(hipEventCreate(&rec.start));
    CULOF_CHECK(hipEventCreate(&rec.stop));
    CULOF_CHECK(hipEventRecord(rec.start, stream));
    std::lock_guard<std::mutex> lock(g_prof_mu);
    slot_ = static_cast<int>(g_prof.size());
    g_prof.push_back(rec);
}

ProfileScope::~ProfileScope() {
    if (slot_ < 0) return;
    hipEvent_t stop;
    {
        std::lock_guard<std::mutex> lock(g_prof_mu);
        if (slot_ >= static_cast<int>(g_prof.size())) return;  // reset meanwhile
        stop = g_prof[slot_].stop;
    }
    (void)hipEventRecord(stop, stream_);
}

namespace {

// ---------------------------------------------------------------------------
// Stream-ordered scratch.
// ---------------------------------------------------------------------------

// Memory the device's default pool keeps after a call's scratch is freed, so
// repeated calls do not return it to the driver and map it in again (about
// 0.4 ms per call for the 9 GB of a 100k x 128 forward with the BF16 screen):
// up to 1/16 of the device's memory, at most 24 GiB (18 GB on a 288 GB
// device; CULOF_POOL_RETAIN_MB overrides). Larger calls release the excess at
// their next synchronisation.
void enable_pool_reuse(int device) {
    static std::mutex mu;
    static std::vector<int> done;
    std::lock_guard<std::mutex> lock(mu);
    if (std::find(done.begin(), done.end(), device) != done.end()) return;
    hipMemPool_t pool;
    if (hipDeviceGetDefaultMemPool(&pool, device) == hipSuccess) {
        size_t free_b = 0, total_b = 0;
        (void)hipMemGetInfo(&free_b, &total_b);
        uint64_t threshold = std::min<uint64_t>(total_b / 16, 24ull << 30);
        if (const char* env = std::getenv("CULOF_POOL_RETAIN_MB")) {
            const long long mb = std::atoll(env);
            if (mb >= 0) threshold = static_cast<uint64_t>(mb) << 20;
        }
        (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &threshold);
    }
    done.push_back(device);
}

// The reference keeps 128 MiB of device memory free when it sizes its tiles
// (src/cuda/pipeline.cu:104-108); so does every call here.
constexpr size_t kFreeMargin = 128ull << 20;

// Device memory a call can still obtain from the stream-ordered allocator:
// free memory plus what the default pool holds idle, less the margin.
size_t device_scratch_available() {
    size_t free_b = 0, total_b = 0;
    CULOF_CHECK(hipMemGetInfo(&free_b, &total_b));
    int device = 0;
    hipMemPool_t pool;
    uint64_t reserved = 0, used = 0;
    if (hipGetDevice(&device) == hipSuccess && hipDeviceGetDefaultMemPool(&pool, device) == hipSuccess &&
        hipMemPoolGetAttribute(pool, hipMemPoolAttrReservedMemCurrent, &reserved) == hipSuccess &&
        hipMemPoolGetAttribute(pool, hipMemPoolAttrUsedMemCurrent, &used) == hipSuccess && reserved > used) {
        free_b += static_cast<size_t>(reserved - used);
    }
    return free_b > kFreeMargin ? free_b - kFreeMargin : 0;
}

// CULOF_SCRATCH_BUDGET (bytes): an upper bound on one call's scratch, below
// what the device has free (to leave memory to other work, or to exercise the
// row tiling on a small budget). 0 when unset.
size_t scratch_budget_env() {
    const char* env = std::getenv("CULOF_SCRATCH_BUDGET");
    if (!env) return 0;
    const long long v = std::atoll(env);
    return v > 0 ? static_cast<size_t>(v) : 0;
}

// A call's device scratch: stream-ordered allocations (hipMallocAsync on the
// call's stream from the default pool), freed when the call returns; or, for
// the no-allocation entry point, consecutive pieces of one fixed device
// region, where running past the region's end is an argument error and no
// allocator is called.
class Scratch {
public:
    explicit Scratch(hipStream_t s) : stream_(s), budget_(scratch_budget_env()) {}
    Scratch(hipStream_t s, void* region, size_t bytes)
        : stream_(s), region_(static_cast<unsigned char*>(region)), cap_(bytes) {}
    ~Scratch() {
        for (void* p : ptrs_) (void)hipFreeAsync(p, stream_);
    }
    Scratch(const Scratch&) = delete;
    Scratch& operator=(const Scratch&) = delete;

    // Bytes get() reserves for `count` elements: whole 256-byte blocks.
    template <typename T>
    static size_t footprint(size_t count) {
        return (std::max<size_t>(count * sizeof(T), 256) + 255) / 256 * 256;
    }

    template <typename T>
    T* get(size_t count) {
        const size_t bytes = footprint<T>(count);
        if (region_) {
            if (bytes > cap_ - used_) {
                std::ostringstream os;
                os << "the call needs more than the " << (cap_ >> 20)
                   << " MiB of scratch the no-allocation entry point holds";
                throw std::invalid_argument(os.str());
            }
            void* p = region_ + used_;
            used_ += bytes;
            return static_cast<T*>(p);
        }
        void* p = nullptr;
        CULOF_CHECK(hipMallocAsync(&p, bytes, stream_));
        ptrs_.push_back(p);
        used_ += bytes;
        return static_cast<T*>(p);
    }

    // Bytes later get() calls may still take (free device memory and the
    // pool's idle blocks less the margin, within CULOF_SCRATCH_BUDGET).
    size_t available() const {
        if (region_) return cap_ - used_;
        size_t avail = device_scratch_available();
        if (budget_) avail = std::min(avail, budget_ > used_ ? budget_ - used_ : size_t{0});
        return avail;
    }

    // True for a fixed region: the pipeline then takes its low-memory paths.
    bool fixed() const { return region_ != nullptr; }

private:
    hipStream_t stream_;
    unsigned char* region_ = nullptr;
    size_t cap_ = 0;
    size_t budget_ = 0;
    size_t used_ = 0;
    std::vector<void*> ptrs_;
};

class Stream {
public:
    Stream() { CULOF_CHECK(hipStreamCreate(&s_)); }
    ~Stream() {
        if (s_) (void)hipStreamDestroy(s_);
    }
    Stream(const Stream&) = delete;
    Stream& operator=(const Stream&) = delete;
    hipStream_t get() const { return s_; }

private:
    hipStream_t s_ = nullptr;
};

void validate(const void* points, int n_points, int n_dims, int k) {
    if (points == nullptr) throw std::invalid_argument("points must not be null");
    if (n_points < 2) throw std::invalid_argument("n_points must be >= 2");
    if (n_dims < 1) throw std::invalid_argument("n_dims must be >= 1");
    if (k < 1 || k >= n_points) {
        std::ostringstream os;
        os << "k must satisfy 1 <= k <= n_points - 1 (got k=" << k << ", n_points=" << n_points
           << ")";
        throw std::invalid_argument(os.str());
    }
}

bool device_present() {
    int count = 0;
    return hipGetDeviceCount(&count) == hipSuccess && count > 0;
}

void require_device() {
    if (!device_present()) throw CudaError("no HIP device available");
}

// Query rows per launch (the reference's tile_rows_for,
// src/cuda/pipeline.cu:99-110): CULOF_TILE_ROWS when set, else the largest of
// rows, rows / 2, rows / 4, ... (down to kMinTileRows) whose scratch,
// bytes_for(tile), fits in `avail`. The reference sizes a rows x n Gram tile;
// here nothing is n wide, and what grows with the tile is the collect
// buffers (about 8-64 KB per query row), so all rows usually go in one launch.
// Every query row is computed on its own, so any tile height gives the same
// bits; a run split into tiles takes the non-symmetric collect pass.
constexpr int kMinTileRows = 256;

template <class Bytes>
int tile_rows_for(int rows, size_t avail, Bytes&& bytes_for) {
    if (const char* env = std::getenv("CULOF_TILE_ROWS")) {
        const int forced = std::atoi(env);
        if (forced > 0) return std::min(forced, rows);
    }
    int tile = rows;
    while (tile > kMinTileRows && bytes_for(tile) > avail) tile = (tile + 1) / 2;
    return std::max(tile, std::min(rows, 1));
}

// Device-resident inputs of the selection stage, built by prepare().
struct Prepared {
    int stride = 0;             // padded row length shared by both row sets
    const float* xs = nullptr;  // key rows: the input, or the z-scored rows
    const float* xc = nullptr;  // centred (and scaled) rows for the distance screen
    const float* norms = nullptr;
    const int* guard = nullptr;
    const double* work = nullptr;  // center_and_norms' work: partials, mean, inv_std, scaled
};

Prepared prepare(const float* points, int n, int d, bool normalize, int form, Scratch& scratch,
                 hipStream_t s) {
    Prepared p;
    p.stride = mfma_padded_dims(d, form);
    const size_t cells = static_cast<size_t>(n) * p.stride;
    float* xc = scratch.get<float>(cells);
    float* norms = scratch.get<float>(n);
    double* work = scratch.get<double>(static_cast<size_t>(kPrepWork) * d);
    int* guard = scratch.get<int>(1);
    // Key rows: the input itself when it already has the padded stride and
    // whole-float4 alignment, else a padded copy (the z-scored rows when
    // normalising).
    const bool in_place = !normalize && p.stride == d && reinterpret_cast<uintptr_t>(points) % 16 == 0;
    float* xs = (normalize || in_place) ? nullptr : scratch.get<float>(cells);
    zero_fill(guard, sizeof(int), s);
    // Centre always; scale only when asked (src/cuda/pipeline.cu:162-165).
    center_and_norms(points, n, d, normalize, xc, p.stride, norms, work, xs, guard, s);
    p.xs = normalize ? xc : (in_place ? points : xs);
    p.xc = xc;
    p.norms = norms;
    p.guard = guard;
    p.work = work;
    return p;
}

// Which kNN path serves (d, k, form): the matrix-core sweep wherever it is
// instantiated (with the candidate order and seeds it is the fastest on
// MI355X for every measured shape), else the VALU sweep, else radix.
// CULOF_KNN_PATH (mfma | valu | radix) pins a path, to benchmark the others.
enum class Path { kMfma, kValu, kRadix };

// CULOF_ORDER=0 streams candidates in input order with unseeded lists (the
// results are the same either way; this is for measuring the difference).
bool ordering_enabled() {
    const char* env = std::getenv("CULOF_ORDER");
    return !(env && env[0] == '0');
}

Path choose_path(int d, int k, int form, int stride) {
    const bool mfma = mfma_supported(d, k, form);
    const bool valu = form == kFormPlain && stride == padded_dims(d) && sweep_supported(d, k);
    if (const char* env = std::getenv("CULOF_KNN_PATH")) {
        const std::string want(env);
        if (want == "radix") return Path::kRadix;
        if (want == "valu" && valu) return Path::kValu;
        if (want == "mfma" && mfma) return Path::kMfma;
    }
    if (mfma) return Path::kMfma;
    if (valu) return Path::kValu;
    return Path::kRadix;
}

// kNN of query rows [row0, row0 + rows). method 0 = the production path for
// `form` (see choose_path), 1 = radix for every row, 2 = VALU sweep.
// Rows a sweep hands back get the exact radix redo.
void knn_rows(const Prepared& p, int n, int d, int k, int form, int row0, int rows, int method,
              int* nbr_idx, float* nbr_d2, float* kdist2, Scratch& scratch, hipStream_t s) {
    if (form & kFormTiesFp64) {
        // FP64 tie-break: select k + 1, keep rows whose boundary is clear, settle
        // the rest in FP64. With k = n - 1 every candidate is a neighbour.
        const int screen = form & kFormMask;
        if (k + 1 > n - 1) {
            knn_rows(p, n, d, k, screen, row0, rows, method, nbr_idx, nbr_d2, kdist2, scratch, s);
            return;
        }
        const size_t n1 = static_cast<size_t>(n) * (k + 1);
        int* idx1 = scratch.get<int>(n1);
        float* d21 = scratch.get<float>(n1);
        float* kd1 = scratch.get<float>(static_cast<size_t>(n));
        int* tie_rows = scratch.get<int>(static_cast<size_t>(rows));
        int* tie_count = scratch.get<int>(1);
        knn_rows(p, n, d, k + 1, screen, row0, rows, method, idx1, d21, kd1, scratch, s);
        // Near-tie rows are collected on the matrix cores like the main pass
        // when that is the path serving it (and scratch can be allocated); the
        // brute-force scan takes the rest.
        TieBand band{};
        const int cap = tie_band_capacity(n, d, p.stride, screen, rows);
        if (method == 0 && cap > 0 && !scratch.fixed() &&
            choose_path(d, k + 1, screen, p.stride) == Path::kMfma) {
            band = TieBand{p.xc, p.norms, p.guard, screen, cap,
                           scratch.get<int>(static_cast<size_t>(cap)), scratch.get<int>(1),
                           scratch.get<double>(static_cast<size_t>(cap)),
                           scratch.get<float>(static_cast<size_t>(n)),
                           scratch.get<unsigned char>(tie_band_scratch_bytes(n, d, screen, cap))};
        }
        tie_break_rows(p.xs, p.stride, n, d, k, row0, rows, idx1, d21, nbr_idx, nbr_d2, kdist2,
                       tie_rows, tie_count, band.rows ? &band : nullptr, s);
        return;
    }
    Path path = Path::kRadix;
    if (method == 0) path = choose_path(d, k, form, p.stride);
    if (method == 2) path = Path::kValu;
    // Candidate order and seeds feed the collect pipeline, whose buffers grow
    // with the rows; a fixed scratch region takes the list sweep instead
    // (no order, per-row lists in registers; the same bits).
    const bool order = path == Path::kMfma && ordering_enabled() && order_supported(p.stride) &&
                       !scratch.fixed();
    // Without the candidate order the matrix cores run the list sweep, which
    // holds k <= kListKMax; larger k then takes radix.
    if (path == Path::kMfma && !order && !mfma_list_supported(d, k, form)) path = Path::kRadix;
    // Only the matrix-core path keeps per-row scratch (its collect buffers).
    auto tile_bytes = [&](int t) {
        if (path != Path::kMfma) return size_t{0};
        size_t b = Scratch::footprint<unsigned char>(mfma_scratch_bytes(n, p.stride, form, t, k, order));
        if (order) {
            b += Scratch::footprint<unsigned char>(order_scratch_bytes(n, p.stride)) +
                 Scratch::footprint<float>(static_cast<size_t>(n));
        }
        return b;
    };
    const size_t fixed = 2 * Scratch::footprint<int>(static_cast<size_t>(rows));
    const size_t avail = scratch.available();
    const int tile = tile_rows_for(rows, avail > fixed ? avail - fixed : 0, tile_bytes);
    if (path == Path::kRadix) {
        for (int r0 = 0; r0 < rows; r0 += tile) {
            radix_knn(p.xs, p.stride, n, d, k, row0 + r0, std::min(tile, rows - r0), nullptr,
                      nullptr, nbr_idx, nbr_d2, kdist2, s);
        }
        return;
    }
    int* fb_count = scratch.get<int>(1);
    int* fb_rows = scratch.get<int>(static_cast<size_t>(rows));
    zero_fill(fb_count, sizeof(int), s);
    void* mfma_scratch = nullptr;
    CandidateOrder ord;
    const int* qorder = nullptr;
    float* tseed = nullptr;
    if (path == Path::kMfma) {
        mfma_scratch = scratch.get<unsigned char>(mfma_scratch_bytes(n, p.stride, form, tile, k, order));
        if (order) {
            // Candidates in locality order, lists seeded from nearby candidates;
            // a full run also takes its queries in that order (coherent tiles).
            ord = build_order(p.xs, n, p.stride,
                              scratch.get<unsigned char>(order_scratch_bytes(n, p.stride)), s);
            tseed = scratch.get<float>(static_cast<size_t>(n));
            seed_thresholds(ord, n, padded_dims(d), k, tseed, s);
            if (row0 == 0 && rows == n) {
                // A full run takes its queries in the order too; its rows with
                // the largest seeds go last, which tightens the tile bounds
                // the symmetric pass prunes with.
                order_large_seeds_last(ord, n, tseed, s);
                qorder = ord.order;
            }
        }
    }
    for (int r0 = 0; r0 < rows; r0 += tile) {
        const int r = std::min(tile, rows - r0);
        if (path == Path::kMfma) {
            mfma_knn(p.xs, p.xc, p.stride, p.norms, p.guard, n, d, k, form, row0 + r0, r,
                     tseed ? &ord : nullptr, qorder ? qorder + r0 : nullptr, tseed, mfma_scratch,
                     nbr_idx, nbr_d2, kdist2, fb_rows, fb_count, rows, s);
        } else {
            sweep_knn(p.xs, p.xc, p.stride, p.norms, p.guard, n, d, k, row0 + r0, r, nbr_idx,
                      nbr_d2, kdist2, fb_rows, fb_count, rows, s);
        }
    }
    radix_knn(p.xs, p.stride, n, d, k, 0, rows, fb_rows, fb_count, nbr_idx, nbr_d2, kdist2, s);
}

void forward_with(const float* points, int n, int d, int k, bool normalize, int form, float* scores,
                  int* nbr_idx, float* nbr_d2, float* kdist2, float* lrd, Scratch& scratch,
                  hipStream_t s) {
    const size_t nn = static_cast<size_t>(n);
    const size_t nk = nn * static_cast<size_t>(k);
    if (!nbr_idx) nbr_idx = scratch.get<int>(nk);
    if (!nbr_d2) nbr_d2 = scratch.get<float>(nk);
    if (!kdist2) kdist2 = scratch.get<float>(nn);
    if (!lrd) lrd = scratch.get<float>(nn);
    const Prepared p = prepare(points, n, d, normalize, screen_form(form), scratch, s);
    knn_rows(p, n, d, k, form, 0, n, 0, nbr_idx, nbr_d2, kdist2, scratch, s);
    compute_lrd(nbr_idx, nbr_d2, kdist2, n, k, lrd, s);
    compute_scores(nbr_idx, lrd, n, k, scores, s);
}

void forward(const float* points, int n, int d, int k, bool normalize, int form, float* scores,
             int* nbr_idx, float* nbr_d2, float* kdist2, float* lrd, hipStream_t s) {
    if (screen_form(form) < 0) throw std::invalid_argument("unknown precision form");
    int device = 0;
    CULOF_CHECK(hipGetDevice(&device));
    enable_pool_reuse(device);
    Scratch scratch(s);
    forward_with(points, n, d, k, normalize, form, scores, nbr_idx, nbr_d2, kdist2, lrd, scratch, s);
}

// Least scratch forward_with takes from a fixed region: every buffer it
// carves there, with the kNN rows in tiles of at most kMinTileRows.
size_t region_floor(const float* points, int n, int d, int k, bool normalize, int form,
                    bool have_outputs) {
    const size_t nn = static_cast<size_t>(n);
    const int screen = screen_form(form);
    const int stride = mfma_padded_dims(d, screen);
    const size_t cells = nn * static_cast<size_t>(stride);
    size_t b = 0;
    if (!have_outputs) {
        b += 2 * Scratch::footprint<int>(nn * static_cast<size_t>(k)) + 2 * Scratch::footprint<float>(nn);
    }
    const bool in_place = !normalize && stride == d && reinterpret_cast<uintptr_t>(points) % 16 == 0;
    b += Scratch::footprint<float>(cells) + Scratch::footprint<float>(nn) +
         Scratch::footprint<double>(static_cast<size_t>(kPrepWork) * d) + Scratch::footprint<int>(1);
    if (!normalize && !in_place) b += Scratch::footprint<float>(cells);
    int k_sel = k;
    if ((form & kFormTiesFp64) && k + 1 <= n - 1) {
        k_sel = k + 1;
        b += 2 * Scratch::footprint<int>(nn * static_cast<size_t>(k_sel)) + 2 * Scratch::footprint<int>(nn) +
             Scratch::footprint<int>(1);
    }
    Path path = choose_path(d, k_sel, screen, stride);
    if (path == Path::kMfma && !mfma_list_supported(d, k_sel, screen)) path = Path::kRadix;
    if (path != Path::kRadix) b += Scratch::footprint<int>(1) + Scratch::footprint<int>(nn);
    if (path == Path::kMfma) {
        b += Scratch::footprint<unsigned char>(
            mfma_scratch_bytes(n, stride, screen, std::min(n, kMinTileRows), k_sel, false));
    }
    return b;
}


// New points scored against a fitted set (scikit-learn's novelty mode,
// LocalOutlierFactor(novelty=True).score_samples): the forward on `points`
// gives the fitted k-distances and densities; each query's k nearest fitted
// points (exact FP32 keys on the same key rows -- the query rows themselves,
// or z-scored with the fitted mean and scale -- none excluded, the
// reference's tie rule and slot order; the VALU sweep, or the radix select
// where the sweep is not instantiated) give its reach distances, density and
// LOF. scores_fit (optional) receives the fitted points' own scores.
void novelty(const float* points, int n, int d, const float* queries, int m, int k, bool normalize,
             int form, float* scores_fit, float* scores_new, int* nbr_idx_new, float* nbr_d2_new,
             hipStream_t s) {
    if (screen_form(form) < 0) throw std::invalid_argument("unknown precision form");
    int device = 0;
    CULOF_CHECK(hipGetDevice(&device));
    enable_pool_reuse(device);
    Scratch scratch(s);
    const size_t nk = static_cast<size_t>(n) * k;
    int* idx = scratch.get<int>(nk);
    float* d2 = scratch.get<float>(nk);
    float* kd = scratch.get<float>(static_cast<size_t>(n));
    float* lrd = scratch.get<float>(static_cast<size_t>(n));
    const Prepared p = prepare(points, n, d, normalize, screen_form(form), scratch, s);
    knn_rows(p, n, d, k, form, 0, n, 0, idx, d2, kd, scratch, s);
    compute_lrd(idx, d2, kd, n, k, lrd, s);
    if (scores_fit) compute_scores(idx, lrd, n, k, scores_fit, s);
    if (m <= 0) return;
    const size_t cells = static_cast<size_t>(m) * p.stride;
    float* qc = scratch.get<float>(cells);
    float* qnorms = scratch.get<float>(static_cast<size_t>(m));
    float* qcopy = normalize ? nullptr : scratch.get<float>(cells);
    apply_rows(queries, m, d, p.work, qc, p.stride, qnorms, qcopy, s);
    const size_t mk = static_cast<size_t>(m) * k;
    int* qidx = nbr_idx_new ? nbr_idx_new : scratch.get<int>(mk);
    float* qd2 = nbr_d2_new ? nbr_d2_new : scratch.get<float>(mk);
    float* qkd = scratch.get<float>(static_cast<size_t>(m));
    float* qlrd = scratch.get<float>(static_cast<size_t>(m));
    const float* qkeys = normalize ? qc : qcopy;
    if (p.stride == padded_dims(d) && sweep_supported(d, k)) {
        // The VALU sweep with the queries' own rows; the rows it hands back
        // (huge norms, fewer than k finite keys) take the exact radix select.
        int* fb_count = scratch.get<int>(1);
        int* fb_rows = scratch.get<int>(static_cast<size_t>(m));
        zero_fill(fb_count, sizeof(int), s);
        sweep_query_knn(qkeys, qc, qnorms, m, p.xs, p.xc, p.stride, p.norms, p.guard, n, d, k, qidx, qd2,
                        qkd, fb_rows, fb_count, m, s);
        radix_query_knn(qkeys, p.xs, p.stride, n, d, k, m, fb_rows, fb_count, qidx, qd2, qkd, s);
    } else {
        radix_query_knn(qkeys, p.xs, p.stride, n, d, k, m, nullptr, nullptr, qidx, qd2, qkd, s);
    }
    compute_lrd(qidx, qd2, kd, m, k, qlrd, s);  // kd: the fitted k-distances, gathered by index
    compute_query_scores(qidx, lrd, qlrd, m, k, scores_new, s);
}

std::vector<float> lof_host(const float* points, int n_points, int n_dims, int k, bool normalize) {
    validate(points, n_points, n_dims, k);
    require_device();
    Stream stream;
    const hipStream_t s = stream.get();
    const size_t n = static_cast<size_t>(n_points);
    std::vector<float> out(n);
    {
        Scratch scratch(s);
        float* d_points = scratch.get<float>(n * n_dims);
        float* d_scores = scratch.get<float>(n);
        CULOF_CHECK(hipMemcpyAsync(d_points, points, n * n_dims * sizeof(float),
                                   hipMemcpyHostToDevice, s));
        forward(d_points, n_points, n_dims, k, normalize, CULOF_FORM_PLAIN, d_scores, nullptr,
                nullptr, nullptr, nullptr, s);
        CULOF_CHECK(hipMemcpyAsync(out.data(), d_scores, n * sizeof(float), hipMemcpyDeviceToHost,
                                   s));
    }
    CULOF_CHECK(hipStreamSynchronize(s));
    return out;
}

std::string describe_device() {
    int device = 0;
    if (hipGetDevice(&device) != hipSuccess) return "no HIP device";
    hipDeviceProp_t prop{};
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) return "unknown device";
    std::ostringstream os;
    os << prop.name << " (" << prop.gcnArchName << ", " << prop.multiProcessorCount << " CUs, "
       << (prop.totalGlobalMem >> 20) << " MiB)";
    return os.str();
}

// ---------------------------------------------------------------------------
// C ABI plumbing: exceptions -> status codes + per-thread message.
// ---------------------------------------------------------------------------

thread_local std::string g_last_error;

template <class F>
int guarded(F&& f) {
    try {
        g_last_error.clear();
        f();
        return CULOF_OK;
    } catch (const std::invalid_argument& e) {
        g_last_error = e.what();
        return CULOF_INVALID_ARGUMENT;
    } catch (const CudaError& e) {
        g_last_error = e.what();
        return device_present() ? CULOF_DEVICE_ERROR : CULOF_NO_DEVICE;
    } catch (const std::exception& e) {
        g_last_error = e.what();
        return CULOF_DEVICE_ERROR;
    } catch (...) {
        g_last_error = "unknown error";
        return CULOF_DEVICE_ERROR;
    }
}

hipStream_t as_stream(void* s) { return static_cast<hipStream_t>(s); }

}  // namespace

int run_guarded(const std::function<void()>& f) { return guarded(f); }

// The forward of culof_lof_device with all scratch in [region, region + bytes)
// (culof_lof_device_noalloc): arguments checked, the region's size checked
// against the least the call needs before anything is launched, no allocator
// called. Calls that share a region must not overlap.
void forward_in_region(const float* points, int n, int d, int k, bool normalize, int form,
                       float* scores, int* nbr_idx, float* nbr_d2, float* kdist2, float* lrd,
                       void* region, size_t bytes, hipStream_t s) {
    validate(points, n, d, k);
    if (!scores) throw std::invalid_argument("scores must not be null");
    if (screen_form(form) < 0) throw std::invalid_argument("unknown precision form");
    require_device();
    const bool have_outputs = nbr_idx && nbr_d2 && kdist2 && lrd;
    const size_t need = region_floor(points, n, d, k, normalize, form, have_outputs);
    if (need > bytes) {
        std::ostringstream os;
        os << "n_points=" << n << ", n_dims=" << d << ", k=" << k << " needs at least "
           << ((need + (1 << 20) - 1) >> 20) << " MiB of scratch; the no-allocation entry point holds "
           << (bytes >> 20) << " MiB (use culof_lof_device)";
        throw std::invalid_argument(os.str());
    }
    Scratch scratch(s, region, bytes);
    forward_with(points, n, d, k, normalize, form, scores, nbr_idx, nbr_d2, kdist2, lrd, scratch, s);
}

}  // namespace detail

// ---------------------------------------------------------------------------
// C++ API (include/culof.h)
// ---------------------------------------------------------------------------

bool cuda_available() { return detail::device_present(); }

std::string device_info() { return detail::describe_device(); }

std::vector<float> lof(const float* points, int n_points, int n_dims, int k, bool normalize) {
    return detail::lof_host(points, n_points, n_dims, k, normalize);
}

}  // namespace culof

// ---------------------------------------------------------------------------
// C ABI (include/culof_c.h)
// ---------------------------------------------------------------------------

using namespace culof;
using namespace culof::detail;

extern "C" {

const char* culof_last_error(void) { return g_last_error.c_str(); }

int culof_device_available(void) { return device_present() ? 1 : 0; }

int culof_device_info(char* buf, size_t capacity) {
    return guarded([&] {
        if (!buf || capacity == 0) throw std::invalid_argument("buffer must not be empty");
        const std::string s = describe_device();
        const size_t len = std::min(capacity - 1, s.size());
        std::memcpy(buf, s.data(), len);
        buf[len] = '\0';
    });
}

int culof_lof(const float* points, int n_points, int n_dims, int k, int normalize,
              float* scores) {
    return guarded([&] {
        if (!scores) throw std::invalid_argument("scores must not be null");
        const std::vector<float> out = lof_host(points, n_points, n_dims, k, normalize != 0);
        std::memcpy(scores, out.data(), out.size() * sizeof(float));
    });
}

int culof_lof_device(const float* points, int n_points, int n_dims, int k, int normalize,
                     int form, float* scores, int* nbr_idx, float* nbr_d2, float* kdist2,
                     float* lrd, void* stream) {
    return guarded([&] {
        validate(points, n_points, n_dims, k);
        if (!scores) throw std::invalid_argument("scores must not be null");
        require_device();
        forward(points, n_points, n_dims, k, normalize != 0, form, scores, nbr_idx, nbr_d2,
                kdist2, lrd, as_stream(stream));
    });
}

int culof_lof_novelty_device(const float* points, int n_points, int n_dims, const float* queries,
                             int n_queries, int k, int normalize, int form, float* scores_fit,
                             float* scores_new, int* nbr_idx_new, float* nbr_d2_new, void* stream) {
    return guarded([&] {
        validate(points, n_points, n_dims, k);
        if (n_queries < 0) throw std::invalid_argument("n_queries must be >= 0");
        if (n_queries > 0 && (!queries || !scores_new))
            throw std::invalid_argument("queries and scores_new must not be null");
        require_device();
        novelty(points, n_points, n_dims, queries, n_queries, k, normalize != 0, form, scores_fit,
                scores_new, nbr_idx_new, nbr_d2_new, as_stream(stream));
    });
}

int culof_stage_center(const float* points, int n_points, int n_dims, int normalize,
                       float* centred, float* norms, void* stream) {
    return guarded([&] {
        validate(points, n_points, n_dims, 1);
        if (!centred || !norms) throw std::invalid_argument("outputs must not be null");
        require_device();
        const hipStream_t s = as_stream(stream);
        Scratch scratch(s);
        double* work = scratch.get<double>(static_cast<size_t>(kPrepWork) * n_dims);
        center_and_norms(points, n_points, n_dims, normalize != 0, centred, n_dims, norms, work,
                         nullptr, nullptr, s);
    });
}

int culof_stage_knn(const float* points, int n_points, int n_dims, int k, int normalize,
                    int form, int row0, int rows, int method, int* nbr_idx, float* nbr_d2,
                    float* kdist2, void* stream) {
    return guarded([&] {
        validate(points, n_points, n_dims, k);
        if (!nbr_idx || !nbr_d2 || !kdist2)
            throw std::invalid_argument("outputs must not be null");
        if (row0 < 0 || rows < 0 || row0 + rows > n_points)
            throw std::invalid_argument("row range outside [0, n_points)");
        const int screen = screen_form(form);
        if (screen < 0) throw std::invalid_argument("unknown precision form");
        if (method < 0 || method > 2) throw std::invalid_argument("method must be 0, 1 or 2");
        const int k_sel = (form & kFormTiesFp64) && k + 1 <= n_points - 1 ? k + 1 : k;
        if (method == 2 && (screen != kFormPlain || !sweep_supported(n_dims, k_sel)))
            throw std::invalid_argument("no VALU sweep instantiation for this (n_dims, k, form)");
        require_device();
        const hipStream_t s = as_stream(stream);
        Scratch scratch(s);
        const Prepared p = prepare(points, n_points, n_dims, normalize != 0, screen, scratch, s);
        knn_rows(p, n_points, n_dims, k, form, row0, rows, method, nbr_idx, nbr_d2, kdist2,
                 scratch, s);
    });
}

int culof_lof_backward_device(const float* points, int n_points, int n_dims, int k,
                              int normalize, const int* nbr_idx, const float* nbr_d2,
                              const float* kdist2, const float* grad_scores, float* grad_points,
                              void* stream) {
    return guarded([&] {
        validate(points, n_points, n_dims, k);
        if (!nbr_idx || !nbr_d2 || !kdist2 || !grad_scores || !grad_points)
            throw std::invalid_argument("inputs and outputs must not be null");
        if (n_dims > kMaxBackwardDims) throw std::invalid_argument("backward supports n_dims <= 256");
        require_device();
        const hipStream_t s = as_stream(stream);
        Scratch scratch(s);
        const float* keyrows = points;
        const double* inv_std = nullptr;
        const double* scaled = nullptr;
        if (normalize) {
            // The z-scored rows and the per-feature scale, exactly as the forward built them.
            float* z = scratch.get<float>(static_cast<size_t>(n_points) * n_dims);
            float* norms = scratch.get<float>(n_points);
            double* work = scratch.get<double>(static_cast<size_t>(kPrepWork) * n_dims);
            center_and_norms(points, n_points, n_dims, true, z, n_dims, norms, work, nullptr, nullptr, s);
            keyrows = z;
            inv_std = work + static_cast<size_t>(257) * n_dims;
            scaled = work + static_cast<size_t>(258) * n_dims;
        }
        void* ws = scratch.get<unsigned char>(backward_scratch_bytes(n_points, n_dims, k));
        lof_backward(keyrows, n_points, n_dims, k, nbr_idx, nbr_d2, kdist2, grad_scores, inv_std,
                     scaled, grad_points, ws, s);
    });
}

int culof_stage_select_rows(const float* keys, int rows, int n_cols, int k, int* nbr_idx,
                            float* nbr_d2, float* kdist2, void* stream) {
    return guarded([&] {
        if (!keys || !nbr_idx || !nbr_d2 || !kdist2)
            throw std::invalid_argument("inputs and outputs must not be null");
        if (rows < 1 || n_cols < 1) throw std::invalid_argument("rows and n_cols must be >= 1");
        if (k < 1 || k > n_cols) throw std::invalid_argument("k must satisfy 1 <= k <= n_cols");
        require_device();
        radix_select_rows(keys, rows, n_cols, k, nbr_idx, nbr_d2, kdist2, as_stream(stream));
    });
}

int culof_stage_lrd(const int* nbr_idx, const float* nbr_d2, const float* kdist2, int n_points,
                    int k, float* lrd, void* stream) {
    return guarded([&] {
        if (!nbr_idx || !nbr_d2 || !kdist2 || !lrd)
            throw std::invalid_argument("inputs and outputs must not be null");
        if (n_points < 1 || k < 1) throw std::invalid_argument("n_points and k must be >= 1");
        require_device();
        compute_lrd(nbr_idx, nbr_d2, kdist2, n_points, k, lrd, as_stream(stream));
    });
}

int culof_stage_scores(const int* nbr_idx, const float* lrd, int n_points, int k, float* scores,
                       void* stream) {
    return guarded([&] {
        if (!nbr_idx || !lrd || !scores)
            throw std::invalid_argument("inputs and outputs must not be null");
        if (n_points < 1 || k < 1) throw std::invalid_argument("n_points and k must be >= 1");
        require_device();
        compute_scores(nbr_idx, lrd, n_points, k, scores, as_stream(stream));
    });
}

int culof_choose_tile_rows(int n_points, size_t budget_bytes) {
    return detail::choose_tile_rows(n_points, budget_bytes);
}

int culof_profile_enable(int on) {
    return guarded([&] {
        if (on) {
            int device = 0;
            CULOF_CHECK(hipGetDevice(&device));
            (void)profile_tile_counter_for(device, true);
        }
        g_prof_on.store(on != 0);
    });
}

int culof_profile_reset(void) {
    return guarded([&] {
        std::lock_guard<std::mutex> lock(g_prof_mu);
        for (ProfRecord& r : g_prof) {
            (void)hipEventSynchronize(r.stop);
            (void)hipEventDestroy(r.start);
            (void)hipEventDestroy(r.stop);
        }
        g_prof.clear();
        int device = 0;
        if (hipGetDevice(&device) == hipSuccess) {
            if (unsigned long long* c = profile_tile_counter_for(device, false)) {
                CULOF_CHECK(hipDeviceSynchronize());
                CULOF_CHECK(hipMemset(c, 0, sizeof(unsigned long long)));
            }
        }
    });
}

int culof_profile_tile_pairs(long long* tile_pairs) {
    return guarded([&] {
        if (!tile_pairs) throw std::invalid_argument("null argument");
        int device = 0;
        CULOF_CHECK(hipGetDevice(&device));
        unsigned long long v = 0;
        if (unsigned long long* c = profile_tile_counter_for(device, false)) {
            CULOF_CHECK(hipDeviceSynchronize());
            CULOF_CHECK(hipMemcpy(&v, c, sizeof(v), hipMemcpyDeviceToHost));
        }
        *tile_pairs = static_cast<long long>(v);
    });
}

int culof_profile_read(const char* name, double* total_ms, long* launches) {
    return guarded([&] {
        if (!name || !total_ms || !launches) throw std::invalid_argument("null argument");
        std::lock_guard<std::mutex> lock(g_prof_mu);
        double sum = 0.0;
        long count = 0;
        for (ProfRecord& r : g_prof) {
            if (r.name.find(name) == std::string::npos) continue;
            CULOF_CHECK(hipEventSynchronize(r.stop));
            float ms = 0.0f;
            CULOF_CHECK(hipEventElapsedTime(&ms, r.start, r.stop));
            sum += ms;
            ++count;
        }
        *total_ms = sum;
        *launches = count;
    });
}

}  // extern "C"
