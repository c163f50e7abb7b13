// libshockhash_amd.so: kernels and C ABI of the MI355X ShockHash seed-search engine.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include <dlfcn.h>

#include "../../include/shockhash_amd.h"
#include "../include/shockhash/GolombMemo.h"
#include "build_common.h"
#include "search_sh2.h"
#include "orientation.h"
#include "ribbon.h"
#include "flat.h"
#include "schedule.h"
#include "sort.h"
#include "strings.h"

using namespace shgpu;

namespace {

// Per-depth start seeds, include/ShockHash.h:47-49.
const uint64_t START_SEED[20] = {
    0x106393c187cae21a, 0x6453cec3f7376937, 0x643e521ddbd2be98, 0x3740c6412f6572cb, 0x717d47562f1ce470,
    0x4cd6eb4c63befb7c, 0x9bfd8c5e18c8da73, 0x082f20e10092a9a3, 0x2ada2ce68d21defc, 0xe33cb4f3e7c6466b,
    0x3980be458c509c59, 0xc466fd9584828e8c, 0x45f0aabe1a61ede6, 0xf6e7b8b33ad9b98d, 0x4ef95e25f4b4983d,
    0x81175195173b92d3, 0x4e50927d8dd15978, 0x1ea2099d1fafae7f, 0x425c8a06fbaaa815, 0xcd4216006c74052a};

static_assert(SHGPU_TUNE_SH2_GENERAL_ONLY == SH2_FLAG_GENERAL_ONLY, "tuning flag of the ShockHash2 leaf search");

constexpr uint32_t MAX_ENCODABLE_BUCKET = 2999;  // memo tables cover sizes < MAX_BUCKET_SIZE = 3000
constexpr uint64_t DEFAULT_MAX_TRIALS = 1ull << 40;  // schedule cap of every node's seed bound

thread_local std::string g_err;

int fail(int code, const std::string &msg) {
    g_err = msg;
    return code;
}

#define HIP_TRY(expr)                                                                              \
    do {                                                                                           \
        hipError_t _e = (expr);                                                                    \
        if (_e != hipSuccess) return fail(SHGPU_ERR_HIP, std::string(#expr ": ") + hipGetErrorString(_e)); \
    } while (0)

struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
    int ensure(size_t want) {
        if (want <= bytes) return SHGPU_OK;
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
        if (!device_malloc(&p, want)) {
            p = nullptr;
            return fail(SHGPU_ERR_NOMEM, "hipMalloc of " + std::to_string(want) + " bytes failed");
        }
        bytes = want;
        return SHGPU_OK;
    }
    template <class T> T *as() const { return (T *)p; }
    ~DevBuf() { if (p) (void)hipFree(p); }
};

}  // namespace

// Device buffers of one construction in flight (the engine's own, or one serving slot's).
// Page-locked host memory (hipHostMalloc), grown on demand: device-to-host copies into it run at full link
// speed instead of through the runtime's staging buffers.
struct PinnedBuf {
    void *p = nullptr;
    size_t bytes = 0;
    int ensure(size_t want) {
        if (want <= bytes) return SHGPU_OK;
        if (p) (void)hipHostFree(p);
        p = nullptr;
        bytes = 0;
        if (!host_malloc_pinned(&p, want)) {
            p = nullptr;
            return fail(SHGPU_ERR_NOMEM, "hipHostMalloc of " + std::to_string(want) + " bytes failed");
        }
        bytes = want;
        return SHGPU_OK;
    }
    ~PinnedBuf() { if (p) (void)hipHostFree(p); }
};

struct BuildBuffers {
    DevBuf firsts_in, firsts_out, seconds_in, tmp, keys, sort_tmp, size_acc, nodes, codes, queue, ctrl, states, help,
        sh2_scratch, staging, plan_tab, plan_sets, plan_bucket, plan_summary, sort_flag, sh2_retry;
    // forest plan: node count tables of plan_variant / plan_leaf (host copy, then the device copy in plan_tab),
    // node capacity of nodes / codes / queue / states
    int plan_variant = -1;
    uint32_t plan_leaf = 0;
    std::vector<uint32_t> plan_counts;
    uint64_t plan_capacity = 0;
    PinnedBuf codes_host;  // codes on their way to the host
    PinnedBuf nodes_host;  // device-planned nodes on their way to the host (SHGPU_NODES_HOST)
    hipEvent_t ev0 = nullptr, ev1 = nullptr, evA = nullptr, evB = nullptr;
    size_t bytes() const {
        const DevBuf *all[] = {&firsts_in, &firsts_out, &seconds_in, &tmp, &keys, &sort_tmp, &size_acc, &nodes,
                               &codes, &queue, &ctrl, &states, &help, &sh2_scratch, &staging, &plan_tab,
                               &plan_sets, &plan_bucket, &plan_summary, &sort_flag, &sh2_retry};
        size_t t = 0;
        for (const DevBuf *d : all) t += d->bytes;
        return t;
    }
    void release() {
        DevBuf *all[] = {&firsts_in, &firsts_out, &seconds_in, &tmp, &keys, &sort_tmp, &size_acc, &nodes,
                         &codes, &queue, &ctrl, &states, &help, &sh2_scratch, &staging, &plan_tab,
                         &plan_sets, &plan_bucket, &plan_summary, &sort_flag, &sh2_retry};
        for (DevBuf *d : all) {
            if (d->p) (void)hipFree(d->p);
            d->p = nullptr;
            d->bytes = 0;
        }
        plan_variant = -1;
        plan_capacity = 0;
    }
    int create_events() {
        if (hipEventCreate(&ev0) != hipSuccess || hipEventCreate(&ev1) != hipSuccess || hipEventCreate(&evA) != hipSuccess ||
            hipEventCreate(&evB) != hipSuccess)
            return fail(SHGPU_ERR_HIP, "hipEventCreate failed");
        return SHGPU_OK;
    }
    void destroy_events() {
        if (ev0) (void)hipEventDestroy(ev0);
        if (ev1) (void)hipEventDestroy(ev1);
        if (evA) (void)hipEventDestroy(evA);
        if (evB) (void)hipEventDestroy(evB);
        ev0 = ev1 = evA = evB = nullptr;
    }
};

// Scratch of shgpu_split_search_noalloc, which allocates nothing: the partition's staging copy of up to
// SHGPU_NOALLOC_MAX_KEYS keys, then the work counter (16 words). Every GPU the library runs on holds its own copy
// (8 MiB) until the process ends; host code reaches it only through the address hipGetSymbolAddress returns for
// the current device (noalloc_scratch_of).
static __device__ __align__(256) uint64_t scratch[SHGPU_NOALLOC_MAX_KEYS + 8];

namespace {
std::mutex g_scratch_mu;
std::unordered_map<int, void *> g_scratch_ptr;  // device -> its copy of `scratch`

// The device address of `scratch` on `device` (current), looked up once per device.
int noalloc_scratch_of(int device, void **out) {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    auto it = g_scratch_ptr.find(device);
    if (it != g_scratch_ptr.end()) {
        *out = it->second;
        return SHGPU_OK;
    }
    void *ptr = nullptr;
    const hipError_t err = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (err != hipSuccess || !ptr)
        return fail(SHGPU_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(err));
    g_scratch_ptr[device] = ptr;
    *out = ptr;
    return SHGPU_OK;
}
}  // namespace

// Copies from pageable host memory of at least UPLOAD_STAGED_MIN bytes go through page-locked chunks: the runtime's
// own path for pageable memory moved ~29 GB/s (160 MB of fingerprints in 5.5 ms), the DMA engine reads page-locked
// memory at the link's rate. UPLOAD_THREADS host threads fill the chunks, two buffers each, while earlier chunks'
// DMAs run.
constexpr size_t UPLOAD_CHUNK = size_t(4) << 20;
constexpr int UPLOAD_THREADS = 8;
constexpr size_t UPLOAD_STAGED_MIN = size_t(16) << 20;

// The uploader's worker threads are started once per engine and wait for jobs; worker t copies chunks t,
// t + UPLOAD_THREADS, ... of a job into its two page-locked buffers in turn, waiting for the event behind a buffer's
// previous DMA before refilling it, and queues each chunk's DMA and a new event on the engine's stream (the runtime
// orders the threads' operations on the one stream).
struct Uploader {
    void *pin[UPLOAD_THREADS][2] = {};
    hipEvent_t ev[UPLOAD_THREADS][2] = {};
    bool ready = false, failed = false;
    std::vector<std::thread> workers;
    std::mutex mu;
    std::condition_variable cv_job, cv_done;
    uint64_t gen = 0;  // job number
    int done = 0;      // workers finished with the current job
    bool stop = false;
    int device = 0;
    hipStream_t stream = nullptr;
    const char *src = nullptr;
    char *dst = nullptr;
    size_t bytes = 0;
    std::atomic<int> err{0};
    std::atomic<int> setup_failed{0};  // a worker could not select the engine's device

    // Allocates the buffers and events and starts the workers on first use (true when they are there).
    bool init(int dev, hipStream_t st) {
        if (ready || failed) return ready;
        device = dev;
        stream = st;
        for (int t = 0; t < UPLOAD_THREADS; t++)
            for (int k = 0; k < 2; k++) {
                if (hipHostMalloc(&pin[t][k], UPLOAD_CHUNK, hipHostMallocDefault) != hipSuccess) pin[t][k] = nullptr;
                if (hipEventCreateWithFlags(&ev[t][k], hipEventDisableTiming) != hipSuccess) ev[t][k] = nullptr;
                // recorded once, so that the first wait on every buffer returns at once
                if (!pin[t][k] || !ev[t][k] || hipEventRecord(ev[t][k], st) != hipSuccess) failed = true;
            }
        if (failed) {
            (void)hipGetLastError();
            release();
            failed = true;
            return false;
        }
        try {
            for (int t = 0; t < UPLOAD_THREADS; t++) workers.emplace_back([this, t] { work(t); });
        } catch (...) {  // no thread: copies take the runtime's own path
            release();
            failed = true;
            return false;
        }
        ready = true;
        return true;
    }

    void work(int t) {
        if (hipSetDevice(device) != hipSuccess) setup_failed = 1;
        uint64_t seen = 0;
        for (;;) {
            {
                std::unique_lock<std::mutex> lk(mu);
                cv_job.wait(lk, [&] { return stop || gen != seen; });
                if (stop) return;
                seen = gen;
            }
            const size_t nchunks = setup_failed.load() ? 0 : (bytes + UPLOAD_CHUNK - 1) / UPLOAD_CHUNK;
            size_t i = 0;
            for (size_t c = (size_t)t; c < nchunks && !err.load(); c += UPLOAD_THREADS, i++) {
                const int k = (int)(i & 1);
                const size_t off = c * UPLOAD_CHUNK, len = std::min(UPLOAD_CHUNK, bytes - off);
                if (hipEventSynchronize(ev[t][k]) != hipSuccess) { err = 1; break; }
                std::memcpy(pin[t][k], src + off, len);
                if (hipMemcpyAsync(dst + off, pin[t][k], len, hipMemcpyHostToDevice, stream) != hipSuccess ||
                    hipEventRecord(ev[t][k], stream) != hipSuccess) {
                    err = 1;
                    break;
                }
            }
            std::lock_guard<std::mutex> lk(mu);
            if (++done == UPLOAD_THREADS) cv_done.notify_one();
        }
    }

    // Queues the copy; returns once every chunk has been copied out of `s` and its DMA queued (0), or 1 on failure.
    int copy(void *d, const void *s, size_t n) {
        std::unique_lock<std::mutex> lk(mu);
        src = (const char *)s;
        dst = (char *)d;
        bytes = n;
        err = 0;
        done = 0;
        gen++;
        cv_job.notify_all();
        cv_done.wait(lk, [&] { return done == UPLOAD_THREADS; });
        return err.load() || setup_failed.load();
    }

    void release() {
        {
            std::lock_guard<std::mutex> lk(mu);
            stop = true;
        }
        cv_job.notify_all();
        for (auto &w : workers) w.join();
        workers.clear();
        for (int t = 0; t < UPLOAD_THREADS; t++)
            for (int k = 0; k < 2; k++) {
                if (ev[t][k]) (void)hipEventSynchronize(ev[t][k]);
                if (pin[t][k]) (void)hipHostFree(pin[t][k]);
                if (ev[t][k]) (void)hipEventDestroy(ev[t][k]);
                pin[t][k] = nullptr;
                ev[t][k] = nullptr;
            }
        ready = false;
    }
};

struct shgpu_engine {
    int device = 0;
    void *noalloc_scratch = nullptr;  // this device's copy of `scratch` (device address)
    hipStream_t stream = nullptr;
    int cu_count = 0;
    int clock_khz = 0;
    char arch[64] = {0};
    // Schedule per variant (ShockHash, SIMDShockHash, ShockHash2). waves_per_cu 0 (the default) takes waves per
    // CU and seeds per lane from the size class of `sched` for the batch's key count.
    shgpu_tuning tuning_v[3] = {{4, 2, 0, 0, DEFAULT_MAX_TRIALS}, {4, 2, 0, 0, DEFAULT_MAX_TRIALS},
                                {4, 2, 0, 0, DEFAULT_MAX_TRIALS}};
    shgpu::Schedule sched = shgpu::Schedule::builtin();
    shgpu_tuning tuning{4, 2, 0, 0, DEFAULT_MAX_TRIALS};  // batched per-node entry points (waves 0: 24 per CU)
    DevBuf counter, sh2_scratch, staging, tmp;  // batched per-node entry points
    BuildBuffers main;                          // shgpu_build_run
    DevBuf retrieval;                           // shgpu_ribbon_build working memory (~42 B per key)
    DevBuf flat;                                // ShockHash2Flat bucketing working memory (~70 B per key)
    Uploader up;                                // page-locked chunks of large uploads from pageable memory
};

struct shgpu_build {
    uint64_t n = 0, nbuckets = 0;
    std::vector<uint64_t> size_acc;
    std::vector<uint64_t> codes;  // one per node, DFS pre-order per bucket, buckets in order
    std::vector<uint64_t> keys;   // filled when desc->copy_keys
    // sort+bucketing, forest plan, seed search and copies, search kernel; the persistent form times the first
    // two and the kernel with HIP events and the third as the rest of the call's wall time
    double ms[4] = {0, 0, 0, 0};
    // The nodes follow from the bucket sizes (of the first plan_buckets buckets); a persistent build plans
    // them on the device, and the host copy is made on first request.
    int variant = 0;
    uint32_t leaf = 0;
    uint64_t max_trials = 0;
    uint64_t plan_buckets = 0;
    mutable std::mutex nodes_mu;
    mutable bool nodes_ready = false;
    mutable std::vector<shgpu_node> nodes;
    // SHGPU_KEYS_DEVICE: the leaf-ordered keys in device memory of its own (freed with the build)
    void *d_keys = nullptr;
    int device = 0;
    shgpu_build() = default;
    shgpu_build(const shgpu_build &) = delete;
    shgpu_build &operator=(const shgpu_build &) = delete;
    ~shgpu_build() {  // frees on the build's device and leaves the caller's current device as it was
        if (!d_keys) return;
        int cur = device;
        (void)hipGetDevice(&cur);
        if (cur != device && hipSetDevice(device) != hipSuccess) return;
        (void)hipFree(d_keys);
        if (cur != device) (void)hipSetDevice(cur);
    }
};

// ===================================================================================== kernels

__global__ void k_fill_u32(uint32_t *p, uint32_t v, uint32_t n) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = v;
}

// Plain streaming copy (bandwidth / counter calibration probe).
__global__ void k_copy_u64(const uint64_t *__restrict__ src, uint64_t *__restrict__ dst, uint64_t n) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

__device__ __forceinline__ uint64_t rotr64(uint64_t x, uint32_t r) { return r ? (x >> r) | (x << (64 - r)) : x; }
__device__ __forceinline__ uint64_t rotl64(uint64_t x, uint32_t r) { return r ? (x << r) | (x >> (64 - r)) : x; }

// Splits the fingerprints into .first rotated right by `rot` (so the top 64 - rot bits, the ones sorted on,
// are the bottom ones) and .second.
__global__ void k_unzip_hashes(const uint64_t *__restrict__ h, uint64_t n, uint64_t *first, uint64_t *second, uint32_t rot) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        first[i] = rotr64(h[2 * i], rot);
        second[i] = h[2 * i + 1];
    }
}

// The sort orders rotated keys r = rotr(first, rot) on their bottom 64 - rot bits (the top bits of first).
// Keys that agree on those bits then form runs in input order, and inside a run comparing r compares the
// remaining bits of first, so a stable insertion sort of every run on r gives the order of a stable sort of
// first on all 64 bits. With rot chosen so that the expected run length is far below one, the runs are
// almost all single keys; a run longer than FIXUP_MAX_RUN (crafted or repeated keys) sets *flag and the
// caller sorts on all 64 bits instead.
constexpr uint32_t FIXUP_MAX_RUN = 64;
__global__ void k_sort_fixup(uint64_t *first, uint64_t *second, uint64_t n, uint32_t rot, uint32_t *flag) {
    const uint64_t mask = (~0ull) >> rot;  // the sorted bits of r
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i + 1 < n; i += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t hi = first[i] & mask;
        if ((first[i + 1] & mask) != hi) continue;           // no run of two or more starts here
        if (i > 0 && (first[i - 1] & mask) == hi) continue;  // inside a run
        uint64_t end = i + 2;
        while (end < n && (first[end] & mask) == hi && end - i <= FIXUP_MAX_RUN) end++;
        if (end - i > FIXUP_MAX_RUN) {
            atomicOr(flag, 1u);
            continue;
        }
        for (uint64_t j = i + 1; j < end; j++) {
            const uint64_t f = first[j], v = second[j];
            uint64_t k = j;
            for (; k > i && first[k - 1] > f; k--) {
                first[k] = first[k - 1];
                second[k] = second[k - 1];
            }
            first[k] = f;
            second[k] = v;
        }
    }
}

// Top key bits a build sorts on: ceil(log2 n) + 6 rounded up to whole 8-bit radix digits, so few keys share
// their sorted bits (at most 1 in 64 on average) and the fix-up has little to do; 64 when full_sort.
// (Measured at 10M keys: 24 bits save a radix pass, 0.1 ms, but put ~45% of keys into runs and the fix-up
// then takes 0.21 ms instead of 0.04 ms.)
static int sort_bits(uint64_t n, bool full_sort) {
    if (full_sort) return 64;
    int lg = 1;
    while (lg < 64 && (1ull << lg) < n) lg++;
    return std::min(64, (lg + 6 + 7) / 8 * 8);
}

// size_acc[b] = number of keys in buckets < b, from firsts (rotated right by rot) sorted ascending (remap128
// is monotone).
__global__ void k_bucket_bounds(const uint64_t *__restrict__ first, uint64_t n, uint64_t nb, uint64_t *size_acc, uint32_t rot) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i <= n; i += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t b = i < n ? __umul64hi(rotl64(first[i], rot), nb) : nb;
        const uint64_t bp = i > 0 ? __umul64hi(rotl64(first[i - 1], rot), nb) : 0;
        const uint64_t lo = i > 0 ? bp + 1 : 0;
        for (uint64_t k = lo; k <= b && k <= nb; k++) size_acc[k] = i;
    }
}

__global__ void k_init_states(NodeState *st, uint64_t *codes, uint32_t n, uint32_t *help, uint32_t nslots) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        st[i].gd = 0;
        st[i].best = SHGPU_NOT_FOUND;
        st[i].finalized = 0;
        codes[i] = SHGPU_NOT_FOUND;  // a node that is never finalised reads as not found
    }
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < nslots; i += gridDim.x * blockDim.x) help[i] = EMPTY_SLOT;
}

// ShockHash2 leaves above 64 keys that outgrew their tables (code SHGPU_NOT_FOUND after the build) as leaf
// tasks, with their node ids; count[0] receives how many.
__global__ void k_sh2_collect_overflow(const shgpu_node *__restrict__ nodes, const uint64_t *__restrict__ codes, uint32_t total,
                                       shgpu_leaf_task *tasks, uint32_t *ids, uint32_t *count) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < total; i += gridDim.x * blockDim.x) {
        const shgpu_node nd = nodes[i];
        if (nd.kind != SHGPU_NODE_LEAF || nd.m <= 64 || codes[i] != SHGPU_NOT_FOUND) continue;
        const uint32_t k = atomicAdd(count, 1u);  // k < total: each node is collected once
        tasks[k] = shgpu_leaf_task{nd.offset, nd.m, 0, 0};
        ids[k] = i;
    }
}

__global__ void k_scatter_codes(const uint64_t *__restrict__ src, const uint32_t *__restrict__ ids, uint32_t n, uint64_t *codes) {
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < n; t += gridDim.x * blockDim.x) codes[ids[t]] = src[t];
}

// ------------------------------------------------------------------ forest plan on the device
// The forest of a persistent build is planned from the bucket sizes where they are computed, so the
// construction runs from the sort to the codes without a host round trip. A bucket's tree depends only on
// its size: the host tabulates the node count of a tree of every size once per (variant, leaf size);
// k_plan_scan turns the bucket sizes into each tree's first node index and its root's queue slot
// (prefix sums over buckets), k_plan_emit writes every tree's nodes in DFS pre-order (the order of
// emit_subtree on the host, which recSplit's recursion fixes: ShockHash.h:303-555).

// One key set of a batch: its bucket prefix sums are size_acc[bbase .. bbase + nbk] (nbk buckets and a
// terminator), its keys start at key_base in the concatenated key array.
struct PlanSet {
    uint64_t key_base;
    uint32_t bbase, nbk;
};

// summary words written by k_plan_scan
constexpr uint32_t SUM_TOTAL = 0, SUM_SPLIT_NODES = 1, SUM_SPLIT_ROOTS = 2, SUM_LEAF_ROOTS = 3, SUM_OVERFLOW = 4,
                   SUM_QSPLIT = 5, SUM_SETS = 8;  // then per set: first node, end node, bucket too large
constexpr uint32_t PLAN_THREADS = 1024, PLAN_STACK = 64;

struct PlanArgs {
    const uint64_t *size_acc;
    const PlanSet *sets;
    uint32_t nsets;
    uint32_t nacc;                   // entries of size_acc over all sets
    const uint32_t *cnt_all;         // nodes of a tree of m keys, m <= MAX_ENCODABLE_BUCKET
    const uint32_t *cnt_split;       // non-leaf nodes of that tree
    const uint32_t *bound_lg;        // log2 of the default seed bound of a node of m keys (seed_bound_log2)
    const float *seed_rate;          // -ln(1 - p) for the probability p that a seed suits a node of m keys
    uint32_t win_leaf, win_lower;    // seeds per window of leaves and of lower aggregation nodes
    uint4 *bucket;                   // per size_acc entry: first node, split roots before, leaf roots before
    uint32_t *summary;
    uint32_t capacity;               // node capacity of nodes / codes / queue / states
    int two_part;                    // split part and leaf part (SHGPU_TUNE_SPLITS_FIRST)
    uint32_t *ctrl;
    uint32_t *help;
    uint32_t nhelp;
    shgpu_node *nodes;
    uint64_t *codes;
    NodeState *state;
    uint32_t *queue;
    int variant;
    uint32_t leaf, lower, upper;
    uint64_t max_trials;
    uint64_t start_seed[20];
};

__device__ __forceinline__ uint32_t plan_set_of(const PlanArgs &p, uint32_t g) {
    uint32_t lo = 0, hi = p.nsets - 1;
    while (lo < hi) {
        const uint32_t mid = (lo + hi + 1) / 2;
        if (p.sets[mid].bbase <= g) lo = mid; else hi = mid - 1;
    }
    return lo;
}

// Keys in the bucket at size_acc entry g of set r (0 for the set's terminator entry).
__device__ __forceinline__ uint32_t plan_bucket_keys(const PlanArgs &p, uint32_t r, uint32_t g) {
    const PlanSet s = p.sets[r];
    if (g >= s.bbase + s.nbk) return 0;
    return (uint32_t)(p.size_acc[g + 1] - p.size_acc[g]);
}

// Inclusive scan of four counters over the block (PLAN_THREADS threads); returns the exclusive prefix and
// the block totals.
__device__ __forceinline__ uint4 plan_block_scan(uint4 v, uint4 &total) {
    __shared__ uint4 wsum[PLAN_THREADS / 64];
    const uint32_t lane = lane_id(), w = threadIdx.x / 64;
    uint4 inc = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t x = __shfl_up(inc.x, d), y = __shfl_up(inc.y, d), z = __shfl_up(inc.z, d), q = __shfl_up(inc.w, d);
        if ((int)lane >= d) { inc.x += x; inc.y += y; inc.z += z; inc.w += q; }
    }
    if (lane == 63) wsum[w] = inc;
    __syncthreads();
    uint4 before = make_uint4(0, 0, 0, 0);
    total = make_uint4(0, 0, 0, 0);
    for (uint32_t k = 0; k < PLAN_THREADS / 64; k++) {
        const uint4 s = wsum[k];
        if (k < w) { before.x += s.x; before.y += s.y; before.z += s.z; before.w += s.w; }
        total.x += s.x; total.y += s.y; total.z += s.z; total.w += s.w;
    }
    return make_uint4(before.x + inc.x - v.x, before.y + inc.y - v.y, before.z + inc.z - v.z, before.w + inc.w - v.w);
}

// One block. Per bucket: (nodes, split nodes, root is a split node, root is a leaf), zero for buckets of
// fewer than 2 keys and for every bucket of a set holding a bucket above MAX_ENCODABLE_BUCKET (that set is
// not searched). Writes the prefix sums per bucket, the summary, the queue control words and the help
// registry; a plan larger than the capacity is reported and planned empty.
__global__ void __launch_bounds__(PLAN_THREADS) k_plan_scan(PlanArgs p) {
    const uint32_t t = threadIdx.x;
    uint32_t *set_out = p.summary + SUM_SETS;
    for (uint32_t r = t; r < p.nsets; r += PLAN_THREADS) set_out[3 * r] = set_out[3 * r + 1] = set_out[3 * r + 2] = 0;
    __syncthreads();
    const uint32_t chunk = (p.nacc + PLAN_THREADS - 1) / PLAN_THREADS;
    const uint32_t c0 = min(p.nacc, t * chunk), c1 = min(p.nacc, c0 + chunk);
    const uint32_t r0 = c0 < c1 ? plan_set_of(p, c0) : 0;
    for (uint32_t g = c0, r = r0; g < c1; g++) {
        while (g > p.sets[r].bbase + p.sets[r].nbk) r++;
        if (plan_bucket_keys(p, r, g) > MAX_ENCODABLE_BUCKET) atomicOr(&set_out[3 * r + 2], 1u);
    }
    __syncthreads();
    auto values = [&](uint32_t r, uint32_t g) {
        const uint32_t m = plan_bucket_keys(p, r, g);
        if (m < 2 || __hip_atomic_load(&set_out[3 * r + 2], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) return make_uint4(0, 0, 0, 0);
        const bool leaf_root = m <= p.leaf;
        return make_uint4(p.cnt_all[m], p.cnt_split[m], leaf_root ? 0u : 1u, leaf_root ? 1u : 0u);
    };
    uint4 sum = make_uint4(0, 0, 0, 0);
    for (uint32_t g = c0, r = r0; g < c1; g++) {
        while (g > p.sets[r].bbase + p.sets[r].nbk) r++;
        const uint4 v = values(r, g);
        sum.x += v.x; sum.y += v.y; sum.z += v.z; sum.w += v.w;
    }
    uint4 total;
    uint4 run = plan_block_scan(sum, total);
    for (uint32_t g = c0, r = r0; g < c1; g++) {
        while (g > p.sets[r].bbase + p.sets[r].nbk) r++;
        p.bucket[g] = make_uint4(run.x, run.z, run.w, 0);
        if (g == p.sets[r].bbase) set_out[3 * r] = run.x;
        if (g == p.sets[r].bbase + p.sets[r].nbk) set_out[3 * r + 1] = run.x;
        const uint4 v = values(r, g);
        run.x += v.x; run.y += v.y; run.z += v.z; run.w += v.w;
    }
    const bool overflow = total.x > p.capacity;
    const uint32_t n_total = overflow ? 0u : total.x;
    // queue: one FIFO part of all nodes, or a split part [0, qsplit) and a leaf part [qsplit, total)
    const uint32_t qsplit = p.two_part ? total.y : n_total;
    if (t == 0) {
        p.summary[SUM_TOTAL] = total.x;
        p.summary[SUM_SPLIT_NODES] = total.y;
        p.summary[SUM_SPLIT_ROOTS] = total.z;
        p.summary[SUM_LEAF_ROOTS] = total.w;
        p.summary[SUM_OVERFLOW] = overflow ? 1u : 0u;
        p.summary[SUM_QSPLIT] = qsplit;
        p.ctrl[C_HEAD] = 0;
        p.ctrl[C_SPLIT_TAIL] = p.two_part ? total.z : total.z + total.w;
        p.ctrl[C_ERROR] = 0;
        p.ctrl[C_FINISHED] = 0;
        p.ctrl[C_LEAF_TAIL] = qsplit + (p.two_part ? total.w : 0u);
        p.ctrl[C_NODES] = n_total;
        p.ctrl[C_WINDOWED] = p.variant == SHGPU_SHOCKHASH2 ? (overflow ? 0u : total.y) : n_total;
        p.ctrl[C_STARTED] = 0;
    }
    for (uint32_t i = t; i < p.nhelp; i += PLAN_THREADS) p.help[i] = EMPTY_SLOT;
}

// The node of m keys at `offset`, depth `level`, the way emit_subtree builds it; its children (sizes >= 2)
// go to (child_off, child_m), and their number is returned.
__device__ __forceinline__ uint32_t plan_node(const PlanArgs &p, uint32_t offset, uint32_t m, uint32_t level,
                                              shgpu_node &nd, uint32_t (&child_off)[4], uint32_t (&child_m)[4]) {
    nd.offset = offset;
    nd.m = (uint16_t)m;
    nd.max_trials = min(p.max_trials, 1ull << p.bound_lg[m]);
    nd.subtree = p.cnt_all[m];
    nd.unit = 0;
    nd.n_children = 0;
    if (m <= p.leaf) {
        nd.kind = SHGPU_NODE_LEAF;
        nd.start_seed = p.variant == SHGPU_SHOCKHASH2 ? 0 : p.variant == SHGPU_SIMD_SHOCKHASH ? p.start_seed[19] : p.start_seed[level];
        return 0;
    }
    uint32_t unit;
    if (m > p.upper) {
        nd.kind = SHGPU_NODE_SPLIT2;
        unit = (uint16_t(m / 2 + p.upper - 1) / p.upper) * p.upper;
        nd.start_seed = p.start_seed[level];
    } else if (m > p.lower) {
        nd.kind = SHGPU_NODE_AGGR_UPPER;
        unit = p.lower;
        nd.start_seed = p.variant == SHGPU_SIMD_SHOCKHASH ? p.start_seed[17] : p.start_seed[level];
    } else {
        nd.kind = SHGPU_NODE_AGGR_LOWER;
        unit = p.leaf;
        nd.start_seed = p.variant == SHGPU_SIMD_SHOCKHASH ? p.start_seed[18] : p.start_seed[level];
    }
    nd.unit = unit;
    uint32_t n = 0, i = 0;
    if (nd.kind == SHGPU_NODE_SPLIT2) {
        child_off[n] = offset;
        child_m[n++] = unit;
        i = unit;
    } else {
        for (; i < m - unit && n < 3; i += unit) {
            child_off[n] = offset + i;
            child_m[n++] = unit;
        }
    }
    if (m - i > 1) {
        child_off[n] = offset + i;
        child_m[n++] = m - i;
    }
    nd.n_children = (uint8_t)n;
    return n;
}

// Waves that may help a node's owner at once (cap), and the completed windows after which any number may
// (all). With q the chance that one window of W seeds holds a suitable seed, the i-th of several windows in
// flight is needed with probability (1 - q)^(i - 1); helpers join only while that is at least HELP_USEFUL for
// the window each would take, so nodes that one window most likely settles (a root split of 2000 keys:
// q ~ 0.7) are not swarmed by waves searching windows past the winner, while long searches (leaves, lower
// aggregations) take many helpers. A search still running after HELP_ALL_AFTER / q windows (a chance of
// e^-8 for a search that can succeed; one over keys that repeat never does) takes every idle wave.
constexpr float HELP_USEFUL = 0.5f, HELP_ALL_AFTER = 8.f;
__device__ __forceinline__ void plan_help_limits(const PlanArgs &p, const shgpu_node &nd, uint32_t &cap, uint32_t &all) {
    const uint32_t W = nd.kind == SHGPU_NODE_LEAF ? p.win_leaf : nd.kind == SHGPU_NODE_AGGR_LOWER ? p.win_lower : 64u;
    const float per_window = p.seed_rate[nd.m] * (float)W;  // -ln(1 - q)
    cap = 0xFFFFFFFFu;
    all = 0;
    if (!(per_window > 0.f)) return;
    const float extra = -__logf(HELP_USEFUL) / per_window;
    const float q = -expm1f(-per_window);
    const float after = HELP_ALL_AFTER / q;
    cap = extra >= 1e6f ? 1000000u : (uint32_t)extra;
    all = after >= 2e9f ? 2000000000u : (uint32_t)after;
}

// One thread per bucket: the bucket's tree in DFS pre-order from its first node index, the node states,
// and the queue slots of its node range (its root's slot gets the root, other slots EMPTY_SLOT; every
// slot of [0, total) belongs to exactly one bucket's range, and the root slots are exactly those the
// tails of k_plan_scan cover).
__global__ void __launch_bounds__(256) k_plan_emit(PlanArgs p) {
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= p.nacc || p.summary[SUM_OVERFLOW]) return;
    const uint32_t r = plan_set_of(p, g);
    const uint32_t m = plan_bucket_keys(p, r, g);
    if (m < 2 || p.summary[SUM_SETS + 3 * r + 2]) return;
    const uint4 b = p.bucket[g];
    const uint32_t qsplit = p.summary[SUM_QSPLIT];
    const uint32_t split_roots = p.summary[SUM_SPLIT_ROOTS], leaf_roots = p.summary[SUM_LEAF_ROOTS];
    const bool leaf_root = m <= p.leaf;
    const uint32_t root_slot = p.two_part ? (leaf_root ? qsplit + b.z : b.y) : b.y + b.z;
    const uint32_t key_base = (uint32_t)p.sets[r].key_base;
    uint32_t st_off[PLAN_STACK], st_m[PLAN_STACK], st_level[PLAN_STACK];
    uint32_t sp = 0, idx = b.x;
    st_off[sp] = key_base + (uint32_t)p.size_acc[g];
    st_m[sp] = m;
    st_level[sp++] = 0;
    while (sp) {
        --sp;
        const uint32_t off = st_off[sp], mm = st_m[sp], level = st_level[sp];
        shgpu_node nd;
        uint32_t co[4], cm[4];
        const uint32_t nc = plan_node(p, off, mm, level, nd, co, cm);
        p.nodes[idx] = nd;
        p.codes[idx] = SHGPU_NOT_FOUND;  // a node that is never finalised reads as not found
        NodeState s;
        s.gd = 0;
        s.best = SHGPU_NOT_FOUND;
        s.finalized = 0;
        plan_help_limits(p, nd, s.help_cap, s.help_all);
        s.helpers = 0;
        p.state[idx] = s;
        const bool root_region = p.two_part ? (idx < split_roots || (idx >= qsplit && idx < qsplit + leaf_roots))
                                                : idx < split_roots + leaf_roots;
        if (!root_region) p.queue[idx] = EMPTY_SLOT;
        idx++;
        for (uint32_t c = nc; c-- > 0;) {  // reversed, so the first child is emitted next
            if (sp == PLAN_STACK) { atomicExch(&p.ctrl[C_ERROR], 1u); return; }
            st_off[sp] = co[c];
            st_m[sp] = cm[c];
            st_level[sp++] = level + 1;
        }
    }
    p.queue[root_slot] = b.x;
}

// ===================================================================================== host helpers

namespace {

struct Strategy {
    int variant;
    uint32_t leaf, lower, upper;
    Strategy(int v, uint32_t L) : variant(v), leaf(L) {
        if (v == SHGPU_SHOCKHASH2) { lower = 4 * L; upper = 3 * lower; }                    // ShockHash2.h:49-56
        else { lower = L * (L > 24 ? 4 : 2); upper = lower * (L > 24 ? 3 : 2); }         // ShockHash.h:63-72
    }
};

// Default seed bound of a node of m keys: 2^(k(m) + SEED_BOUND_SLACK) seeds, where k(m) is the Golomb-Rice
// parameter of the node's code (include/ShockHash.h:53-61,79-127; include/ShockHash2.h:33-47,63-115). Codes
// are geometric with mean about 2^k(m), so a search that runs 2^24 times its mean has failed for good (its
// keys repeat: three equal keys in one ShockHash leaf never form a pseudoforest, and the reference then loops
// forever, include/SIMDShockHash.hpp:131,171). The schedule's max_trials caps it further. Returned as log2
// per node size, evaluated with libm at run time (a bound, not part of the stored format).
constexpr uint32_t SEED_BOUND_SLACK = 24;

// Estimated probability that a seed splits a node of m > leaf keys into the strategy's parts (the multinomial
// estimate of include/ShockHash.h:79-127, include/ShockHash2.h:63-115).
double split_success_probability(const Strategy &st, uint32_t m) {
    uint32_t fanout, unit;
    if (m > st.upper) { unit = st.upper * (uint16_t(m / 2 + st.upper - 1) / st.upper); fanout = 2; }
    else if (m > st.lower) { unit = st.lower; fanout = uint16_t(m + st.lower - 1) / st.lower; }
    else { unit = st.leaf; fanout = uint16_t(m + st.leaf - 1) / st.leaf; }
    double sqrt_prod = std::sqrt((double)(m - (fanout - 1) * unit));
    for (uint32_t i = 0; i + 1 < fanout; i++) sqrt_prod *= std::sqrt((double)unit);
    return std::sqrt((double)m) / (std::pow(2 * M_PI, (fanout - 1.) / 2) * sqrt_prod);
}
std::vector<uint32_t> seed_bound_log2(const Strategy &st) {
    std::vector<uint32_t> lg(MAX_ENCODABLE_BUCKET + 1, 0);
    for (uint32_t m = 0; m <= MAX_ENCODABLE_BUCKET; m++) {
        uint32_t k;
        if (m <= st.leaf) {
            k = st.variant == SHGPU_SHOCKHASH2 ? shockhash::bij_memo2[std::min<uint32_t>(m, shockhash::MAX_LEAF_SIZE2)]
                                               : shockhash::bij_memo[std::min<uint32_t>(m, shockhash::MAX_LEAF_SIZE)];
        } else {
            const double p = split_success_probability(st, m);
            const double len = std::ceil(std::log2(-std::log((std::sqrt(5) + 1) / 2) / std::log1p(-p)));
            k = len > 0 ? (uint32_t)std::min(len, 40.0) : 0u;
        }
        lg[m] = std::min<uint32_t>(63, k + SEED_BOUND_SLACK);
    }
    return lg;
}

// Per node size m: -ln(1 - p) as the bits of a float, p the probability that a seed suits a node of m keys:
// a split node's from the multinomial estimate its Golomb parameter comes from, a leaf's from that parameter k
// (codes are about geometric with mean 2^k / ln 2). It sizes how many waves help a node's search (plan_help_cap).
std::vector<uint32_t> seed_rates(const Strategy &st) {
    std::vector<uint32_t> out(MAX_ENCODABLE_BUCKET + 1, 0);
    for (uint32_t m = 2; m <= MAX_ENCODABLE_BUCKET; m++) {
        double p;
        if (m <= st.leaf) {
            const uint32_t k = st.variant == SHGPU_SHOCKHASH2 ? shockhash::bij_memo2[std::min<uint32_t>(m, shockhash::MAX_LEAF_SIZE2)]
                                                               : shockhash::bij_memo[std::min<uint32_t>(m, shockhash::MAX_LEAF_SIZE)];
            p = std::min(1.0, std::log(2.0) / std::ldexp(1.0, (int)k));
        } else {
            p = split_success_probability(st, m);
        }
        const float r = p >= 1.0 ? 1e30f : (float)-std::log1p(-p);
        std::memcpy(&out[m], &r, 4);
    }
    return out;
}

// Per node size: min(cap, 2^seed_bound_log2), the max_trials of every node of that size.
std::vector<uint64_t> node_bounds(const Strategy &st, uint64_t cap) {
    const std::vector<uint32_t> lg = seed_bound_log2(st);
    std::vector<uint64_t> b(lg.size());
    for (size_t m = 0; m < lg.size(); m++) b[m] = std::min<uint64_t>(cap, 1ull << lg[m]);
    return b;
}

// Emits the subtree of a node of m keys at `offset` in DFS pre-order (recSplit order,
// ShockHash.h:303-555 / SIMDShockHash.hpp:563-617 / ShockHash2.h:259-394). bounds[m]: the max_trials of a
// node of m keys.
void emit_subtree(std::vector<shgpu_node> &out, const Strategy &st, uint32_t offset, uint32_t m, int level,
                  const std::vector<uint64_t> &bounds) {
    const size_t me = out.size();
    out.push_back(shgpu_node{});
    shgpu_node nd{};
    nd.offset = offset;
    nd.m = (uint16_t)m;
    nd.max_trials = bounds[m];
    if (m <= st.leaf) {
        nd.kind = SHGPU_NODE_LEAF;
        nd.start_seed = st.variant == SHGPU_SHOCKHASH2 ? 0
                        : st.variant == SHGPU_SIMD_SHOCKHASH ? START_SEED[19] : START_SEED[level];
        nd.subtree = 1;
        out[me] = nd;
        return;
    }
    uint32_t unit;
    if (m > st.upper) {
        nd.kind = SHGPU_NODE_SPLIT2;
        unit = (uint16_t(m / 2 + st.upper - 1) / st.upper) * st.upper;
        nd.start_seed = START_SEED[level];
    } else if (m > st.lower) {
        nd.kind = SHGPU_NODE_AGGR_UPPER;
        unit = st.lower;
        nd.start_seed = st.variant == SHGPU_SIMD_SHOCKHASH ? START_SEED[17] : START_SEED[level];
    } else {
        nd.kind = SHGPU_NODE_AGGR_LOWER;
        unit = st.leaf;
        nd.start_seed = st.variant == SHGPU_SIMD_SHOCKHASH ? START_SEED[18] : START_SEED[level];
    }
    nd.unit = unit;
    uint32_t nchild = 0, i = 0;
    if (nd.kind == SHGPU_NODE_SPLIT2) {  // ShockHash.h:479-480
        emit_subtree(out, st, offset, unit, level + 1, bounds);
        nchild++;
        i = unit;
    } else {  // ShockHash.h:510-514,543-547
        for (; i < m - unit; i += unit) {
            emit_subtree(out, st, offset + i, unit, level + 1, bounds);
            nchild++;
        }
    }
    if (m - i > 1) {  // the last part only when it has at least 2 keys
        emit_subtree(out, st, offset + i, m - i, level + 1, bounds);
        nchild++;
    }
    nd.n_children = (uint8_t)nchild;
    nd.subtree = (uint32_t)(out.size() - me);
    out[me] = nd;
}

// Every non-root node must be published exactly once by its parent (children of node i are
// i + 1, then each next sibling follows the previous sibling's subtree).
int check_forest(const std::vector<shgpu_node> &nodes, uint32_t n_roots) {
    std::vector<uint8_t> seen(nodes.size(), 0);
    uint64_t pushed = 0;
    for (size_t i = 0; i < nodes.size(); i++) {
        uint64_t child = i + 1;
        for (uint32_t c = 0; c < nodes[i].n_children; c++) {
            if (child >= nodes.size() || seen[child]) return fail(SHGPU_ERR_ARG, "inconsistent forest");
            seen[child] = 1;
            pushed++;
            child += nodes[child].subtree;
        }
    }
    if (pushed + n_roots != nodes.size()) return fail(SHGPU_ERR_ARG, "inconsistent forest: node count");
    return SHGPU_OK;
}

int check_variant(int variant, uint32_t leaf) {
    if (variant == SHGPU_SHOCKHASH || variant == SHGPU_SIMD_SHOCKHASH) {
        if (leaf < 1 || leaf > 60) return fail(SHGPU_ERR_UNSUPPORTED, "ShockHash leaf size must be in [1, 60]");
        return SHGPU_OK;
    }
    if (variant == SHGPU_SHOCKHASH2) {
        if (leaf < 1 || leaf > SH2_MAX_LEAF)
            return fail(SHGPU_ERR_UNSUPPORTED, "ShockHash2 leaf size must be in [1, 128] (MAX_LEAF_SIZE2)");
        return SHGPU_OK;
    }
    return fail(SHGPU_ERR_ARG, "unknown variant");
}

// Resident waves per CU of a persistent build: the schedule's value, or by the batch's key count when it is 0.
// Persistent grid of a schedule: CUs x waves per CU (a tuning resolved by resolve_tuning names its count; the
// batched per-node entry points default to 24). Small builds are set by their longest chain of dependent searches,
// and fewer resident waves (fewer idle pollers) leave each searching wave more of its SIMD and of the memory
// system; large ones need the waves to hide latency: the size classes of the schedule carry that choice.
int grid_for(const shgpu_engine *e, const shgpu_tuning &t) {
    return std::max(1, e->cu_count * (int)(t.waves_per_cu ? t.waves_per_cu : 24u) / WAVES_PER_BLOCK);
}
// Batched per-node entry points: the engine-wide schedule, 24 waves per CU unless it names a count.
int grid_for(const shgpu_engine *e) { return grid_for(e, e->tuning); }

hipStream_t pick(shgpu_engine *e, void *stream) { return stream ? (hipStream_t)stream : e->stream; }

int reset_counter(shgpu_engine *e, hipStream_t s) {
    int rc = e->counter.ensure(64);
    if (rc) return rc;
    k_fill_u32<<<1, 64, 0, s>>>(e->counter.as<uint32_t>(), 0, 16);
    HIP_TRY(hipGetLastError());
    return SHGPU_OK;
}

// P(`balls` keys hashed uniformly into `bins` cells leave no cell empty), by inclusion-exclusion.
double cover_probability(uint32_t balls, uint32_t bins) {
    double sum = 0, binom = 1;
    for (uint32_t i = 0; i <= bins; i++) {
        sum += (i & 1 ? -1.0 : 1.0) * binom * std::pow(1.0 - (double)i / bins, (double)balls);
        binom = binom * (double)(bins - i) / (double)(i + 1);
    }
    return std::max(sum, 0.0);
}

// Scratch layout of the ShockHash2 leaf searches of a batch whose largest leaf has max_m keys. Up to 64 keys:
// the fixed tables (1024 half-seeds, 4096 candidates). Above: two keys per lane, with tables sized from the
// largest leaf. Its code (pairTriangular of two pairElegant codes) is about s^4 / 2 for s half-seeds and is
// geometric with mean about 2^k / ln 2, k = bij_memo2[m] the Golomb parameter the reference chose for it, so
// P(s > x * s_mean) = exp(-x^4) with s_mean = (2^(k+1) / ln 2)^(1/4) (~110 at m = 64, ~15,000 at m = 128).
// The tables hold x = 2 * scale mean half-seeds (a leaf needs more with probability exp(-16) for scale 1,
// and the host searches those again with scale 2, exp(-256)) and the candidates among their pairs: each pair of
// half-seeds covers all ceil(m/2) cells with the probability that m keys cover them.
struct Sh2Layout {
    bool wide = false;
    Sh2Caps caps = SH2_NARROW_CAPS;
    size_t wave_bytes = SH2_SCRATCH_BYTES;
};
Sh2Layout sh2_layout(uint32_t max_m, uint32_t scale = 1, bool tight = false) {
    Sh2Layout l;
    if (max_m <= 64) return l;
    l.wide = true;
    const uint32_t k = shockhash::bij_memo2[std::min<uint32_t>(max_m, SH2_MAX_LEAF)];
    const double s_mean = std::pow(std::ldexp(1.0, (int)k + 1) / std::log(2.0), 0.25);
    const double x = tight ? 0.25 : 2.0 * scale;  // tight: a test knob that makes leaves overflow
    const double hs = std::min(65472.0, std::max(1024.0, std::ceil(x * s_mean / 64) * 64));
    const double lambda = hs * hs * cover_probability(max_m, (max_m + 1) / 2);
    const double cand = std::min(4194304.0, std::max(4096.0, std::ceil((1.1 * lambda + 64) / 64) * 64));
    l.caps = Sh2Caps{(uint32_t)hs, (uint32_t)cand};
    l.wave_bytes = sh2_scratch_bytes(2, l.caps);
    return l;
}
// Waves that search wide leaves share at most SH2_WIDE_BUDGET bytes of scratch.
constexpr size_t SH2_WIDE_BUDGET = 16ull << 30;
int sh2_grid_cap(const Sh2Layout &l, int grid) {
    if (!l.wide) return grid;
    const size_t blocks = SH2_WIDE_BUDGET / (l.wave_bytes * WAVES_PER_BLOCK);
    return std::max(1, std::min(grid, (int)std::min<size_t>(blocks, 1u << 30)));
}

int ensure_sh2_scratch(shgpu_engine *e, size_t waves, const Sh2Layout &l) { return e->sh2_scratch.ensure(waves * l.wave_bytes); }

static int load_default_schedule(shgpu_engine *e);  // defined with the schedule API below

// A variant's schedule for a batch of n keys: its tuning, with waves per CU and seeds per lane from the size class
// when the tuning leaves waves_per_cu at 0.
shgpu_tuning resolve_tuning(const shgpu_engine *e, int variant, uint64_t n) {
    shgpu_tuning t = e->tuning_v[variant];
    if (t.waves_per_cu == 0) {
        const shgpu_schedule_class &c = e->sched.lookup((uint32_t)variant, n);
        t.waves_per_cu = c.waves_per_cu;
        t.sh1_seeds_per_lane = c.sh1_seeds_per_lane;
        t.split_seeds_per_lane = c.split_seeds_per_lane;
    }
    return t;
}

template <int VARIANT>
int launch_build_v(const shgpu_tuning &t, bool wide, const BuildArgs &a, int grid, hipStream_t s) {
    const hipError_t err = launch_build_variant<VARIANT>(t.sh1_seeds_per_lane, t.split_seeds_per_lane, wide, a, grid, s);
    if (err != hipSuccess) return fail(SHGPU_ERR_HIP, std::string("k_build launch: ") + hipGetErrorString(err));
    return SHGPU_OK;
}

template <int VARIANT>
int launch_node_v(shgpu_engine *, bool wide, const BuildArgs &a, uint32_t nid, uint32_t buf, uint32_t waves, hipStream_t s) {
    const hipError_t err = launch_node_variant<VARIANT>(wide, a, nid, buf, waves, s);
    if (err != hipSuccess) return fail(SHGPU_ERR_HIP, std::string("k_node launch: ") + hipGetErrorString(err));
    return SHGPU_OK;
}

}  // namespace

// ===================================================================================== C ABI

extern "C" {

const char *shgpu_last_error(void) { return g_err.c_str(); }

const char *shgpu_status_string(int status) {
    switch (status) {
        case SHGPU_OK: return "ok";
        case SHGPU_ERR_ARG: return "invalid argument";
        case SHGPU_ERR_HIP: return "HIP runtime error";
        case SHGPU_ERR_NOMEM: return "device out of memory";
        case SHGPU_ERR_SEARCH_LIMIT: return "seed search bound exceeded (duplicate keys?)";
        case SHGPU_ERR_BUCKET_TOO_LARGE: return "bucket larger than the encodable maximum";
        case SHGPU_ERR_UNSUPPORTED: return "unsupported configuration";
        case SHGPU_ERR_NO_DEVICE: return "no usable HIP device";
        case SHGPU_ERR_ADMISSION: return "request exceeds the serving budget or the server stopped";
        default: return "unknown status";
    }
}

int shgpu_engine_create(int device, shgpu_engine **out) {
    if (!out) return fail(SHGPU_ERR_ARG, "out is NULL");
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= device || device < 0)
        return fail(SHGPU_ERR_NO_DEVICE, "no HIP device " + std::to_string(device));
    HIP_TRY(hipSetDevice(device));
    hipDeviceProp_t prop;
    HIP_TRY(hipGetDeviceProperties(&prop, device));
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        return fail(SHGPU_ERR_NO_DEVICE, std::string("device is ") + prop.gcnArchName + ", library is built for gfx950");
    void *scratch_dev = nullptr;
    int rc = noalloc_scratch_of(device, &scratch_dev);
    if (rc) return rc;
    shgpu_engine *e = new shgpu_engine();
    e->device = device;
    e->noalloc_scratch = scratch_dev;
    e->cu_count = prop.multiProcessorCount;
    e->clock_khz = prop.clockRate;
    std::snprintf(e->arch, sizeof(e->arch), "%s", prop.gcnArchName);
    if ((rc = load_default_schedule(e))) {
        delete e;
        return rc;
    }
    if (hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking) != hipSuccess || e->main.create_events()) {
        delete e;
        return fail(SHGPU_ERR_HIP, "hipStreamCreate / hipEventCreate failed");
    }
    *out = e;
    return SHGPU_OK;
}

void shgpu_engine_destroy(shgpu_engine *e) {
    if (!e) return;
    (void)hipSetDevice(e->device);
    (void)hipStreamSynchronize(e->stream);
    e->up.release();
    (void)hipStreamDestroy(e->stream);
    e->main.destroy_events();
    delete e;
}

int shgpu_engine_info(shgpu_engine *e, int *cu_count, int *clock_khz, char *arch, size_t arch_len) {
    if (!e) return fail(SHGPU_ERR_ARG, "engine is NULL");
    if (cu_count) *cu_count = e->cu_count;
    if (clock_khz) *clock_khz = e->clock_khz;
    if (arch && arch_len) std::snprintf(arch, arch_len, "%s", e->arch);
    return SHGPU_OK;
}

int shgpu_get_tuning(shgpu_engine *e, shgpu_tuning *out) {
    if (!e || !out) return fail(SHGPU_ERR_ARG, "NULL argument");
    *out = e->tuning;
    return SHGPU_OK;
}

static int check_tuning(const shgpu_tuning *t) {
    auto okS = [](uint32_t s) { return s == 1 || s == 2 || s == 4; };
    if (!okS(t->sh1_seeds_per_lane) || !okS(t->split_seeds_per_lane))
        return fail(SHGPU_ERR_ARG, "seeds per lane must be 1, 2 or 4");
    if (t->waves_per_cu > 32) return fail(SHGPU_ERR_ARG, "waves_per_cu must be in [0, 32] (0: by batch size)");
    if (t->max_trials == 0) return fail(SHGPU_ERR_ARG, "max_trials must be positive");
    if (t->flags & ~(SHGPU_TUNE_SH2_GENERAL_ONLY | SHGPU_TUNE_SPLITS_FIRST | SHGPU_TUNE_CHECK_PLAN | SHGPU_TUNE_SH2_TIGHT_TABLES))
        return fail(SHGPU_ERR_ARG, "unknown tuning flags");
    return SHGPU_OK;
}

int shgpu_set_tuning(shgpu_engine *e, const shgpu_tuning *t) {
    if (!e || !t) return fail(SHGPU_ERR_ARG, "NULL argument");
    int rc = check_tuning(t);
    if (rc) return rc;
    e->tuning = *t;
    for (auto &v : e->tuning_v) v = *t;
    return SHGPU_OK;
}

int shgpu_get_variant_tuning(shgpu_engine *e, int variant, shgpu_tuning *out) {
    if (!e || !out || variant < 0 || variant > 2) return fail(SHGPU_ERR_ARG, "bad argument");
    *out = e->tuning_v[variant];
    return SHGPU_OK;
}

int shgpu_set_variant_tuning(shgpu_engine *e, int variant, const shgpu_tuning *t) {
    if (!e || !t || variant < 0 || variant > 2) return fail(SHGPU_ERR_ARG, "bad argument");
    int rc = check_tuning(t);
    if (rc) return rc;
    e->tuning_v[variant] = *t;
    return SHGPU_OK;
}

int shgpu_sync(shgpu_engine *e) {
    if (!e) return fail(SHGPU_ERR_ARG, "engine is NULL");
    HIP_TRY(hipSetDevice(e->device));
    HIP_TRY(hipStreamSynchronize(e->stream));
    return SHGPU_OK;
}

int shgpu_sh1_leaf_search(shgpu_engine *e, const uint64_t *d_keys, const shgpu_leaf_task *d_tasks, uint32_t n,
                          uint64_t *d_seeds, void *stream) {
    if (!e || (n && (!d_keys || !d_tasks || !d_seeds))) return fail(SHGPU_ERR_ARG, "NULL argument");
    if (n == 0) return SHGPU_OK;
    HIP_TRY(hipSetDevice(e->device));
    hipStream_t s = pick(e, stream);
    int rc = reset_counter(e, s);
    if (rc) return rc;
    const int grid = std::min<int>(grid_for(e), (int)((n + WAVES_PER_BLOCK - 1) / WAVES_PER_BLOCK));
    uint32_t *ctr = e->counter.as<uint32_t>();
    HIP_TRY(launch_sh1_batch(e->tuning.sh1_seeds_per_lane, grid, s, d_keys, d_tasks, n, d_seeds, ctr));
    return SHGPU_OK;
}

int shgpu_sh2_leaf_search(shgpu_engine *e, const uint64_t *d_keys, const shgpu_leaf_task *d_tasks, uint32_t n,
                          uint64_t *d_codes, void *stream) {
    if (!e || (n && (!d_keys || !d_tasks || !d_codes))) return fail(SHGPU_ERR_ARG, "NULL argument");
    if (n == 0) return SHGPU_OK;
    HIP_TRY(hipSetDevice(e->device));
    hipStream_t s = pick(e, stream);
    // the table layout follows the largest leaf: read the tasks' sizes
    std::vector<shgpu_leaf_task> h(n);
    HIP_TRY(hipMemcpyAsync(h.data(), d_tasks, n * sizeof(shgpu_leaf_task), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    uint32_t max_m = 0;
    for (const auto &t : h) max_m = std::max(max_m, t.m);
    if (max_m > SH2_MAX_LEAF) return fail(SHGPU_ERR_ARG, "ShockHash2 leaf task with more than 128 keys");
    const Sh2Layout l = sh2_layout(max_m, 1, (e->tuning.flags & SHGPU_TUNE_SH2_TIGHT_TABLES) != 0);
    int rc = reset_counter(e, s);
    if (rc) return rc;
    const int grid = std::min<int>(sh2_grid_cap(l, grid_for(e)), (int)((n + WAVES_PER_BLOCK - 1) / WAVES_PER_BLOCK));
    rc = ensure_sh2_scratch(e, (size_t)grid * WAVES_PER_BLOCK, l);
    if (rc) return rc;
    HIP_TRY(launch_sh2_batch(l.wide, grid, s, d_keys, d_tasks, n, d_codes, e->counter.as<uint32_t>(), e->sh2_scratch.as<uint8_t>(),
                             l.wave_bytes, l.caps, e->tuning.flags));
    if (!l.wide) return SHGPU_OK;
    // leaves above 64 keys that outgrew the tables: searched again with tables twice as large
    std::vector<uint64_t> codes(n);
    HIP_TRY(hipMemcpyAsync(codes.data(), d_codes, n * 8, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    std::vector<shgpu_leaf_task> again;
    std::vector<uint32_t> ids;
    for (uint32_t i = 0; i < n; i++)
        if (codes[i] == SHGPU_NOT_FOUND && h[i].m > 64) { again.push_back(h[i]); ids.push_back(i); }
    if (again.empty()) return SHGPU_OK;
    const uint32_t k = (uint32_t)again.size();
    const Sh2Layout l2 = sh2_layout(max_m, 2);
    const size_t off_out = sh2_align256((size_t)k * sizeof(shgpu_leaf_task));
    if ((rc = e->staging.ensure(off_out + (size_t)k * 8))) return rc;
    HIP_TRY(hipMemcpyAsync(e->staging.p, again.data(), k * sizeof(shgpu_leaf_task), hipMemcpyHostToDevice, s));
    if ((rc = reset_counter(e, s))) return rc;
    const int grid2 = std::min<int>(sh2_grid_cap(l2, grid_for(e)), (int)((k + WAVES_PER_BLOCK - 1) / WAVES_PER_BLOCK));
    if ((rc = ensure_sh2_scratch(e, (size_t)grid2 * WAVES_PER_BLOCK, l2))) return rc;
    uint64_t *out2 = (uint64_t *)(e->staging.as<uint8_t>() + off_out);
    HIP_TRY(launch_sh2_batch(true, grid2, s, d_keys, (const shgpu_leaf_task *)e->staging.p, k, out2, e->counter.as<uint32_t>(),
                             e->sh2_scratch.as<uint8_t>(), l2.wave_bytes, l2.caps, e->tuning.flags & SHGPU_TUNE_SH2_GENERAL_ONLY));
    std::vector<uint64_t> codes2(k);
    HIP_TRY(hipMemcpyAsync(codes2.data(), out2, k * 8, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    for (uint32_t t = 0; t < k; t++) HIP_TRY(hipMemcpyAsync(d_codes + ids[t], &codes2[t], 8, hipMemcpyHostToDevice, s));
    HIP_TRY(hipStreamSynchronize(s));
    return SHGPU_OK;
}

int shgpu_split_search(shgpu_engine *e, uint64_t *d_keys, const shgpu_split_task *d_tasks, uint32_t n, uint64_t *d_seeds,
                       void *stream) {
    if (!e || (n && (!d_keys || !d_tasks || !d_seeds))) return fail(SHGPU_ERR_ARG, "NULL argument");
    if (n == 0) return SHGPU_OK;
    HIP_TRY(hipSetDevice(e->device));
    hipStream_t s = pick(e, stream);
    // The partition needs a scratch copy as large as the furthest key a task touches; the caller's
    // tasks are on the device, so size it from the largest key offset reachable (tmp grows to the
    // key array extent the caller declares through the tasks; read them back to size it).
    std::vector<shgpu_split_task> h(n);
    HIP_TRY(hipMemcpyAsync(h.data(), d_tasks, n * sizeof(shgpu_split_task), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    uint64_t extent = 0;
    for (const auto &t : h) {
        if (t.m > 65535) return fail(SHGPU_ERR_ARG, "split task with more than 65535 keys");
        const uint32_t fanout = t.kind == SHGPU_NODE_SPLIT2 ? 2 : (t.unit ? (t.m + t.unit - 1) / t.unit : 0);
        if (t.unit == 0 || fanout < 2 || fanout > 4 || t.unit >= t.m) return fail(SHGPU_ERR_ARG, "split task fanout must be in [2, 4]");
        extent = std::max<uint64_t>(extent, (uint64_t)t.offset + t.m);
    }
    int rc = e->tmp.ensure(extent * 8 + 64);
    if (rc) return rc;
    rc = reset_counter(e, s);
    if (rc) return rc;
    const int grid = std::min<int>(grid_for(e), (int)((n + WAVES_PER_BLOCK - 1) / WAVES_PER_BLOCK));
    uint32_t *ctr = e->counter.as<uint32_t>();
    HIP_TRY(launch_split_batch(false, e->tuning.split_seeds_per_lane, grid, s, d_keys, e->tmp.as<uint64_t>(), d_tasks, n,
                               d_seeds, ctr, extent));
    return SHGPU_OK;
}

int shgpu_split_search_noalloc(shgpu_engine *e, uint64_t *d_keys, uint64_t n_keys, const shgpu_split_task *d_tasks,
                               uint32_t n, uint64_t *d_seeds, void *stream) {
    if (!e || (n && (!d_keys || !d_tasks || !d_seeds))) return fail(SHGPU_ERR_ARG, "NULL argument");
    if (n_keys > SHGPU_NOALLOC_MAX_KEYS)
        return fail(SHGPU_ERR_ARG, "shgpu_split_search_noalloc: " + std::to_string(n_keys) + " keys exceed its scratch of " +
                                       std::to_string(SHGPU_NOALLOC_MAX_KEYS) + " keys");
    if (n == 0) return SHGPU_OK;
    HIP_TRY(hipSetDevice(e->device));
    uint64_t *scratch_dev = (uint64_t *)e->noalloc_scratch;  // this device's copy of `scratch`
    if (!scratch_dev) return fail(SHGPU_ERR_HIP, "no-allocation scratch of this device is unavailable");
    uint32_t *ctr = (uint32_t *)(scratch_dev + SHGPU_NOALLOC_MAX_KEYS);
    hipStream_t s = pick(e, stream);
    k_fill_u32<<<1, 64, 0, s>>>(ctr, 0, 16);  // a kernel, not a memset: callers may capture this stream in a graph
    HIP_TRY(hipGetLastError());
    const int grid = std::min<int>(grid_for(e), (int)((n + WAVES_PER_BLOCK - 1) / WAVES_PER_BLOCK));
    HIP_TRY(launch_split_batch(true, e->tuning.split_seeds_per_lane, grid, s, d_keys, scratch_dev, d_tasks, n, d_seeds, ctr,
                               n_keys));
    return SHGPU_OK;
}

struct BuildInput {
    const uint64_t *d_hashes;  // n fingerprints {first, second} in device memory
    uint64_t n;
};

// Node counts of a tree of every size m <= MAX_ENCODABLE_BUCKET (all nodes, non-leaf nodes): the shape of
// emit_subtree, by sizes in increasing order (children are smaller than their parent).
static int subtree_counts(const Strategy &st, std::vector<uint32_t> &all, std::vector<uint32_t> &split) {
    all.assign(MAX_ENCODABLE_BUCKET + 1, 0);
    split.assign(MAX_ENCODABLE_BUCKET + 1, 0);
    all[1] = 1;  // a part of one key (leaf size 1) is a leaf node; roots and last parts need 2 keys
    for (uint32_t m = 2; m <= MAX_ENCODABLE_BUCKET; m++) {
        if (m <= st.leaf) {
            all[m] = 1;
            continue;
        }
        uint32_t a = 1, sp = 1, nc = 0;
        auto child = [&](uint32_t c) {
            a += all[c];
            sp += split[c];
            nc++;
        };
        if (m > st.upper) {
            const uint32_t unit = (uint16_t(m / 2 + st.upper - 1) / st.upper) * st.upper;
            child(unit);
            if (m - unit > 1) child(m - unit);
        } else {
            const uint32_t unit = m > st.lower ? st.lower : st.leaf;
            uint32_t i = 0;
            for (; i < m - unit; i += unit) child(unit);
            if (m - i > 1) child(m - i);
        }
        if (nc > 4) return fail(SHGPU_ERR_UNSUPPORTED, "tree node with more than 4 children");
        all[m] = a;
        split[m] = sp;
    }
    return SHGPU_OK;
}

// Host copy of a build's nodes (DFS pre-order per bucket, offsets within the build's keys).
static int host_plan(const shgpu_build &b, std::vector<shgpu_node> &out) {
    const Strategy st(b.variant, b.leaf);
    const std::vector<uint64_t> bounds = node_bounds(st, b.max_trials);
    out.clear();
    out.reserve(b.codes.size());
    const uint64_t nb = b.plan_buckets;
    for (uint64_t k = 0; k < nb; k++) {
        const uint64_t m = b.size_acc[k + 1] - b.size_acc[k];
        if (m > 1) emit_subtree(out, st, (uint32_t)b.size_acc[k], (uint32_t)m, 0, bounds);
    }
    if (out.size() != b.codes.size()) return fail(SHGPU_ERR_HIP, "host plan and device plan differ in node count");
    return SHGPU_OK;
}

// Searches again, with tables twice as large (sh2_layout scale 2), the ShockHash2 leaves above 64 keys whose
// search outgrew the build's tables, and writes their codes into a.codes. The stream is idle on entry (the
// caller has synchronised it). A leaf that fails again keeps SHGPU_NOT_FOUND (repeated keys).
static int sh2_retry_overflowed(shgpu_engine *e, BuildBuffers &B, hipStream_t s, const BuildArgs &a, uint32_t total,
                                uint32_t leaf, uint32_t flags) {
    const size_t off_tasks = 256, off_ids = off_tasks + sh2_align256((size_t)total * sizeof(shgpu_leaf_task)),
                 off_out = off_ids + sh2_align256((size_t)total * 4), bytes = off_out + (size_t)total * 8 + 64;
    int rc = B.sh2_retry.ensure(bytes);
    if (rc) return rc;
    uint8_t *base = B.sh2_retry.as<uint8_t>();
    uint32_t *count = (uint32_t *)base;  // [0] overflowed leaves, [16] the batch's work counter
    shgpu_leaf_task *tasks = (shgpu_leaf_task *)(base + off_tasks);
    uint32_t *ids = (uint32_t *)(base + off_ids);
    uint64_t *out = (uint64_t *)(base + off_out);
    k_fill_u32<<<1, 64, 0, s>>>(count, 0, 64);
    HIP_TRY(hipGetLastError());
    k_sh2_collect_overflow<<<std::max(1u, std::min(1024u, (total + 255) / 256)), 256, 0, s>>>(a.nodes, a.codes, total, tasks, ids, count);
    HIP_TRY(hipGetLastError());
    uint32_t n = 0;
    HIP_TRY(hipMemcpyAsync(&n, count, 4, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    if (n == 0) return SHGPU_OK;
    const Sh2Layout l = sh2_layout(leaf, 2);
    const int grid = std::min(sh2_grid_cap(l, grid_for(e)), (int)((n + WAVES_PER_BLOCK - 1) / WAVES_PER_BLOCK));
    if ((rc = B.sh2_scratch.ensure((size_t)grid * WAVES_PER_BLOCK * l.wave_bytes))) return rc;
    HIP_TRY(launch_sh2_batch(true, grid, s, a.keys, tasks, n, out, count + 16, B.sh2_scratch.as<uint8_t>(), l.wave_bytes, l.caps,
                             flags & SHGPU_TUNE_SH2_GENERAL_ONLY));
    k_scatter_codes<<<std::max(1u, std::min(1024u, (n + 255) / 256)), 256, 0, s>>>(out, ids, n, a.codes);
    HIP_TRY(hipGetLastError());
    return SHGPU_OK;
}

// The construction pipeline for a batch of independent key sets on one buffer set and stream: each
// set is sorted and bucketed on its own, their forests share one node array, queue and persistent
// kernel (the trees are independent, so every set's codes equal those of a build of that set alone).
// The persistent grid uses 1/grid_div of the variant's schedule (serving slots share the device).
// The persistent form plans the forest on the device and runs from the sort to the codes without a host
// round trip; the plain form plans on the host and launches one kernel per node.
// status[i] receives the outcome of set i; the return value is the first failure (SHGPU_OK if none).
static int run_batch(shgpu_engine *e, BuildBuffers &B, hipStream_t s, uint32_t grid_div, const shgpu_build_desc *d,
                     const BuildInput *in, size_t count, shgpu_build **out, int *status, bool full_sort = false) {
    using clk = std::chrono::steady_clock;
    if (!e || !d || !out || !in || !status || count == 0) return fail(SHGPU_ERR_ARG, "NULL argument");
    int rc = SHGPU_OK;
    auto fail_all = [&](int code) {
        for (size_t r = 0; r < count; r++) status[r] = code;
        return code;
    };
    for (size_t r = 0; r < count; r++) out[r] = nullptr;
    fail_all(SHGPU_OK);
    if ((rc = check_variant(d->variant, d->leaf_size))) return fail_all(rc);
    if (d->bucket_size == 0) return fail_all(fail(SHGPU_ERR_ARG, "bucket_size must be positive"));
    uint64_t N = 0, NB = 0;
    std::vector<uint64_t> base(count), nbk(count), bbase(count);
    for (size_t r = 0; r < count; r++) {
        if (in[r].n && !in[r].d_hashes) return fail_all(fail(SHGPU_ERR_ARG, "NULL argument"));
        base[r] = N;
        bbase[r] = NB;
        nbk[r] = std::max<uint64_t>(1, (in[r].n + d->bucket_size - 1) / d->bucket_size);
        N += in[r].n;
        NB += nbk[r] + 1;
    }
    if (N >= (1ull << 32)) return fail_all(fail(SHGPU_ERR_UNSUPPORTED, "more than 2^32 - 1 keys"));
    if (NB >= (1ull << 31)) return fail_all(fail(SHGPU_ERR_UNSUPPORTED, "too many buckets"));
    if (hipSetDevice(e->device) != hipSuccess) return fail_all(fail(SHGPU_ERR_HIP, "hipSetDevice failed"));
#define BATCH_HIP(call)                                                                                        \
    do {                                                                                                       \
        hipError_t err_ = (call);                                                                              \
        if (err_ != hipSuccess) return fail_all(fail(SHGPU_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(err_))); \
    } while (0)
    auto t0 = clk::now();
    const shgpu_tuning tune = resolve_tuning(e, d->variant, N);
    BATCH_HIP(hipEventRecord(B.evA, s));
    std::vector<std::unique_ptr<shgpu_build>> bs(count);
    for (size_t r = 0; r < count; r++) {
        bs[r].reset(new shgpu_build());
        bs[r]->n = in[r].n;
        bs[r]->nbuckets = nbk[r];
        bs[r]->plan_buckets = nbk[r];
        bs[r]->variant = d->variant;
        bs[r]->leaf = d->leaf_size;
        bs[r]->max_trials = tune.max_trials;
    }

    // 1. partition: sort by .first, emit .second in order (Sorter.hpp:15-36), bucket bounds. The radix sort
    // covers the top key bits only and k_sort_fixup orders the rare runs that share them.
    if ((rc = B.firsts_in.ensure(N * 8 + 8)) || (rc = B.firsts_out.ensure(N * 8 + 8)) ||
        (rc = B.seconds_in.ensure(N * 8 + 8)) || (rc = B.keys.ensure(N * 8 + 64)) || (rc = B.tmp.ensure(N * 8 + 64)) ||
        (rc = B.size_acc.ensure(NB * 8)) || (rc = B.sort_flag.ensure(64)))
        return fail_all(rc);
    uint64_t nmax = 0;
    for (size_t r = 0; r < count; r++) nmax = std::max(nmax, in[r].n);
    size_t sort_bytes = 0;
    BATCH_HIP(shgpu_sort_pairs_u64(nullptr, sort_bytes, B.firsts_in.as<uint64_t>(), B.firsts_out.as<uint64_t>(),
                                   B.seconds_in.as<uint64_t>(), B.keys.as<uint64_t>(), (int)nmax, 64, s));
    if ((rc = B.sort_tmp.ensure(sort_bytes + 16))) return fail_all(rc);
    k_fill_u32<<<1, 64, 0, s>>>(B.sort_flag.as<uint32_t>(), 0u, 1);
    BATCH_HIP(hipGetLastError());
    for (size_t r = 0; r < count; r++) {
        const uint64_t n = in[r].n, o = base[r];
        const int bits = sort_bits(n, full_sort);
        const uint32_t rot = (uint32_t)(64 - bits);
        if (n) {
            k_unzip_hashes<<<1024, 256, 0, s>>>(in[r].d_hashes, n, B.firsts_in.as<uint64_t>() + o, B.seconds_in.as<uint64_t>() + o, rot);
            BATCH_HIP(hipGetLastError());
            size_t bytes = sort_bytes;
            BATCH_HIP(shgpu_sort_pairs_u64(B.sort_tmp.p, bytes, B.firsts_in.as<uint64_t>() + o, B.firsts_out.as<uint64_t>() + o,
                                           B.seconds_in.as<uint64_t>() + o, B.keys.as<uint64_t>() + o, (int)n, bits, s));
            if (rot > 0) {
                k_sort_fixup<<<1024, 256, 0, s>>>(B.firsts_out.as<uint64_t>() + o, B.keys.as<uint64_t>() + o, n, rot,
                                                  B.sort_flag.as<uint32_t>());
                BATCH_HIP(hipGetLastError());
            }
        }
        k_bucket_bounds<<<1024, 256, 0, s>>>(B.firsts_out.as<uint64_t>() + o, n, nbk[r], B.size_acc.as<uint64_t>() + bbase[r], rot);
        BATCH_HIP(hipGetLastError());
    }
    BATCH_HIP(hipEventRecord(B.evB, s));

    // ShockHash2 leaves above 64 keys need tables that grow with the leaf size (fewer waves)
    const Sh2Layout sh2l = d->variant == SHGPU_SHOCKHASH2
                               ? sh2_layout(d->leaf_size, 1, (tune.flags & SHGPU_TUNE_SH2_TIGHT_TABLES) != 0)
                               : Sh2Layout{};
    const int grid = sh2_grid_cap(sh2l, std::max(1, grid_for(e, tune) / (int)std::max(1u, grid_div)));
    const uint32_t nslots = (uint32_t)grid * WAVES_PER_BLOCK;
    std::vector<uint64_t> acc_all(NB);
    std::vector<uint64_t> codes;
    std::vector<uint64_t> node_begin(count, 0), node_end(count, 0);
    uint32_t ctrl_error = 0;
    uint32_t sort_rerun = 0;  // a run of keys sharing the sorted bits was too long: sort on all 64 bits
    const uint64_t *codes_pinned = nullptr;  // persistent form: codes arrive in pinned memory (n_codes of them)
    const shgpu_node *nodes_pinned = nullptr;  // ... and with SHGPU_NODES_HOST the device-planned nodes
    uint64_t n_codes = 0;
    bool timed_kernel = false;
    BuildArgs a;
    a.keys = B.keys.as<uint64_t>();
    a.tmp = B.tmp.as<uint64_t>();
    a.split_hash = d->variant == SHGPU_SIMD_SHOCKHASH ? SHGPU_SPLIT_SIMD : SHGPU_SPLIT_SUX;
    a.flags = tune.flags;
    a.rot_leaf = d->rotation_fitting && d->variant != SHGPU_SHOCKHASH2 ? d->leaf_size : 0u;
    a.scratch = nullptr;
    a.sh2_caps = sh2l.caps;
    a.sh2_wave_bytes = sh2l.wave_bytes;
    a.nslots = nslots;
    const Strategy st(d->variant, d->leaf_size);
    auto t1 = clk::now(), t2 = t1;

    if (d->mode == SHGPU_MODE_PERSISTENT) {
        // 2. forest plan on the device (node count tables per variant and leaf size, uploaded once).
        if (B.plan_variant != d->variant || B.plan_leaf != d->leaf_size) {
            std::vector<uint32_t> all, split;
            if ((rc = subtree_counts(st, all, split))) return fail_all(rc);
            const std::vector<uint32_t> lg = seed_bound_log2(st);
            const std::vector<uint32_t> rate = seed_rates(st);
            B.plan_counts = all;
            B.plan_counts.insert(B.plan_counts.end(), split.begin(), split.end());
            B.plan_counts.insert(B.plan_counts.end(), lg.begin(), lg.end());
            B.plan_counts.insert(B.plan_counts.end(), rate.begin(), rate.end());
            if ((rc = B.plan_tab.ensure(B.plan_counts.size() * 4))) return fail_all(rc);
            BATCH_HIP(hipMemcpyAsync(B.plan_tab.p, B.plan_counts.data(), B.plan_counts.size() * 4, hipMemcpyHostToDevice, s));
            B.plan_variant = d->variant;
            B.plan_leaf = d->leaf_size;
        }
        // node capacity: the largest tree of a bucket up to 1.3x the set's mean bucket size, per bucket
        uint64_t cap = 1024;
        for (size_t r = 0; r < count; r++) {
            const uint64_t mean = in[r].n / nbk[r] + 1;
            const uint64_t hi = std::min<uint64_t>(MAX_ENCODABLE_BUCKET, mean + mean * 3 / 10 + 8);
            uint32_t mx = 0;
            for (uint64_t m = 0; m <= hi; m++) mx = std::max(mx, B.plan_counts[m]);
            cap += nbk[r] * mx;
        }
        cap = std::max<uint64_t>(cap, B.plan_capacity);
        // queue entries carry a node id below 2^31 - 2 (bit 31 names the key buffer, 0xFFFFFFFE and up are marks)
        if (cap >= 0x7FFFFFFEull) return fail_all(fail(SHGPU_ERR_ARG, "forest too large for one build"));
        std::vector<PlanSet> sets(count);
        for (size_t r = 0; r < count; r++) sets[r] = PlanSet{base[r], (uint32_t)bbase[r], (uint32_t)nbk[r]};
        if ((rc = B.plan_sets.ensure(count * sizeof(PlanSet))) || (rc = B.plan_bucket.ensure(NB * sizeof(uint4))) ||
            (rc = B.plan_summary.ensure((SUM_SETS + 3 * count) * 4)) || (rc = B.ctrl.ensure(CTRL_WORDS * 4)) ||
            (rc = B.help.ensure((size_t)nslots * 4 + 64)))
            return fail_all(rc);
        BATCH_HIP(hipMemcpyAsync(B.plan_sets.p, sets.data(), count * sizeof(PlanSet), hipMemcpyHostToDevice, s));
        if (d->variant == SHGPU_SHOCKHASH2) {
            if ((rc = B.sh2_scratch.ensure((size_t)grid * WAVES_PER_BLOCK * sh2l.wave_bytes))) return fail_all(rc);
            a.scratch = B.sh2_scratch.as<uint8_t>();
        }
        std::vector<uint32_t> summary(SUM_SETS + 3 * count);
        for (int attempt = 0;; attempt++) {
            if ((rc = B.nodes.ensure(cap * sizeof(shgpu_node))) || (rc = B.codes.ensure(cap * 8)) ||
                (rc = B.queue.ensure(cap * 4)) || (rc = B.states.ensure(cap * sizeof(NodeState))))
                return fail_all(rc);
            B.plan_capacity = cap;
            PlanArgs p;
            p.size_acc = B.size_acc.as<uint64_t>();
            p.sets = B.plan_sets.as<PlanSet>();
            p.nsets = (uint32_t)count;
            p.nacc = (uint32_t)NB;
            p.cnt_all = B.plan_tab.as<uint32_t>();
            p.cnt_split = B.plan_tab.as<uint32_t>() + MAX_ENCODABLE_BUCKET + 1;
            p.bound_lg = B.plan_tab.as<uint32_t>() + 2 * (MAX_ENCODABLE_BUCKET + 1);
            p.seed_rate = (const float *)(B.plan_tab.as<uint32_t>() + 3 * (MAX_ENCODABLE_BUCKET + 1));
            p.win_leaf = 64u * tune.sh1_seeds_per_lane;
            p.win_lower = 64u * tune.split_seeds_per_lane;
            p.bucket = B.plan_bucket.as<uint4>();
            p.summary = B.plan_summary.as<uint32_t>();
            p.capacity = (uint32_t)std::min<uint64_t>(cap, 0xFFFFFFFFull);
            p.two_part = two_part_queue(tune.flags);
            p.ctrl = B.ctrl.as<uint32_t>();
            p.help = B.help.as<uint32_t>();
            p.nhelp = nslots;
            p.nodes = B.nodes.as<shgpu_node>();
            p.codes = B.codes.as<uint64_t>();
            p.state = B.states.as<NodeState>();
            p.queue = B.queue.as<uint32_t>();
            p.variant = d->variant;
            p.leaf = st.leaf;
            p.lower = st.lower;
            p.upper = st.upper;
            p.max_trials = tune.max_trials;
            for (int i = 0; i < 20; i++) p.start_seed[i] = START_SEED[i];
            k_plan_scan<<<1, PLAN_THREADS, 0, s>>>(p);
            BATCH_HIP(hipGetLastError());
            k_plan_emit<<<(unsigned)((NB + 255) / 256), 256, 0, s>>>(p);
            BATCH_HIP(hipGetLastError());

            // 3. seed search: the persistent kernel reads the node count from the control words.
            a.nodes = p.nodes;
            a.codes = B.codes.as<uint64_t>();
            a.queue = p.queue;
            a.ctrl = p.ctrl;
            a.state = p.state;
            a.help = p.help;
            a.total = 0;
            BATCH_HIP(hipEventRecord(B.ev0, s));
            if (d->variant == SHGPU_SHOCKHASH) rc = launch_build_v<SHGPU_SHOCKHASH>(tune, false, a, grid, s);
            else if (d->variant == SHGPU_SIMD_SHOCKHASH) rc = launch_build_v<SHGPU_SIMD_SHOCKHASH>(tune, false, a, grid, s);
            else rc = launch_build_v<SHGPU_SHOCKHASH2>(tune, sh2l.wide, a, grid, s);
            if (rc) return fail_all(rc);
            BATCH_HIP(hipEventRecord(B.ev1, s));
            timed_kernel = true;
            BATCH_HIP(hipMemcpyAsync(summary.data(), p.summary, summary.size() * 4, hipMemcpyDeviceToHost, s));
            BATCH_HIP(hipMemcpyAsync(acc_all.data(), B.size_acc.p, NB * 8, hipMemcpyDeviceToHost, s));
            BATCH_HIP(hipMemcpyAsync(&ctrl_error, a.ctrl + C_ERROR, 4, hipMemcpyDeviceToHost, s));
            BATCH_HIP(hipMemcpyAsync(&sort_rerun, B.sort_flag.p, 4, hipMemcpyDeviceToHost, s));
            BATCH_HIP(hipStreamSynchronize(s));
            if (sort_rerun) return run_batch(e, B, s, grid_div, d, in, count, out, status, true);
            if (!summary[SUM_OVERFLOW]) break;
            if (attempt > 0) return fail_all(fail(SHGPU_ERR_HIP, "forest plan: node capacity overflow after resizing"));
            cap = (uint64_t)summary[SUM_TOTAL] + 1024;  // exact size, then run the plan and the search again
        }
        const uint32_t total = summary[SUM_TOTAL];
        if (sh2l.wide && total && !ctrl_error &&
            (rc = sh2_retry_overflowed(e, B, s, a, total, d->leaf_size, tune.flags)))
            return fail_all(rc);
        if ((rc = B.codes_host.ensure((size_t)total * 8 + 8))) return fail_all(rc);
        codes_pinned = (const uint64_t *)B.codes_host.p;
        n_codes = total;
        if (total) BATCH_HIP(hipMemcpyAsync(B.codes_host.p, a.codes, (size_t)total * 8, hipMemcpyDeviceToHost, s));
        if ((d->copy_keys & SHGPU_NODES_HOST) && total) {
            if ((rc = B.nodes_host.ensure((size_t)total * sizeof(shgpu_node)))) return fail_all(rc);
            nodes_pinned = (const shgpu_node *)B.nodes_host.p;
            BATCH_HIP(hipMemcpyAsync(B.nodes_host.p, a.nodes, (size_t)total * sizeof(shgpu_node), hipMemcpyDeviceToHost, s));
        }
        for (size_t r = 0; r < count; r++) {
            node_begin[r] = summary[SUM_SETS + 3 * r];
            node_end[r] = summary[SUM_SETS + 3 * r + 1];
            if (summary[SUM_SETS + 3 * r + 2]) status[r] = fail(SHGPU_ERR_BUCKET_TOO_LARGE, "a bucket holds more than 2999 keys");
        }
        if (total && ctrl_error) return fail_all(fail(SHGPU_ERR_HIP, "persistent build: a queue slot was never filled (inconsistent forest)"));
        if (tune.flags & SHGPU_TUNE_CHECK_PLAN) {  // the device plan against the host plan, node by node
            std::vector<shgpu_node> dev(total);
            if (total) BATCH_HIP(hipMemcpyAsync(dev.data(), a.nodes, (size_t)total * sizeof(shgpu_node), hipMemcpyDeviceToHost, s));
            BATCH_HIP(hipStreamSynchronize(s));
            for (size_t r = 0; r < count; r++) {
                if (status[r] != SHGPU_OK) continue;
                shgpu_build &b = *bs[r];
                b.size_acc.assign(acc_all.begin() + bbase[r], acc_all.begin() + bbase[r] + nbk[r] + 1);
                b.codes.assign(node_end[r] - node_begin[r], 0);
                std::vector<shgpu_node> host;
                if ((rc = host_plan(b, host))) return fail_all(rc);
                for (size_t i = 0; i < host.size(); i++) {
                    shgpu_node dn = dev[node_begin[r] + i];
                    dn.offset -= (uint32_t)base[r];
                    if (std::memcmp(&dn, &host[i], sizeof(shgpu_node)) != 0)
                        return fail_all(fail(SHGPU_ERR_HIP, "device plan differs from the host plan at node " + std::to_string(i)));
                }
            }
        }
    } else {
        // 2. forest on the host (the plain form launches one kernel per node from it).
        BATCH_HIP(hipMemcpyAsync(acc_all.data(), B.size_acc.p, NB * 8, hipMemcpyDeviceToHost, s));
        BATCH_HIP(hipMemcpyAsync(&sort_rerun, B.sort_flag.p, 4, hipMemcpyDeviceToHost, s));
        BATCH_HIP(hipStreamSynchronize(s));
        if (sort_rerun) return run_batch(e, B, s, grid_div, d, in, count, out, status, true);
        t1 = clk::now();
        std::vector<shgpu_node> nodes;
        const std::vector<uint64_t> bounds = node_bounds(st, tune.max_trials);
        uint32_t roots = 0;
        for (size_t r = 0; r < count; r++) {
            node_begin[r] = node_end[r] = nodes.size();
            const uint64_t *acc = acc_all.data() + bbase[r];
            uint64_t limit_buckets = nbk[r];
            if (d->plain_bucket_limit) limit_buckets = std::min<uint64_t>(nbk[r], d->plain_bucket_limit);
            bs[r]->plan_buckets = limit_buckets;
            for (uint64_t k = 0; k < nbk[r]; k++) {
                const uint64_t m = acc[k + 1] - acc[k];
                if (m > MAX_ENCODABLE_BUCKET) {
                    status[r] = fail(SHGPU_ERR_BUCKET_TOO_LARGE, "bucket " + std::to_string(k) + " holds " + std::to_string(m) + " keys");
                    break;
                }
            }
            if (status[r] != SHGPU_OK) continue;  // this set is not searched
            for (uint64_t k = 0; k < limit_buckets; k++) {
                const uint64_t m = acc[k + 1] - acc[k];
                if (m > 1) {
                    roots++;
                    emit_subtree(nodes, st, (uint32_t)(base[r] + acc[k]), (uint32_t)m, 0, bounds);
                }
            }
            node_end[r] = nodes.size();
        }
        const uint32_t total = (uint32_t)nodes.size();
        if ((rc = check_forest(nodes, roots))) return fail_all(rc);
        if ((rc = B.nodes.ensure((size_t)total * sizeof(shgpu_node) + 64)) || (rc = B.codes.ensure((size_t)total * 8 + 64)) ||
            (rc = B.ctrl.ensure(CTRL_WORDS * 4)))
            return fail_all(rc);
        if (total) BATCH_HIP(hipMemcpyAsync(B.nodes.p, nodes.data(), (size_t)total * sizeof(shgpu_node), hipMemcpyHostToDevice, s));
        t2 = clk::now();
        a.nodes = B.nodes.as<shgpu_node>();
        a.codes = B.codes.as<uint64_t>();
        a.queue = nullptr;
        a.ctrl = B.ctrl.as<uint32_t>();
        a.state = nullptr;
        a.help = nullptr;
        a.total = total;
        if (d->variant == SHGPU_SHOCKHASH2) {
            if ((rc = B.sh2_scratch.ensure(sh2l.wave_bytes))) return fail_all(rc);
            a.scratch = B.sh2_scratch.as<uint8_t>();
        }
        // 3. seed search: one launch per node, DFS order (every node's own state is its window counter).
        if ((rc = B.states.ensure((size_t)total * sizeof(NodeState) + 64))) return fail_all(rc);
        a.state = B.states.as<NodeState>();
        if (total) {
            k_init_states<<<256, 256, 0, s>>>(a.state, a.codes, total, nullptr, 0);
            BATCH_HIP(hipGetLastError());
        }
        const uint32_t waves = (uint32_t)e->cu_count;  // one wave per CU joins every node's windows
        // key buffer of every node: roots in keys, each level partitions into the other buffer
        std::vector<uint8_t> buf(total, 0);
        for (uint32_t nid = 0; nid < total; nid++) {
            uint32_t c = nid + 1;
            for (uint32_t k = 0; k < nodes[nid].n_children; k++, c += nodes[c].subtree) buf[c] = buf[nid] ^ 1;
        }
        for (uint32_t nid = 0; nid < total; nid++) {
            const bool sh2_leaf = d->variant == SHGPU_SHOCKHASH2 && nodes[nid].kind == SHGPU_NODE_LEAF;
            const uint32_t wv = sh2_leaf ? 1u : waves;
            if (d->variant == SHGPU_SHOCKHASH) rc = launch_node_v<SHGPU_SHOCKHASH>(e, false, a, nid, buf[nid], wv, s);
            else if (d->variant == SHGPU_SIMD_SHOCKHASH) rc = launch_node_v<SHGPU_SIMD_SHOCKHASH>(e, false, a, nid, buf[nid], wv, s);
            else rc = launch_node_v<SHGPU_SHOCKHASH2>(e, sh2l.wide, a, nid, buf[nid], wv, s);
            if (rc) return fail_all(rc);
        }
        if (sh2l.wide && total) {
            BATCH_HIP(hipStreamSynchronize(s));
            if ((rc = sh2_retry_overflowed(e, B, s, a, total, d->leaf_size, tune.flags))) return fail_all(rc);
        }
        codes.resize(total);
        if (total) BATCH_HIP(hipMemcpyAsync(codes.data(), a.codes, (size_t)total * 8, hipMemcpyDeviceToHost, s));
        for (size_t r = 0; r < count; r++) {
            if (status[r] != SHGPU_OK) continue;
            bs[r]->nodes.assign(nodes.begin() + node_begin[r], nodes.begin() + node_end[r]);
            for (auto &nd : bs[r]->nodes) nd.offset -= (uint32_t)base[r];  // offsets within this set's keys
            bs[r]->nodes_ready = true;
        }
    }
    std::vector<uint64_t> keys_all;
    if ((d->copy_keys & SHGPU_KEYS_HOST) && N) {
        keys_all.resize(N);
        BATCH_HIP(hipMemcpyAsync(keys_all.data(), B.keys.p, N * 8, hipMemcpyDeviceToHost, s));
    }
    if (d->copy_keys & SHGPU_KEYS_DEVICE) {
        for (size_t r = 0; r < count; r++) {
            if (status[r] != SHGPU_OK) continue;
            shgpu_build &b = *bs[r];
            b.device = e->device;
            if (!device_malloc(&b.d_keys, b.n * 8 + 8)) {
                b.d_keys = nullptr;
                return fail_all(fail(SHGPU_ERR_NOMEM, "hipMalloc of the build's device keys failed"));
            }
            if (b.n) BATCH_HIP(hipMemcpyAsync(b.d_keys, B.keys.as<uint64_t>() + base[r], b.n * 8, hipMemcpyDeviceToDevice, s));
        }
    }
    BATCH_HIP(hipStreamSynchronize(s));
#undef BATCH_HIP
    auto t3 = clk::now();
    if (codes_pinned) codes.assign(codes_pinned, codes_pinned + n_codes);
    float kms = 0, sort_ms = 0, plan_ms = 0;
    if (timed_kernel && (hipEventElapsedTime(&kms, B.ev0, B.ev1) != hipSuccess || hipEventElapsedTime(&sort_ms, B.evA, B.evB) != hipSuccess ||
                         hipEventElapsedTime(&plan_ms, B.evB, B.ev0) != hipSuccess))
        return fail_all(fail(SHGPU_ERR_HIP, "hipEventElapsedTime failed"));
    const double wall_ms = std::chrono::duration<double, std::milli>(t3 - t0).count();
    int first_rc = SHGPU_OK;
    for (size_t r = 0; r < count; r++) {
        shgpu_build &b = *bs[r];
        if (status[r] == SHGPU_OK) {
            b.size_acc.assign(acc_all.begin() + bbase[r], acc_all.begin() + bbase[r] + nbk[r] + 1);
            b.codes.assign(codes.begin() + node_begin[r], codes.begin() + node_end[r]);
            for (size_t i = 0; i < b.codes.size(); i++)
                if (b.codes[i] == SHGPU_NOT_FOUND) {
                    status[r] = fail(SHGPU_ERR_SEARCH_LIMIT, "node " + std::to_string(i) + " exhausted its seed bound");
                    break;
                }
        }
        if (status[r] != SHGPU_OK) {
            if (first_rc == SHGPU_OK) first_rc = status[r];
            continue;
        }
        if (!keys_all.empty()) b.keys.assign(keys_all.begin() + base[r], keys_all.begin() + base[r] + b.n);
        if (nodes_pinned) {  // the device plan, offsets made relative to this set's keys
            b.nodes.assign(nodes_pinned + node_begin[r], nodes_pinned + node_end[r]);
            for (auto &nd : b.nodes) nd.offset -= (uint32_t)base[r];
            b.nodes_ready = true;
        }
        if (timed_kernel) {
            b.ms[0] = sort_ms;
            b.ms[1] = plan_ms;
            b.ms[2] = std::max(0.0, wall_ms - sort_ms - plan_ms);
            b.ms[3] = kms;
        } else {
            b.ms[0] = std::chrono::duration<double, std::milli>(t1 - t0).count();
            b.ms[1] = std::chrono::duration<double, std::milli>(t2 - t1).count();
            b.ms[2] = std::chrono::duration<double, std::milli>(t3 - t2).count();
        }
        out[r] = bs[r].release();
    }
    return first_rc;
}

static int run_build(shgpu_engine *e, BuildBuffers &B, hipStream_t s, uint32_t grid_div, const shgpu_build_desc *d,
                     const uint64_t *d_hashes, uint64_t n, shgpu_build **out) {
    if (!out) return fail(SHGPU_ERR_ARG, "NULL argument");
    const BuildInput in{d_hashes, n};
    int st = SHGPU_OK;
    return run_batch(e, B, s, grid_div, d, &in, 1, out, &st);
}

int shgpu_build_run(shgpu_engine *e, const shgpu_build_desc *d, const uint64_t *d_hashes, uint64_t n, shgpu_build **out) {
    if (!e) return fail(SHGPU_ERR_ARG, "NULL argument");
    return run_build(e, e->main, e->stream, 1, d, d_hashes, n, out);
}

int shgpu_leaf_orientation(shgpu_engine *e, int variant, const uint64_t *d_keys, const shgpu_orient_task *d_tasks,
                           uint64_t n, uint64_t *d_out_keys, uint8_t *d_out_bits) {
    if (!e || (n && (!d_keys || !d_tasks || !d_out_keys || !d_out_bits))) return fail(SHGPU_ERR_ARG, "NULL argument");
    if (variant != SHGPU_SHOCKHASH && variant != SHGPU_SIMD_SHOCKHASH && variant != SHGPU_SHOCKHASH2)
        return fail(SHGPU_ERR_ARG, "unknown variant");
    if (hipSetDevice(e->device) != hipSuccess) return fail(SHGPU_ERR_HIP, "hipSetDevice failed");
    std::string err;
    const int rc = shgpu_leaf_orientation_impl(e->stream, variant == SHGPU_SHOCKHASH2, d_keys, d_tasks, n, d_out_keys,
                                               d_out_bits, err);
    return rc == SHGPU_OK ? SHGPU_OK : fail(rc, err);
}

int shgpu_hash_strings(shgpu_engine *e, const uint8_t *d_data, const uint64_t *d_offsets, uint64_t n, uint64_t *d_hashes) {
    if (!e || (n && (!d_data || !d_offsets || !d_hashes))) return fail(SHGPU_ERR_ARG, "NULL argument");
    if (hipSetDevice(e->device) != hipSuccess) return fail(SHGPU_ERR_HIP, "hipSetDevice failed");
    std::string err;
    const int rc = shgpu_hash_strings_impl(e->stream, d_data, d_offsets, n, d_hashes, err);
    return rc == SHGPU_OK ? SHGPU_OK : fail(rc, err);
}

int shgpu_ribbon_build(shgpu_engine *e, const uint64_t *d_keys, const uint8_t *d_bits, uint64_t n, uint32_t *h_offsets,
                       uint8_t *h_seeds, uint64_t *h_words, uint64_t words_capacity, uint64_t *words_used) {
    if (!e || !h_offsets || !h_seeds || !h_words || !words_used || (n && (!d_keys || !d_bits)))
        return fail(SHGPU_ERR_ARG, "NULL argument");
    if (hipSetDevice(e->device) != hipSuccess) return fail(SHGPU_ERR_HIP, "hipSetDevice failed");
    std::string err;
    const int rc = shgpu_ribbon_build_impl(e->stream, d_keys, d_bits, n, h_offsets, h_seeds, h_words, words_capacity,
                                           words_used, err, &e->retrieval.p, &e->retrieval.bytes);
    return rc == SHGPU_OK ? SHGPU_OK : fail(rc, err);
}

int shgpu_flat_layers(shgpu_engine *e, const uint64_t *d_mhcs, uint64_t n, uint32_t k, const uint32_t *mapping,
                      uint32_t range, uint8_t *thresholds, uint32_t *free_counts, uint64_t *layer_bases, uint32_t *layers,
                      uint64_t *bumped_mhcs, uint64_t *bumped_index, uint64_t *n_bumped) {
    if (!e || !mapping || !layer_bases || !layers || !n_bumped || (n && (!d_mhcs || !bumped_mhcs || !bumped_index)) ||
        (n >= k && (!thresholds || !free_counts)))
        return fail(SHGPU_ERR_ARG, "NULL argument");
    if (hipSetDevice(e->device) != hipSuccess) return fail(SHGPU_ERR_HIP, "hipSetDevice failed");
    std::string err;
    const int rc = shgpu_flat_layers_impl(e->stream, d_mhcs, n, k, mapping, range, thresholds, free_counts, layer_bases,
                                          layers, bumped_mhcs, bumped_index, n_bumped, err, &e->flat.p, &e->flat.bytes);
    return rc == SHGPU_OK ? SHGPU_OK : fail(rc, err);
}

int shgpu_flat_bucket_keys(shgpu_engine *e, const uint64_t *d_mhcs, uint64_t n, uint32_t k, const uint32_t *mapping,
                           uint32_t range, const uint64_t *layer_bases, uint32_t layers, const uint8_t *thresholds,
                           const uint64_t *fb_index, const uint64_t *fb_bucket, uint64_t n_fb, uint64_t *d_bucket_keys) {
    if (!e || !mapping || !layer_bases || (n && !d_mhcs) || (n_fb && (!fb_index || !fb_bucket)) ||
        (n >= k && (!thresholds || !d_bucket_keys)))
        return fail(SHGPU_ERR_ARG, "NULL argument");
    if (hipSetDevice(e->device) != hipSuccess) return fail(SHGPU_ERR_HIP, "hipSetDevice failed");
    std::string err;
    const int rc = shgpu_flat_bucket_keys_impl(e->stream, d_mhcs, n, k, mapping, range, layer_bases, layers, thresholds,
                                               fb_index, fb_bucket, n_fb, d_bucket_keys, err, &e->flat.p, &e->flat.bytes);
    return rc == SHGPU_OK ? SHGPU_OK : fail(rc, err);
}

int shgpu_plan_forest(int variant, uint32_t leaf_size, const uint64_t *size_acc, uint64_t nbuckets, shgpu_node *out,
                      uint64_t capacity, uint64_t *count) {
    if (!size_acc || !count) return fail(SHGPU_ERR_ARG, "NULL argument");
    int rc = check_variant(variant, leaf_size);
    if (rc) return rc;
    const Strategy st(variant, leaf_size);
    const std::vector<uint64_t> bounds = node_bounds(st, DEFAULT_MAX_TRIALS);
    std::vector<shgpu_node> nodes;
    uint32_t roots = 0;
    for (uint64_t k = 0; k < nbuckets; k++) {
        const uint64_t m = size_acc[k + 1] - size_acc[k];
        if (m > MAX_ENCODABLE_BUCKET) return fail(SHGPU_ERR_BUCKET_TOO_LARGE, "bucket too large");
        if (m > 1) { roots++; emit_subtree(nodes, st, (uint32_t)size_acc[k], (uint32_t)m, 0, bounds); }
    }
    if ((rc = check_forest(nodes, roots))) return rc;
    *count = nodes.size();
    if (out) {
        if (capacity < nodes.size()) return fail(SHGPU_ERR_ARG, "capacity too small");
        std::memcpy(out, nodes.data(), nodes.size() * sizeof(shgpu_node));
    }
    return SHGPU_OK;
}

namespace {

// Builds the sample under candidate schedules in two stages: waves per CU in {4, 8, 12, 16, 24} at the default
// seeds per lane (leaf 4, split 2), then every seeds-per-lane pair (leaf {1, 2, 4} for ShockHash1 leaves,
// split {1, 2, 4}) at the fastest wave count. Each candidate runs twice (once when a build takes over a second) and
// must give the first candidate's codes. The variant's tuning is restored afterwards. best: the candidate with the
// shortest search kernel.
static int sweep_schedules(shgpu_engine *e, const shgpu_build_desc *desc, const uint64_t *d_hashes, uint64_t n,
                           shgpu_tuning &best, double &best_ms) {
    int rc = check_variant(desc->variant, desc->leaf_size);
    if (rc) return rc;
    const int v = desc->variant;
    const shgpu_tuning base = e->tuning_v[v];
    shgpu_build_desc d = *desc;
    d.mode = SHGPU_MODE_PERSISTENT;
    d.copy_keys = 0;
    std::vector<uint64_t> ref;
    bool have_ref = false;
    int reps = 2;
    best = base;
    best_ms = 1e30;
    auto run = [&](uint32_t waves, uint32_t leaf_s, uint32_t split_s) -> int {
        shgpu_tuning t = base;
        t.sh1_seeds_per_lane = leaf_s;
        t.split_seeds_per_lane = split_s;
        t.waves_per_cu = waves;
        e->tuning_v[v] = t;
        double ms = 1e30;
        for (int rep = 0; rep < reps; rep++) {
            shgpu_build *b = nullptr;
            int rc2 = shgpu_build_run(e, &d, d_hashes, n, &b);
            if (rc2) return rc2;
            ms = std::min(ms, b->ms[3]);
            const bool same = !have_ref || b->codes == ref;
            if (!have_ref) {
                ref = b->codes;
                have_ref = true;
                if (b->ms[3] > 1000.0) reps = 1;
            }
            shgpu_build_free(b);
            if (!same) return fail(SHGPU_ERR_HIP, "autotune: the schedule (" + std::to_string(waves) + " waves per CU) changed the codes");
        }
        if (ms < best_ms) {
            best_ms = ms;
            best = t;
        }
        return SHGPU_OK;
    };
    const uint32_t waves[5] = {4, 8, 12, 16, 24}, seeds[3] = {1, 2, 4};
    for (uint32_t wv : waves)
        if ((rc = run(wv, 4, 2))) break;
    const uint32_t bw = best.waves_per_cu;
    const int nleaf = v == SHGPU_SHOCKHASH2 ? 1 : 3;  // ShockHash2 leaves are not windowed
    for (int li = 0; li < nleaf && !rc; li++)
        for (uint32_t ss : seeds) {
            const uint32_t ls = nleaf == 1 ? 4 : seeds[li];
            if (ls == 4 && ss == 2) continue;  // timed in the first stage
            if ((rc = run(bw, ls, ss))) break;
        }
    e->tuning_v[v] = base;
    return rc;
}

// Directory of the loaded library, "" when unknown.
static std::string library_dir() {
    Dl_info info;
    if (dladdr((void *)&shgpu_engine_create, &info) && info.dli_fname) {
        const std::string p = info.dli_fname;
        const size_t k = p.rfind('/');
        return k == std::string::npos ? std::string(".") : p.substr(0, k);
    }
    return "";
}

static bool read_text(const std::string &path, std::string &out) {
    std::ifstream f(path, std::ios::binary);
    if (!f) return false;
    std::ostringstream ss;
    ss << f.rdbuf();
    out = ss.str();
    return (bool)f || f.eof();
}

static int load_schedule_file(shgpu_engine *e, const std::string &path) {
    std::string text, err;
    if (!read_text(path, text)) return fail(SHGPU_ERR_ARG, "schedule: cannot read " + path);
    std::vector<shgpu_schedule_class> cls;
    int rc = shgpu::parse_schedule_json(text, cls, err);
    if (!rc) rc = e->sched.replace(cls, err);
    if (rc) return fail(rc, path + ": " + err);
    e->sched.source = path;
    return SHGPU_OK;
}

// SHGPU_SCHEDULE (a path, or "none" for the built-in table), else schedule_<arch>.json beside the library or in its
// parent directory, else the built-in table.
static int load_default_schedule(shgpu_engine *e) {
    const char *env = std::getenv("SHGPU_SCHEDULE");
    if (env && *env) return std::strcmp(env, "none") == 0 ? SHGPU_OK : load_schedule_file(e, env);
    std::string arch = e->arch;
    arch = arch.substr(0, arch.find(':'));
    const std::string dir = library_dir();
    if (dir.empty()) return SHGPU_OK;
    for (const std::string &p : {dir + "/schedule_" + arch + ".json", dir + "/../schedule_" + arch + ".json"}) {
        std::ifstream f(p);
        if (f) return load_schedule_file(e, p);
    }
    return SHGPU_OK;
}

}  // namespace

int shgpu_autotune(shgpu_engine *e, const shgpu_build_desc *desc, const uint64_t *d_hashes, uint64_t n,
                   shgpu_tuning *best, double *best_ms) {
    if (!e || !desc || (n && !d_hashes)) return fail(SHGPU_ERR_ARG, "NULL argument");
    shgpu_tuning t;
    double ms = 0;
    int rc = sweep_schedules(e, desc, d_hashes, n, t, ms);
    if (rc) return rc;
    e->tuning_v[desc->variant] = t;
    if (best) *best = t;
    if (best_ms) *best_ms = ms;
    return SHGPU_OK;
}

int shgpu_autotune_class(shgpu_engine *e, const shgpu_build_desc *desc, const uint64_t *d_hashes, uint64_t n,
                         uint64_t max_keys, shgpu_schedule_class *chosen) {
    if (!e || !desc || (n && !d_hashes)) return fail(SHGPU_ERR_ARG, "NULL argument");
    if (max_keys == 0) return fail(SHGPU_ERR_ARG, "max_keys must be positive");
    shgpu_tuning t;
    double ms = 0;
    int rc = sweep_schedules(e, desc, d_hashes, n, t, ms);
    if (rc) return rc;
    shgpu_schedule_class c{};
    c.variant = (uint32_t)desc->variant;
    c.waves_per_cu = t.waves_per_cu;
    c.sh1_seeds_per_lane = t.sh1_seeds_per_lane;
    c.split_seeds_per_lane = t.split_seeds_per_lane;
    c.max_keys = max_keys;
    c.tuned_ms = ms;
    c.sample_keys = n;
    e->sched.put(c);
    if (chosen) *chosen = c;
    return SHGPU_OK;
}

int shgpu_schedule_get(shgpu_engine *e, shgpu_schedule_class *out, uint32_t capacity, uint32_t *count) {
    if (!e || !count) return fail(SHGPU_ERR_ARG, "NULL argument");
    *count = (uint32_t)e->sched.classes.size();
    if (out) {
        if (capacity < *count) return fail(SHGPU_ERR_ARG, "capacity too small");
        std::memcpy(out, e->sched.classes.data(), *count * sizeof(shgpu_schedule_class));
    }
    return SHGPU_OK;
}

int shgpu_schedule_set(shgpu_engine *e, const shgpu_schedule_class *classes, uint32_t count) {
    if (!e || (count && !classes)) return fail(SHGPU_ERR_ARG, "NULL argument");
    std::string err;
    const int rc = e->sched.replace(std::vector<shgpu_schedule_class>(classes, classes + count), err);
    if (rc) return fail(rc, err);
    e->sched.source = "set";
    return SHGPU_OK;
}

int shgpu_schedule_save(shgpu_engine *e, const char *path) {
    if (!e || !path) return fail(SHGPU_ERR_ARG, "NULL argument");
    std::string arch = e->arch;
    arch = arch.substr(0, arch.find(':'));
    std::ofstream f(path, std::ios::binary | std::ios::trunc);
    if (!f) return fail(SHGPU_ERR_ARG, std::string("schedule: cannot write ") + path);
    f << e->sched.to_json(arch);
    f.close();
    if (!f) return fail(SHGPU_ERR_ARG, std::string("schedule: write failed: ") + path);
    return SHGPU_OK;
}

int shgpu_schedule_load(shgpu_engine *e, const char *path) {
    if (!e || !path) return fail(SHGPU_ERR_ARG, "NULL argument");
    return load_schedule_file(e, path);
}

const char *shgpu_schedule_source(shgpu_engine *e) { return e ? e->sched.source.c_str() : ""; }

int shgpu_schedule_parse(const char *json, shgpu_schedule_class *out, uint32_t capacity, uint32_t *count) {
    if (!json || !count) return fail(SHGPU_ERR_ARG, "NULL argument");
    std::vector<shgpu_schedule_class> cls;
    std::string err;
    const int rc = shgpu::parse_schedule_json(json, cls, err);
    if (rc) return fail(rc, err);
    *count = (uint32_t)cls.size();
    if (out) {
        if (capacity < *count) return fail(SHGPU_ERR_ARG, "capacity too small");
        std::memcpy(out, cls.data(), cls.size() * sizeof(shgpu_schedule_class));
    }
    return SHGPU_OK;
}

// Host memory the runtime does not know as page-locked (or registered).
static bool pageable(const void *p) {
    hipPointerAttribute_t attr{};
    if (hipPointerGetAttributes(&attr, p) != hipSuccess) {
        (void)hipGetLastError();
        return true;
    }
    return attr.type != hipMemoryTypeHost && attr.type != hipMemoryTypeDevice && attr.type != hipMemoryTypeManaged;
}

// Queues the copy of `bytes` host bytes to device memory on the engine's stream; the source may be reused once this
// returns. Large pageable sources go through the uploader's page-locked chunks.
static int upload(shgpu_engine *e, void *dst, const void *src, size_t bytes) {
    if (bytes < UPLOAD_STAGED_MIN || !pageable(src) || !e->up.init(e->device, e->stream)) {
        HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, e->stream));
        return SHGPU_OK;
    }
    if (e->up.copy(dst, src, bytes)) {
        // the whole range again by the runtime's own path (chunks already queued carry the same bytes); a real
        // error (a bad destination) fails here too
        (void)hipGetLastError();
        HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, e->stream));
    }
    return SHGPU_OK;
}

int shgpu_build_run_host(shgpu_engine *e, const shgpu_build_desc *desc, const uint64_t *h_hashes, uint64_t n,
                         shgpu_build **out) {
    if (!e || !desc || !out || (n && !h_hashes)) return fail(SHGPU_ERR_ARG, "NULL argument");
    HIP_TRY(hipSetDevice(e->device));
    int rc = e->staging.ensure(n * 16 + 16);
    if (rc) return rc;
    if (n) {
        rc = upload(e, e->staging.p, h_hashes, n * 16);
        if (rc) return rc;
    }
    return shgpu_build_run(e, desc, e->staging.as<uint64_t>(), n, out);
}

int shgpu_device_memory(shgpu_engine *e, size_t *free_bytes, size_t *total_bytes) {
    if (!e || !free_bytes || !total_bytes) return fail(SHGPU_ERR_ARG, "NULL argument");
    HIP_TRY(hipSetDevice(e->device));
    HIP_TRY(hipMemGetInfo(free_bytes, total_bytes));
    return SHGPU_OK;
}

int shgpu_device_alloc(shgpu_engine *e, size_t bytes, void **d_out) {
    if (!e || !d_out) return fail(SHGPU_ERR_ARG, "NULL argument");
    HIP_TRY(hipSetDevice(e->device));
    if (!device_malloc(d_out, bytes ? bytes : 1)) return fail(SHGPU_ERR_NOMEM, "hipMalloc of " + std::to_string(bytes) + " bytes failed");
    return SHGPU_OK;
}

int shgpu_device_free(shgpu_engine *e, void *d) {
    if (!e) return fail(SHGPU_ERR_ARG, "NULL argument");
    HIP_TRY(hipSetDevice(e->device));
    HIP_TRY(hipStreamSynchronize(e->stream));
    if (d) HIP_TRY(hipFree(d));
    return SHGPU_OK;
}

int shgpu_host_alloc(shgpu_engine *e, size_t bytes, void **h_out) {
    if (!e || !h_out) return fail(SHGPU_ERR_ARG, "NULL argument");
    *h_out = nullptr;
    HIP_TRY(hipSetDevice(e->device));
    if (hipHostMalloc(h_out, bytes ? bytes : 1, hipHostMallocDefault) != hipSuccess) {
        *h_out = nullptr;
        (void)hipGetLastError();
        return fail(SHGPU_ERR_NOMEM, "hipHostMalloc of " + std::to_string(bytes) + " bytes failed");
    }
    return SHGPU_OK;
}

int shgpu_host_free(shgpu_engine *e, void *h) {
    if (!e) return fail(SHGPU_ERR_ARG, "NULL argument");
    HIP_TRY(hipSetDevice(e->device));
    HIP_TRY(hipStreamSynchronize(e->stream));
    if (h) HIP_TRY(hipHostFree(h));
    return SHGPU_OK;
}

int shgpu_copy_to_device(shgpu_engine *e, void *d_dst, const void *h_src, size_t bytes) {
    if (!e || (bytes && (!d_dst || !h_src))) return fail(SHGPU_ERR_ARG, "NULL argument");
    HIP_TRY(hipSetDevice(e->device));
    if (bytes) {
        const int rc = upload(e, d_dst, h_src, bytes);
        if (rc) return rc;
    }
    HIP_TRY(hipStreamSynchronize(e->stream));
    return SHGPU_OK;
}

int shgpu_device_copy(shgpu_engine *e, void *d_dst, const void *d_src, size_t bytes) {
    if (!e || (bytes && (!d_dst || !d_src)) || bytes % 8) return fail(SHGPU_ERR_ARG, "bad argument");
    HIP_TRY(hipSetDevice(e->device));
    if (bytes) {
        k_copy_u64<<<e->cu_count * 8, 256, 0, e->stream>>>((const uint64_t *)d_src, (uint64_t *)d_dst, bytes / 8);
        HIP_TRY(hipGetLastError());
    }
    HIP_TRY(hipStreamSynchronize(e->stream));
    return SHGPU_OK;
}

int shgpu_copy_to_host(shgpu_engine *e, void *h_dst, const void *d_src, size_t bytes) {
    if (!e || (bytes && (!d_src || !h_dst))) return fail(SHGPU_ERR_ARG, "NULL argument");
    HIP_TRY(hipSetDevice(e->device));
    if (bytes) HIP_TRY(hipMemcpyAsync(h_dst, d_src, bytes, hipMemcpyDeviceToHost, e->stream));
    HIP_TRY(hipStreamSynchronize(e->stream));
    return SHGPU_OK;
}

void shgpu_build_free(shgpu_build *b) { delete b; }
uint64_t shgpu_build_nbuckets(const shgpu_build *b) { return b ? b->nbuckets : 0; }
uint64_t shgpu_build_nnodes(const shgpu_build *b) { return b ? b->codes.size() : 0; }

int shgpu_build_bucket_sizes(const shgpu_build *b, uint64_t *size_acc) {
    if (!b || !size_acc) return fail(SHGPU_ERR_ARG, "NULL argument");
    std::memcpy(size_acc, b->size_acc.data(), b->size_acc.size() * 8);
    return SHGPU_OK;
}

int shgpu_build_nodes(const shgpu_build *b, shgpu_node *nodes, uint64_t *codes) {
    if (!b) return fail(SHGPU_ERR_ARG, "NULL argument");
    if (nodes) {
        std::lock_guard<std::mutex> lock(b->nodes_mu);
        if (!b->nodes_ready) {
            int rc = host_plan(*b, b->nodes);
            if (rc) return rc;
            b->nodes_ready = true;
        }
        std::memcpy(nodes, b->nodes.data(), b->nodes.size() * sizeof(shgpu_node));
    }
    if (codes) std::memcpy(codes, b->codes.data(), b->codes.size() * 8);
    return SHGPU_OK;
}

int shgpu_build_keys(const shgpu_build *b, uint64_t *keys) {
    if (!b || !keys) return fail(SHGPU_ERR_ARG, "NULL argument");
    if (b->keys.size() != b->n) return fail(SHGPU_ERR_ARG, "build was run without copy_keys");
    std::memcpy(keys, b->keys.data(), b->n * 8);
    return SHGPU_OK;
}

int shgpu_build_device_keys(const shgpu_build *b, const uint64_t **d_keys) {
    if (!b || !d_keys) return fail(SHGPU_ERR_ARG, "NULL argument");
    if (!b->d_keys) return fail(SHGPU_ERR_ARG, "build was run without SHGPU_KEYS_DEVICE");
    *d_keys = (const uint64_t *)b->d_keys;
    return SHGPU_OK;
}

int shgpu_build_timings(const shgpu_build *b, double *ms4) {
    if (!b || !ms4) return fail(SHGPU_ERR_ARG, "NULL argument");
    for (int i = 0; i < 4; i++) ms4[i] = b->ms[i];
    return SHGPU_OK;
}

}  // extern "C"

// ===================================================================================== serving

namespace {
struct Request {
    uint64_t id = 0;
    const uint64_t *hashes = nullptr;  // caller-owned fingerprints (valid until taken)
    bool on_device = false;            // hashes is a device pointer (already resident in HBM)
    uint64_t n = 0;
    size_t need = 0;                   // device bytes reserved while running
    std::chrono::steady_clock::time_point t_submit, t_start, t_done;
    int status = 0;                    // 0 queued, 1 running, 2 done, < 0: -(error code)
    std::string error;
    shgpu_build *result = nullptr;
};
}  // namespace

struct shgpu_server {
    shgpu_engine *e = nullptr;
    shgpu_server_config cfg{};
    struct Slot {
        BuildBuffers B;
        size_t held = 0;  // B.bytes() after the slot's last request (read and written under mtx)
        hipStream_t stream = nullptr;
        std::thread worker;
    };
    std::vector<std::unique_ptr<Slot>> slots;
    std::mutex mtx;
    std::condition_variable cv_work, cv_done;
    std::deque<Request *> waiting;
    std::unordered_map<uint64_t, Request *> requests;
    uint64_t next_id = 1;
    size_t reserved = 0, peak_reserved = 0;
    uint64_t running = 0;
    bool stop = false;

    // Device bytes a build of n keys holds while it runs: sort buffers and keys (5 x 8 B), staged
    // fingerprints (16 B), radix-sort scratch (about 24 B), forest records (node, code, queue slot,
    // search state) and a fixed margin.
    size_t need_for(uint64_t n) const {
        const uint64_t nodes = n / std::max<uint32_t>(1, cfg.leaf_size) * 2 + 64;
        return n * (5 * 8 + 16 + 24) + nodes * (sizeof(shgpu_node) + 8 + 4 + sizeof(NodeState)) + (8u << 20);
    }
    // Bytes every slot holds regardless of the request: the ShockHash2 per-wave leaf scratch.
    size_t fixed_bytes() const {
        return cfg.variant == SHGPU_SHOCKHASH2 ? cfg.slots * slot_waves() * sh2_layout(cfg.leaf_size).wave_bytes : 0;
    }
    size_t slot_waves() const {
        const int grid = std::max(1, grid_for(e, resolve_tuning(e, cfg.variant, ~0ull)) / (int)cfg.slots);
        return (size_t)(cfg.variant == SHGPU_SHOCKHASH2 ? sh2_grid_cap(sh2_layout(cfg.leaf_size), grid) : grid) * WAVES_PER_BLOCK;
    }
    size_t allocated() const {
        size_t t = 0;
        for (auto &sl : slots) t += sl->held;
        return t;
    }

    uint64_t batch_keys() const { return cfg.max_batch_keys ? cfg.max_batch_keys : (8ull << 20); }
    uint32_t batch_requests() const { return cfg.max_batch_requests ? cfg.max_batch_requests : 256u; }

    // A free slot takes every waiting request, in arrival order, that fits the budget and the batch
    // limits, and builds them together in one persistent kernel.
    void work(Slot &slot) {
        (void)hipSetDevice(e->device);
        std::vector<Request *> batch;
        std::vector<BuildInput> in;
        std::vector<shgpu_build *> res;
        std::vector<int> st;
        for (;;) {
            batch.clear();
            {
                std::unique_lock<std::mutex> lock(mtx);
                cv_work.wait(lock, [&] {
                    return stop || (!waiting.empty() && reserved + waiting.front()->need <= cfg.hbm_budget_bytes);
                });
                if (stop && waiting.empty()) return;
                if (waiting.empty()) continue;
                uint64_t keys = 0;
                const auto now = std::chrono::steady_clock::now();
                while (!waiting.empty() && batch.size() < batch_requests()) {
                    Request *r = waiting.front();
                    if (reserved + r->need > cfg.hbm_budget_bytes) break;
                    if (!batch.empty() && keys + r->n > batch_keys()) break;
                    waiting.pop_front();
                    reserved += r->need;
                    keys += r->n;
                    r->status = 1;
                    r->t_start = now;
                    batch.push_back(r);
                }
                peak_reserved = std::max(peak_reserved, reserved);
                running += batch.size();
            }
            int rc = SHGPU_OK;
            std::string err;
            // stage the requests that came from host memory
            uint64_t staged = 0;
            for (Request *r : batch)
                if (!r->on_device) staged += r->n;
            if (staged) rc = slot.B.staging.ensure(staged * 16 + 16);
            in.clear();
            staged = 0;
            for (Request *r : batch) {
                BuildInput x{r->hashes, r->n};
                if (!r->on_device && rc == SHGPU_OK) {
                    x.d_hashes = slot.B.staging.as<uint64_t>() + 2 * staged;
                    if (r->n && hipMemcpyAsync((void *)x.d_hashes, r->hashes, r->n * 16, hipMemcpyHostToDevice, slot.stream) != hipSuccess)
                        rc = fail(SHGPU_ERR_HIP, "hipMemcpyAsync of request fingerprints failed");
                    staged += r->n;
                }
                in.push_back(x);
            }
            res.assign(batch.size(), nullptr);
            st.assign(batch.size(), rc);
            if (rc == SHGPU_OK) {
                shgpu_build_desc d{};
                d.variant = cfg.variant;
                d.leaf_size = cfg.leaf_size;
                d.bucket_size = cfg.bucket_size;
                d.mode = SHGPU_MODE_PERSISTENT;
                d.copy_keys = cfg.copy_keys;
                rc = run_batch(e, slot.B, slot.stream, cfg.slots, &d, in.data(), in.size(), res.data(), st.data());
            }
            if (rc != SHGPU_OK) err = g_err;
            {
                std::lock_guard<std::mutex> lock(mtx);
                const auto now = std::chrono::steady_clock::now();
                for (size_t i = 0; i < batch.size(); i++) {
                    Request *r = batch[i];
                    r->t_done = now;
                    r->result = res[i];
                    r->status = st[i] == SHGPU_OK ? 2 : -st[i];
                    if (st[i] != SHGPU_OK) r->error = err.empty() ? shgpu_status_string(st[i]) : err;
                    reserved -= r->need;
                }
                running -= batch.size();
                // keep the device footprint within the budget: a slot whose buffers grew beyond what the
                // budget allows gives them back
                slot.held = slot.B.bytes();
                if (allocated() > cfg.hbm_budget_bytes) {
                    slot.B.release();
                    slot.held = 0;
                }
            }
            cv_done.notify_all();
            cv_work.notify_all();
        }
    }
};

extern "C" {

int shgpu_server_create(shgpu_engine *e, const shgpu_server_config *cfg, shgpu_server **out) {
    if (!e || !cfg || !out) return fail(SHGPU_ERR_ARG, "NULL argument");
    *out = nullptr;
    int rc = check_variant(cfg->variant, cfg->leaf_size);
    if (rc) return rc;
    if (cfg->slots < 1 || cfg->slots > 16) return fail(SHGPU_ERR_ARG, "slots must be in [1, 16]");
    if (cfg->bucket_size == 0) return fail(SHGPU_ERR_ARG, "bucket_size must be positive");
    HIP_TRY(hipSetDevice(e->device));
    auto *sv = new shgpu_server();
    sv->e = e;
    sv->cfg = *cfg;
    sv->reserved = sv->peak_reserved = sv->fixed_bytes();
    if (sv->reserved + sv->need_for(0) > cfg->hbm_budget_bytes) {
        const std::string msg = "HBM budget below the slots' fixed scratch (" + std::to_string(sv->reserved) + " bytes)";
        delete sv;
        return fail(SHGPU_ERR_ADMISSION, msg);
    }
    for (uint32_t i = 0; i < cfg->slots; i++) {
        auto sl = std::make_unique<shgpu_server::Slot>();
        const bool stream_ok = hipStreamCreateWithFlags(&sl->stream, hipStreamNonBlocking) == hipSuccess;
        if (!stream_ok || sl->B.create_events()) {
            // no worker has started yet (they start below): release what the slots created so far
            if (stream_ok) (void)hipStreamDestroy(sl->stream);
            sl->B.destroy_events();
            for (auto &x : sv->slots) {
                (void)hipStreamDestroy(x->stream);
                x->B.destroy_events();
            }
            delete sv;
            return fail(SHGPU_ERR_HIP, "stream or event creation failed");
        }
        sv->slots.push_back(std::move(sl));
    }
    for (auto &sl : sv->slots) {
        shgpu_server::Slot *p = sl.get();
        sl->worker = std::thread([sv, p] { sv->work(*p); });
    }
    *out = sv;
    return SHGPU_OK;
}

void shgpu_server_destroy(shgpu_server *sv) {
    if (!sv) return;
    {
        std::lock_guard<std::mutex> lock(sv->mtx);
        sv->stop = true;
        for (Request *r : sv->waiting) { r->status = -SHGPU_ERR_ADMISSION; r->error = "server stopped"; }
        sv->waiting.clear();
    }
    sv->cv_work.notify_all();
    for (auto &sl : sv->slots) sl->worker.join();
    (void)hipSetDevice(sv->e->device);
    for (auto &sl : sv->slots) {
        (void)hipStreamSynchronize(sl->stream);
        (void)hipStreamDestroy(sl->stream);
        sl->B.destroy_events();
        sl->B.release();
    }
    for (auto &kv : sv->requests) { shgpu_build_free(kv.second->result); delete kv.second; }
    delete sv;
}

static int server_submit(shgpu_server *sv, const uint64_t *hashes, bool on_device, uint64_t n, uint64_t *id) {
    if (!sv || !id || (n && !hashes)) return fail(SHGPU_ERR_ARG, "NULL argument");
    auto *r = new Request();
    r->hashes = hashes;
    r->on_device = on_device;
    r->n = n;
    r->need = sv->need_for(n) - (on_device ? n * 16 : 0);
    if (sv->fixed_bytes() + r->need > sv->cfg.hbm_budget_bytes) {
        delete r;
        return fail(SHGPU_ERR_ADMISSION, "request needs more device memory than the server's budget");
    }
    if (n >= (1ull << 32)) {
        delete r;
        return fail(SHGPU_ERR_UNSUPPORTED, "more than 2^32 - 1 keys");
    }
    {
        std::lock_guard<std::mutex> lock(sv->mtx);
        if (sv->stop) {
            delete r;
            return fail(SHGPU_ERR_ADMISSION, "server stopped");
        }
        r->id = sv->next_id++;
        r->t_submit = std::chrono::steady_clock::now();
        sv->requests[r->id] = r;
        sv->waiting.push_back(r);
        *id = r->id;
    }
    sv->cv_work.notify_all();
    return SHGPU_OK;
}

int shgpu_server_submit(shgpu_server *sv, const uint64_t *h_hashes, uint64_t n, uint64_t *id) {
    return server_submit(sv, h_hashes, false, n, id);
}

int shgpu_server_submit_device(shgpu_server *sv, const uint64_t *d_hashes, uint64_t n, uint64_t *id) {
    return server_submit(sv, d_hashes, true, n, id);
}

int shgpu_server_wait(shgpu_server *sv, uint64_t id, double timeout_s, int *status) {
    if (!sv) return fail(SHGPU_ERR_ARG, "NULL argument");
    std::unique_lock<std::mutex> lock(sv->mtx);
    auto it = sv->requests.find(id);
    if (it == sv->requests.end()) return fail(SHGPU_ERR_ARG, "unknown request id");
    Request *r = it->second;
    auto done = [&] { return r->status >= 2 || r->status < 0; };
    if (timeout_s < 0) sv->cv_done.wait(lock, done);
    else sv->cv_done.wait_for(lock, std::chrono::duration<double>(timeout_s), done);
    if (status) *status = r->status;
    return SHGPU_OK;
}

int shgpu_server_take(shgpu_server *sv, uint64_t id, shgpu_build **out, double *latency_ms, double *queue_ms) {
    if (!sv || !out) return fail(SHGPU_ERR_ARG, "NULL argument");
    std::lock_guard<std::mutex> lock(sv->mtx);
    auto it = sv->requests.find(id);
    if (it == sv->requests.end()) return fail(SHGPU_ERR_ARG, "unknown request id");
    Request *r = it->second;
    if (r->status >= 0 && r->status < 2) return fail(SHGPU_ERR_ARG, "request not finished");
    const int st = r->status;
    std::string err = r->error;
    *out = r->result;
    if (latency_ms) *latency_ms = std::chrono::duration<double, std::milli>(r->t_done - r->t_submit).count();
    if (queue_ms) *queue_ms = std::chrono::duration<double, std::milli>(r->t_start - r->t_submit).count();
    sv->requests.erase(it);
    delete r;
    if (st < 0) return fail(-st, err);
    return SHGPU_OK;
}

int shgpu_server_request_bytes(const shgpu_server *sv, uint64_t n, uint64_t *bytes) {
    if (!sv || !bytes) return fail(SHGPU_ERR_ARG, "NULL argument");
    *bytes = sv->need_for(n);
    return SHGPU_OK;
}

int shgpu_server_info(shgpu_server *sv, uint64_t *queued, uint64_t *running, uint64_t *reserved_bytes,
                      uint64_t *peak_reserved_bytes) {
    if (!sv) return fail(SHGPU_ERR_ARG, "NULL argument");
    std::lock_guard<std::mutex> lock(sv->mtx);
    if (queued) *queued = sv->waiting.size();
    if (running) *running = sv->running;
    if (reserved_bytes) *reserved_bytes = sv->reserved;
    if (peak_reserved_bytes) *peak_reserved_bytes = sv->peak_reserved;
    return SHGPU_OK;
}

}  // extern "C"
