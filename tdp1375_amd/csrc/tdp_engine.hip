// Host engine of the MI355X proof-search path: slots, streams, cancellation,
// device memory, launch orchestration and the C ABI (include/tdp1375.h).
//
// Replaces, for one GPU per slot:
//   slot_manager.cu:13-175          slots, per-slot streams, cancel protocol
//   sorting_search.cu:617-699       per-slot buffers, wave count
//   sorting_search.cu:701-856       per-search orchestration and outcomes
//   pi_filter.cu:157-260            filter host side
//   cuda_error.cuh, cuda_memory*.cuh  error taxonomy, OOM retry with backoff
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <sstream>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "../../include/tdp1375.h"
#include "tdp_common.h"
#include "tdp_filter.h"
#include "tdp_canon.h"
#include "tdp_search.h"
#include "tdp_learn.h"

namespace tdp {
namespace {

// ---------------------------------------------------------------- errors
// Same taxonomy as cuda_error.cuh:44-76, mapped to status codes at the C ABI.
class HipException : public std::runtime_error {
public:
    HipException(hipError_t code, const char *file, int line)
        : std::runtime_error(build(code, file, line)) {}

private:
    static std::string build(hipError_t code, const char *file, int line) {
        std::ostringstream oss;
        oss << file << ':' << line << ": HIP error " << hipGetErrorName(code) << " ("
            << hipGetErrorString(code) << ')';
        return oss.str();
    }
};
class GpuMemoryExhausted : public std::runtime_error {
public:
    explicit GpuMemoryExhausted(const std::string &m) : std::runtime_error(m) {}
};
class QueueOverflow : public std::runtime_error {
public:
    explicit QueueOverflow(const std::string &m) : std::runtime_error(m) {}
};
class Cancelled : public std::runtime_error {
public:
    explicit Cancelled(const std::string &m) : std::runtime_error(m) {}
};

#define HIP_CHECK(x)                                                  \
    do {                                                              \
        hipError_t e_ = (x);                                          \
        if (e_ != hipSuccess) throw HipException(e_, __FILE__, __LINE__); \
    } while (0)

thread_local std::string g_last_error;
thread_local double g_filter_ms = 0.0;

int fail(int code, const char *msg) {
    g_last_error = msg;
    return code;
}

template <typename F>
int guarded(F &&f) {
    try {
        f();
        return TDP_OK;
    } catch (const Cancelled &e) {
        return fail(TDP_ERR_CANCELLED, e.what());
    } catch (const QueueOverflow &e) {
        return fail(TDP_ERR_QUEUE_OVERFLOW, e.what());
    } catch (const GpuMemoryExhausted &e) {
        return fail(TDP_ERR_OUT_OF_MEMORY, e.what());
    } catch (const HipException &e) {
        return fail(TDP_ERR_HIP, e.what());
    } catch (const std::invalid_argument &e) {
        return fail(TDP_ERR_INVALID_ARGUMENT, e.what());
    } catch (const std::exception &e) {
        return fail(TDP_ERR_INTERNAL, e.what());
    }
}

// ---------------------------------------------------------------- memory
// cuda_memory.cuh:60-99: retry hipMalloc on out-of-memory with exponential
// backoff (100 ms doubling to 5 s) until the retry timeout has passed
// (gGpuMemoryRetryTimeoutMs, cuda_memory_pressure.cuh:10: 60 s by default,
// negative values act as 0; tdp_set_alloc_timeout).  Retries are counted
// (gGpuMemoryRetries).
std::atomic<int64_t> g_alloc_timeout_ms{60000};
std::atomic<uint64_t> g_alloc_retries{0};

void *device_alloc(size_t bytes, const char *label) {
    if (bytes == 0) return nullptr;
    void *p = nullptr;
    const auto start = std::chrono::steady_clock::now();
    int backoff = 100;
    const int64_t timeout_ms = std::max<int64_t>(g_alloc_timeout_ms.load(std::memory_order_relaxed), 0);
    while (true) {
        const hipError_t st = hipMalloc(&p, bytes);
        if (st == hipSuccess) return p;
        if (st != hipErrorOutOfMemory) HIP_CHECK(st);
        (void)hipGetLastError();
        const auto el = std::chrono::duration_cast<std::chrono::milliseconds>(
                            std::chrono::steady_clock::now() - start).count();
        if (el >= timeout_ms) {
            throw GpuMemoryExhausted(std::string("HIP out of memory while allocating ") +
                                     std::to_string(bytes) + " bytes for " + label);
        }
        g_alloc_retries.fetch_add(1, std::memory_order_relaxed);
        std::this_thread::sleep_for(std::chrono::milliseconds(backoff));
        backoff = std::min(backoff * 2, 5000);
    }
}

struct DevFree {
    int device;
    void operator()(void *p) const noexcept {
        if (!p) return;
        int cur = -1;
        (void)hipGetDevice(&cur);
        if (cur != device) (void)hipSetDevice(device);
        (void)hipFree(p);
        if (cur >= 0 && cur != device) (void)hipSetDevice(cur);
    }
};
using DevPtr = std::unique_ptr<void, DevFree>;

DevPtr make_dev(size_t bytes, const char *label) {
    int dev = 0;
    HIP_CHECK(hipGetDevice(&dev));
    return DevPtr(device_alloc(bytes, label), DevFree{dev});
}

struct HostFree {
    void operator()(void *p) const noexcept {
        if (p) (void)hipHostFree(p);
    }
};
using HostPtr = std::unique_ptr<void, HostFree>;

HostPtr make_pinned(size_t bytes) {
    void *p = nullptr;
    HIP_CHECK(hipHostMalloc(&p, bytes, hipHostMallocDefault));
    return HostPtr(p);
}

template <typename T>
T *as(const DevPtr &p) { return static_cast<T *>(p.get()); }
template <typename T>
T *as(const HostPtr &p) { return static_cast<T *>(p.get()); }

uint32_t floor_pow2(uint64_t v) {
    if (v == 0) return 0;
    uint64_t p = 1;
    while (p * 2 <= v && p < (1ULL << 31)) p *= 2;
    return (uint32_t)p;
}

// ---------------------------------------------------------------- predicates
// Bit tables of the reference's float32 predicates (sorting_search.cu:321-362)
// evaluated on the host: accept[d] bit e <=> satisfies_rate(e, d),
// reach[d] bit e <=> can_still_reach(e, d).  The arithmetic is IEEE binary32
// as on the device (no contraction: each product is rounded before use).
bool satisfies_rate_f32(int e, int d, int n, int e0, float min_rate, bool full) {
    if (d <= 0) return false;
    if (full && e != n) return false;
    const int gained = e - e0;
    volatile float ratio = (float)gained / (float)d;
    return gained > 0 && ratio >= min_rate;
}

bool can_still_reach_f32(int e, int d, int n, int e0, float min_rate, int D, bool full) {
    const int gained = e - e0;
    const int moves_left = D - d;
    if (moves_left < 0) return false;
    const int maxp = n - e;
    if (full) {
        const int lower = (maxp + 1) / 2;
        if (d + lower > D) return false;
        const int optimistic = gained + 2 * moves_left;
        volatile float half = std::ceil((float)maxp / 2.0f);
        volatile float prod = min_rate * ((float)d + half);
        if ((float)optimistic < std::ceil(prod)) return false;
    } else {
        const int optimistic = gained + 2 * moves_left;
        volatile float prod = min_rate * (float)D;
        const int required = (int)std::floor(prod);
        if (optimistic < required) return false;
        const int lower = (maxp + 1) / 2;
        const int best = d + lower;
        if (best > 0) {
            volatile float ratio = (float)(gained + maxp) / (float)best;
            if (ratio < min_rate) return false;
        }
    }
    return true;
}

void predicate_tables(int n, int e0, float min_rate, int D, bool full, uint64_t *acc, uint64_t *rch) {
    for (int d = 0; d < 16; ++d) {
        uint64_t a = 0, r = 0;
        for (int e = 0; e <= n && e < 64; ++e) {
            if (satisfies_rate_f32(e, d, n, e0, min_rate, full)) a |= 1ULL << e;
            if (can_still_reach_f32(e, d, n, e0, min_rate, D, full)) r |= 1ULL << e;
        }
        acc[d] = a;
        rch[d] = r;
    }
}

int count_odd_cycles(const uint8_t *spi, int n) {
    uint64_t seen = 0;
    int c = 0;
    for (int s = 0; s < n; ++s) {
        if (seen >> s & 1) continue;
        int len = 0, x = s;
        do {
            seen |= 1ULL << x;
            len++;
            x = spi[x];
        } while (x != s);
        c += len & 1;
    }
    return c;
}

// ---------------------------------------------------------------- triples
struct TripleTable {
    std::vector<uint32_t> packed;   // all n in 3..63, concatenated
    uint32_t offset[64] = {0};
    TripleTable() {
        for (int n = 0; n <= kMaxN; ++n) {
            offset[n] = (uint32_t)packed.size();
            for (int i = 0; i < n - 2; ++i)
                for (int j = i + 1; j < n - 1; ++j)
                    for (int k = j + 1; k < n; ++k) packed.push_back(pack_triple(i, j, k));
        }
    }
    void unpack(int n, uint32_t t, int32_t *ijk) const {
        const uint32_t p = packed[offset[n] + t];
        ijk[0] = p & 0xFF;
        ijk[1] = (p >> 8) & 0xFF;
        ijk[2] = (p >> 16) & 0xFF;
    }
};
const TripleTable &triples() {
    static TripleTable t;
    return t;
}

struct DeviceShared {

    int waves_max = 0;           // resident waves of search_kernel
    int waves_max_learned = 0;   // resident waves of search_learned_kernel
    int cus = 0;
};
std::mutex g_dev_mutex;
std::vector<std::unique_ptr<DeviceShared>> g_dev;

DeviceShared &device_shared(int device) {
    std::lock_guard<std::mutex> lk(g_dev_mutex);
    if ((size_t)device >= g_dev.size()) g_dev.resize(device + 1);
    if (!g_dev[device]) {
        auto d = std::make_unique<DeviceShared>();



        HIP_CHECK(hipDeviceGetAttribute(&d->cus, hipDeviceAttributeMultiprocessorCount, device));
        int blocks = 0;
        HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, search_kernel, 256, 0));
        d->waves_max = std::max(1, blocks) * d->cus * 4;
        blocks = 0;
        HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, search_learned_kernel, 256, 0));
        d->waves_max_learned = std::max(1, blocks) * d->cus * 4;
        g_dev[device] = std::move(d);
    }
    return *g_dev[device];
}

// ---------------------------------------------------------------- learned ordering
// One move-ordering model per process, on the device that was current when it
// was first set: fp32 master parameters, Adam moments and the bf16 copies the
// MFMA passes read (no reference counterpart).
struct LearnedModel {
    int device = -1;
    DevPtr params, grads, m, v, w1t, b1, w2, b2, loss;
    bool ready = false;
    int step = 0;
};
std::mutex g_model_mutex;
LearnedModel g_model;

void ensure_model_buffers() {
    if (g_model.device >= 0) {
        HIP_CHECK(hipSetDevice(g_model.device));
        return;
    }
    int dev = 0;
    HIP_CHECK(hipGetDevice(&dev));
    g_model.device = dev;
    g_model.params = make_dev(kMlpParams * 4, "mlp.params");
    g_model.grads = make_dev(kMlpParams * 4, "mlp.grads");
    g_model.m = make_dev(kMlpParams * 4, "mlp.adam");
    g_model.v = make_dev(kMlpParams * 4, "mlp.adam");
    g_model.w1t = make_dev(kIn * kHidden * 2, "mlp.w1");
    g_model.b1 = make_dev(kHidden * 4, "mlp.b1");
    g_model.w2 = make_dev(kHidden * 2, "mlp.w2");
    g_model.b2 = make_dev(4, "mlp.b2");
    g_model.loss = make_dev(4, "mlp.loss");
}

MlpWeights model_weights() {
    MlpWeights M;
    M.w1t = as<__bf16>(g_model.w1t);
    M.b1 = as<float>(g_model.b1);
    M.w2 = as<__bf16>(g_model.w2);
    M.b2 = as<float>(g_model.b2);
    return M;
}

void model_set(const float *w1, const float *b1, const float *w2, float b2) {
    std::vector<float> p(kMlpParams);
    std::memcpy(p.data(), w1, sizeof(float) * kIn * kHidden);
    std::memcpy(p.data() + kIn * kHidden, b1, sizeof(float) * kHidden);
    std::memcpy(p.data() + kIn * kHidden + kHidden, w2, sizeof(float) * kHidden);
    p[kMlpParams - 1] = b2;
    ensure_model_buffers();
    HIP_CHECK(hipMemcpy(g_model.params.get(), p.data(), kMlpParams * 4, hipMemcpyHostToDevice));
    HIP_CHECK(hipMemset(g_model.m.get(), 0, kMlpParams * 4));
    HIP_CHECK(hipMemset(g_model.v.get(), 0, kMlpParams * 4));
    hipLaunchKernelGGL(mlp_publish_kernel, dim3((kMlpParams + 255) / 256), dim3(256), 0, 0, as<float>(g_model.params),
                       as<__bf16>(g_model.w1t), as<float>(g_model.b1), as<__bf16>(g_model.w2), as<float>(g_model.b2));
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipDeviceSynchronize());
    g_model.ready = true;
    g_model.step = 0;
}

void require_model() {
    if (!g_model.ready) throw std::invalid_argument("no move-ordering model is set (tdp_mlp_set_weights)");
    HIP_CHECK(hipSetDevice(g_model.device));
}

// Device copies of a row set (features, labels, group starts).
struct RowSet {
    DevPtr X, labels, gstart, logits, hpre, dy;
    int rows = 0, groups = 0;
};

RowSet upload_rows(const uint32_t *X, const uint8_t *labels, const int32_t *gstart, int N, int G) {
    RowSet r;
    r.rows = N;
    r.groups = G;
    r.X = make_dev((size_t)std::max(N, 1) * 16, "mlp.rows");
    HIP_CHECK(hipMemcpy(r.X.get(), X, (size_t)N * 16, hipMemcpyHostToDevice));
    r.logits = make_dev((size_t)std::max(N, 1) * 4, "mlp.logits");
    if (labels) {
        r.labels = make_dev((size_t)std::max(N, 1), "mlp.labels");
        HIP_CHECK(hipMemcpy(r.labels.get(), labels, (size_t)N, hipMemcpyHostToDevice));
        r.gstart = make_dev((size_t)(G + 1) * 4, "mlp.groups");
        HIP_CHECK(hipMemcpy(r.gstart.get(), gstart, (size_t)(G + 1) * 4, hipMemcpyHostToDevice));
        r.hpre = make_dev((size_t)std::max(N, 1) * kHidden * 4, "mlp.hidden");
        r.dy = make_dev((size_t)std::max(N, 1) * 4, "mlp.dy");
    }
    return r;
}

// forward + loss + backward over rows [r0, r0 + n) holding groups
// [g0, g0 + ng) (group starts relative to the set); gradients into g_model.grads
void model_backward(const RowSet &rs, int r0, int n, int g0, int ng, float scale, float loss_scale) {
    const MlpWeights M = model_weights();
    const uint32_t *X = as<uint32_t>(rs.X) + (size_t)r0 * 4;
    float *hpre = as<float>(rs.hpre) + (size_t)r0 * kHidden;
    float *logits = as<float>(rs.logits);
    float *dy = as<float>(rs.dy);
    const int waves = (n + 15) / 16;
    hipLaunchKernelGGL(mlp_forward_kernel, dim3((waves + 3) / 4), dim3(256), 0, 0, X, n, M, hpre, logits + r0);
    hipLaunchKernelGGL(group_softmax_kernel, dim3((ng + 255) / 256), dim3(256), 0, 0, logits,
                       as<uint8_t>(rs.labels), as<int32_t>(rs.gstart) + g0, ng, scale, loss_scale, dy,
                       as<float>(g_model.loss));
    float *g = as<float>(g_model.grads);
    HIP_CHECK(hipMemsetAsync(g, 0, (size_t)kIn * kHidden * 4, 0));
    const int rows_per_chunk = 4096;
    const int chunks = (n + rows_per_chunk - 1) / rows_per_chunk;
    hipLaunchKernelGGL(mlp_dw1_kernel, dim3(128 * chunks / 4), dim3(256), 0, 0, X, n, hpre, dy + r0, M.w2,
                       rows_per_chunk, g);
    hipLaunchKernelGGL(mlp_column_kernel, dim3(kHidden), dim3(256), 0, 0, n, hpre, dy + r0, M.w2,
                       g + kIn * kHidden, g + kIn * kHidden + kHidden, g + kMlpParams - 1);
    HIP_CHECK(hipGetLastError());
}

uint64_t splitmix(uint64_t &s) {
    uint64_t z = (s += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

// Launch tuning of the work-sharing mode (tdp_set_tuning); 0 = default.
struct Tuning {
    int waves_per_cu = 0;     // default: the occupancy limit of search_kernel
    int share_interval = 8;   // expansions between donation checks (power of two)
    int share_max = 8;        // nodes per donation
    int idle_backoff = 16;    // longest idle sleep, x 127 s_sleep units
    int max_active = 1024;    // configurations per persistent launch
    int deadline_ms = 1000;   // TDP_FLAG_SOFT_DEADLINE: time budget of each launch
    int batch = 12;           // nodes a wave expands before one dedup pass (plain ordering)
};
std::mutex g_tuning_mutex;
Tuning g_tuning;

// Nodes per expansion batch: the tuning's, at most what the kernel's LDS node
// pool holds at this node size, and at most kLearnedBatchMax with the learned
// ordering (which keeps each batch node's cycle table).
int batch_nodes(int batch, int node_words, bool learned) {
    int b = std::min(batch, kNodePoolWords / node_words);
    if (learned) b = std::min(b, kLearnedBatchMax);
    return std::max(b, 1);
}

Tuning current_tuning() {
    std::lock_guard<std::mutex> lk(g_tuning_mutex);
    return g_tuning;
}

// ---------------------------------------------------------------- slots
struct CaseInput {
    const uint8_t *pi, *spi;
    int n, e0, D, full;
    float min_rate;
};

struct Slot {
    int device = -1;
    hipStream_t stream = nullptr;
    hipStream_t cancel_stream = nullptr;
    DevPtr cancel_flag;        // int on the device, polled by the kernel
    HostPtr cancel_one;        // pinned int holding 1 (slot_manager.cu:86-87)
    std::atomic<bool> cancelled{false};
    std::atomic<int64_t> serve_ticket{-1};   // serving: request << 16 | slot of the running search
    std::mutex run_mutex;      // one search at a time per slot

    int64_t queue_budget = 0;
    int64_t dedup_size = 0;
    uint64_t dd_buckets = 0;   // dedup_size / dd_width lines of kDedupBucket tags
    uint32_t dd_width = 0;     // tags used per line: min(8, dedup_size)
    uint32_t epoch = 0;

    DevPtr queue_mem;          // stacks + inbox nodes + spill pool nodes
    size_t queue_bytes = 0;
    DevPtr pool_seq;
    size_t pool_seq_cap = 0;
    DevPtr share_lines;        // inbox lines + idle bits (parallel mode)
    size_t share_lines_cap = 0;
    DevPtr ctrl;
    DevPtr dd_tag, dd_key;

    DevPtr cfgs, results;
    HostPtr h_cfgs, h_results, h_ctrl;
    size_t cfg_cap = 0;

    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    hipGraphExec_t graph_exec = nullptr;   // TDP_FLAG_GRAPH: one launch round, updated per round
    int graph_kernel = -1;                 // search kernel in graph_exec: 0 plain, 1 learned
    tdp_launch_stats stats{};

    // configurations of the last batch, reused by TDP_FLAG_RESIDENT
    int last_batch_count = -1;
    std::vector<CaseInput> last_cases;
    std::vector<int> last_index;
    std::vector<uint8_t> undecided;   // per case of the last launch (soft deadline)
};

std::mutex g_slots_mutex;
std::vector<std::unique_ptr<Slot>> g_slots;

Slot &get_slot(int slot) {
    std::lock_guard<std::mutex> lk(g_slots_mutex);
    if (slot < 0 || (size_t)slot >= g_slots.size() || !g_slots[slot]) throw std::invalid_argument("invalid slot");
    return *g_slots[slot];
}

void destroy_slot(Slot &s) {
    if (s.device >= 0) (void)hipSetDevice(s.device);
    if (s.stream) (void)hipStreamSynchronize(s.stream);
    if (s.cancel_stream) (void)hipStreamSynchronize(s.cancel_stream);
    s.queue_mem.reset();
    s.share_lines.reset();
    s.pool_seq.reset();
    s.ctrl.reset();
    s.dd_tag.reset();
    s.dd_key.reset();
    s.cfgs.reset();
    s.results.reset();
    s.cancel_flag.reset();
    s.h_cfgs.reset();
    s.h_results.reset();
    s.h_ctrl.reset();
    s.cancel_one.reset();
    if (s.graph_exec) (void)hipGraphExecDestroy(s.graph_exec);
    s.graph_exec = nullptr;
    if (s.ev0) (void)hipEventDestroy(s.ev0);
    if (s.ev1) (void)hipEventDestroy(s.ev1);
    if (s.stream) (void)hipStreamDestroy(s.stream);
    if (s.cancel_stream) (void)hipStreamDestroy(s.cancel_stream);
    s.stream = s.cancel_stream = nullptr;
    s.ev0 = s.ev1 = nullptr;
}

void destroy_all_slots() {
    std::lock_guard<std::mutex> lk(g_slots_mutex);
    int cur = -1;
    (void)hipGetDevice(&cur);
    for (auto &s : g_slots)
        if (s) destroy_slot(*s);
    g_slots.clear();
    if (cur >= 0) (void)hipSetDevice(cur);
}

void serve_stop();
void serve_start(int64_t queue_bytes, int64_t dedup, int slots, int max_n, uint32_t flags, int lifetime_s);

// TDP_SERVE=1 in the environment of initSlots: the slots become handles of
// one serving session on device 0 that holds all of their queue memory, so
// concurrent searchForSortingSeq calls share one persistent launch.
bool serving_requested() {
    const char *e = std::getenv("TDP_SERVE");
    return e && std::atoi(e) != 0;
}

void init_slots(int total, int devices, int64_t budget, int64_t dedup) {
    if (total <= 0) throw std::invalid_argument("totalSlots must be positive");
    if (devices <= 0) throw std::invalid_argument("devicesCount must be positive");
    if (budget < 0) throw std::invalid_argument("queueBytesBudget must be non-negative");
    serve_stop();
    destroy_all_slots();
    const bool serving = serving_requested();
    int count = 0;
    HIP_CHECK(hipGetDeviceCount(&count));
    std::lock_guard<std::mutex> lk(g_slots_mutex);
    int cur = -1;
    (void)hipGetDevice(&cur);
    // a slot that fails half-way (e.g. out of memory) is torn down before
    // the error propagates, and so are the slots already built: a failed
    // initSlots leaves no slots, as destroySlots would
    std::unique_ptr<Slot> s;
    try {
    for (int i = 0; i < total; ++i) {
        s = std::make_unique<Slot>();
        s->device = i % devices;
        HIP_CHECK(hipSetDevice(s->device));
        HIP_CHECK(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
        HIP_CHECK(hipStreamCreateWithFlags(&s->cancel_stream, hipStreamNonBlocking));
        HIP_CHECK(hipEventCreate(&s->ev0));
        HIP_CHECK(hipEventCreate(&s->ev1));
        s->cancel_flag = make_dev(sizeof(int), "sort.cancelFlag");
        HIP_CHECK(hipMemset(s->cancel_flag.get(), 0, sizeof(int)));
        s->cancel_one = make_pinned(sizeof(int));
        *as<int>(s->cancel_one) = 1;
        s->queue_budget = serving ? 0 : budget;
        s->dedup_size = serving ? 0 : std::max<int64_t>(0, dedup);
        (void)device_shared(s->device);
        s->queue_bytes = serving ? 0 : (size_t)budget;
        s->queue_mem = make_dev(std::max<size_t>(s->queue_bytes, 256), "sort.nodes");
        s->ctrl = make_dev(sizeof(ControlLine) * kNumControl, "sort.control");
        s->h_ctrl = make_pinned(sizeof(ControlLine) * kNumControl);
        if (s->dedup_size > 0) {
            // dedupTableSize entries as lines of 8 tags (a size that is not a
            // multiple of 8 loses its last partial line; below 8 entries one
            // line holds them all)
            s->dd_width = (uint32_t)std::min<int64_t>(kDedupBucket, s->dedup_size);
            s->dd_buckets = (uint64_t)s->dedup_size / s->dd_width;
            const size_t lines = (size_t)s->dd_buckets;
            s->dd_tag = make_dev(lines * kDedupBucket * 8, "sort.dedupTable");
            s->dd_key = make_dev(lines * kDedupBucket * 64, "sort.dedupTable");
            HIP_CHECK(hipMemset(s->dd_tag.get(), 0, lines * kDedupBucket * 8));
        }
        s->epoch = 0;
        g_slots.push_back(std::move(s));
    }
    if (serving) {
        // the session holds what the slots would have held: their queue
        // budgets and their dedup tables (each reference slot has its own)
        int lifetime = 86400, inflight = 256;
        if (const char *e = std::getenv("TDP_SERVE_LIFETIME_S")) lifetime = std::max(1, std::atoi(e));
        if (const char *e = std::getenv("TDP_SERVE_SLOTS")) inflight = std::max(1, std::min(65536, std::atoi(e)));
        HIP_CHECK(hipSetDevice(0));
        serve_start(budget * (int64_t)total, std::max<int64_t>(0, dedup) * (int64_t)total, inflight, kMaxN, 0u,
                    lifetime);
    }
    } catch (...) {
        if (s) destroy_slot(*s);
        for (auto &o : g_slots)
            if (o) destroy_slot(*o);
        g_slots.clear();
        if (cur >= 0) (void)hipSetDevice(cur);
        throw;
    }
    if (cur >= 0) (void)hipSetDevice(cur);
}

void ensure_config_capacity(Slot &s, size_t count) {
    if (count <= s.cfg_cap) return;
    size_t cap = std::max<size_t>(count, 64);
    s.cfgs = make_dev(cap * sizeof(ConfigDesc), "sort.configs");
    s.results = make_dev(cap * sizeof(ConfigResult), "sort.results");
    s.h_cfgs = make_pinned(cap * sizeof(ConfigDesc));
    s.h_results = make_pinned(cap * sizeof(ConfigResult));
    s.cfg_cap = cap;
}


void validate_case(const CaseInput &c) {
    if (c.n > kMaxN) throw std::invalid_argument("performGpuSortingSearch: n exceeds MAX_N");
    if (c.D <= 0 || c.D > kMaxDepth) throw std::invalid_argument("performGpuSortingSearch: invalid maximumDepth");
    if (c.n < 0) throw std::invalid_argument("performGpuSortingSearch: inconsistent input sizes");
    // The kernel indexes LDS tables by symbol: both inputs must be permutations
    // of 0..n-1 (the reference leaves this undefined on the device).
    uint64_t seen_pi = 0, seen_spi = 0;
    for (int i = 0; i < c.n; ++i) {
        if (c.pi[i] >= c.n || c.spi[i] >= c.n)
            throw std::invalid_argument("performGpuSortingSearch: pi and spi must be permutations of 0..n-1");
        seen_pi |= 1ULL << c.pi[i];
        seen_spi |= 1ULL << c.spi[i];
    }
    const uint64_t full_mask = (c.n == 64) ? ~0ULL : ((1ULL << c.n) - 1ULL);
    if (seen_pi != full_mask || seen_spi != full_mask)
        throw std::invalid_argument("performGpuSortingSearch: pi and spi must be permutations of 0..n-1");
}

void fill_config(ConfigDesc &d, const CaseInput &c) {
    std::memset(&d, 0, sizeof(d));
    d.n = (uint32_t)c.n;
    d.max_depth = (uint32_t)c.D;
    d.e_start = (uint32_t)count_odd_cycles(c.spi, c.n);
    d.triple_count = (uint32_t)triple_count(c.n);
    d.flags = c.full ? 1u : 0u;
    predicate_tables(c.n, c.e0, c.min_rate, c.D, c.full != 0, d.accept, d.reach);
    std::memcpy(d.spi0, c.spi, (size_t)c.n);
    std::memcpy(d.pi0, c.pi, (size_t)c.n);
}

bool cancel_requested(Slot &s) { return s.cancelled.load(std::memory_order_acquire); }

// Device watchdog: a launch that runs longer than TDP_KERNEL_TIMEOUT_S seconds
// (default 3600) stops itself and reports an error instead of running on.
unsigned long long kernel_timeout_ticks() {
    double secs = 3600.0;
    if (const char *e = std::getenv("TDP_KERNEL_TIMEOUT_S")) secs = std::atof(e);
    if (secs <= 0) return 0ULL;
    return (unsigned long long)(secs * 1e8);   // s_memrealtime runs at 100 MHz
}

// One persistent launch over `cases` on slot s.  Returns per-case results in
// s.h_results; throws on cancel / overflow.
void run_launch(Slot &s, const std::vector<CaseInput> &cases, uint32_t flags, bool single_search) {
    const int count = (int)cases.size();
    const bool resident = (flags & TDP_FLAG_RESIDENT) != 0;
    const bool learned = (flags & TDP_FLAG_LEARNED) != 0;
    const bool soft = (flags & TDP_FLAG_SOFT_DEADLINE) != 0;
    const bool graph = (flags & TDP_FLAG_GRAPH) != 0;
    s.undecided.assign((size_t)count, 0);
    HIP_CHECK(hipSetDevice(s.device));
    DeviceShared &dev = device_shared(s.device);
    ensure_config_capacity(s, (size_t)count);
    int max_n = 3;
    for (int c = 0; c < count; ++c) {
        if (!resident) fill_config(as<ConfigDesc>(s.h_cfgs)[c], cases[c]);
        max_n = std::max(max_n, cases[c].n);
    }
    const uint32_t nw = (uint32_t)node_words_for(max_n);
    const size_t node_bytes = (size_t)nw * 4;

    int waves;
    bool donate;
    // every wave of a persistent launch must be resident at once: the grid
    // is the launched kernel's occupancy limit at most
    const int resident_waves = (flags & TDP_FLAG_LEARNED) ? dev.waves_max_learned : dev.waves_max;
    if (flags & TDP_FLAG_SINGLE_WAVE) {
        waves = 1;
        donate = false;
    } else if (flags & TDP_FLAG_DETERMINISTIC) {
        waves = std::min(count, resident_waves);
        donate = false;
    } else {
        waves = resident_waves;
        donate = true;
    }
    // queue memory (queueBytesBudget, sorting_search.cu:655-658): the
    // parallel mode sets aside one inbox of share_max nodes per wave and a
    // quarter of the rest for the spill pool; the remainder is split evenly
    // into private stacks (powers of two)
    const Tuning tune = current_tuning();
    if (!(flags & TDP_FLAG_SINGLE_WAVE) && tune.waves_per_cu > 0)
        waves = std::min(waves, std::max(1, tune.waves_per_cu * dev.cus));
    uint32_t share_max = donate ? (uint32_t)tune.share_max : 0;
    // the inboxes take at most a quarter of the budget: smaller hand-offs
    // first, and a budget too small even for one node each runs without
    // sharing
    while (share_max > 1 && (size_t)waves * share_max * node_bytes * 4 > s.queue_bytes) share_max /= 2;
    size_t inbox_bytes = (size_t)waves * share_max * node_bytes;
    if (donate && inbox_bytes * 4 > s.queue_bytes) {
        donate = false;
        share_max = 0;
        inbox_bytes = 0;
    }
    const size_t pool_bytes = donate ? (s.queue_bytes - inbox_bytes) / 4 : 0;
    const size_t stack_bytes = s.queue_bytes - inbox_bytes - pool_bytes;
    uint32_t stack_cap = floor_pow2(stack_bytes / ((size_t)waves * node_bytes));
    if (stack_cap < 2) {
        // the budget cannot hold two nodes per wave: fewer waves
        waves = std::max<int>(1, (int)(stack_bytes / (2 * node_bytes)));
        stack_cap = floor_pow2(stack_bytes / ((size_t)std::max(1, waves) * node_bytes));
    }
    if (stack_cap < 2) throw QueueOverflow("GPU sorting search queue overflow");
    if (share_max > stack_cap / 2) share_max = stack_cap / 2;
    if (share_max == 0) donate = false;
    uint32_t pool_cap = donate ? floor_pow2(pool_bytes / node_bytes) : 0;
    if (pool_cap < 64) pool_cap = 0;
    if (pool_cap && s.pool_seq_cap < pool_cap) {
        s.pool_seq = make_dev((size_t)pool_cap * 8, "sort.poolSeq");
        s.pool_seq_cap = pool_cap;
    }
    const uint32_t idle_words = (uint32_t)((waves + 63) / 64);
    const size_t lines_needed = ((size_t)waves + idle_words) * 128;
    if (donate && s.share_lines_cap < lines_needed) {
        s.share_lines = make_dev(lines_needed, "sort.shareLines");
        s.share_lines_cap = lines_needed;
    }

    SearchParams P{};
    P.cfgs = as<ConfigDesc>(s.cfgs);
    P.results = as<ConfigResult>(s.results);
    P.num_configs = count;
    P.num_waves = waves;
    P.stacks = as<uint32_t>(s.queue_mem);
    P.stack_cap = stack_cap;
    P.node_words = nw;
    P.spi_words = (uint32_t)spi_words_for(max_n);
    P.inbox_nodes = donate ? P.stacks + (size_t)waves * stack_cap * nw : nullptr;
    P.inbox = donate ? as<unsigned long long>(s.share_lines) : nullptr;
    P.idle_bits = donate ? P.inbox + (size_t)waves * 16 : nullptr;
    P.idle_words = donate ? idle_words : 0;
    // pool after the stacks and inboxes; pool_cap 1 with no pool: every
    // spill attempt finds it full
    P.pool_nodes = P.stacks + (size_t)waves * (stack_cap + (donate ? share_max : 0)) * nw;
    P.pool_seq = pool_cap ? as<unsigned long long>(s.pool_seq) : nullptr;
    P.pool_cap = pool_cap ? pool_cap : 1;
    P.ctrl = as<ControlLine>(s.ctrl);
    P.dd_tag = as<unsigned long long>(s.dd_tag);
    P.dd_key = as<uint32_t>(s.dd_key);
    P.dd_buckets = s.dd_buckets;
    P.dd_width = s.dd_width;
    P.epoch = s.epoch ? s.epoch : 1;
    P.cancel_flag = as<int>(s.cancel_flag);
    P.donate = donate ? 1 : 0;
    P.share_mask = (uint32_t)tune.share_interval - 1;
    // the fixed expansion order of the deterministic modes takes one node at
    // a time; a batch fits the kernel's LDS node pool (and, with the learned
    // ordering, its per-node cycle tables)
    P.batch = donate ? batch_nodes(tune.batch, nw, learned) : 1;
    P.share_max = share_max;
    P.idle_backoff = (uint32_t)tune.idle_backoff;
    P.timeout_ticks = soft ? (unsigned long long)tune.deadline_ms * 100000ULL : kernel_timeout_ticks();
    if (learned) {
        std::lock_guard<std::mutex> lk(g_model_mutex);
        if (!g_model.ready) throw std::invalid_argument("no move-ordering model is set (tdp_mlp_set_weights)");
        if (g_model.device != s.device) throw std::invalid_argument("the move-ordering model is on another device");
        P.mlp = model_weights();
    }

    if (!resident)
        HIP_CHECK(hipMemcpyAsync(s.cfgs.get(), s.h_cfgs.get(), (size_t)count * sizeof(ConfigDesc),
                                 hipMemcpyHostToDevice, s.stream));

    // Admission control (work-sharing mode): at most max_active
    // configurations per persistent launch, launched back to back.  All
    // configurations of a launch share the dedup table, and the entries of a
    // finished round are freed at once by the next epoch, so bounding the
    // configurations in flight keeps the table from thrashing.
    const int round = (donate && tune.max_active > 0) ? std::min(count, tune.max_active) : count;
    tdp_launch_stats acc{};
    acc.waves = waves;
    acc.node_bytes = node_bytes;
    acc.key_bytes = (uint64_t)spi_words_for(max_n) * 4;
    int stop_flag = 0;
    bool cancelled = false;
    const int blocks = (waves + 3) / 4;
    for (int base = 0; base < count; base += round) {
        const int sub = std::min(round, count - base);
        // dedup epoch (replaces the per-search table memset, sorting_search.cu:744-747)
        if (s.dedup_size > 0) {
            s.epoch++;
            if (s.epoch > 0xFFFF) {
                HIP_CHECK(hipMemsetAsync(s.dd_tag.get(), 0, (size_t)s.dd_buckets * kDedupBucket * 8, s.stream));
                s.epoch = 1;
            }
        }
        P.cfgs = as<ConfigDesc>(s.cfgs) + base;
        P.results = as<ConfigResult>(s.results) + base;
        P.num_configs = sub;
        P.epoch = s.epoch ? s.epoch : 1;
        // one thread per configuration, wave (inbox line), pool slot and control line
        const size_t reset_n = std::max<size_t>({(size_t)sub, (size_t)waves, (size_t)pool_cap, (size_t)kNumControl});
        const auto launch_search = [&] {
            if (learned)
                hipLaunchKernelGGL(search_learned_kernel, dim3(blocks), dim3(256), 0, s.stream, P);
            else
                hipLaunchKernelGGL(search_kernel, dim3(blocks), dim3(256), 0, s.stream, P);
            HIP_CHECK(hipGetLastError());
        };
        if (graph) {
            // The round (reset kernel, search kernel, control-line copy) as a
            // graph: captured each round and applied to the slot's executable
            // graph by hipGraphExecUpdate (same topology, new parameters), so
            // it is instantiated once per slot (again when the search kernel
            // function changes).  The graph holds no memset node.  The events
            // bracket the whole graph, so in this mode kernel_ms is the time
            // of the round (reset kernel + search kernel + control copy).
            HIP_CHECK(hipStreamBeginCapture(s.stream, hipStreamCaptureModeThreadLocal));
            hipGraph_t g = nullptr;
            try {
                hipLaunchKernelGGL(reset_kernel, dim3((unsigned)((reset_n + 255) / 256)), dim3(256), 0, s.stream, P);
                HIP_CHECK(hipGetLastError());
                launch_search();
                HIP_CHECK(hipMemcpyAsync(s.h_ctrl.get(), s.ctrl.get(), sizeof(ControlLine) * kNumControl,
                                         hipMemcpyDeviceToHost, s.stream));
            } catch (...) {
                (void)hipStreamEndCapture(s.stream, &g);
                if (g) (void)hipGraphDestroy(g);
                throw;
            }
            HIP_CHECK(hipStreamEndCapture(s.stream, &g));
            bool updated = false;
            if (s.graph_exec && s.graph_kernel != (learned ? 1 : 0)) {
                // the search node runs another kernel function: instantiate
                // anew rather than rely on an update changing the function
                (void)hipGraphExecDestroy(s.graph_exec);
                s.graph_exec = nullptr;
            }
            if (s.graph_exec) {
                hipGraphNode_t bad = nullptr;
                hipGraphExecUpdateResult ur = hipGraphExecUpdateError;
                updated = hipGraphExecUpdate(s.graph_exec, g, &bad, &ur) == hipSuccess &&
                          ur == hipGraphExecUpdateSuccess;
                if (!updated) {
                    (void)hipGetLastError();
                    (void)hipGraphExecDestroy(s.graph_exec);
                    s.graph_exec = nullptr;
                }
            }
            hipError_t ie = hipSuccess;
            if (!updated) ie = hipGraphInstantiate(&s.graph_exec, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if (ie != hipSuccess) s.graph_exec = nullptr;
            HIP_CHECK(ie);
            s.graph_kernel = learned ? 1 : 0;
            HIP_CHECK(hipEventRecord(s.ev0, s.stream));
            HIP_CHECK(hipGraphLaunch(s.graph_exec, s.stream));
            HIP_CHECK(hipEventRecord(s.ev1, s.stream));
        } else {
            hipLaunchKernelGGL(reset_kernel, dim3((unsigned)((reset_n + 255) / 256)), dim3(256), 0, s.stream, P);
            HIP_CHECK(hipGetLastError());
            HIP_CHECK(hipEventRecord(s.ev0, s.stream));
            launch_search();
            HIP_CHECK(hipEventRecord(s.ev1, s.stream));
            HIP_CHECK(hipMemcpyAsync(s.h_ctrl.get(), s.ctrl.get(), sizeof(ControlLine) * kNumControl,
                                     hipMemcpyDeviceToHost, s.stream));
        }

        // wait: short spin, then poll with a small sleep; stop polling once a
        // cancel was requested (the kernel sees the device flag and drains)
        const auto t0 = std::chrono::steady_clock::now();
        while (!cancel_requested(s)) {
            const hipError_t st = hipStreamQuery(s.stream);
            if (st == hipSuccess) break;
            if (st != hipErrorNotReady) HIP_CHECK(st);
            const auto us = std::chrono::duration_cast<std::chrono::microseconds>(
                                std::chrono::steady_clock::now() - t0).count();
            if (us > 200) std::this_thread::sleep_for(std::chrono::microseconds(us > 20000 ? 500 : 50));
        }
        HIP_CHECK(hipStreamSynchronize(s.stream));
        float ms = 0.f;
        HIP_CHECK(hipEventElapsedTime(&ms, s.ev0, s.ev1));
        const ControlLine *ctl = as<ControlLine>(s.h_ctrl);
        acc.kernel_ms += ms;
        acc.donations += ctl[kCtlDonations].v;
        acc.steals += ctl[kCtlSteals].v;
        acc.pops += ctl[kCtlPops].v;
        acc.pushes += ctl[kCtlPushes].v;
        acc.probes += ctl[kCtlProbes].v;
        acc.key_reads += ctl[kCtlKeyReads].v;
        acc.key_writes += ctl[kCtlKeyWrites].v;
        acc.evictions += ctl[kCtlEvictions].v;
        acc.spills += ctl[kCtlSpills].v;
        acc.spill_takes += ctl[kCtlSpillTakes].v;
        for (int k = 0; k < kNumPhases; ++k) acc.phase_ticks[k] += ctl[kCtlPhase0 + k].v;
        if (soft && ctl[kCtlStop].v == 2) {
            // the launch's time budget ran out: what is not found is
            // undecided, and the next round runs with a fresh budget
            HIP_CHECK(hipMemcpyAsync(s.h_results.get(), s.results.get(), (size_t)count * sizeof(ConfigResult),
                                     hipMemcpyDeviceToHost, s.stream));
            HIP_CHECK(hipStreamSynchronize(s.stream));
            const ConfigResult *r = as<ConfigResult>(s.h_results);
            for (int c = base; c < base + sub; ++c) s.undecided[c] = (!r[c].found && r[c].live != 0) ? 1 : 0;
        } else if (ctl[kCtlStop].v != 0) {
            stop_flag = (int)ctl[kCtlStop].v;
            acc.overflow_info = ctl[kCtlOverflowInfo].v;
        }
        cancelled = cancel_requested(s);
        if (stop_flag || cancelled) break;
    }
    acc.overflowed = stop_flag != 0;
    s.stats = acc;
    HIP_CHECK(hipMemcpyAsync(s.h_results.get(), s.results.get(), (size_t)count * sizeof(ConfigResult),
                             hipMemcpyDeviceToHost, s.stream));
    HIP_CHECK(hipStreamSynchronize(s.stream));

    const ConfigResult *res = as<ConfigResult>(s.h_results);
    bool all_found = true;
    for (int c = 0; c < count; ++c) all_found &= res[c].found != 0;
    if (stop_flag == 2)
        throw std::runtime_error("GPU sorting search watchdog timeout (TDP_KERNEL_TIMEOUT_S)");
    // sorting_search.cu:841-846
    if (!all_found && cancelled) throw Cancelled("GPU sorting search cancelled");
    if (!all_found && stop_flag != 0) throw QueueOverflow("GPU sorting search queue overflow");
    (void)single_search;
}

void copy_counters(const ConfigResult &r, tdp_counters *out) {
    out->expansions = r.counters[kCntExpansions];
    out->generated = r.counters[kCntGenerated];
    out->admitted = r.counters[kCntAdmitted];
    out->dominated = r.counters[kCntDominated];
    out->improved = r.counters[kCntImproved];
    out->inserted = r.counters[kCntInserted];
    out->uninserted = r.counters[kCntUninserted];
    out->roots_live = r.counters[kCntRootsLive];
}

void clear_cancel(Slot &s) {
    // slot_manager.cu:127-144: clear the host flag, let a pending cancel copy
    // land, then clear the device flag on the slot stream
    s.cancelled.store(false, std::memory_order_release);
    HIP_CHECK(hipStreamSynchronize(s.cancel_stream));
    HIP_CHECK(hipMemsetAsync(s.cancel_flag.get(), 0, sizeof(int), s.stream));
}

// ---------------------------------------------------------------- serving
// A device-resident request loop (serve_kernel): one persistent launch whose
// waves take requests as they arrive, share their work and the session's
// queue memory and dedup table, and reply through pinned host memory.  Host
// threads submit requests concurrently (tdp_serve_search blocks until its
// reply), and tdp_search routes to the session while one runs, so
// concurrent searchForSortingSeq calls of ForkJoin workers share one launch.
// Admission: a request waits for a free device slot, a free request-ring
// entry and, unless nothing else is in flight, for room in the queue memory:
// the roots of its start node (C(n,3) nodes, all pushed by the wave that
// takes it, spilling to the shared pool when they do not fit) must fit beside
// the roots of the requests in flight within half of the spill pool.  The
// number of slots bounds the requests in flight, and with them the demand on
// the shared dedup table (a table too full for the requests in flight admits
// children without entries, which costs repeated work).
struct ServeSession {
    int device = -1;
    hipStream_t stream = nullptr;
    bool learned = false;
    int slots = 0;
    uint32_t ring = 0;
    int max_n = 0;
    size_t node_bytes = 0;
    int64_t reserve_limit = 0, reserved = 0;
    DevPtr queue_mem, share_lines, pool_seq, ctrl, dd_tag, dd_key, cfgs, results, slot_state, req_index,
        ring_slot;
    HostPtr host_req, host_reply, host_ctl;
    ServeRequest *h_req = nullptr;
    ServeReply *h_reply = nullptr;
    unsigned long long *h_ctl = nullptr;
    SearchParams P{};
    std::mutex mu;
    std::condition_variable cv;
    uint64_t next_request = 0;
    uint64_t cancels = 0;
    std::vector<int> free_slots;
    std::vector<uint64_t> slot_request;   // request in the slot, ~0 when free
    std::vector<int64_t> slot_need;
    std::vector<uint8_t> slot_replied;    // the submitting thread has read the reply
    std::vector<uint32_t> slot_uses;
    bool ended = false;                   // the launch is over (stop, overflow, watchdog)
    unsigned long long end_reason = 0;
    uint64_t completed = 0;
};
std::mutex g_serve_mutex;                 // start / stop
std::shared_ptr<ServeSession> g_serve;    // the running session, if any

std::shared_ptr<ServeSession> current_serve() {
    std::lock_guard<std::mutex> lk(g_serve_mutex);
    return g_serve;
}

template <typename T>
T load_acquire(const T *p) { return __atomic_load_n(p, __ATOMIC_ACQUIRE); }
template <typename T>
void store_release(T *p, T v) { __atomic_store_n(p, v, __ATOMIC_RELEASE); }

HostPtr make_pinned_coherent(size_t bytes) {
    void *p = nullptr;
    HIP_CHECK(hipHostMalloc(&p, bytes, hipHostMallocCoherent | hipHostMallocMapped));
    std::memset(p, 0, bytes);
    return HostPtr(p);
}

template <typename T>
T *device_view(const HostPtr &h) {
    void *d = nullptr;
    HIP_CHECK(hipHostGetDevicePointer(&d, h.get(), 0));
    return static_cast<T *>(d);
}

// The launch has ended when its stream is idle; records why (stop word).
bool serve_ended(ServeSession &S) {
    if (S.ended) return true;
    const hipError_t st = hipStreamQuery(S.stream);
    if (st == hipErrorNotReady) return false;
    (void)hipGetLastError();
    unsigned long long why = 0;
    (void)hipMemcpy(&why, &as<ControlLine>(S.ctrl)[kCtlStop].v, sizeof(why), hipMemcpyDeviceToHost);
    S.end_reason = why;
    S.ended = true;
    return true;
}

[[noreturn]] void serve_fail(const ServeSession &S) {
    if (S.end_reason == 1) throw QueueOverflow("GPU sorting search queue overflow");
    if (S.end_reason == 2) throw std::runtime_error("GPU sorting search watchdog timeout (serving session lifetime)");
    throw std::runtime_error("serving session stopped");
}

// Slots whose request has drained and whose reply was read go back to the
// free list (caller holds S.mu).
void serve_reclaim(ServeSession &S) {
    for (int s = 0; s < S.slots; ++s) {
        const uint64_t r = S.slot_request[s];
        if (r == ~0ULL || !S.slot_replied[s]) continue;
        if (load_acquire(&S.h_reply[s].drained) != r + 1) continue;
        S.slot_request[s] = ~0ULL;
        S.reserved -= S.slot_need[s];
        S.slot_need[s] = 0;
        S.free_slots.push_back(s);
        S.completed++;
    }
}

void serve_start(int64_t queue_bytes, int64_t dedup, int slots, int max_n, uint32_t flags, int lifetime_s) {
    if (slots < 1 || slots > kMaxBatchConfigs) throw std::invalid_argument("serving slots must be in 1..65536");
    if (max_n < 3 || max_n > kMaxN) throw std::invalid_argument("serving max_n must be in 3..63");
    if (queue_bytes <= 0) throw std::invalid_argument("queueBytesBudget must be positive");
    if (lifetime_s <= 0) throw std::invalid_argument("serving lifetime must be positive");
    std::lock_guard<std::mutex> lk(g_serve_mutex);
    if (g_serve) throw std::invalid_argument("a serving session is already running");
    auto S = std::make_shared<ServeSession>();
    HIP_CHECK(hipGetDevice(&S->device));
    DeviceShared &dev = device_shared(S->device);
    S->learned = (flags & TDP_FLAG_LEARNED) != 0;
    S->slots = slots;
    S->ring = (uint32_t)slots;
    S->max_n = max_n;
    const uint32_t nw = (uint32_t)node_words_for(max_n);
    S->node_bytes = (size_t)nw * 4;
    const Tuning tune = current_tuning();
    // the feeder takes the first block; the search waves the rest (all of
    // them resident: the grid is the serving kernel's occupancy limit)
    int occ_blocks = 0;
    HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(
        &occ_blocks, S->learned ? serve_learned_kernel : serve_kernel, 256, 0));
    int waves = std::max(1, occ_blocks) * dev.cus * 4 - 4;
    if (tune.waves_per_cu > 0) waves = std::min(waves, std::max(4, tune.waves_per_cu * dev.cus));
    waves = std::max(4, waves - waves % 4);
    uint32_t share_max = (uint32_t)tune.share_max;
    while (share_max > 1 && (size_t)waves * share_max * S->node_bytes * 4 > (size_t)queue_bytes) share_max /= 2;
    const size_t inbox_bytes = (size_t)waves * share_max * S->node_bytes;
    if (inbox_bytes * 4 > (size_t)queue_bytes) throw QueueOverflow("GPU sorting search queue overflow");
    const size_t pool_bytes = ((size_t)queue_bytes - inbox_bytes) / 4;
    const size_t stack_bytes = (size_t)queue_bytes - inbox_bytes - pool_bytes;
    const uint32_t stack_cap = floor_pow2(stack_bytes / ((size_t)waves * S->node_bytes));
    if (stack_cap < 2) throw QueueOverflow("GPU sorting search queue overflow");
    if (share_max > stack_cap / 2) share_max = stack_cap / 2;
    uint32_t pool_cap = floor_pow2(pool_bytes / S->node_bytes);
    if (pool_cap < 64) pool_cap = 0;
    // the roots of a request go onto the stack of the wave that takes it and
    // spill to the pool when they do not fit; the roots of the requests in
    // flight are kept within half of the pool
    S->reserve_limit = (int64_t)std::max<size_t>(pool_bytes / 2, S->node_bytes);

    S->queue_mem = make_dev((size_t)queue_bytes, "serve.nodes");
    const uint32_t idle_words = (uint32_t)((waves + 63) / 64);
    S->share_lines = make_dev(((size_t)waves + idle_words) * 128, "serve.shareLines");
    if (pool_cap) S->pool_seq = make_dev((size_t)pool_cap * 8, "serve.poolSeq");
    S->ctrl = make_dev(sizeof(ControlLine) * kNumControl, "serve.control");
    S->cfgs = make_dev(sizeof(ConfigDesc) * (size_t)slots, "serve.configs");
    S->results = make_dev(sizeof(ConfigResult) * (size_t)slots, "serve.results");
    S->slot_state = make_dev(8 * (size_t)slots, "serve.slotState");
    S->req_index = make_dev(8 * (size_t)slots, "serve.requestIndex");
    S->ring_slot = make_dev(4 * (size_t)S->ring, "serve.ringSlot");
    uint64_t buckets = 0;
    uint32_t width = 0;
    if (dedup > 0) {
        width = (uint32_t)std::min<int64_t>(kDedupBucket, dedup);
        buckets = (uint64_t)dedup / width;
        S->dd_tag = make_dev((size_t)buckets * kDedupBucket * 8, "serve.dedupTable");
        S->dd_key = make_dev((size_t)buckets * kDedupBucket * 64, "serve.dedupTable");
        HIP_CHECK(hipMemset(S->dd_tag.get(), 0, (size_t)buckets * kDedupBucket * 8));
    }
    HIP_CHECK(hipMemset(S->slot_state.get(), 0, 8 * (size_t)slots));
    HIP_CHECK(hipMemset(S->req_index.get(), 0xFF, 8 * (size_t)slots));
    S->host_req = make_pinned_coherent(sizeof(ServeRequest) * S->ring);
    S->host_reply = make_pinned_coherent(sizeof(ServeReply) * (size_t)slots);
    S->host_ctl = make_pinned_coherent(8 * kServeHostWords);
    S->h_req = as<ServeRequest>(S->host_req);
    S->h_reply = as<ServeReply>(S->host_reply);
    S->h_ctl = as<unsigned long long>(S->host_ctl);

    SearchParams &P = S->P;
    P.cfgs = as<ConfigDesc>(S->cfgs);
    P.results = as<ConfigResult>(S->results);
    P.num_configs = slots;
    P.num_waves = waves;
    P.stacks = as<uint32_t>(S->queue_mem);
    P.stack_cap = stack_cap;
    P.node_words = nw;
    P.spi_words = (uint32_t)spi_words_for(max_n);
    P.inbox_nodes = P.stacks + (size_t)waves * stack_cap * nw;
    P.inbox = as<unsigned long long>(S->share_lines);
    P.idle_bits = P.inbox + (size_t)waves * 16;
    P.idle_words = idle_words;
    P.pool_nodes = P.stacks + (size_t)waves * (stack_cap + share_max) * nw;
    P.pool_seq = pool_cap ? as<unsigned long long>(S->pool_seq) : nullptr;
    P.pool_cap = pool_cap ? pool_cap : 1;
    P.ctrl = as<ControlLine>(S->ctrl);
    P.dd_tag = as<unsigned long long>(S->dd_tag);
    P.dd_key = as<uint32_t>(S->dd_key);
    P.dd_buckets = buckets;
    P.dd_width = width;
    P.epoch = 1;
    P.cancel_flag = nullptr;
    P.donate = 1;
    P.share_mask = (uint32_t)tune.share_interval - 1;
    P.batch = batch_nodes(tune.batch, nw, S->learned);
    P.share_max = share_max;
    P.idle_backoff = (uint32_t)tune.idle_backoff;
    P.timeout_ticks = (unsigned long long)lifetime_s * 100000000ULL;
    if (S->learned) {
        std::lock_guard<std::mutex> mlk(g_model_mutex);
        if (!g_model.ready) throw std::invalid_argument("no move-ordering model is set (tdp_mlp_set_weights)");
        if (g_model.device != S->device) throw std::invalid_argument("the move-ordering model is on another device");
        P.mlp = model_weights();
    }
    P.serve = 1;
    P.slot_state = as<unsigned long long>(S->slot_state);
    P.req_index = as<unsigned long long>(S->req_index);
    P.host_req = device_view<ServeRequest>(S->host_req);
    P.host_reply = device_view<ServeReply>(S->host_reply);
    P.host_ctl = device_view<unsigned long long>(S->host_ctl);
    P.ring_slot = as<uint32_t>(S->ring_slot);
    P.ring = S->ring;

    S->free_slots.resize(slots);
    for (int i = 0; i < slots; ++i) S->free_slots[i] = slots - 1 - i;
    S->slot_request.assign(slots, ~0ULL);
    S->slot_need.assign(slots, 0);
    S->slot_replied.assign(slots, 0);
    S->slot_uses.assign(slots, 0);

    HIP_CHECK(hipStreamCreateWithFlags(&S->stream, hipStreamNonBlocking));
    const size_t reset_n = std::max<size_t>({(size_t)slots, (size_t)waves, (size_t)pool_cap, (size_t)kNumControl});
    hipLaunchKernelGGL(reset_kernel, dim3((unsigned)((reset_n + 255) / 256)), dim3(256), 0, S->stream, P);
    HIP_CHECK(hipGetLastError());
    const int blocks = waves / 4 + 1;
    if (S->learned)
        hipLaunchKernelGGL(serve_learned_kernel, dim3(blocks), dim3(256), 0, S->stream, P);
    else
        hipLaunchKernelGGL(serve_kernel, dim3(blocks), dim3(256), 0, S->stream, P);
    const hipError_t le = hipGetLastError();
    if (le != hipSuccess) {
        (void)hipStreamDestroy(S->stream);
        HIP_CHECK(le);
    }
    g_serve = S;
}

void serve_stop() {
    std::shared_ptr<ServeSession> S;
    {
        std::lock_guard<std::mutex> lk(g_serve_mutex);
        S = g_serve;
        g_serve.reset();
    }
    if (!S) return;
    HIP_CHECK(hipSetDevice(S->device));
    store_release(&S->h_ctl[kServeHostStop], 1ULL);
    {
        std::lock_guard<std::mutex> lk(S->mu);
        S->cv.notify_all();
    }
    HIP_CHECK(hipStreamSynchronize(S->stream));
    {
        std::lock_guard<std::mutex> lk(S->mu);
        (void)serve_ended(*S);
        S->ended = true;
        S->cv.notify_all();
    }
    (void)hipStreamDestroy(S->stream);
    S->stream = nullptr;
}

void serve_cancel(ServeSession &S, int64_t ticket);
int serve_try_admit(ServeSession &S, const ConfigDesc &d, int64_t need, uint64_t *out_request);

// One request through the session; blocks until its reply.  *cancel_ticket
// (if given) receives request << 16 | slot for tdp_serve_cancel while the
// request runs.
void serve_search(ServeSession &S, const CaseInput &c, uint32_t flags, int32_t *out_moves, int32_t *out_depth,
                  tdp_counters *out_counters, std::atomic<int64_t> *cancel_ticket) {
    if (c.n > S.max_n) throw std::invalid_argument("performGpuSortingSearch: n exceeds the serving session's max_n");
    (void)flags;
    ConfigDesc d;
    fill_config(d, c);
    const int64_t need = (int64_t)triple_count(c.n) * (int64_t)S.node_bytes;
    uint64_t r;
    int slot;
    {
        std::unique_lock<std::mutex> lk(S.mu);
        while (true) {
            if (serve_ended(S)) serve_fail(S);
            serve_reclaim(S);
            slot = serve_try_admit(S, d, need, &r);
            if (slot >= 0) break;
            S.cv.wait_for(lk, std::chrono::microseconds(200));
        }
    }
    if (cancel_ticket) {
        // a cancel that arrived before the request was published
        int64_t expect = -1;
        if (!cancel_ticket->compare_exchange_strong(expect, (int64_t)((r << 16) | (uint64_t)slot)))
            serve_cancel(S, (int64_t)((r << 16) | (uint64_t)slot));
    }
    const auto t0 = std::chrono::steady_clock::now();
    bool ok = false;
    while (true) {
        if (load_acquire(&S.h_reply[slot].seq) == r + 1) {
            ok = true;
            break;
        }
        if (cancel_ticket && cancel_ticket->load() == -2) break;   // cancelled from another thread
        const auto us = std::chrono::duration_cast<std::chrono::microseconds>(
                            std::chrono::steady_clock::now() - t0).count();
        if (us > 2000) {
            std::lock_guard<std::mutex> lk(S.mu);
            if (serve_ended(S)) break;
        }
        std::this_thread::sleep_for(std::chrono::microseconds(us < 200 ? 5 : us < 20000 ? 50 : 500));
    }
    const bool cancelled = cancel_ticket && cancel_ticket->exchange(-1) == -2;
    ServeReply rep{};
    if (ok) std::memcpy(&rep, &S.h_reply[slot], sizeof(rep));
    {
        std::lock_guard<std::mutex> lk(S.mu);
        S.slot_replied[slot] = 1;
        S.cv.notify_all();
        if (!ok && S.ended) serve_fail(S);
    }
    if (!ok || cancelled || rep.found == 2) throw Cancelled("GPU sorting search cancelled");
    *out_depth = 0;
    if (out_counters) {
        out_counters->expansions = rep.counters[kCntExpansions];
        out_counters->generated = rep.counters[kCntGenerated];
        out_counters->admitted = rep.counters[kCntAdmitted];
        out_counters->dominated = rep.counters[kCntDominated];
        out_counters->improved = rep.counters[kCntImproved];
        out_counters->inserted = rep.counters[kCntInserted];
        out_counters->uninserted = rep.counters[kCntUninserted];
        out_counters->roots_live = rep.counters[kCntRootsLive];
    }
    if (rep.found != 1 || rep.depth <= 0) return;
    for (int m = 0; m < rep.depth; ++m) triples().unpack(c.n, rep.moves[m], out_moves + 3 * m);
    *out_depth = rep.depth;
}

// Admission test and publication of one request (caller holds S.mu):
// returns its slot, or -1 when it cannot be admitted now.
int serve_try_admit(ServeSession &S, const ConfigDesc &d, int64_t need, uint64_t *out_request) {
    const bool ring_free = S.next_request < S.ring ||
                           load_acquire(&S.h_ctl[kServeHostConsumed]) >= S.next_request - S.ring + 1;
    const bool room = S.reserved == 0 || S.reserved + need <= S.reserve_limit;
    if (S.free_slots.empty() || !ring_free || !room) return -1;
    const int slot = S.free_slots.back();
    S.free_slots.pop_back();
    if (++S.slot_uses[slot] >= 0xFFFFu)
        throw std::runtime_error("serving slot generations exhausted: restart the serving session");
    const uint64_t r = S.next_request++;
    S.slot_request[slot] = r;
    S.slot_need[slot] = need;
    S.slot_replied[slot] = 0;
    S.reserved += need;
    ServeRequest &q = S.h_req[r % S.ring];
    q.desc = d;
    q.slot = (uint32_t)slot;
    store_release<unsigned long long>(&q.seq, r + 1);
    *out_request = r;
    return slot;
}

// A whole batch through the session from one caller: requests are admitted
// as slots and queue memory free up (continuous batching) while the replies
// of those in flight are collected.
void serve_batch(ServeSession &S, const std::vector<CaseInput> &cases, const std::vector<int> &index,
                 tdp_case_result *out) {
    const size_t count = cases.size();
    std::vector<ConfigDesc> descs(count);
    std::vector<int64_t> need(count);
    for (size_t c = 0; c < count; ++c) {
        if (cases[c].n > S.max_n)
            throw std::invalid_argument("performGpuSortingSearch: n exceeds the serving session's max_n");
        fill_config(descs[c], cases[c]);
        need[c] = (int64_t)triple_count(cases[c].n) * (int64_t)S.node_bytes;
    }
    struct Flight { size_t c; int slot; uint64_t r; };
    std::vector<Flight> flight;
    size_t next = 0, done = 0;
    while (done < count) {
        bool progress = false;
        {
            std::lock_guard<std::mutex> lk(S.mu);
            if (serve_ended(S)) serve_fail(S);
            serve_reclaim(S);
            while (next < count) {
                uint64_t r = 0;
                const int slot = serve_try_admit(S, descs[next], need[next], &r);
                if (slot < 0) break;
                flight.push_back({next, slot, r});
                ++next;
                progress = true;
            }
        }
        for (size_t q = 0; q < flight.size();) {
            const Flight f = flight[q];
            if (load_acquire(&S.h_reply[f.slot].seq) != f.r + 1) {
                ++q;
                continue;
            }
            ServeReply rep;
            std::memcpy(&rep, &S.h_reply[f.slot], sizeof(rep));
            {
                std::lock_guard<std::mutex> lk(S.mu);
                S.slot_replied[f.slot] = 1;
            }
            tdp_case_result &o = out[index[f.c]];
            o.found = rep.found == 1 ? 1 : 0;
            o.depth = rep.found == 1 ? rep.depth : 0;
            for (int m = 0; m < o.depth; ++m) triples().unpack(cases[f.c].n, rep.moves[m], &o.moves[m][0]);
            o.counters.expansions = rep.counters[kCntExpansions];
            o.counters.generated = rep.counters[kCntGenerated];
            o.counters.admitted = rep.counters[kCntAdmitted];
            o.counters.dominated = rep.counters[kCntDominated];
            o.counters.improved = rep.counters[kCntImproved];
            o.counters.inserted = rep.counters[kCntInserted];
            o.counters.uninserted = rep.counters[kCntUninserted];
            o.counters.roots_live = rep.counters[kCntRootsLive];
            if (rep.found == 2) throw Cancelled("GPU sorting search cancelled");
            flight[q] = flight.back();
            flight.pop_back();
            ++done;
            progress = true;
        }
        if (!progress) std::this_thread::sleep_for(std::chrono::microseconds(20));
    }
    std::lock_guard<std::mutex> lk(S.mu);
    serve_reclaim(S);
    S.cv.notify_all();
}

// Publishes a cancellation of request ticket (request << 16 | slot) to the
// feeder (a ring of kServeCancelEntries entries, consumed within a feeder
// poll).
void serve_cancel(ServeSession &S, int64_t ticket) {
    std::lock_guard<std::mutex> lk(S.mu);
    const uint64_t k = S.cancels++;
    store_release(&S.h_ctl[kServeHostCancelRing + (k % kServeCancelEntries)], (unsigned long long)ticket);
    store_release(&S.h_ctl[kServeHostCancels], (unsigned long long)S.cancels);
}


int search_impl(int slot, const uint8_t *pi, const uint8_t *spi, int n, int e0, float min_rate, int D,
                int full, uint32_t flags, int32_t *out_moves, int32_t *out_depth, tdp_counters *out_counters) {
    return guarded([&] {
        if (!pi || !spi || !out_moves || !out_depth) throw std::invalid_argument("null argument");
        CaseInput c{pi, spi, n, e0, D, full, min_rate};
        if (n > kMaxN) throw std::invalid_argument("performGpuSortingSearch: n exceeds MAX_N");
        if (D <= 0 || D > kMaxDepth) throw std::invalid_argument("performGpuSortingSearch: invalid maximumDepth");
        Slot &s = get_slot(slot);
        std::lock_guard<std::mutex> run(s.run_mutex);
        HIP_CHECK(hipSetDevice(s.device));
        if (auto S = current_serve(); S && S->device == s.device) {
            // the serving session takes the request (concurrent calls share it)
            *out_depth = 0;
            if (out_counters) std::memset(out_counters, 0, sizeof(*out_counters));
            if (triple_count(n) == 0) return;
            validate_case(c);
            s.serve_ticket.store(-1);
            serve_search(*S, c, flags, out_moves, out_depth, out_counters, &s.serve_ticket);
            return;
        }
        clear_cancel(s);
        *out_depth = 0;
        if (out_counters) std::memset(out_counters, 0, sizeof(*out_counters));
        if (triple_count(n) == 0) return;   // sorting_search.cu:718-719
        validate_case(c);
        std::vector<CaseInput> cases{c};
        run_launch(s, cases, flags, true);
        const ConfigResult &r = as<ConfigResult>(s.h_results)[0];
        if (out_counters) copy_counters(r, out_counters);
        if (!r.found || r.depth <= 0) return;
        for (int m = 0; m < r.depth; ++m) triples().unpack(n, r.moves[m], out_moves + 3 * m);
        *out_depth = r.depth;
    });
}


// ---------------------------------------------------------------- filter
struct FilterWorkspace {
    int device = -1;
    DevPtr mask, counts, offsets, total, out, moves;
    size_t mask_words = 0, blocks = 0, out_bytes = 0, moves_cap = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
};
std::mutex g_filter_mutex;
std::vector<std::unique_ptr<FilterWorkspace>> g_filter_ws;

// pi_filter.cu:165-166 and the argument checks of both filter entry points
void validate_filter_args(const uint8_t *cycle, int L, const uint8_t *triple, int tl, int n, int64_t out_cap) {
    if (n > kMaxN) throw std::invalid_argument("n exceeds PI_FILTER_MAX_N");
    if (n < 1) throw std::invalid_argument("n must be at least 1");
    if (n > 21) throw std::invalid_argument("n too large: (n-1)! exceeds 64-bit rank enumeration");
    if (L < 0 || tl < 0 || L > 64 || tl > 64) throw std::invalid_argument("invalid cycle or triple length");
    if ((L > 0 && !cycle) || (tl > 0 && !triple)) throw std::invalid_argument("null argument");
    for (int i = 0; i < L; ++i)
        if (cycle[i] >= n) throw std::invalid_argument("cycle symbol out of range");
    for (int i = 0; i < tl; ++i)
        if (triple[i] >= n) throw std::invalid_argument("triple symbol out of range");
    if (out_cap < 0) throw std::invalid_argument("negative output capacity");
}

// Parameters common to both filter entry points (everything but the buffers).
FilterParams filter_params(const uint8_t *cycle, int L, const uint8_t *triple, int tl, int n, int skip) {
    FilterParams F{};
    F.n = n;
    F.cycle_len = L;
    F.triple_len = tl;
    F.skip_2move = skip ? 1 : 0;
    std::memcpy(F.cycle, cycle, (size_t)L);
    std::memcpy(F.triple, triple, (size_t)tl);
    F.fact[0] = 1;
    for (int i = 1; i < 24; ++i) F.fact[i] = F.fact[i - 1] * (uint64_t)(i < 21 ? i : 1);
    const uint64_t group = (uint64_t)filter_group_size(n);
    for (int i = 0; i < 24; ++i) F.gfact[i] = i >= filter_tail(n) ? F.fact[i] / group : 0;
    F.rank32 = F.fact[n - 1] < (1ULL << 32) ? 1 : 0;
    return F;
}

void filter_impl(const uint8_t *cycle, int L, const uint8_t *triple, int tl, int n, int skip, uint8_t *out,
                 int64_t out_cap, int64_t *out_total, int64_t *out_written) {
    validate_filter_args(cycle, L, triple, tl, n, out_cap);

    FilterParams F = filter_params(cycle, L, triple, tl, n, skip);
    // 2-move candidates of the cycle (pi_filter.cu:86-115): triples of cycle
    // indices i<j<k whose arcs (taken through the symbol -> index map as the
    // reference does) turn `before` into `before + 2` odd cycles.
    uint8_t cidx[64] = {0};
    for (int i = 0; i < L; ++i) cidx[cycle[i]] = (uint8_t)i;
    auto dist = [&](int a, int b) {
        const int ai = cidx[a], bi = cidx[b];
        return bi >= ai ? bi - ai : (L - ai) + bi;
    };
    std::vector<uint32_t> moves;
    const int before = (L % 2 == 1) ? 1 : 0;
    for (int i = 0; i < L - 2; ++i)
        for (int j = i + 1; j < L - 1; ++j)
            for (int k = j + 1; k < L; ++k) {
                const int a = cycle[i], b = cycle[j], c = cycle[k];
                const int after = (dist(a, b) % 2 == 1) + (dist(b, c) % 2 == 1) + (dist(c, a) % 2 == 1);
                if (after - before == 2) moves.push_back((uint32_t)a | ((uint32_t)b << 8) | ((uint32_t)c << 16));
            }
    F.move_count = (int)moves.size();
    const uint64_t total = F.fact[n - 1];

    int device = 0;
    HIP_CHECK(hipGetDevice(&device));
    std::lock_guard<std::mutex> lk(g_filter_mutex);
    if ((size_t)device >= g_filter_ws.size()) g_filter_ws.resize(device + 1);
    if (!g_filter_ws[device]) {
        auto w = std::make_unique<FilterWorkspace>();
        w->device = device;
        HIP_CHECK(hipStreamCreateWithFlags(&w->stream, hipStreamNonBlocking));
        HIP_CHECK(hipEventCreate(&w->ev0));
        HIP_CHECK(hipEventCreate(&w->ev1));
        w->total = make_dev(8, "filter.total");
        g_filter_ws[device] = std::move(w);
    }
    FilterWorkspace &W = *g_filter_ws[device];
    // ranks per launch chunk: 2^27 groups (a multiple of the group size, so
    // every chunk but the last holds whole groups; the total is one too)
    const uint64_t group = (uint64_t)filter_group_size(n);
    const uint64_t kChunk = group << 27;
    const uint64_t chunk = std::min<uint64_t>(total, kChunk);
    const uint64_t chunk_groups = (chunk + group - 1) / group;
    F.split_log2 = filter_split_log2(chunk_groups);
    const size_t blocks = (size_t)(((chunk_groups << F.split_log2) + 255) / 256);
    if (W.mask_words < chunk_groups) {
        W.mask = make_dev((size_t)chunk_groups * 4, "filter.mask");
        W.mask_words = (size_t)chunk_groups;
    }
    if (W.blocks < blocks) {
        W.counts = make_dev(blocks * 4, "filter.counts");
        W.offsets = make_dev(blocks * 8, "filter.offsets");
        W.blocks = blocks;
    }
    const uint64_t cap_rows = (uint64_t)std::min<int64_t>(out_cap, (int64_t)total);
    const size_t out_bytes = std::max<size_t>((size_t)cap_rows * (size_t)n, 1);
    if (W.out_bytes < out_bytes) {
        W.out = make_dev(out_bytes, "filter.output");
        W.out_bytes = out_bytes;
    }
    if (W.moves_cap < std::max<size_t>(moves.size(), 1)) {
        W.moves = make_dev(std::max<size_t>(moves.size(), 1) * 4, "filter.moves");
        W.moves_cap = std::max<size_t>(moves.size(), 1);
    }
    if (!moves.empty())
        HIP_CHECK(hipMemcpyAsync(W.moves.get(), moves.data(), moves.size() * 4, hipMemcpyHostToDevice, W.stream));
    F.moves = as<uint32_t>(W.moves);
    F.mask = as<uint32_t>(W.mask);
    F.block_counts = as<uint32_t>(W.counts);
    F.out = as<uint8_t>(W.out);
    F.out_cap = cap_rows;
    HIP_CHECK(hipMemsetAsync(W.total.get(), 0, 8, W.stream));
    HIP_CHECK(hipEventRecord(W.ev0, W.stream));
    for (uint64_t off = 0; off < total; off += chunk) {
        F.rank_offset = off;
        F.rank_end = std::min(total, off + chunk);
        const uint64_t nb = ((((F.rank_end - off + group - 1) / group) << F.split_log2) + 255) / 256;
        if (filter_nibbles(n))
            hipLaunchKernelGGL(filter_mask_kernel, dim3((unsigned)nb), dim3(256), 0, W.stream, F);
        else
            hipLaunchKernelGGL(filter_mask_any_kernel, dim3((unsigned)nb), dim3(256), 0, W.stream, F);
        HIP_CHECK(hipGetLastError());
        hipLaunchKernelGGL(filter_scan_kernel, dim3(1), dim3(1024), 0, W.stream, as<uint32_t>(W.counts), nb,
                           as<uint64_t>(W.offsets), as<uint64_t>(W.total));
        HIP_CHECK(hipGetLastError());
        if (cap_rows > 0) {
            if (filter_nibbles(n))
                hipLaunchKernelGGL(filter_write_kernel, dim3((unsigned)nb), dim3(256), 0, W.stream, F,
                                   (const uint64_t *)as<uint64_t>(W.offsets));
            else
                hipLaunchKernelGGL(filter_write_any_kernel, dim3((unsigned)nb), dim3(256), 0, W.stream, F,
                                   (const uint64_t *)as<uint64_t>(W.offsets));
            HIP_CHECK(hipGetLastError());
        }
    }
    HIP_CHECK(hipEventRecord(W.ev1, W.stream));
    uint64_t count = 0;
    HIP_CHECK(hipMemcpyAsync(&count, W.total.get(), 8, hipMemcpyDeviceToHost, W.stream));
    HIP_CHECK(hipStreamSynchronize(W.stream));
    const uint64_t rows = std::min<uint64_t>(count, cap_rows);
    if (rows > 0)
        HIP_CHECK(hipMemcpy(out, W.out.get(), (size_t)rows * (size_t)n, hipMemcpyDeviceToHost));
    float ms = 0.f;
    HIP_CHECK(hipEventElapsedTime(&ms, W.ev0, W.ev1));
    g_filter_ms = ms;
    if (out_total) *out_total = (int64_t)count;
    if (out_written) *out_written = (int64_t)rows;
}

// ---------------------------------------------------------------- allocation-free filter
// tdp_filter_noalloc runs the filter's passes on the caller's stream with all
// of its working memory in `scratch`, a file-scope device array: rank chunks
// of 2^21 groups (8 MiB of masks, 128 KiB of block counts, 256 KiB of offsets)
// and the 2-move table (at most C(64,3) entries) with its count.  Each device holds
// its own copy of the array from library load to process end; host code
// reaches it only through the address hipGetSymbolAddress returns for the
// current device, resolved once per device on the first call there.
constexpr uint64_t kNoAllocGroups = 1ULL << 21;                          // rank groups per pass
constexpr size_t kNoAllocBlocks = (size_t)(kNoAllocGroups * kFilterMaxThreadsPerGroup / 256);   // mask kernel blocks
constexpr size_t kNoAllocMaxMoves = 64 * 63 * 62 / 6;                    // C(64,3)
constexpr size_t kScratchMaskOff = 0;
constexpr size_t kScratchCountsOff = kScratchMaskOff + kNoAllocGroups * 4;
constexpr size_t kScratchOffsetsOff = kScratchCountsOff + kNoAllocBlocks * 4;
constexpr size_t kScratchMovesOff = kScratchOffsetsOff + kNoAllocBlocks * 8;
constexpr size_t kScratchMoveCountOff = kScratchMovesOff + ((kNoAllocMaxMoves * 4 + 255) & ~(size_t)255);
constexpr size_t kScratchBytes = kScratchMoveCountOff + 256;

__device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_scratch_mutex;
std::vector<unsigned char *> g_scratch_ptr;   // per device; nullptr until resolved

unsigned char *scratch_for(int device) {
    std::lock_guard<std::mutex> lk(g_scratch_mutex);
    if ((size_t)device >= g_scratch_ptr.size()) g_scratch_ptr.resize(device + 1, nullptr);
    if (!g_scratch_ptr[device]) {
        void *ptr = nullptr;
        HIP_CHECK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        if (!ptr) throw std::runtime_error("no device address for the filter scratch array");
        g_scratch_ptr[device] = static_cast<unsigned char *>(ptr);
    }
    return g_scratch_ptr[device];
}

void filter_noalloc_impl(const uint8_t *cycle, int L, const uint8_t *triple, int tl, int n, int skip,
                         uint8_t *out_rows, int64_t out_cap, uint64_t *out_total, hipStream_t stream) {
    validate_filter_args(cycle, L, triple, tl, n, out_cap);
    if (!out_total) throw std::invalid_argument("null argument");
    if (out_cap > 0 && !out_rows) throw std::invalid_argument("null argument");
    // every working buffer of this call fits the scratch array by
    // construction: one pass takes at most kNoAllocGroups rank groups and a cycle
    // of at most 64 symbols has at most C(64,3) 2-move candidates
    const size_t max_moves = L >= 3 ? (size_t)L * (L - 1) * (L - 2) / 6 : 0;
    if (max_moves > kNoAllocMaxMoves) throw std::invalid_argument("cycle longer than the scratch array allows");
    int device = 0;
    HIP_CHECK(hipGetDevice(&device));
    unsigned char *base = scratch_for(device);
    FilterParams F = filter_params(cycle, L, triple, tl, n, skip);
    uint32_t *moves = reinterpret_cast<uint32_t *>(base + kScratchMovesOff);
    uint32_t *move_count = reinterpret_cast<uint32_t *>(base + kScratchMoveCountOff);
    F.moves = moves;
    F.move_count_dev = move_count;
    F.mask = reinterpret_cast<uint32_t *>(base + kScratchMaskOff);
    F.block_counts = reinterpret_cast<uint32_t *>(base + kScratchCountsOff);
    uint64_t *offsets = reinterpret_cast<uint64_t *>(base + kScratchOffsetsOff);
    F.out = out_rows;
    F.out_cap = (uint64_t)out_cap;
    const uint64_t total = F.fact[n - 1];
    // the set-up pass also clears the caller's total, which the scans then
    // accumulate (no memset: the sequence may be captured into a graph)
    hipLaunchKernelGGL(filter_moves_kernel, dim3(1), dim3(64), 0, stream, F, moves, move_count, out_total);
    HIP_CHECK(hipGetLastError());
    const uint64_t group = (uint64_t)filter_group_size(n);
    const uint64_t chunk = group * kNoAllocGroups;   // whole groups per pass
    F.split_log2 = filter_split_log2(std::min<uint64_t>(kNoAllocGroups, (total + group - 1) / group));
    for (uint64_t off = 0; off < total; off += chunk) {
        F.rank_offset = off;
        F.rank_end = std::min(total, off + chunk);
        const uint64_t nb = ((((F.rank_end - off + group - 1) / group) << F.split_log2) + 255) / 256;
        if (filter_nibbles(n))
            hipLaunchKernelGGL(filter_mask_kernel, dim3((unsigned)nb), dim3(256), 0, stream, F);
        else
            hipLaunchKernelGGL(filter_mask_any_kernel, dim3((unsigned)nb), dim3(256), 0, stream, F);
        HIP_CHECK(hipGetLastError());
        hipLaunchKernelGGL(filter_scan_kernel, dim3(1), dim3(1024), 0, stream, F.block_counts, nb, offsets,
                           out_total);
        HIP_CHECK(hipGetLastError());
        if (out_cap > 0) {
            if (filter_nibbles(n))
                hipLaunchKernelGGL(filter_write_kernel, dim3((unsigned)nb), dim3(256), 0, stream, F,
                                   (const uint64_t *)offsets);
            else
                hipLaunchKernelGGL(filter_write_any_kernel, dim3((unsigned)nb), dim3(256), 0, stream, F,
                                   (const uint64_t *)offsets);
            HIP_CHECK(hipGetLastError());
        }
    }
}

void host_eval_children(const uint8_t *pi, const uint8_t *spi, int n, int32_t *out) {
    uint32_t info[64];
    bool seen[64] = {false};
    for (int s = 0; s < n; ++s) {
        if (seen[s]) continue;
        std::vector<int> cyc;
        int x = s;
        while (!seen[x]) {
            seen[x] = true;
            cyc.push_back(x);
            x = spi[x];
        }
        const int L = (int)cyc.size();
        for (int q = 0; q < L; ++q) info[cyc[q]] = cycle_info(cyc[0], q, L);
    }
    const int e = count_odd_cycles(spi, n);
    int t = 0;
    for (int i = 0; i < n - 2; ++i)
        for (int j = i + 1; j < n - 1; ++j)
            for (int k = j + 1; k < n; ++k) out[t++] = e + parity_delta(info[pi[i]], info[pi[j]], info[pi[k]]);
}

}  // namespace
}  // namespace tdp

using namespace tdp;

extern "C" {

const char *tdp_last_error(void) { return g_last_error.c_str(); }

int tdp_set_launch_deadline(int milliseconds) {
    return guarded([&] {
        if (milliseconds <= 0) throw std::invalid_argument("deadline must be positive");
        std::lock_guard<std::mutex> lk(g_tuning_mutex);
        g_tuning.deadline_ms = milliseconds;
    });
}

int tdp_set_expansion_batch(int nodes) {
    return guarded([&] {
        if (nodes < 0 || nodes > kExpansionBatchMax) throw std::invalid_argument("batch must be in 0..16");
        std::lock_guard<std::mutex> lk(g_tuning_mutex);
        g_tuning.batch = nodes ? nodes : Tuning().batch;
    });
}

int tdp_set_tuning(int waves_per_cu, int share_interval, int share_max, int idle_backoff, int max_active) {
    return guarded([&] {
        if (waves_per_cu < 0 || share_interval < 0 || share_max < 0 || idle_backoff < 0 || max_active < 0)
            throw std::invalid_argument("tuning values must be non-negative");
        if (share_interval && (share_interval & (share_interval - 1)))
            throw std::invalid_argument("share_interval must be a power of two");
        Tuning t;
        t.deadline_ms = current_tuning().deadline_ms;
        t.batch = current_tuning().batch;
        if (waves_per_cu) t.waves_per_cu = waves_per_cu;
        if (share_interval) t.share_interval = share_interval;
        if (share_max) t.share_max = share_max;
        if (idle_backoff) t.idle_backoff = idle_backoff;
        if (max_active) t.max_active = max_active;
        std::lock_guard<std::mutex> lk(g_tuning_mutex);
        g_tuning = t;
    });
}

int tdp_set_alloc_timeout(int64_t milliseconds) {
    g_alloc_timeout_ms.store(milliseconds, std::memory_order_relaxed);
    return TDP_OK;
}

uint64_t tdp_alloc_retries(void) { return g_alloc_retries.load(std::memory_order_relaxed); }

int tdp_device_count(void) {
    int c = 0;
    if (hipGetDeviceCount(&c) != hipSuccess) {
        (void)hipGetLastError();
        return 0;
    }
    return c;
}

int tdp_init_slots(int total_slots, int devices_count, int64_t queue_bytes_budget, int dedup_table_size) {
    return guarded([&] { init_slots(total_slots, devices_count, queue_bytes_budget, dedup_table_size); });
}

int tdp_destroy_slots(void) {
    return guarded([&] {
        serve_stop();
        destroy_all_slots();
    });
}

int tdp_serve_start(int64_t queue_bytes, int dedup_table_size, int slots, int max_n, uint32_t flags,
                    int lifetime_s) {
    return guarded([&] { serve_start(queue_bytes, dedup_table_size, slots, max_n, flags, lifetime_s); });
}

int tdp_serve_stop(void) {
    return guarded([&] { serve_stop(); });
}

int tdp_serve_search(const uint8_t *pi, const uint8_t *spi, int n, int initial_even_cycles, float min_rate,
                     int max_moves, int full_sorting, int32_t *out_moves, int32_t *out_depth,
                     tdp_counters *out_counters) {
    return guarded([&] {
        if (!pi || !spi || !out_moves || !out_depth) throw std::invalid_argument("null argument");
        if (n > kMaxN) throw std::invalid_argument("performGpuSortingSearch: n exceeds MAX_N");
        if (max_moves <= 0 || max_moves > kMaxDepth)
            throw std::invalid_argument("performGpuSortingSearch: invalid maximumDepth");
        auto S = current_serve();
        if (!S) throw std::invalid_argument("no serving session is running (tdp_serve_start)");
        HIP_CHECK(hipSetDevice(S->device));
        CaseInput c{pi, spi, n, initial_even_cycles, max_moves, full_sorting, min_rate};
        *out_depth = 0;
        if (out_counters) std::memset(out_counters, 0, sizeof(*out_counters));
        if (triple_count(n) == 0) return;
        validate_case(c);
        serve_search(*S, c, 0u, out_moves, out_depth, out_counters, nullptr);
    });
}

int tdp_serve_batch(int count, const uint8_t *pis, const uint8_t *spis, const int32_t *ns, const int32_t *e0s,
                    const float *min_rates, const int32_t *max_moves, const int32_t *full_sorting,
                    tdp_case_result *out) {
    return guarded([&] {
        if (count < 0) throw std::invalid_argument("negative batch size");
        if (count == 0) return;
        if (!pis || !spis || !ns || !e0s || !min_rates || !max_moves || !full_sorting || !out)
            throw std::invalid_argument("null argument");
        auto S = current_serve();
        if (!S) throw std::invalid_argument("no serving session is running (tdp_serve_start)");
        HIP_CHECK(hipSetDevice(S->device));
        std::memset(out, 0, sizeof(tdp_case_result) * (size_t)count);
        std::vector<CaseInput> cases;
        std::vector<int> index;
        for (int c = 0; c < count; ++c) {
            CaseInput ci{pis + 64 * (size_t)c, spis + 64 * (size_t)c, ns[c], e0s[c], max_moves[c], full_sorting[c],
                         min_rates[c]};
            validate_case(ci);
            if (triple_count(ci.n) == 0) continue;
            cases.push_back(ci);
            index.push_back(c);
        }
        serve_batch(*S, cases, index, out);
    });
}

int tdp_serve_stats(uint64_t *completed, uint64_t *in_flight, int64_t *reserved_bytes) {
    return guarded([&] {
        auto S = current_serve();
        if (!S) throw std::invalid_argument("no serving session is running (tdp_serve_start)");
        std::lock_guard<std::mutex> lk(S->mu);
        serve_reclaim(*S);
        if (completed) *completed = S->completed;
        if (in_flight) *in_flight = (uint64_t)(S->slots - (int)S->free_slots.size());
        if (reserved_bytes) *reserved_bytes = S->reserved;
    });
}

int tdp_search(int slot, const uint8_t *pi, const uint8_t *spi, int n, int initial_even_cycles, float min_rate,
               int max_moves, int full_sorting, int32_t *out_moves, int32_t *out_depth) {
    return search_impl(slot, pi, spi, n, initial_even_cycles, min_rate, max_moves, full_sorting, 0u, out_moves,
                       out_depth, nullptr);
}

int tdp_search_ex(int slot, const uint8_t *pi, const uint8_t *spi, int n, int initial_even_cycles, float min_rate,
                  int max_moves, int full_sorting, uint32_t flags, int32_t *out_moves, int32_t *out_depth,
                  tdp_counters *out_counters) {
    return search_impl(slot, pi, spi, n, initial_even_cycles, min_rate, max_moves, full_sorting, flags, out_moves,
                       out_depth, out_counters);
}

int tdp_search_batch(int slot, int count, const uint8_t *pis, const uint8_t *spis, const int32_t *ns,
                     const int32_t *e0s, const float *min_rates, const int32_t *max_moves,
                     const int32_t *full_sorting, uint32_t flags, tdp_case_result *out) {
    return guarded([&] {
        if (count < 0) throw std::invalid_argument("negative batch size");
        if (count > kMaxBatchConfigs) throw std::invalid_argument("batch larger than 65536 configurations");
        if (count == 0) return;
        if (!pis || !spis || !ns || !e0s || !min_rates || !max_moves || !full_sorting || !out)
            throw std::invalid_argument("null argument");
        Slot &s = get_slot(slot);
        std::lock_guard<std::mutex> run(s.run_mutex);
        HIP_CHECK(hipSetDevice(s.device));
        clear_cancel(s);
        std::vector<CaseInput> cases;
        std::vector<int> index;   // batch position of each launched case
        std::memset(out, 0, sizeof(tdp_case_result) * (size_t)count);
        if (flags & TDP_FLAG_RESIDENT) {
            if (count != s.last_batch_count || s.last_cases.empty())
                throw std::invalid_argument("resident batch does not match the previous batch on this slot");
            cases = s.last_cases;
            index = s.last_index;
        } else {
            for (int c = 0; c < count; ++c) {
                CaseInput ci{pis + 64 * (size_t)c, spis + 64 * (size_t)c, ns[c], e0s[c], max_moves[c],
                             full_sorting[c], min_rates[c]};
                validate_case(ci);
                if (triple_count(ci.n) == 0) continue;
                cases.push_back(ci);
                index.push_back(c);
            }
        }
        if (cases.empty()) return;
        s.last_batch_count = -1;
        run_launch(s, cases, flags, false);
        if (!(flags & TDP_FLAG_RESIDENT)) {
            s.last_cases = cases;
            for (auto &ci : s.last_cases) ci.pi = ci.spi = nullptr;   // only n is kept
            s.last_index = index;
        }
        s.last_batch_count = count;
        const ConfigResult *res = as<ConfigResult>(s.h_results);
        for (size_t q = 0; q < cases.size(); ++q) {
            tdp_case_result &o = out[index[q]];
            o.found = res[q].found ? 1 : 0;
            o.depth = res[q].found ? res[q].depth : (s.undecided[q] ? -1 : 0);
            for (int m = 0; m < o.depth; ++m) triples().unpack(cases[q].n, res[q].moves[m], &o.moves[m][0]);
            copy_counters(res[q], &o.counters);
        }
    });
}

int tdp_last_launch_stats(int slot, tdp_launch_stats *out) {
    return guarded([&] {
        if (!out) throw std::invalid_argument("null argument");
        *out = get_slot(slot).stats;
    });
}

void tdp_cancel(int slot) {
    // slot_manager.cu:100-125: host flag, then an async copy of a pinned 1
    // into the device flag on the slot's cancel stream.  Never throws.
    try {
        Slot &s = get_slot(slot);
        if (auto S = current_serve(); S && S->device == s.device) {
            // serving: the running request is cancelled through the feeder;
            // one not yet published is cancelled when it is
            const int64_t t = s.serve_ticket.exchange(-2);
            if (t >= 0) serve_cancel(*S, t);
            return;
        }
        s.cancelled.store(true, std::memory_order_release);
        int cur = -1;
        (void)hipGetDevice(&cur);
        (void)hipSetDevice(s.device);
        (void)hipMemcpyAsync(s.cancel_flag.get(), s.cancel_one.get(), sizeof(int), hipMemcpyHostToDevice,
                             s.cancel_stream);
        if (cur >= 0 && cur != s.device) (void)hipSetDevice(cur);
    } catch (...) {
    }
}

int tdp_filter(const uint8_t *cycle, int cycle_len, const uint8_t *triple, int triple_len, int n,
               int skip_2move_check, uint8_t *out, int64_t out_rows_cap, int64_t *out_total, int64_t *out_written) {
    return guarded([&] {
        if (out_rows_cap > 0 && !out) throw std::invalid_argument("null argument");
        filter_impl(cycle, cycle_len, triple, triple_len, n, skip_2move_check, out, out_rows_cap, out_total,
                    out_written);
    });
}

double tdp_filter_last_kernel_ms(void) { return g_filter_ms; }

int tdp_filter_noalloc(const uint8_t *cycle, int cycle_len, const uint8_t *triple, int triple_len, int n,
                       int skip_2move_check, uint8_t *out_rows_dev, int64_t out_rows_cap, uint64_t *out_total_dev,
                       void *stream) {
    return guarded([&] {
        filter_noalloc_impl(cycle, cycle_len, triple, triple_len, n, skip_2move_check, out_rows_dev, out_rows_cap,
                            out_total_dev, static_cast<hipStream_t>(stream));
    });
}

int tdp_eval_children(const uint8_t *pi, const uint8_t *spi, int n, int32_t *out) {
    return guarded([&] {
        if (n < 0 || n > kMaxN) throw std::invalid_argument("performGpuSortingSearch: n exceeds MAX_N");
        const int T = triple_count(n);
        if (T == 0) return;
        CaseInput c{pi, spi, n, 0, 1, 0, 1.0f};
        validate_case(c);
        int dev = 0;
        HIP_CHECK(hipGetDevice(&dev));
        DeviceShared &ds = device_shared(dev);
        DevPtr d_pi = make_dev(64, "eval.pi"), d_spi = make_dev(64, "eval.spi"), d_out = make_dev((size_t)T * 4, "eval.out");
        HIP_CHECK(hipMemcpy(d_pi.get(), pi, (size_t)n, hipMemcpyHostToDevice));
        HIP_CHECK(hipMemcpy(d_spi.get(), spi, (size_t)n, hipMemcpyHostToDevice));
        const int e = count_odd_cycles(spi, n);
        hipLaunchKernelGGL(eval_children_kernel, dim3(1), dim3(64), 0, nullptr, as<uint8_t>(d_pi), as<uint8_t>(d_spi),
                           n, T, e, as<int32_t>(d_out), 0, 0, 1, nullptr);
        HIP_CHECK(hipGetLastError());
        HIP_CHECK(hipDeviceSynchronize());
        HIP_CHECK(hipMemcpy(out, d_out.get(), (size_t)T * 4, hipMemcpyDeviceToHost));
    });
}

int tdp_eval_pair(const uint8_t *pi_a, const uint8_t *spi_a, const uint8_t *pi_b, const uint8_t *spi_b, int n,
                  int e_a, int e_b, uint32_t goal_a, uint32_t admit_a, uint32_t goal_b, uint32_t admit_b,
                  int32_t *out) {
    return guarded([&] {
        if (n < 3 || n > 32) throw std::invalid_argument("tdp_eval_pair: n must be in 3..32");
        if (!out) throw std::invalid_argument("null argument");
        CaseInput ca{pi_a, spi_a, n, 0, 1, 0, 1.0f}, cb{pi_b, spi_b, n, 0, 1, 0, 1.0f};
        validate_case(ca);
        validate_case(cb);
        const size_t rec = 4 + (size_t)kStageEntries;
        DevPtr d_in = make_dev(256, "eval.pair.in"), d_out = make_dev(2 * rec * 4, "eval.pair.out");
        uint8_t host[256] = {0};
        std::memcpy(host, pi_a, (size_t)n);
        std::memcpy(host + 64, spi_a, (size_t)n);
        std::memcpy(host + 128, pi_b, (size_t)n);
        std::memcpy(host + 192, spi_b, (size_t)n);
        HIP_CHECK(hipMemcpy(d_in.get(), host, 256, hipMemcpyHostToDevice));
        const uint8_t *b = as<uint8_t>(d_in);
        hipLaunchKernelGGL(eval_pair_kernel, dim3(1), dim3(64), 0, nullptr, b, b + 64, b + 128, b + 192, n, e_a, e_b,
                           goal_a, admit_a, goal_b, admit_b, as<int32_t>(d_out));
        HIP_CHECK(hipGetLastError());
        HIP_CHECK(hipDeviceSynchronize());
        HIP_CHECK(hipMemcpy(out, d_out.get(), 2 * rec * 4, hipMemcpyDeviceToHost));
    });
}

int tdp_eval_features(const uint8_t *pi, const uint8_t *spi, int n, int e0, int depth, int max_depth,
                      uint32_t *out) {
    return guarded([&] {
        if (n < 0 || n > kMaxN) throw std::invalid_argument("performGpuSortingSearch: n exceeds MAX_N");
        if (!out) throw std::invalid_argument("null argument");
        const int T = triple_count(n);
        if (T == 0) return;
        CaseInput c{pi, spi, n, 0, 1, 0, 1.0f};
        validate_case(c);
        DevPtr d_pi = make_dev(64, "eval.pi"), d_spi = make_dev(64, "eval.spi");
        DevPtr d_out = make_dev((size_t)T * 16, "eval.features");
        HIP_CHECK(hipMemcpy(d_pi.get(), pi, (size_t)n, hipMemcpyHostToDevice));
        HIP_CHECK(hipMemcpy(d_spi.get(), spi, (size_t)n, hipMemcpyHostToDevice));
        const int e = count_odd_cycles(spi, n);
        hipLaunchKernelGGL(eval_children_kernel, dim3(1), dim3(64), 0, nullptr, as<uint8_t>(d_pi), as<uint8_t>(d_spi),
                           n, T, e, nullptr, e0, depth, max_depth, as<uint32_t>(d_out));
        HIP_CHECK(hipGetLastError());
        HIP_CHECK(hipDeviceSynchronize());
        HIP_CHECK(hipMemcpy(out, d_out.get(), (size_t)T * 16, hipMemcpyDeviceToHost));
    });
}

int tdp_canonical_batch(int count, const uint8_t *spis, const uint8_t *pis, const int32_t *ns, float *out_signatures,
                        int32_t *out_hashes, int32_t *out_index, uint8_t *out_canonical_spi) {
    return guarded([&] {
        if (count < 0) throw std::invalid_argument("count must be non-negative");
        if (count == 0) return;
        if (!spis || !pis || !ns || !out_signatures || !out_hashes || !out_index || !out_canonical_spi)
            throw std::invalid_argument("null argument");
        // Configuration.java takes spi without fixed points (every symbol in a
        // cycle of the string form) and pi a cycle over the same symbols
        int max_n = 1;
        for (int c = 0; c < count; ++c) {
            const int n = ns[c];
            if (n < 1 || n > kMaxCanonN) throw std::invalid_argument("configuration size must be in 1..64");
            max_n = std::max(max_n, n);
            const uint8_t *spi = spis + (size_t)c * kMaxCanonN, *pi = pis + (size_t)c * kMaxCanonN;
            uint64_t seen_s = 0, seen_p = 0;
            for (int x = 0; x < n; ++x) {
                if (spi[x] >= n || pi[x] >= n) throw std::invalid_argument("spi and pi must be permutations of 0..n-1");
                if (spi[x] == x && n > 1) throw std::invalid_argument("spi has a fixed point");
                seen_s |= 1ULL << spi[x];
                seen_p |= 1ULL << pi[x];
            }
            const uint64_t all = n == 64 ? ~0ULL : ((1ULL << n) - 1);
            if (seen_s != all || seen_p != all) throw std::invalid_argument("spi and pi must be permutations of 0..n-1");
        }
        const size_t rows = (size_t)count * kMaxCanonN;
        DevPtr d_spi = make_dev(rows, "canon.spi"), d_pi = make_dev(rows, "canon.pi");
        DevPtr d_n = make_dev((size_t)count * 4, "canon.n");
        DevPtr d_sig = make_dev(rows * 4, "canon.signature"), d_hash = make_dev((size_t)count * 4, "canon.hash");
        DevPtr d_idx = make_dev((size_t)count * 4, "canon.index"), d_out = make_dev(rows, "canon.spi_out");
        HIP_CHECK(hipMemcpy(d_spi.get(), spis, rows, hipMemcpyHostToDevice));
        HIP_CHECK(hipMemcpy(d_pi.get(), pis, rows, hipMemcpyHostToDevice));
        HIP_CHECK(hipMemcpy(d_n.get(), ns, (size_t)count * 4, hipMemcpyHostToDevice));
        HIP_CHECK(hipMemset(d_sig.get(), 0, rows * 4));
        HIP_CHECK(hipMemset(d_out.get(), 0, rows));
        launch_canon(as<uint8_t>(d_spi), as<uint8_t>(d_pi), as<int32_t>(d_n), count, max_n, as<float>(d_sig),
                     as<int32_t>(d_hash), as<int32_t>(d_idx), as<uint8_t>(d_out), nullptr);
        HIP_CHECK(hipGetLastError());
        HIP_CHECK(hipDeviceSynchronize());
        HIP_CHECK(hipMemcpy(out_signatures, d_sig.get(), rows * 4, hipMemcpyDeviceToHost));
        HIP_CHECK(hipMemcpy(out_hashes, d_hash.get(), (size_t)count * 4, hipMemcpyDeviceToHost));
        HIP_CHECK(hipMemcpy(out_index, d_idx.get(), (size_t)count * 4, hipMemcpyDeviceToHost));
        HIP_CHECK(hipMemcpy(out_canonical_spi, d_out.get(), rows, hipMemcpyDeviceToHost));
    });
}

int tdp_host_eval_children(const uint8_t *pi, const uint8_t *spi, int n, int32_t *out) {
    return guarded([&] {
        if (n < 0 || n > kMaxN) throw std::invalid_argument("performGpuSortingSearch: n exceeds MAX_N");
        CaseInput c{pi, spi, n, 0, 1, 0, 1.0f};
        validate_case(c);
        host_eval_children(pi, spi, n, out);
    });
}

int tdp_host_predicate_tables(int n, int e0, float min_rate, int max_depth, int full, uint64_t *acc,
                              uint64_t *rch) {
    return guarded([&] {
        if (n < 0 || n > kMaxN) throw std::invalid_argument("performGpuSortingSearch: n exceeds MAX_N");
        if (max_depth <= 0 || max_depth > kMaxDepth)
            throw std::invalid_argument("performGpuSortingSearch: invalid maximumDepth");
        predicate_tables(n, e0, min_rate, max_depth, full != 0, acc, rch);
    });
}

int tdp_mlp_set_weights(const float *w1, const float *b1, const float *w2, float b2) {
    return guarded([&] {
        if (!w1 || !b1 || !w2) throw std::invalid_argument("null argument");
        std::lock_guard<std::mutex> lk(g_model_mutex);
        model_set(w1, b1, w2, b2);
    });
}

int tdp_mlp_get_weights(float *w1, float *b1, float *w2, float *b2) {
    return guarded([&] {
        if (!w1 || !b1 || !w2 || !b2) throw std::invalid_argument("null argument");
        std::lock_guard<std::mutex> lk(g_model_mutex);
        require_model();
        std::vector<float> p(kMlpParams);
        HIP_CHECK(hipMemcpy(p.data(), g_model.params.get(), kMlpParams * 4, hipMemcpyDeviceToHost));
        std::memcpy(w1, p.data(), sizeof(float) * kIn * kHidden);
        std::memcpy(b1, p.data() + kIn * kHidden, sizeof(float) * kHidden);
        std::memcpy(w2, p.data() + kIn * kHidden + kHidden, sizeof(float) * kHidden);
        *b2 = p[kMlpParams - 1];
    });
}

int tdp_mlp_init(uint64_t seed) {
    return guarded([&] {
        // He-normal first layer, small second layer (Box-Muller on SplitMix64)
        std::vector<float> w1(kIn * kHidden), b1(kHidden, 0.f), w2(kHidden);
        uint64_t st = seed;
        auto normal = [&]() {
            const double u1 = ((splitmix(st) >> 11) + 1.0) / 9007199254740993.0;
            const double u2 = (splitmix(st) >> 11) / 9007199254740992.0;
            return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
        };
        for (auto &x : w1) x = (float)(normal() * std::sqrt(2.0 / 16.0));
        for (auto &x : w2) x = (float)(normal() * 0.05);
        std::lock_guard<std::mutex> lk(g_model_mutex);
        model_set(w1.data(), b1.data(), w2.data(), 0.f);
    });
}

int tdp_mlp_forward(const uint32_t *X, int N, float *logits) {
    return guarded([&] {
        if (N < 0 || (N > 0 && (!X || !logits))) throw std::invalid_argument("null argument");
        if (N == 0) return;
        std::lock_guard<std::mutex> lk(g_model_mutex);
        require_model();
        RowSet rs = upload_rows(X, nullptr, nullptr, N, 0);
        const int waves = (N + 15) / 16;
        hipLaunchKernelGGL(mlp_forward_kernel, dim3((waves + 3) / 4), dim3(256), 0, 0, as<uint32_t>(rs.X), N,
                           model_weights(), nullptr, as<float>(rs.logits));
        HIP_CHECK(hipGetLastError());
        HIP_CHECK(hipMemcpy(logits, rs.logits.get(), (size_t)N * 4, hipMemcpyDeviceToHost));
    });
}

int tdp_mlp_gradients(const uint32_t *X, const uint8_t *labels, const int32_t *group_start, int G, float *grads,
                      float *loss) {
    return guarded([&] {
        if (G <= 0 || !X || !labels || !group_start || !grads || !loss) throw std::invalid_argument("null argument");
        const int N = group_start[G];
        if (N <= 0) throw std::invalid_argument("empty row set");
        std::lock_guard<std::mutex> lk(g_model_mutex);
        require_model();
        RowSet rs = upload_rows(X, labels, group_start, N, G);
        HIP_CHECK(hipMemset(g_model.loss.get(), 0, 4));
        model_backward(rs, 0, N, 0, G, 1.0f / (float)G, 1.0f / (float)G);
        HIP_CHECK(hipDeviceSynchronize());
        HIP_CHECK(hipMemcpy(grads, g_model.grads.get(), kMlpParams * 4, hipMemcpyDeviceToHost));
        HIP_CHECK(hipMemcpy(loss, g_model.loss.get(), 4, hipMemcpyDeviceToHost));
    });
}

int tdp_mlp_train(const uint32_t *X, const uint8_t *labels, const int32_t *group_start, int G, int epochs,
                  float lr, int groups_per_batch, uint64_t seed, float *epoch_loss) {
    return guarded([&] {
        if (G <= 0 || epochs <= 0 || groups_per_batch <= 0 || !X || !labels || !group_start)
            throw std::invalid_argument("invalid training arguments");
        const int N = group_start[G];
        std::lock_guard<std::mutex> lk(g_model_mutex);
        require_model();
        std::vector<int> order(G);
        for (int g = 0; g < G; ++g) order[g] = g;
        std::vector<uint32_t> Xp((size_t)N * 4);
        std::vector<uint8_t> Lp(N);
        std::vector<int32_t> Gp(G + 1);
        uint64_t st = seed;
        for (int ep = 0; ep < epochs; ++ep) {
            // groups in a new random order each epoch (Fisher-Yates)
            for (int g = G - 1; g > 0; --g) std::swap(order[g], order[(int)(splitmix(st) % (uint64_t)(g + 1))]);
            int r = 0;
            for (int q = 0; q < G; ++q) {
                const int g = order[q];
                Gp[q] = r;
                for (int x = group_start[g]; x < group_start[g + 1]; ++x, ++r) {
                    std::memcpy(&Xp[(size_t)r * 4], &X[(size_t)x * 4], 16);
                    Lp[r] = labels[x];
                }
            }
            Gp[G] = r;
            RowSet rs = upload_rows(Xp.data(), Lp.data(), Gp.data(), N, G);
            HIP_CHECK(hipMemset(g_model.loss.get(), 0, 4));
            for (int g0 = 0; g0 < G; g0 += groups_per_batch) {
                const int ng = std::min(groups_per_batch, G - g0);
                const int r0 = Gp[g0], n = Gp[g0 + ng] - r0;
                if (n <= 0) continue;
                // gradients: mean over the batch's groups; loss: mean over the epoch
                model_backward(rs, r0, n, g0, ng, 1.0f / (float)ng, 1.0f / (float)G);
                g_model.step++;
                const float b1c = 1.0f - std::pow(0.9f, (float)g_model.step);
                const float b2c = 1.0f - std::pow(0.999f, (float)g_model.step);
                hipLaunchKernelGGL(mlp_adam_kernel, dim3((kMlpParams + 255) / 256), dim3(256), 0, 0,
                                   as<float>(g_model.params), as<float>(g_model.grads), as<float>(g_model.m),
                                   as<float>(g_model.v), kMlpParams, lr, b1c, b2c, as<__bf16>(g_model.w1t),
                                   as<float>(g_model.b1), as<__bf16>(g_model.w2), as<float>(g_model.b2));
                HIP_CHECK(hipGetLastError());
            }
            HIP_CHECK(hipDeviceSynchronize());
            float l = 0.f;
            HIP_CHECK(hipMemcpy(&l, g_model.loss.get(), 4, hipMemcpyDeviceToHost));
            if (epoch_loss) epoch_loss[ep] = l;
        }
    });
}

}  // extern "C"
