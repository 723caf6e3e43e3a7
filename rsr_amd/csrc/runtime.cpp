// runtime.cpp -- C ABI implementation: contexts, device-resident datasets and program
// batches, lowering cache, and the evaluation entry points.
#include <hip/hip_runtime.h>
#include <atomic>
#include <float.h>
#include <math.h>
#include <string.h>

#include <algorithm>
#include <cstdarg>
#include <map>
#include <memory>
#include <mutex>
#include <numeric>
#include <string>
#include <thread>
#include <vector>

#include "../../include/rsr_amd.h"
#include "kernels.h"
#include "lower.h"
#include "jobs.h"
#include "runtime_internal.h"

using namespace rsr;

// ---------------------------------------------------------------------------------
// errors
// ---------------------------------------------------------------------------------
static thread_local std::string g_last_error;

int rsr::set_error(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return code;
}

int rsr::hip_error(hipError_t e, const char* what) {
    return set_error(e == hipErrorOutOfMemory ? RSR_E_OOM : RSR_E_HIP, "%s: %s", what, hipGetErrorString(e));
}

#define HIP_TRY(expr, what)                              \
    do {                                                 \
        hipError_t e_ = (expr);                          \
        if (e_ != hipSuccess) return hip_error(e_, what); \
    } while (0)

extern "C" const char* rsr_error_string(int code) {
    switch (code) {
        case RSR_OK: return "ok";
        case RSR_E_INVALID_ARG: return "invalid argument";
        case RSR_E_HIP: return "HIP runtime error";
        case RSR_E_OOM: return "device out of memory";
        case RSR_E_BUDGET: return "HBM budget exceeded";
        case RSR_E_PROGRAM: return "program rejected by lowering";
        case RSR_E_NO_DEVICE: return "no HIP device";
        case RSR_E_UNSUPPORTED: return "unsupported";
        case RSR_E_QUEUE_FULL: return "request queue full";
        case RSR_E_TIMEOUT: return "timed out";
        case RSR_E_NOT_READY: return "not ready";
        default: return "unknown error";
    }
}

extern "C" const char* rsr_last_error_message(void) { return g_last_error.c_str(); }
extern "C" const char* rsr_version(void) { return "rsr_amd 0.1 (gfx950)"; }

// ---------------------------------------------------------------------------------
// device buffers with budget accounting
// ---------------------------------------------------------------------------------
// Block sizes: powers of two from 256 B up to 1 MiB, then multiples of 2 MiB, so that
// released blocks are reused by later requests of similar size.
static size_t block_size(size_t n) {
    if (n <= (1u << 20)) {
        size_t b = 256;
        while (b < n) b <<= 1;
        return b;
    }
    const size_t g = (size_t)2 << 20;
    return (n + g - 1) / g * g;
}

// Give cached blocks back to the device (all work on the stream is finished first).
void rsr::trim_pool(rsr_context* c) {
    if (c->pool.empty()) return;
    (void)hipStreamSynchronize(c->stream);
    for (auto& kv : c->pool) (void)hipFree(kv.second);
    c->used -= c->pool_bytes;
    c->pool.clear();
    c->pool_bytes = 0;
}

// Every buffer of a context is used only by work on the context's one stream, so a
// released block can be handed to the next request at once: stream order puts its new
// use after every earlier use.
int DevBuf::alloc(rsr_context* c, size_t n) {
    release();
    if (n == 0) return RSR_OK;
    const size_t sz = block_size(n);
    auto it = c->pool.find(sz);
    if (it != c->pool.end()) {
        ptr = it->second;
        c->pool.erase(it);
        c->pool_bytes -= sz;
        ctx = c;
        bytes = sz;
        return RSR_OK;
    }
    if (c->used + sz > c->budget) trim_pool(c);
    if (c->used + sz > c->budget)
        return set_error(RSR_E_BUDGET, "allocation of %zu bytes exceeds the HBM budget (%llu of %llu used)", sz,
                         (unsigned long long)c->used, (unsigned long long)c->budget);
    hipError_t e = hipMalloc(&ptr, sz);
    if (e != hipSuccess) {
        trim_pool(c);
        e = hipMalloc(&ptr, sz);
    }
    if (e != hipSuccess) {
        ptr = nullptr;
        return hip_error(e, "hipMalloc");
    }
    ctx = c;
    bytes = sz;
    c->used += sz;
    return RSR_OK;
}

void DevBuf::release() {
    if (ptr && ctx) {
        ctx->pool.emplace(bytes, ptr);
        ctx->pool_bytes += bytes;
    } else if (ptr) {
        (void)hipFree(ptr);
    }
    ptr = nullptr;
    bytes = 0;
    ctx = nullptr;
}

int DevBuf::upload(rsr_context* c, const void* src, size_t n) {
    int rc = alloc(c, n);
    if (rc) return rc;
    if (n) HIP_TRY(hipMemcpyAsync(ptr, src, n, hipMemcpyHostToDevice, c->stream), "upload");
    return RSR_OK;
}

int rsr::ensure_host_stage(rsr_context* c, size_t need) {
    if (c->host_stage_cap >= need) return RSR_OK;
    if (c->host_stage) (void)hipHostFree(c->host_stage);
    c->host_stage = nullptr;
    c->host_stage_cap = 0;
    const size_t cap = std::max<size_t>((need + 255) / 256 * 256 * 2, (size_t)1 << 20);
    if (hipHostMalloc((void**)&c->host_stage, cap, hipHostMallocDefault) != hipSuccess) {
        c->host_stage = nullptr;
        (void)hipGetLastError();
        return RSR_E_HIP;
    }
    c->host_stage_cap = cap;
    return RSR_OK;
}

int rsr::ensure_scratch(rsr_context* c, DevBuf& b, size_t n) {
    if (b.bytes >= n) return RSR_OK;
    return b.alloc(c, n + n / 4);
}

// ---------------------------------------------------------------------------------
// context
// ---------------------------------------------------------------------------------
extern "C" int rsr_context_create(int device_ordinal, uint64_t hbm_budget_bytes, rsr_context** out) {
    if (!out) return set_error(RSR_E_INVALID_ARG, "out is null");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return set_error(RSR_E_NO_DEVICE, "no HIP device");
    if (device_ordinal < 0 || device_ordinal >= ndev)
        return set_error(RSR_E_INVALID_ARG, "device %d out of range (%d devices)", device_ordinal, ndev);
    HIP_TRY(hipSetDevice(device_ordinal), "hipSetDevice");
    auto c = std::make_unique<rsr_context>();
    c->device = device_ordinal;
    HIP_TRY(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking), "hipStreamCreate");
    size_t free_b = 0, total_b = 0;
    HIP_TRY(hipMemGetInfo(&free_b, &total_b), "hipMemGetInfo");
    c->budget = hbm_budget_bytes ? hbm_budget_bytes : (uint64_t)(free_b * 0.9);
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device_ordinal) == hipSuccess) {
        c->num_cus = prop.multiProcessorCount;
        c->max_lds = prop.sharedMemPerBlock;
    }
    int rc = c->queue.alloc(c.get(), rsr::kQueueInts * sizeof(int));
    if (rc) return rc;
    HIP_TRY(hipMemsetAsync(c->queue.ptr, 0, rsr::kQueueInts * sizeof(int), c->stream), "work queue");
    HIP_TRY(hipStreamSynchronize(c->stream), "work queue");
    *out = c.release();
    return RSR_OK;
}

// A fresh pair of events for the next main-kernel launch when profiling is on.
const KernelTimer* rsr_context::next_timer() {
    if (!profiling) return nullptr;
    hipEvent_t a, b;
    if (hipEventCreate(&a) != hipSuccess) return nullptr;
    if (hipEventCreate(&b) != hipSuccess) {
        (void)hipEventDestroy(a);
        return nullptr;
    }
    prof_events.emplace_back(a, b);
    timer_slot.begin = a;
    timer_slot.end = b;
    return &timer_slot;
}

static void drain_profile(rsr_context* ctx) {
    (void)hipStreamSynchronize(ctx->stream);
    for (auto& ev : ctx->prof_events) {
        float ms = 0.0f;
        if (hipEventElapsedTime(&ms, ev.first, ev.second) == hipSuccess) {
            ctx->prof_ms += ms;
            ctx->prof_launches++;
        }
        (void)hipEventDestroy(ev.first);
        (void)hipEventDestroy(ev.second);
    }
    ctx->prof_events.clear();
}

extern "C" int rsr_context_profile(rsr_context* ctx, int enable) {
    if (!ctx) return set_error(RSR_E_INVALID_ARG, "ctx is null");
    std::lock_guard<std::mutex> g(ctx->mu);
    HIP_TRY(hipSetDevice(ctx->device), "hipSetDevice");
    drain_profile(ctx);
    ctx->profiling = enable != 0;
    ctx->prof_ms = 0.0;
    ctx->prof_launches = 0;
    return RSR_OK;
}

extern "C" int rsr_context_profile_read(rsr_context* ctx, double* kernel_ms, int64_t* launches) {
    if (!ctx) return set_error(RSR_E_INVALID_ARG, "ctx is null");
    std::lock_guard<std::mutex> g(ctx->mu);
    HIP_TRY(hipSetDevice(ctx->device), "hipSetDevice");
    drain_profile(ctx);
    if (kernel_ms) *kernel_ms = ctx->prof_ms;
    if (launches) *launches = ctx->prof_launches;
    return RSR_OK;
}

extern "C" int rsr_context_synchronize(rsr_context* ctx) {
    if (!ctx) return set_error(RSR_E_INVALID_ARG, "ctx is null");
    HIP_TRY(hipSetDevice(ctx->device), "hipSetDevice");
    HIP_TRY(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    return RSR_OK;
}

extern "C" int rsr_context_destroy(rsr_context* ctx) {
    if (!ctx) return RSR_OK;
    if (ctx->live_servers.load() > 0)
        return set_error(RSR_E_INVALID_ARG, "the context still has a server; destroy the server first");
    if (ctx->live_objects.load() > 0)
        return set_error(RSR_E_INVALID_ARG, "the context still has %d datasets or program batches; destroy them first",
                         ctx->live_objects.load());
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    drain_profile(ctx);
    ctx->scratch_act.release();
    ctx->scratch_partials.release();
    ctx->scratch_gpartials.release();
    ctx->scratch_state.release();
    ctx->scratch_gstate.release();
    ctx->scratch_out.release();
    ctx->scratch_out2.release();
    ctx->scratch_gspill.release();
    ctx->queue.release();
    rsr::trim_pool(ctx);
    if (ctx->host_counters) (void)hipHostFree(ctx->host_counters);
    if (ctx->host_stage) (void)hipHostFree(ctx->host_stage);
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return RSR_OK;
}

extern "C" int rsr_context_memory(const rsr_context* ctx, uint64_t* used_bytes, uint64_t* budget_bytes) {
    if (!ctx) return set_error(RSR_E_INVALID_ARG, "ctx is null");
    if (used_bytes) *used_bytes = ctx->used;
    if (budget_bytes) *budget_bytes = ctx->budget;
    return RSR_OK;
}

extern "C" int rsr_context_set_tuning(rsr_context* ctx, int ppl, int group) {
    if (!ctx) return set_error(RSR_E_INVALID_ARG, "ctx is null");
    if (ppl == 0 && group == 0) {  // back to per-workload tuning
        std::lock_guard<std::mutex> g(ctx->mu);
        ctx->auto_tune = true;
        ctx->tune_version++;
        return RSR_OK;
    }
    if (ppl < 1 || ppl > 4) return set_error(RSR_E_INVALID_ARG, "points_per_lane must be 1, 2, 3 or 4");
    if (group < 1 || group > 4096) return set_error(RSR_E_INVALID_ARG, "programs_per_item out of range");
    std::lock_guard<std::mutex> g(ctx->mu);
    ctx->tuning.ppl = ppl;
    ctx->tuning.group = group;
    ctx->auto_tune = false;
    ctx->tune_version++;
    return RSR_OK;
}

static int ceil_log2(int64_t v) {
    int k = 0;
    while (((int64_t)1 << k) < v) ++k;
    return k;
}

uint64_t rsr::tuning_key(int num_features, int64_t count8, int n_progs, int summation) {
    return ((uint64_t)(summation & 1) << 24) | ((uint64_t)num_features << 16) | ((uint64_t)ceil_log2(count8) << 8) |
           (uint64_t)ceil_log2(n_progs);
}

// Heuristic, fitted to timings of every tuning (points per lane 2 to 4 x programs per item
// 1 to 32) on MI355X over the four workload types at 13 shapes (400 to 10^6 points, 1,000 to
// 16,384 programs), in reference order; it is within 3% of the best tuning at each.
//   - Large evaluations (200,000 or more (tile of 192 points, program) pairs): three points
//     per lane (four waves per SIMD) and 8 programs per work item.
//   - Mid-size ones (8,000 or more pairs of 256-point tiles): four points per lane and the
//     most programs per item, up to 4, that still gives 8,000 work items (an item of 8
//     programs was up to 37% slower there: too few items for the chip).
//   - Smaller ones: two points per lane, the same rule for the programs per item.
static EvalTuning heuristic_tuning(int num_features, int64_t count8, int n_progs) {
    (void)num_features;
    auto tiles = [&](int ppl) { return (count8 + 64 * ppl - 1) / (64 * ppl); };
    EvalTuning t;
    const int64_t np = std::max(n_progs, 1);
    if (tiles(3) * np >= 200000) {
        t.ppl = 3;
        t.group = 8;
        return t;
    }
    t.ppl = tiles(4) * np >= 8000 ? 4 : 2;
    t.group = 1;
    for (int g = 4; g >= 2; g /= 2)
        if (tiles(t.ppl) * ((np + g - 1) / g) >= 8000) {
            t.group = g;
            break;
        }
    return t;
}

EvalTuning rsr::pick_tuning(const rsr_context* ctx, const EvalStep& st) {
    if (!ctx->auto_tune) return ctx->tuning;
    auto it = ctx->tune_cache.find(tuning_key(st.D.num_features, st.D.count8, st.P.n_progs, ctx->summation));
    if (it != ctx->tune_cache.end()) return it->second;
    return heuristic_tuning(st.D.num_features, st.D.count8, st.P.n_progs);
}

extern "C" int rsr_context_set_summation(rsr_context* ctx, int mode) {
    if (!ctx) return set_error(RSR_E_INVALID_ARG, "ctx is null");
    if (mode != RSR_SUM_KAHAN && mode != RSR_SUM_REFERENCE) return set_error(RSR_E_INVALID_ARG, "bad summation mode");
    std::lock_guard<std::mutex> g(ctx->mu);
    ctx->summation = mode;
    ctx->tune_version++;
    return RSR_OK;
}

extern "C" int rsr_context_set_launch_mode(rsr_context* ctx, int mode) {
    if (!ctx) return set_error(RSR_E_INVALID_ARG, "ctx is null");
    if (mode != RSR_LAUNCH_ONESHOT && mode != RSR_LAUNCH_PERSISTENT && mode != RSR_LAUNCH_AUTO)
        return set_error(RSR_E_INVALID_ARG, "bad launch mode");
    std::lock_guard<std::mutex> g(ctx->mu);
    ctx->launch_mode = mode;
    ctx->tune_version++;
    return RSR_OK;
}

uint64_t rsr::next_object_id() {
    static std::atomic<uint64_t> next{1};
    return next.fetch_add(1);
}

rsr::EvalGraph::~EvalGraph() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (host) (void)hipHostFree(host);
}

extern "C" int rsr_context_set_graphs(rsr_context* ctx, int enable) {
    if (!ctx) return set_error(RSR_E_INVALID_ARG, "ctx is null");
    std::lock_guard<std::mutex> g(ctx->mu);
    ctx->graphs = enable != 0;
    return RSR_OK;
}

extern "C" int rsr_context_set_partials_limit(rsr_context* ctx, uint64_t bytes) {
    if (!ctx) return set_error(RSR_E_INVALID_ARG, "ctx is null");
    if (bytes < 4096) return set_error(RSR_E_INVALID_ARG, "partials limit below 4096 bytes");
    std::lock_guard<std::mutex> g(ctx->mu);
    ctx->arena_bytes = (size_t)bytes;
    ctx->tune_version++;
    return RSR_OK;
}

extern "C" int rsr_context_get_tuning(const rsr_context* ctx, int* ppl, int* group) {
    if (!ctx) return set_error(RSR_E_INVALID_ARG, "ctx is null");
    if (ppl) *ppl = ctx->auto_tune ? 0 : ctx->tuning.ppl;
    if (group) *group = ctx->auto_tune ? 0 : ctx->tuning.group;
    return RSR_OK;
}

extern "C" int rsr_context_query_tuning(rsr_context* ctx, rsr_programs* p, const rsr_dataset* ds, int* ppl,
                                        int* group, int* persistent) {
    if (!ctx || !p || !ds) return set_error(RSR_E_INVALID_ARG, "null argument");
    std::lock_guard<std::mutex> g(ctx->mu);
    HIP_TRY(hipSetDevice(ctx->device), "hipSetDevice");
    Lowering* L;
    int rc = get_lowering(p, ds->num_features, false, &L);
    if (rc) return rc;
    EvalStep probe;
    probe.D = dev_view(ds);
    probe.P = dev_view(p, L);
    probe.P.n_progs = p->n - L->n_invalid;
    EvalTuning t = pick_tuning(ctx, probe);
    while (t.ppl > 1 && eval_mse_lds_bytes(L->spill_slots, t.ppl) > ctx->max_lds) t.ppl /= 2;
    if (ppl) *ppl = t.ppl;
    if (group) *group = t.group;
    if (persistent) *persistent = ctx->queue_for(t) != nullptr ? 1 : 0;
    return RSR_OK;
}

// ---------------------------------------------------------------------------------
// datasets
// ---------------------------------------------------------------------------------
static int dataset_upload(rsr_context* ctx, std::unique_ptr<rsr_dataset>& d) {
    d->ctx = ctx;
    // a multiple of every tile size (64 x 1, 2, 3 or 4 points): whole tiles stay in bounds
    const int64_t npad = std::max<int64_t>(768, (d->count8 + 767) / 768 * 768);
    d->npad = npad;
    // host SoA padded to npad with zeros (tile padding is masked in the kernels)
    std::vector<float> f((size_t)std::max(d->num_features, 1) * npad, 0.0f), t((size_t)npad, 0.0f);
    for (int k = 0; k < d->num_features; ++k)
        memcpy(&f[(size_t)k * npad], &d->soa[(size_t)k * d->count8], (size_t)d->count8 * sizeof(float));
    memcpy(t.data(), d->target.data(), (size_t)d->count8 * sizeof(float));
    int rc = d->dfeat.upload(ctx, f.data(), f.size() * sizeof(float));
    if (rc) return rc;
    rc = d->dtarget.upload(ctx, t.data(), t.size() * sizeof(float));
    if (rc) return rc;
    HIP_TRY(hipStreamSynchronize(ctx->stream), "dataset upload");
    return RSR_OK;
}

DevDataset rsr::dev_view(const rsr_dataset* d) {
    DevDataset v;
    v.feat = (const float*)d->dfeat.ptr;
    v.target = (const float*)d->dtarget.ptr;
    v.num_features = d->num_features;
    v.num_samples = d->num_samples;
    v.count8 = d->count8;
    v.npad = d->npad;
    return v;
}

extern "C" int rsr_dataset_create(rsr_context* ctx, const float* rows, const float* y, int64_t num_samples,
                                  const uint8_t* feature_types, int n_feature_types, int normalize,
                                  rsr_dataset** out) {
    if (!ctx || !out || (!rows && num_samples) || (!y && num_samples) || num_samples < 0 || n_feature_types < 0)
        return set_error(RSR_E_INVALID_ARG, "invalid dataset arguments");
    int F = 0;
    for (int i = 0; i < n_feature_types; ++i) {
        int s = vt_size(feature_types[i]);
        F += s ? s : 1;  // dataset.rs:47-54: Int/Bool count one column
    }
    if (F > kRegs) return set_error(RSR_E_UNSUPPORTED, "%d feature components exceed the 32-register file", F);
    std::lock_guard<std::mutex> g(ctx->mu);
    HIP_TRY(hipSetDevice(ctx->device), "hipSetDevice");
    auto d = std::make_unique<rsr_dataset>();
    d->num_features = F;
    d->num_samples = num_samples;
    d->count8 = (num_samples + 7) / 8 * 8;
    d->feature_types.assign(feature_types, feature_types + n_feature_types);
    d->means.assign(F, 0.0f);
    d->stds.assign(F, 1.0f);
    const float ns = (float)num_samples;
    if (normalize && num_samples > 0) {  // dataset.rs:63-94, f32 left-to-right sums
        for (int f = 0; f < F; ++f) {
            float sum = 0.0f;
            for (int64_t i = 0; i < num_samples; ++i) sum += rows[i * F + f];
            const float m = sum / ns;
            float var = 0.0f;
            for (int64_t i = 0; i < num_samples; ++i) {
                const float dd = rows[i * F + f] - m;
                var += dd * dd;
            }
            var = var / ns;
            d->means[f] = m;
            d->stds[f] = var < 1e-9f ? 1.0f : sqrtf(var);
        }
        float ts = 0.0f;
        for (int64_t i = 0; i < num_samples; ++i) ts += y[i];
        d->tmean = ts / ns;
        float tv = 0.0f;
        for (int64_t i = 0; i < num_samples; ++i) {
            const float dd = y[i] - d->tmean;
            tv += dd * dd;
        }
        tv = tv / ns;
        d->tstd = tv < 1e-9f ? 1.0f : sqrtf(tv);
    }
    d->soa.assign((size_t)F * d->count8, 0.0f);
    d->target.assign((size_t)d->count8, 0.0f);
    for (int f = 0; f < F; ++f)
        for (int64_t i = 0; i < num_samples; ++i)
            d->soa[(size_t)f * d->count8 + i] = (rows[i * F + f] - d->means[f]) / d->stds[f];
    for (int64_t i = 0; i < num_samples; ++i) d->target[i] = (y[i] - d->tmean) / d->tstd;
    int rc = dataset_upload(ctx, d);
    if (rc) return rc;
    ctx->live_datasets.insert(d->id);
    ctx->live_objects++;
    *out = d.release();
    return RSR_OK;
}

extern "C" int rsr_dataset_create_aosoa(rsr_context* ctx, const float* feature_flat, const float* target_batches,
                                        int num_features, int64_t num_batches, int64_t num_samples,
                                        rsr_dataset** out) {
    if (!ctx || !out || num_features < 0 || num_batches < 0 || num_samples < 0 || num_samples > num_batches * 8 ||
        (num_batches && (!target_batches || (num_features && !feature_flat))))
        return set_error(RSR_E_INVALID_ARG, "invalid AoSoA dataset arguments");
    if (num_features > kRegs) return set_error(RSR_E_UNSUPPORTED, "too many feature components");
    std::lock_guard<std::mutex> g(ctx->mu);
    HIP_TRY(hipSetDevice(ctx->device), "hipSetDevice");
    auto d = std::make_unique<rsr_dataset>();
    const int F = num_features;
    d->num_features = F;
    d->num_samples = num_samples;
    d->count8 = num_batches * 8;
    d->means.assign(F, 0.0f);
    d->stds.assign(F, 1.0f);
    d->soa.assign((size_t)F * d->count8, 0.0f);
    d->target.assign((size_t)d->count8, 0.0f);
    for (int64_t b = 0; b < num_batches; ++b)
        for (int l = 0; l < 8; ++l) {
            for (int f = 0; f < F; ++f) d->soa[(size_t)f * d->count8 + b * 8 + l] = feature_flat[(b * F + f) * 8 + l];
            d->target[b * 8 + l] = target_batches[b * 8 + l];
        }
    int rc = dataset_upload(ctx, d);
    if (rc) return rc;
    ctx->live_datasets.insert(d->id);
    ctx->live_objects++;
    *out = d.release();
    return RSR_OK;
}

extern "C" int rsr_dataset_subset(rsr_context* ctx, const rsr_dataset* src, int64_t target_samples,
                                  rsr_dataset** out) {
    if (!ctx || !src || !out || target_samples < 0) return set_error(RSR_E_INVALID_ARG, "invalid subset arguments");
    std::lock_guard<std::mutex> g(ctx->mu);
    HIP_TRY(hipSetDevice(ctx->device), "hipSetDevice");
    auto d = std::make_unique<rsr_dataset>();
    const int F = src->num_features;
    const int64_t samples = std::min(target_samples, src->num_samples);
    d->num_features = F;
    d->feature_types = src->feature_types;
    d->means = src->means;
    d->stds = src->stds;
    d->tmean = src->tmean;
    d->tstd = src->tstd;
    if (samples == src->num_samples) {  // clone (dataset.rs:362-364)
        d->num_samples = src->num_samples;
        d->count8 = src->count8;
        d->soa = src->soa;
        d->target = src->target;
    } else {
        d->num_samples = samples;
        d->count8 = (samples + 7) / 8 * 8;
        const double step = ((double)src->num_samples - 1.0) / std::max((double)samples - 1.0, 1.0);
        d->soa.assign((size_t)F * d->count8, 0.0f);
        d->target.assign((size_t)d->count8, 0.0f);
        for (int64_t i = 0; i < d->count8; ++i) {
            const int64_t orig = (int64_t)llround((double)i * step);  // f64 round half away from zero
            if (orig >= src->num_samples) continue;
            for (int f = 0; f < F; ++f) d->soa[(size_t)f * d->count8 + i] = src->soa[(size_t)f * src->count8 + orig];
            d->target[i] = src->target[orig];
        }
    }
    int rc = dataset_upload(ctx, d);
    if (rc) return rc;
    ctx->live_datasets.insert(d->id);
    ctx->live_objects++;
    *out = d.release();
    return RSR_OK;
}

extern "C" int rsr_dataset_info(const rsr_dataset* ds, int* num_features, int64_t* num_samples,
                                int64_t* num_batches) {
    if (!ds) return set_error(RSR_E_INVALID_ARG, "ds is null");
    if (num_features) *num_features = ds->num_features;
    if (num_samples) *num_samples = ds->num_samples;
    if (num_batches) *num_batches = ds->count8 / 8;
    return RSR_OK;
}

extern "C" int rsr_dataset_stats(const rsr_dataset* ds, float* stats_out) {
    if (!ds || !stats_out) return set_error(RSR_E_INVALID_ARG, "null argument");
    const int F = ds->num_features;
    for (int f = 0; f < F; ++f) {
        stats_out[f] = ds->means[f];
        stats_out[F + f] = ds->stds[f];
    }
    stats_out[2 * F] = ds->tmean;
    stats_out[2 * F + 1] = ds->tstd;
    return RSR_OK;
}

extern "C" int rsr_dataset_destroy(rsr_dataset* ds) {
    if (!ds) return RSR_OK;
    rsr_context* ctx = ds->ctx;
    std::lock_guard<std::mutex> g(ctx->mu);
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    ctx->live_datasets.erase(ds->id);
    delete ds;
    ctx->live_objects--;
    return RSR_OK;
}

// ---------------------------------------------------------------------------------
// program batches
// ---------------------------------------------------------------------------------
RefProgram rsr::ref_program(const rsr_programs* p, int i) {
    RefProgram r;
    r.ops = p->ops.data() + p->code_off[i];
    r.args = p->args.data() + p->code_off[i];
    r.n_code = (int)(p->code_off[i + 1] - p->code_off[i]);
    r.const_types = p->ctypes.data() + p->const_off[i];
    r.n_consts = (int)(p->const_off[i + 1] - p->const_off[i]);
    r.const_values = p->values.data() + p->value_off[i];
    return r;
}

static void build_params(rsr_programs* p) {
    p->params.clear();
    p->param_off.assign(p->n + 1, 0);
    for (int i = 0; i < p->n; ++i) {
        p->param_off[i] = (uint32_t)p->params.size();
        flatten_params(ref_program(p, i), p->params);
    }
    p->param_off[p->n] = (uint32_t)p->params.size();
    p->params.push_back(0.0f);  // never empty
}

extern "C" int rsr_programs_create(rsr_context* ctx, int n, const uint8_t* ops, const uint16_t* args,
                                   const int64_t* code_offsets, const uint8_t* const_types,
                                   const int64_t* const_offsets, const float* const_values,
                                   const int64_t* value_offsets, rsr_programs** out) {
    if (!ctx || !out || n < 0 || (n > 0 && (!code_offsets || !const_offsets || !value_offsets)))
        return set_error(RSR_E_INVALID_ARG, "invalid program batch arguments");
    auto p = std::make_unique<rsr_programs>();
    p->ctx = ctx;
    p->n = n;
    const int64_t ncode = n ? code_offsets[n] - code_offsets[0] : 0;
    const int64_t ncons = n ? const_offsets[n] - const_offsets[0] : 0;
    if (ncode < 0 || ncons < 0 || (ncode && (!ops || !args)) || (ncons && !const_types))
        return set_error(RSR_E_INVALID_ARG, "invalid offsets");
    p->code_off.resize(n + 1);
    p->const_off.resize(n + 1);
    p->value_off.resize(n + 1);
    for (int i = 0; i <= n; ++i) {
        p->code_off[i] = code_offsets[i] - code_offsets[0];
        p->const_off[i] = const_offsets[i] - const_offsets[0];
    }
    for (int i = 0; i < n; ++i) {
        if (p->code_off[i + 1] < p->code_off[i] || p->const_off[i + 1] < p->const_off[i])
            return set_error(RSR_E_INVALID_ARG, "offsets of program %d are decreasing", i);
    }
    p->ops.assign(ops, ops + ncode);
    p->args.assign(args, args + ncode);
    p->ctypes.assign(const_types, const_types + ncons);
    // payload layout: one contiguous run per program, value_offsets relative to the first
    int64_t nvals = 0;
    for (int i = 0; i < n; ++i) {
        p->value_off[i] = nvals;
        for (int64_t c = p->const_off[i]; c < p->const_off[i + 1]; ++c) {
            int s = vt_size(p->ctypes[c]);
            nvals += s ? s : 1;
        }
    }
    p->value_off[n] = nvals;
    if (nvals && !const_values) return set_error(RSR_E_INVALID_ARG, "const_values is null");
    p->values.resize(nvals + 1);
    for (int i = 0; i < n; ++i) {
        const int64_t len = p->value_off[i + 1] - p->value_off[i];
        if (len) memcpy(&p->values[p->value_off[i]], const_values + (value_offsets[i] - value_offsets[0]),
                        (size_t)len * sizeof(float));
    }
    build_params(p.get());
    {
        std::lock_guard<std::mutex> g(ctx->mu);
        HIP_TRY(hipSetDevice(ctx->device), "hipSetDevice");
        int rc = p->dparams.upload(ctx, p->params.data(), p->params.size() * sizeof(float));
        if (rc) return rc;
        HIP_TRY(hipStreamSynchronize(ctx->stream), "params upload");
    }
    ctx->live_objects++;
    *out = p.release();
    return RSR_OK;
}

extern "C" int rsr_programs_count(const rsr_programs* p, int* n, int64_t* n_values) {
    if (!p) return set_error(RSR_E_INVALID_ARG, "progs is null");
    if (n) *n = p->n;
    if (n_values) *n_values = p->value_off.empty() ? 0 : p->value_off[p->n];
    return RSR_OK;
}

extern "C" int rsr_programs_get_constants(const rsr_programs* p, float* out) {
    if (!p || !out) return set_error(RSR_E_INVALID_ARG, "null argument");
    memcpy(out, p->values.data(), (size_t)p->value_off[p->n] * sizeof(float));
    return RSR_OK;
}

int rsr::upload_params(rsr_programs* p, rsr_context* on) {
    rsr_context* ctx = on ? on : p->ctx;
    HIP_TRY(hipMemcpyAsync(p->dparams.ptr, p->params.data(), p->params.size() * sizeof(float), hipMemcpyHostToDevice,
                           ctx->stream),
            "params upload");
    return RSR_OK;
}

extern "C" int rsr_programs_set_constants(rsr_programs* p, const float* values) {
    if (!p || !values) return set_error(RSR_E_INVALID_ARG, "null argument");
    memcpy(p->values.data(), values, (size_t)p->value_off[p->n] * sizeof(float));
    build_params(p);
    std::lock_guard<std::mutex> g(p->ctx->mu);
    HIP_TRY(hipSetDevice(p->ctx->device), "hipSetDevice");
    int rc = upload_params(p);
    if (rc) return rc;
    HIP_TRY(hipStreamSynchronize(p->ctx->stream), "params upload");
    return RSR_OK;
}

// Lower (or fetch the cached lowering of) a batch for a feature count.
int rsr::get_lowering(rsr_programs* p, int F, bool dual, Lowering** out) {
    const int key = F * 2 + (dual ? 1 : 0);
    auto it = p->lowerings.find(key);
    if (it != p->lowerings.end()) {
        *out = it->second.get();
        return RSR_OK;
    }
    rsr_context* ctx = p->ctx;
    auto L = std::make_unique<Lowering>();
    L->status.resize(p->n);
    L->big.assign(p->n, 0);
    L->nparams.resize(p->n);
    L->differentiable.resize(p->n);
    std::vector<UInstr> code;
    std::vector<ProgDesc> descs(p->n);
    std::vector<uint8_t> active(p->n, 0);
    std::vector<uint32_t> cost(p->n, 0);
    // The programs lower independently: up to 16 host threads, a contiguous range each.
    std::vector<Lowered> lws((size_t)p->n);
    {
        const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
        const int nt = std::min<int>({16, (int)hw, (p->n + 255) / 256});
        auto work = [&](int t, int nthreads) {
            const int a = (int)((int64_t)p->n * t / nthreads), b = (int)((int64_t)p->n * (t + 1) / nthreads);
            for (int i = a; i < b; ++i) lws[(size_t)i] = lower_program(ref_program(p, i), F, dual);
        };
        if (nt <= 1) {
            work(0, 1);
        } else {
            std::vector<std::thread> pool;
            for (int t = 1; t < nt; ++t) pool.emplace_back(work, t, nt);
            work(0, nt);
            for (auto& th : pool) th.join();
        }
    }
    for (int i = 0; i < p->n; ++i) {
        Lowered& lw = lws[(size_t)i];
        L->status[i] = lw.status;
        L->nparams[i] = (uint8_t)std::min<uint32_t>(lw.n_params, 32);
        L->differentiable[i] = lw.differentiable ? 1 : 0;
        descs[i].code_off = (uint32_t)code.size();
        descs[i].param_off = p->param_off[i];
        if (lw.status == PROG_OK) {
            descs[i].code_len = (uint32_t)lw.code.size();
            descs[i].out_reg = lw.out_reg;
            code.insert(code.end(), lw.code.begin(), lw.code.end());
            active[i] = 1;
            cost[i] = (uint32_t)lw.code.size();
            L->total_uops += lw.code.size();
            L->max_params = std::max<int>(L->max_params, L->nparams[i]);
            L->spill_slots = std::max<int>(L->spill_slots, std::min<int>((int)lw.spill_slots, kMaxSpillSlots));
            if ((int)lw.spill_slots > kMaxSpillSlots) {
                descs[i].out_reg |= kBigProg;
                L->big[i] = 1;
                L->n_big++;
                L->gslots = std::max<int>(L->gslots, (int)lw.spill_slots - kMaxSpillSlots);
            }
        } else {
            descs[i].code_len = 0;
            descs[i].out_reg = 0;
            L->n_invalid++;
        }
    }
    code.push_back(UInstr{0u, 0u, 0u, 0u});  // prefetch padding
    // scheduling order: most expensive first, so groups are balanced and the tail is short;
    // the big programs last (their own launches take the last n_big positions)
    L->order.resize(p->n);
    std::iota(L->order.begin(), L->order.end(), 0);
    std::stable_sort(L->order.begin(), L->order.end(), [&](int a, int b) {
        if (L->big[a] != L->big[b]) return L->big[a] < L->big[b];
        return cost[a] > cost[b];
    });
    L->active_host = active;
    L->desc = descs;
    // one host block, one device buffer and one copy for the five arrays (256-byte aligned
    // parts): small batches are dominated by the per-copy cost of the runtime
    auto al = [](size_t b) { return (b + 255) / 256 * 256; };
    const size_t b_code = al(code.size() * sizeof(UInstr)), b_desc = al(std::max<size_t>(1, descs.size()) * sizeof(ProgDesc)),
                 b_order = al(std::max<size_t>(1, L->order.size()) * sizeof(int32_t)),
                 b_active = al(std::max<size_t>(1, active.size())), b_np = al(std::max<size_t>(1, L->nparams.size()));
    std::vector<char> blob(b_code + b_desc + b_order + b_active + b_np, 0);
    size_t at = 0;
    auto put = [&](const void* src, size_t bytes, size_t span) {
        if (bytes) memcpy(blob.data() + at, src, bytes);
        at += span;
        return at - span;
    };
    const size_t o_code = put(code.data(), code.size() * sizeof(UInstr), b_code);
    const size_t o_desc = put(descs.data(), descs.size() * sizeof(ProgDesc), b_desc);
    const size_t o_order = put(L->order.data(), L->order.size() * sizeof(int32_t), b_order);
    const size_t o_active = put(active.data(), active.size(), b_active);
    const size_t o_np = put(L->nparams.data(), L->nparams.size(), b_np);
    int rc;
    if ((rc = L->dbuf.upload(ctx, blob.data(), blob.size()))) return rc;
    const char* base = (const char*)L->dbuf.ptr;
    L->dcode = (const UInstr*)(base + o_code);
    L->ddesc = (const ProgDesc*)(base + o_desc);
    L->dorder = (const int32_t*)(base + o_order);
    L->dactive = (const uint8_t*)(base + o_active);
    L->dnparams = (const uint8_t*)(base + o_np);
    HIP_TRY(hipStreamSynchronize(ctx->stream), "lowering upload");
    *out = L.get();
    p->lowerings[key] = std::move(L);
    return RSR_OK;
}

DevPrograms rsr::dev_view(const rsr_programs* p, const Lowering* L) {
    DevPrograms v;
    v.code = L->dcode;
    v.progs = L->ddesc;
    v.params = (const float*)p->dparams.ptr;
    v.order = L->dorder;
    v.n_progs = p->n;
    v.spill_slots = L->spill_slots;
    v.n_big = L->n_big;
    v.gslots = L->gslots;
    return v;
}

extern "C" int rsr_programs_status(rsr_context* ctx, rsr_programs* p, const rsr_dataset* ds, int32_t* status_out) {
    if (!ctx || !p || !ds || !status_out) return set_error(RSR_E_INVALID_ARG, "null argument");
    std::lock_guard<std::mutex> g(ctx->mu);
    HIP_TRY(hipSetDevice(ctx->device), "hipSetDevice");
    Lowering* L;
    int rc = get_lowering(p, ds->num_features, false, &L);
    if (rc) return rc;
    memcpy(status_out, L->status.data(), (size_t)p->n * sizeof(int32_t));
    return RSR_OK;
}

extern "C" int rsr_programs_destroy(rsr_programs* p) {
    if (!p) return RSR_OK;
    rsr_context* ctx = p->ctx;
    std::lock_guard<std::mutex> g(ctx->mu);
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    delete p;
    ctx->live_objects--;
    return RSR_OK;
}

// ---------------------------------------------------------------------------------
// evaluation
// ---------------------------------------------------------------------------------
extern "C" int rsr_eval_mse(rsr_context* ctx, rsr_programs* p, const rsr_dataset* ds, float* mse_out) {
    if (!ctx || !p || !ds || (!mse_out && p->n)) return set_error(RSR_E_INVALID_ARG, "null argument");
    if (p->n == 0) return RSR_OK;
    std::lock_guard<std::mutex> g(ctx->mu);
    HIP_TRY(hipSetDevice(ctx->device), "hipSetDevice");
    if (ctx->graphs) {
        const int rc = eval_mse_graph(ctx, p, ds, mse_out);
        if (rc != RSR_E_UNSUPPORTED) return rc;
    }
    std::unique_ptr<Job> job;
    int rc = make_eval_job(ctx, p, ds, nullptr, mse_out, &job);
    if (rc) return rc;
    return run_job(ctx, *job);
}

extern "C" int rsr_autotune(rsr_context* ctx, rsr_programs* p, const rsr_dataset* ds, int* points_per_lane,
                            int* programs_per_item) {
    if (!ctx || !p || !ds) return set_error(RSR_E_INVALID_ARG, "null argument");
    if (p->n == 0) return set_error(RSR_E_INVALID_ARG, "empty program batch");
    std::lock_guard<std::mutex> g(ctx->mu);
    HIP_TRY(hipSetDevice(ctx->device), "hipSetDevice");
    return autotune(ctx, p, ds, points_per_lane, programs_per_item);
}

extern "C" int rsr_eval_points(rsr_context* ctx, rsr_programs* p, int first, int count, const rsr_dataset* ds,
                               float* out) {
    if (!ctx || !p || !ds || !out || first < 0 || count < 0 || first + count > p->n)
        return set_error(RSR_E_INVALID_ARG, "invalid eval_points arguments");
    if (count == 0) return RSR_OK;
    std::lock_guard<std::mutex> g(ctx->mu);
    HIP_TRY(hipSetDevice(ctx->device), "hipSetDevice");
    Lowering* L;
    int rc = get_lowering(p, ds->num_features, false, &L);
    if (rc) return rc;
    for (int i = first; i < first + count; ++i)
        if (L->status[i] != PROG_OK) return set_error(RSR_E_PROGRAM, "program %d rejected (%d)", i, L->status[i]);
    const size_t n = (size_t)count * ds->count8;
    if ((rc = ensure_scratch(ctx, ctx->scratch_out, n * sizeof(float)))) return rc;
    bool any_big = false;
    for (int i = first; i < first + count; ++i) any_big = any_big || L->big[i];
    BigSpill big;
    if (any_big) {
        big.floats = big_points_floats(L->gslots, 1);
        if ((rc = ensure_scratch(ctx, ctx->scratch_gspill, big.floats * sizeof(float)))) return rc;
        big.ptr = (float*)ctx->scratch_gspill.ptr;
    }
    HIP_TRY(launch_eval_points(dev_view(p, L), dev_view(ds), first, count, any_big, (float*)ctx->scratch_out.ptr,
                               ctx->stream, big),
            "eval_points launch");
    HIP_TRY(hipMemcpyAsync(out, ctx->scratch_out.ptr, n * sizeof(float), hipMemcpyDeviceToHost, ctx->stream),
            "points download");
    HIP_TRY(hipStreamSynchronize(ctx->stream), "eval_points");
    return RSR_OK;
}

extern "C" int rsr_eval_mse_grad(rsr_context* ctx, rsr_programs* p, const rsr_dataset* ds, float* mse_out,
                                 float* grad_out) {
    if (!ctx || !p || !ds || (p->n && (!mse_out || !grad_out))) return set_error(RSR_E_INVALID_ARG, "null argument");
    if (p->n == 0) return RSR_OK;
    std::lock_guard<std::mutex> g(ctx->mu);
    HIP_TRY(hipSetDevice(ctx->device), "hipSetDevice");
    std::unique_ptr<Job> job;
    int rc = make_grad_job(ctx, p, ds, nullptr, mse_out, grad_out, &job);
    if (rc) return rc;
    return run_job(ctx, *job);
}

extern "C" int rsr_eval_points_dual(rsr_context* ctx, rsr_programs* p, int prog, int active_param,
                                    const rsr_dataset* ds, float* val_out, float* grad_out) {
    if (!ctx || !p || !ds || !val_out || !grad_out || prog < 0 || prog >= p->n || active_param < 0)
        return set_error(RSR_E_INVALID_ARG, "invalid eval_points_dual arguments");
    std::lock_guard<std::mutex> g(ctx->mu);
    HIP_TRY(hipSetDevice(ctx->device), "hipSetDevice");
    Lowering* L;
    int rc = get_lowering(p, ds->num_features, true, &L);
    if (rc) return rc;
    if (L->status[prog] != PROG_OK) return set_error(RSR_E_PROGRAM, "program %d rejected (%d)", prog, L->status[prog]);
    const size_t n = (size_t)ds->count8;
    if ((rc = ensure_scratch(ctx, ctx->scratch_out, 2 * n * sizeof(float)))) return rc;
    float* dv = (float*)ctx->scratch_out.ptr;
    BigSpill big;
    if (L->big[prog]) {
        big.floats = big_points_floats(L->gslots, 2);
        if ((rc = ensure_scratch(ctx, ctx->scratch_gspill, big.floats * sizeof(float)))) return rc;
        big.ptr = (float*)ctx->scratch_gspill.ptr;
    }
    HIP_TRY(launch_eval_points_dual(dev_view(p, L), dev_view(ds), prog, active_param, L->big[prog] != 0, dv, dv + n,
                                    ctx->stream, big),
            "eval_points_dual launch");
    HIP_TRY(hipMemcpyAsync(val_out, dv, n * sizeof(float), hipMemcpyDeviceToHost, ctx->stream), "download");
    HIP_TRY(hipMemcpyAsync(grad_out, dv + n, n * sizeof(float), hipMemcpyDeviceToHost, ctx->stream), "download");
    HIP_TRY(hipStreamSynchronize(ctx->stream), "eval_points_dual");
    return RSR_OK;
}
