// runtime_internal.h -- host-side objects behind the C ABI handles.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <map>
#include <memory>
#include <mutex>
#include <atomic>
#include <unordered_set>
#include <vector>

#include "kernels.h"
#include "lower.h"

struct rsr_context;

namespace rsr {

// Device allocation charged against the context's HBM budget.
struct DevBuf {
    void* ptr = nullptr;
    size_t bytes = 0;
    rsr_context* ctx = nullptr;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { release(); }
    int alloc(rsr_context* c, size_t n);
    int upload(rsr_context* c, const void* src, size_t n);
    void release();
};

// One lowering of a program batch (per feature count, primal or dual).
struct Lowering {
    std::vector<int32_t> status;
    std::vector<uint8_t> nparams;         // min(param_count, 32)
    std::vector<uint8_t> differentiable;  // no Inverse ops
    std::vector<int32_t> order;           // cost-sorted scheduling order
    std::vector<uint8_t> active_host;
    std::vector<ProgDesc> desc;           // host copy of the program descriptors
    std::vector<uint8_t> big;             // spills beyond the LDS slots (kBigProg)
    int n_invalid = 0;
    int max_params = 0;
    int spill_slots = 0;                  // LDS spill slots (max over the batch, <= kMaxSpillSlots)
    int n_big = 0;                        // big programs; the scheduling order puts them last
    int gslots = 0;                       // device-memory spill slots of the big programs (max)
    uint64_t total_uops = 0;
    // the device copies of the micro-code, descriptors, scheduling order, active flags and
    // parameter counts: one buffer, one upload
    DevBuf dbuf;
    const UInstr* dcode = nullptr;
    const ProgDesc* ddesc = nullptr;
    const int32_t* dorder = nullptr;
    const uint8_t* dactive = nullptr;
    const uint8_t* dnparams = nullptr;
};

int set_error(int code, const char* fmt, ...);
int hip_error(hipError_t e, const char* what);
uint64_t next_object_id();  // process-unique, never reused

// A captured evaluation of one program batch over one dataset (graph mode).  The graph owns
// its outputs, scheduling order and summation arena, so the context's later allocations
// cannot invalidate them.  The buffers of the dataset, of the batch's lowering and of its
// parameters belong to those objects: the batch's buffers live as long as the batch (its
// parameters are rewritten in place), and a graph is dropped once its dataset is destroyed.
struct EvalGraph {
    uint64_t ds_id = 0;
    uint64_t last_use = 0;  // replay order, for the per-batch bound
    int summation = 0, ppl = 0, group = 0;
    bool persistent = false;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    DevBuf d_mse, d_order, arena;
    float* host = nullptr;  // pinned copy of the MSEs
    ~EvalGraph();
};

}  // namespace rsr

struct rsr_context {
    int device = 0;
    int num_cus = 256;
    size_t max_lds = 65536;  // dynamic LDS bytes per workgroup
    std::vector<uint8_t> host_act;  // staging for active masks (lives until the stream syncs)
    hipStream_t stream = nullptr;
    uint64_t budget = 0;
    uint64_t used = 0;  // device bytes held, including cached free blocks
    std::multimap<size_t, void*> pool;  // released blocks by size
    uint64_t pool_bytes = 0;
    std::recursive_mutex rmu;
    std::mutex mu;
    rsr::EvalTuning tuning;     // the fixed tuning (auto_tune off)
    bool auto_tune = true;      // per-workload tuning: autotuned choice, else the heuristic
    std::map<uint64_t, rsr::EvalTuning> tune_cache;  // rsr_autotune results by workload shape
    // bumped whenever the summation, tuning, tuning cache or partials limit change (server
    // lanes copy the settings only then)
    std::atomic<uint64_t> tune_version{0};
    int summation = 1;  // 1 (default): reference order (bit-identical to the CPU evaluator); 0: pairwise + Kahan
    bool graphs = false;  // rsr_eval_mse replays a captured graph per (programs, dataset, settings)
    rsr::DevBuf scratch_partials, scratch_gpartials, scratch_state, scratch_gstate, scratch_out, scratch_out2, scratch_act;
    rsr::DevBuf scratch_gspill;  // device spill areas of the big-program launches
    size_t arena_bytes = (size_t)4 << 30;  // summation partials / terms per launch round
    // The device work queue of the persistent evaluation launch (zero between launches) and
    // the launch form the evaluations use.
    rsr::DevBuf queue;
    int launch_mode = 2;  // RSR_LAUNCH_AUTO
    // the work queue of a launch at tuning t, or null for the one-shot form
    int* queue_for(const rsr::EvalTuning& t) const {
        return launch_mode == 1 || (launch_mode == 2 && t.persistent) ? (int*)queue.ptr : nullptr;
    }
    // pinned host mirror of the jobs' counters a tick reads back (each job owns its device pair)
    int* host_counters = nullptr;
    size_t host_counters_cap = 0;
    // pinned host staging for the results of the jobs a server lane finishes in one tick
    // (downloaded together, one synchronisation)
    char* host_stage = nullptr;
    size_t host_stage_cap = 0;
    std::atomic<int> live_servers{0};  // a context is not destroyed under a running server
    std::unordered_set<uint64_t> live_datasets;  // ids of the context's datasets (guarded by mu)
    uint64_t graph_clock = 0;                     // graph replays, for the graphs' last_use
    std::atomic<int> live_objects{0};  // ... nor while datasets or program batches use it
    // profiling: events around every main evaluation kernel, accumulated on read
    bool profiling = false;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> prof_events;
    double prof_ms = 0.0;
    int64_t prof_launches = 0;
    const rsr::KernelTimer* next_timer();
    rsr::KernelTimer timer_slot;
};

struct rsr_dataset {
    rsr_context* ctx = nullptr;
    uint64_t id = rsr::next_object_id();
    int num_features = 0;
    int64_t num_samples = 0;
    int64_t count8 = 0;
    int64_t npad = 0;
    std::vector<uint8_t> feature_types;
    std::vector<float> soa;     // host copy [F][count8]
    std::vector<float> target;  // [count8]
    std::vector<float> means, stds;
    float tmean = 0.0f, tstd = 1.0f;
    rsr::DevBuf dfeat, dtarget;
};

struct rsr_programs {
    rsr_context* ctx = nullptr;
    int n = 0;
    std::vector<uint8_t> ops;
    std::vector<uint16_t> args;
    std::vector<uint8_t> ctypes;
    std::vector<float> values;  // constant payloads (create-time layout, compacted)
    std::vector<int64_t> code_off, const_off, value_off;
    std::vector<float> params;  // flat parameters of all programs
    std::vector<uint32_t> param_off;
    rsr::DevBuf dparams;
    std::map<int, std::unique_ptr<rsr::Lowering>> lowerings;
    std::vector<std::unique_ptr<rsr::EvalGraph>> graphs;
};

namespace rsr {
// The evaluation tuning (points per lane, programs per work item) for an MSE step: the
// context's fixed tuning, or in auto mode the autotuned choice for the step's workload
// shape (feature count, point count and program count, each in powers of two, and the
// summation mode), else a
// heuristic fitted to a sweep of every tuning over the four workload types.
EvalTuning pick_tuning(const rsr_context* ctx, const EvalStep& st);
uint64_t tuning_key(int num_features, int64_t count8, int n_progs, int summation);
void trim_pool(rsr_context* c);
int ensure_scratch(rsr_context* c, DevBuf& b, size_t n);
// The context's pinned host staging block, grown to at least `need` bytes (the old block
// freed: callers synchronise before they return, so nothing is in flight into it).
// RSR_E_HIP when pinned memory cannot be had; the callers then download directly.
int ensure_host_stage(rsr_context* c, size_t need);
DevDataset dev_view(const rsr_dataset* d);
DevPrograms dev_view(const rsr_programs* p, const Lowering* L);
RefProgram ref_program(const rsr_programs* p, int i);
int get_lowering(rsr_programs* p, int F, bool dual, Lowering** out);
// on: the context whose stream carries the copy (default: the batch's own)
int upload_params(rsr_programs* p, rsr_context* on = nullptr);
}  // namespace rsr
