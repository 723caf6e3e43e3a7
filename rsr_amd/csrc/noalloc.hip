// noalloc.hip -- rsr_eval_mse_noalloc: compute_mse of a program batch with no device
// allocation at all, for callers that cannot allocate (code under HIP graph capture, tight
// serving loops, harnesses that forbid it).
//
// The evaluation's scratch -- the persistent launch's work queue, the per-tile Kahan partials
// (or, in reference summation, the per-point terms) and the device spill areas of big
// programs -- lives in one file-scope
// device array of kScratchBytes per GPU, held until the process ends.  Host code never
// passes the array's name to a kernel or a copy: its per-device address comes from
// hipGetSymbolAddress, taken once per device on the entry point's first call there.
// Every call on one GPU shares that array (its work queue included), so calls that overlap
// on one GPU race: calls through one context are serialised by the context (one lock, one
// stream); calls through two contexts on the same GPU must not overlap.
#include <float.h>
#include <math.h>
#include <mutex>

#include "../../include/rsr_amd.h"
#include "kernels.h"
#include "runtime_internal.h"

namespace rsr {
namespace {

constexpr size_t kScratchBytes = (size_t)64 << 20;
constexpr int kMaxDevices = 64;

static __device__ __align__(256) float scratch[kScratchBytes / sizeof(float)];

// Per-device address of `scratch` (null until taken), guarded by dev_mu.
float* dev_scratch[kMaxDevices] = {nullptr};
std::mutex dev_mu;

int scratch_for_current_device(int device, float** out) {
    if (device < 0 || device >= kMaxDevices) return set_error(RSR_E_INVALID_ARG, "device ordinal %d", device);
    std::lock_guard<std::mutex> g(dev_mu);
    if (dev_scratch[device] == nullptr) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || p == nullptr) return hip_error(e != hipSuccess ? e : hipErrorInvalidSymbol, "scratch");
        dev_scratch[device] = (float*)p;
    }
    *out = dev_scratch[device];
    return RSR_OK;
}

// NaN for every program the lowering rejected (the evaluation launch skips them).
// Rejected programs get NaN; over a dataset without points every evaluated program gets
// f32::MAX (the reference's 0 / 0 is not finite, evaluator.rs:173).
__global__ void k_fill_rejected(float* __restrict__ out, const uint8_t* __restrict__ active, int n, int empty) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && !active[i]) out[i] = __builtin_nanf("");
    else if (i < n && empty) out[i] = FLT_MAX;
}

}  // namespace
}  // namespace rsr

using namespace rsr;

extern "C" RSR_API int rsr_eval_mse_noalloc(rsr_context* ctx, rsr_programs* p, const rsr_dataset* ds,
                                            float* mse_out_device) {
    if (!ctx || !p || !ds || (!mse_out_device && p->n)) return set_error(RSR_E_INVALID_ARG, "null argument");
    if (p->n == 0) return RSR_OK;
    std::lock_guard<std::mutex> g(ctx->mu);
    hipError_t e = hipSetDevice(ctx->device);
    if (e != hipSuccess) return hip_error(e, "hipSetDevice");
    float* base = nullptr;
    int rc = scratch_for_current_device(ctx->device, &base);
    if (rc) return rc;
    // the lowering must exist already: lowering uploads micro-code (an allocation)
    auto it = p->lowerings.find(ds->num_features * 2);
    if (it == p->lowerings.end())
        return set_error(RSR_E_INVALID_ARG,
                         "the batch is not lowered for %d features yet: call rsr_programs_status (or any evaluation) "
                         "with this dataset first",
                         ds->num_features);
    const Lowering* L = it->second.get();
    const bool terms = ctx->summation == RSR_SUM_REFERENCE;
    EvalStep st;
    st.P = dev_view(p, L);
    st.D = dev_view(ds);
    st.active = L->dactive;
    st.mse_out = mse_out_device;
    // the tuning rsr_eval_mse would take for this batch (so the bits agree)
    EvalTuning t;
    {
        EvalStep probe = st;
        probe.P.n_progs = p->n - L->n_invalid;
        t = pick_tuning(ctx, probe);
    }
    st.group = t.group;
    while (t.ppl > 1 && eval_mse_lds_bytes(L->spill_slots, t.ppl) > ctx->max_lds) t.ppl /= 2;
    if (eval_mse_lds_bytes(L->spill_slots, t.ppl) > ctx->max_lds)
        return set_error(RSR_E_UNSUPPORTED, "spill area exceeds the workgroup LDS limit");
    const size_t part = mse_partials_floats(st, t.ppl, terms);
    const size_t spill = big_spill_floats(&st, 1, t.ppl, false);
    // the array: the work queue (kQueueInts ints, zero at load and after every launch; 256
    // bytes), the partials or terms, then the spill areas (256-byte aligned)
    const size_t head = 64;
    const size_t cap = kScratchBytes / sizeof(float) - head;
    const size_t part_aligned = (part + 63) / 64 * 64;
    if (part_aligned + spill > cap)
        return set_error(RSR_E_INVALID_ARG,
                         "the evaluation needs %zu bytes of scratch, above the %zu bytes of rsr_eval_mse_noalloc; "
                         "use rsr_eval_mse",
                         (head + part_aligned + spill) * sizeof(float), kScratchBytes);
    int* queue = ctx->queue_for(t) ? (int*)base : nullptr;
    st.partials = base + head;
    BigSpill big;
    if (spill) {
        big.ptr = base + head + part_aligned;
        big.floats = cap - part_aligned;
    }
    hipLaunchKernelGGL(k_fill_rejected, dim3((unsigned)((p->n + 255) / 256)), dim3(256), 0, ctx->stream,
                       mse_out_device, st.active, p->n, ds->num_samples == 0 ? 1 : 0);
    if ((e = hipGetLastError()) != hipSuccess) return hip_error(e, "rejected-program fill");
    if ((e = launch_mse_steps(&st, 1, t, terms, ctx->stream, nullptr, big, queue)) != hipSuccess)
        return hip_error(e, "no-alloc evaluation launch");
    if (L->n_invalid) return set_error(RSR_E_PROGRAM, "%d program(s) rejected by lowering", L->n_invalid);
    return RSR_OK;
}
