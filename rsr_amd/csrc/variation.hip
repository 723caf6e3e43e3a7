// variation.hip -- device-side variation: the search loop's generator, mutations and
// crossover (operators/{generator,mutation,crossover}.rs) run on the GPU, one thread per
// offspring, over parents resident in device memory.  Each offspring draws from its own
// Xoshiro256++ stream keyed by its task's seed, so every offspring of a generation is
// produced at once and the result does not depend on the launch shape.
#include <string.h>

#include <mutex>
#include <vector>

#include "../../include/rsr_amd.h"
#include "runtime_internal.h"
#include "variation.h"
#include "variation_host.h"

#define HIP_TRY(expr, what)                               \
    do {                                                  \
        hipError_t e_ = (expr);                           \
        if (e_ != hipSuccess) return hip_error(e_, what); \
    } while (0)

namespace rsr {
namespace {

struct VarySet {  // device views of one launch
    const rsr_vary_task* tasks;
    int n;
    // parents (rsr_programs_create layout)
    const uint8_t* ops;
    const uint16_t* args;
    const int64_t* co;
    const uint8_t* ct;
    const int64_t* cof;
    const float* cv;
    const int64_t* vo;
    // registry and operators
    const uint8_t* var_type;
    const uint8_t* var_idx;
    int n_vars;
    const uint8_t* allowed;
    int n_allowed;
    // fixed-capacity output slots (one per task, worked in place)
    uint8_t* o_ops;
    uint16_t* o_args;
    uint8_t* o_ct;
    float* o_cv;
    int32_t* o_meta;  // [n][8]: len, constants, payload floats, status, offsets of the packed copy
    // the packed copy: each task reserves its exact extent (bump: nodes, constants, values)
    uint8_t* p_ops;
    uint16_t* p_args;
    uint8_t* p_ct;
    float* p_cv;
    int* bump;
};

__device__ VProg parent_view(const VarySet& S, int i) {
    VProg p;
    const int64_t c0 = S.co[i], k0 = S.cof[i];
    p.op = S.ops + c0;
    p.arg = S.args + c0;
    p.ctype = S.ct + k0;
    p.cval = S.cv + S.vo[i];
    p.len = (int)(S.co[i + 1] - c0);
    p.nconst = (int)(S.cof[i + 1] - k0);
    return p;
}

// One offspring per wave (lane 0; the other lanes leave at once).  Each offspring is a
// serial walk with its own control flow: 64 offspring in one wave's lanes would run the
// union of their paths one after the other, where one per wave spreads them over the CUs.
__global__ __launch_bounds__(64) void k_vary(const VarySet S) {
    if (threadIdx.x != 0) return;
    const int t = blockIdx.x;
    if (t >= S.n) return;
    const rsr_vary_task task = S.tasks[t];
    VRng r = vrng_seed(task.seed);
    VOut o;
    o.op = S.o_ops + (int64_t)t * kVMaxNodes;
    o.arg = S.o_args + (int64_t)t * kVMaxNodes;
    o.ctype = S.o_ct + (int64_t)t * kVMaxNodes;
    o.cval = S.o_cv + (int64_t)t * kVMaxVals;
    o.len = o.nconst = o.nval = 0;
    o.overflow = false;
    VEnv env;
    env.var_type = S.var_type;
    env.var_idx = S.var_idx;
    env.n_vars = S.n_vars;
    env.ops = S.allowed;
    env.n_ops = S.n_allowed;
    int kind = task.kind;
    if (kind == RSR_VARY_MUTATE) {  // island.rs:156-179: point, constant or subtree, uniformly
        const uint32_t m = vrng_below(r, 3);
        kind = m == 0 ? RSR_VARY_POINT : m == 1 ? RSR_VARY_CONSTANT : RSR_VARY_SUBTREE;
    }
    switch (kind) {
        case RSR_VARY_GENERATE: v_generate(o, r, env, task.type, task.depth); break;
        case RSR_VARY_CROSSOVER:
            v_crossover(o, r, parent_view(S, task.a), parent_view(S, task.b), task.max_size);
            break;
        case RSR_VARY_POINT: v_point_mutation(o, r, env, parent_view(S, task.a)); break;
        case RSR_VARY_CONSTANT: v_constant_perturbation(o, r, parent_view(S, task.a)); break;
        default: v_subtree_mutation(o, r, env, parent_view(S, task.a), task.max_size, task.depth); break;
    }
    int32_t* meta = S.o_meta + 8 * (int64_t)t;
    meta[0] = o.len;
    meta[1] = o.nconst;
    meta[2] = o.nval;
    meta[3] = o.overflow ? 1 : (kind << 8);
    if (o.overflow) return;
    // pack: the host copies back only the used extents, and reassembles in task order
    const int at_n = atomicAdd(&S.bump[0], o.len), at_k = atomicAdd(&S.bump[1], o.nconst),
              at_v = atomicAdd(&S.bump[2], o.nval);
    meta[4] = at_n;
    meta[5] = at_k;
    meta[6] = at_v;
    for (int i = 0; i < o.len; ++i) {
        S.p_ops[at_n + i] = o.op[i];
        S.p_args[at_n + i] = o.arg[i];
    }
    for (int i = 0; i < o.nconst; ++i) S.p_ct[at_k + i] = o.ctype[i];
    for (int i = 0; i < o.nval; ++i) S.p_cv[at_v + i] = o.cval[i];
}

bool needs_parent(int kind) { return kind != RSR_VARY_GENERATE; }

// The packed offspring's four arrays (code, operands, constant types, values), each at the
// start of its own capacity-sized region, gathered into one contiguous block for one
// download: bytes [0, n0) from a, [n0, n0 + n1) from b, and so on.
struct Gather4 {
    const uint8_t* src[4];
    int64_t end[4];  // cumulative byte counts
    uint8_t* dst;
};
__global__ __launch_bounds__(256) void k_gather4(const Gather4 G) {
    const int64_t total = G.end[3];
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
        const int s = i < G.end[0] ? 0 : i < G.end[1] ? 1 : i < G.end[2] ? 2 : 3;
        const int64_t b = s == 0 ? 0 : G.end[s - 1];
        G.dst[i] = G.src[s][i - b];
    }
}

}  // namespace

int vary(rsr_context* ctx, const std::vector<rsr_vary_task>& tasks, const Wire& parents, const VaryEnv& env,
         Wire& out, std::vector<uint8_t>* applied) {
    out.clear();
    const int n = (int)tasks.size();
    if (n == 0) return RSR_OK;
    if (n > 65536) return set_error(RSR_E_INVALID_ARG, "%d tasks: at most 65536 per call", n);
    const int np = parents.n();
    for (int t = 0; t < n; ++t) {
        const rsr_vary_task& k = tasks[t];
        if (k.kind > RSR_VARY_MUTATE) return set_error(RSR_E_INVALID_ARG, "task %d: unknown kind %d", t, k.kind);
        if (needs_parent(k.kind) && (k.a < 0 || k.a >= np)) return set_error(RSR_E_INVALID_ARG, "task %d: parent a", t);
        if (k.kind == RSR_VARY_CROSSOVER && (k.b < 0 || k.b >= np))
            return set_error(RSR_E_INVALID_ARG, "task %d: parent b", t);
        if ((k.kind == RSR_VARY_GENERATE || k.kind == RSR_VARY_SUBTREE || k.kind == RSR_VARY_MUTATE) &&
            (k.depth < 0 || k.depth > 5))
            return set_error(RSR_E_INVALID_ARG, "task %d: depth %d outside [0, 5]", t, k.depth);
        if (k.kind == RSR_VARY_GENERATE && (k.type == VtBool || k.type == VtInt || k.type > VtMat3))
            return set_error(RSR_E_INVALID_ARG, "task %d: target type %d", t, k.type);
    }
    // the kernel walks parents by their LoadConst nodes and payload sizes: every parent must
    // be well formed (opcodes, one constant per LoadConst node, payloads of its types)
    for (int i = 0; i < np; ++i) {
        if (parents.len(i) > kVMaxNodes)
            return set_error(RSR_E_INVALID_ARG, "parent %d has %lld nodes, above %d", i, (long long)parents.len(i),
                             kVMaxNodes);
        int64_t nk = 0, nv = 0;
        for (int64_t k = parents.co[i]; k < parents.co[i + 1]; ++k) {
            if (parents.ops[k] > NegF) return set_error(RSR_E_INVALID_ARG, "parent %d: opcode %d", i, parents.ops[k]);
            nk += v_is_const(parents.ops[k]);
        }
        if (nk != parents.cof[i + 1] - parents.cof[i])
            return set_error(RSR_E_INVALID_ARG, "parent %d: %lld constant nodes, %lld constants", i, (long long)nk,
                             (long long)(parents.cof[i + 1] - parents.cof[i]));
        for (int64_t k = parents.cof[i]; k < parents.cof[i + 1]; ++k) {
            if (parents.ct[k] > VtMat3) return set_error(RSR_E_INVALID_ARG, "parent %d: constant type", i);
            nv += v_payload(parents.ct[k]);
        }
        if (nv != parents.vo[i + 1] - parents.vo[i]) return set_error(RSR_E_INVALID_ARG, "parent %d: payload size", i);
    }
    for (size_t i = 0; i < env.ops.size(); ++i)
        if (env.ops[i] < AddF || env.ops[i] >= NegF)
            return set_error(RSR_E_INVALID_ARG, "allowed operator %d is not an operator", env.ops[i]);
    for (size_t i = 0; i < env.var_type.size(); ++i)
        if (env.var_type[i] > VtMat3) return set_error(RSR_E_INVALID_ARG, "variable %zu: type", i);

    std::lock_guard<std::mutex> g(ctx->mu);
    hipError_t e = hipSetDevice(ctx->device);
    if (e != hipSuccess) return hip_error(e, "hipSetDevice");
    // one device buffer: the inputs, the output slots, their packed copy and the metadata
    auto al = [](size_t b) { return (b + 255) / 256 * 256; };
    const size_t b_tasks = al(n * sizeof(rsr_vary_task)), b_ops = al(parents.ops.size() + 1),
                 b_args = al(2 * parents.args.size() + 2), b_co = al(8 * parents.co.size()),
                 b_ct = al(parents.ct.size() + 1), b_cof = al(8 * parents.cof.size()),
                 b_cv = al(4 * parents.cv.size() + 4), b_vo = al(8 * parents.vo.size()),
                 b_vt = al(env.var_type.size() + 1), b_vi = al(env.var_idx.size() + 1), b_al = al(env.ops.size() + 1),
                 b_bump = al(3 * sizeof(int)), b_oops = al((size_t)n * kVMaxNodes),
                 b_oargs = al((size_t)n * kVMaxNodes * 2), b_oct = al((size_t)n * kVMaxNodes),
                 b_ocv = al((size_t)n * kVMaxVals * 4), b_meta = al((size_t)n * 32);
    // the uploaded inputs end with the zeroed extents counter, which the kernel's metadata
    // follows, so one download returns both; then the output slots and their packed copy
    const size_t in_bytes = b_tasks + b_ops + b_args + b_co + b_ct + b_cof + b_cv + b_vo + b_vt + b_vi + b_al + b_bump;
    const size_t slots = b_oops + b_oargs + b_oct + b_ocv;
    const size_t total = in_bytes + b_meta + 2 * slots;
    DevBuf buf;
    int rc = buf.alloc(ctx, total);
    if (rc) return rc;
    std::vector<char> host(in_bytes, 0);
    size_t off = 0;
    auto put = [&](const void* src, size_t bytes, size_t span) {
        const size_t at = off;
        if (bytes) memcpy(host.data() + at, src, bytes);
        off += span;
        return (char*)buf.ptr + at;
    };
    VarySet S;
    S.tasks = (const rsr_vary_task*)put(tasks.data(), n * sizeof(rsr_vary_task), b_tasks);
    S.n = n;
    S.ops = (const uint8_t*)put(parents.ops.data(), parents.ops.size(), b_ops);
    S.args = (const uint16_t*)put(parents.args.data(), 2 * parents.args.size(), b_args);
    S.co = (const int64_t*)put(parents.co.data(), 8 * parents.co.size(), b_co);
    S.ct = (const uint8_t*)put(parents.ct.data(), parents.ct.size(), b_ct);
    S.cof = (const int64_t*)put(parents.cof.data(), 8 * parents.cof.size(), b_cof);
    S.cv = (const float*)put(parents.cv.data(), 4 * parents.cv.size(), b_cv);
    S.vo = (const int64_t*)put(parents.vo.data(), 8 * parents.vo.size(), b_vo);
    S.var_type = (const uint8_t*)put(env.var_type.data(), env.var_type.size(), b_vt);
    S.var_idx = (const uint8_t*)put(env.var_idx.data(), env.var_idx.size(), b_vi);
    S.n_vars = (int)env.var_type.size();
    S.allowed = (const uint8_t*)put(env.ops.data(), env.ops.size(), b_al);
    S.n_allowed = (int)env.ops.size();
    S.bump = (int*)put(nullptr, 0, b_bump);  // zero (the staging buffer is zero-filled)
    S.o_meta = (int32_t*)((char*)buf.ptr + in_bytes);
    char* o = (char*)buf.ptr + in_bytes + b_meta;
    S.o_ops = (uint8_t*)o;
    S.o_args = (uint16_t*)(o + b_oops);
    S.o_ct = (uint8_t*)(o + b_oops + b_oargs);
    S.o_cv = (float*)(o + b_oops + b_oargs + b_oct);
    char* pk = o + slots;
    S.p_ops = (uint8_t*)pk;
    S.p_args = (uint16_t*)(pk + b_oops);
    S.p_ct = (uint8_t*)(pk + b_oops + b_oargs);
    S.p_cv = (float*)(pk + b_oops + b_oargs + b_oct);
    HIP_TRY(hipMemcpyAsync(buf.ptr, host.data(), in_bytes, hipMemcpyHostToDevice, ctx->stream), "variation upload");
    hipLaunchKernelGGL(k_vary, dim3((unsigned)n), dim3(64), 0, ctx->stream, S);
    HIP_TRY(hipGetLastError(), "variation launch");
    // the extents counter and the metadata: one download
    std::vector<int32_t> head(b_bump / 4 + (size_t)n * 8);
    HIP_TRY(hipMemcpyAsync(head.data(), S.bump, head.size() * 4, hipMemcpyDeviceToHost, ctx->stream),
            "variation meta");
    HIP_TRY(hipStreamSynchronize(ctx->stream), "variation");
    const int used[3] = {head[0], head[1], head[2]};
    const int32_t* meta = head.data() + b_bump / 4;
    for (int t = 0; t < n; ++t)
        if (meta[8 * (size_t)t + 3] & 1)
            return set_error(RSR_E_UNSUPPORTED, "task %d: offspring above %d nodes", t, kVMaxNodes);
    // the packed extents only, gathered into the (no longer needed) output slots and
    // downloaded in one copy: code, operands (2-byte aligned), types, values (4-byte aligned)
    const int64_t n_ops = used[0], n_args = 2 * (int64_t)used[0], n_ct = used[1], n_cv = 4 * (int64_t)used[2];
    const int64_t a_args = (n_ops + 1) / 2 * 2, a_ct = a_args + n_args, a_cv = (a_ct + n_ct + 3) / 4 * 4;
    const int64_t packed = a_cv + n_cv;
    if ((size_t)packed > slots) return set_error(RSR_E_HIP, "variation: packed offspring exceed their slots");
    std::vector<float> blk((size_t)packed / 4 + 1);
    if (packed > 0) {
        Gather4 G;
        G.src[0] = S.p_ops;
        G.src[1] = (const uint8_t*)S.p_args;
        G.src[2] = S.p_ct;
        G.src[3] = (const uint8_t*)S.p_cv;
        // a part's padding bytes (at most 3) are read past its used extent, inside the
        // buffer, and never read back below
        G.end[0] = a_args;
        G.end[1] = a_ct;
        G.end[2] = a_cv;
        G.end[3] = packed;
        G.dst = (uint8_t*)o;
        hipLaunchKernelGGL(k_gather4, dim3((unsigned)std::min<int64_t>((packed + 255) / 256, 1024)), dim3(256), 0,
                           ctx->stream, G);
        HIP_TRY(hipGetLastError(), "variation gather");
        HIP_TRY(hipMemcpyAsync(blk.data(), o, (size_t)packed, hipMemcpyDeviceToHost, ctx->stream), "variation download");
        HIP_TRY(hipStreamSynchronize(ctx->stream), "variation download");
    }
    const uint8_t* hb = (const uint8_t*)blk.data();
    const uint8_t* ops = hb;
    const uint16_t* args = (const uint16_t*)(hb + a_args);
    const uint8_t* ct = hb + a_ct;
    const float* cv = (const float*)(hb + a_cv);
    if (applied) applied->assign((size_t)n, 0);
    for (int t = 0; t < n; ++t) {  // in task order, whatever order the tasks packed in
        const int32_t* m = meta + 8 * (size_t)t;
        out.ops.insert(out.ops.end(), ops + m[4], ops + m[4] + m[0]);
        out.args.insert(out.args.end(), args + m[4], args + m[4] + m[0]);
        out.ct.insert(out.ct.end(), ct + m[5], ct + m[5] + m[1]);
        out.cv.insert(out.cv.end(), cv + m[6], cv + m[6] + m[2]);
        out.co.push_back((int64_t)out.ops.size());
        out.cof.push_back((int64_t)out.ct.size());
        out.vo.push_back((int64_t)out.cv.size());
        if (applied) (*applied)[t] = (uint8_t)(m[3] >> 8);
    }
    return RSR_OK;
}

}  // namespace rsr

using namespace rsr;

extern "C" RSR_API int rsr_vary(rsr_context* ctx, int n_tasks, const rsr_vary_task* tasks, int n_parents,
                                const uint8_t* ops, const uint16_t* args, const int64_t* code_offsets,
                                const uint8_t* const_types, const int64_t* const_offsets, const float* const_values,
                                const int64_t* value_offsets, int n_vars, const uint8_t* var_types,
                                const uint8_t* var_indices, int n_allowed, const uint8_t* allowed_ops,
                                int64_t code_capacity, int64_t const_capacity, int64_t value_capacity,
                                uint8_t* ops_out, uint16_t* args_out, int64_t* code_offsets_out,
                                uint8_t* const_types_out, int64_t* const_offsets_out, float* const_values_out,
                                int64_t* value_offsets_out, uint8_t* kind_out) {
    if (!ctx || n_tasks < 0 || n_parents < 0 || n_vars < 0 || n_allowed < 0 || (n_tasks && !tasks) ||
        (n_parents && (!code_offsets || !const_offsets || !value_offsets)) || (n_vars && (!var_types || !var_indices)) ||
        (n_allowed && !allowed_ops) || !code_offsets_out || !const_offsets_out || !value_offsets_out)
        return set_error(RSR_E_INVALID_ARG, "null argument");
    Wire parents;
    for (int i = 0; i < n_parents; ++i) parents.add(ops, args, code_offsets, const_types, const_offsets, const_values,
                                                     value_offsets, i);
    VaryEnv env;
    env.var_type.assign(var_types, var_types + n_vars);
    env.var_idx.assign(var_indices, var_indices + n_vars);
    env.ops.assign(allowed_ops, allowed_ops + n_allowed);
    Wire out;
    std::vector<uint8_t> kinds;
    int rc = vary(ctx, std::vector<rsr_vary_task>(tasks, tasks + n_tasks), parents, env, out, &kinds);
    if (rc) return rc;
    if ((int64_t)out.ops.size() > code_capacity || (int64_t)out.ct.size() > const_capacity ||
        (int64_t)out.cv.size() > value_capacity)
        return set_error(RSR_E_INVALID_ARG, "outputs need %zu code, %zu constant and %zu value entries",
                         out.ops.size(), out.ct.size(), out.cv.size());
    if (!out.ops.empty() && (!ops_out || !args_out)) return set_error(RSR_E_INVALID_ARG, "null code output");
    if (!out.ct.empty() && !const_types_out) return set_error(RSR_E_INVALID_ARG, "null constant output");
    if (!out.cv.empty() && !const_values_out) return set_error(RSR_E_INVALID_ARG, "null value output");
    if (!out.ops.empty()) {
        memcpy(ops_out, out.ops.data(), out.ops.size());
        memcpy(args_out, out.args.data(), 2 * out.args.size());
    }
    if (!out.ct.empty()) memcpy(const_types_out, out.ct.data(), out.ct.size());
    if (!out.cv.empty()) memcpy(const_values_out, out.cv.data(), 4 * out.cv.size());
    memcpy(code_offsets_out, out.co.data(), 8 * out.co.size());
    memcpy(const_offsets_out, out.cof.data(), 8 * out.cof.size());
    memcpy(value_offsets_out, out.vo.data(), 8 * out.vo.size());
    if (kind_out) memcpy(kind_out, kinds.data(), kinds.size());
    return RSR_OK;
}

// The variation streams on the host (the kernel's own functions): raw Xoshiro256++ outputs
// from a given state, the state seed_from_u64 makes of a seed, and the samplers.
extern "C" RSR_API int rsr_xoshiro256pp(const uint64_t* state, int n, uint64_t* out) {
    if (!state || n < 0 || (n && !out)) return set_error(RSR_E_INVALID_ARG, "null argument");
    VRng r;
    for (int i = 0; i < 4; ++i) r.s[i] = state[i];
    for (int i = 0; i < n; ++i) out[i] = vrng_next(r);
    return RSR_OK;
}

extern "C" RSR_API int rsr_xoshiro_seed(uint64_t seed, uint64_t* state_out) {
    if (!state_out) return set_error(RSR_E_INVALID_ARG, "null argument");
    const VRng r = vrng_seed(seed);
    for (int i = 0; i < 4; ++i) state_out[i] = r.s[i];
    return RSR_OK;
}

extern "C" RSR_API int rsr_vary_samples(uint64_t seed, int kind, int n, float a, float b, uint32_t bound,
                                        float* out) {
    if (n < 0 || (n && !out) || kind < 0 || kind > 3 || (kind == 2 && bound == 0) || (kind == 3 && !(a < b)))
        return set_error(RSR_E_INVALID_ARG, "invalid sampler arguments");
    VRng r = vrng_seed(seed);
    for (int i = 0; i < n; ++i) {
        switch (kind) {
            case 0: out[i] = vrng_f32(r); break;
            case 1: out[i] = vrng_bool(r) ? 1.0f : 0.0f; break;
            case 2: out[i] = (float)vrng_below(r, bound); break;
            default: out[i] = vrng_range(r, a, b); break;
        }
    }
    return RSR_OK;
}
