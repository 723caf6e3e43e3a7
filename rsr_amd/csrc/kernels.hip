// kernels.hip -- fitness-evaluation kernels for gfx950.
//
// Work decomposition (evaluation): a work item is (point tile, program group).  The
// four waves of a 256-thread workgroup take four consecutive groups of the same tile,
// so the tile's feature columns are fetched from HBM once and served to the other
// waves from L2/L1.  Each wave loads its 64 * PPL points' features straight into the
// low registers of its register file, where they stay resident while the wave
// interprets every program of its group: a program never re-reads the dataset.  Per
// (program, tile) the wave writes one partial sum of squared errors; a finalize kernel
// adds the partials of each program in a fixed order with Kahan-compensated FP32, so
// results are bitwise reproducible run to run.
#include <float.h>
#include "interp.h"
#include "kernels.h"

namespace rsr {

static constexpr int kWaves = 4;  // waves per workgroup

// Load feature columns [0, F) of the lane's points into the register files.
template <int PPL>
__device__ __forceinline__ void load_features(Files<PPL>& R, const float* __restrict__ feat, int F, int64_t npad,
                                              int64_t base, int lane) {
#pragma unroll
    for (int f = 0; f < kRegs; ++f) {
        if (f < F) {
            sfor<PPL>([&](auto P) {
                v32f t = fget<P(), PPL>(R);
                t[f] = feat[(int64_t)f * npad + base + P() * 64 + lane];
                fget<P(), PPL>(R) = t;
            });
        }
    }
}

template <int PPL>
__global__ __launch_bounds__(256) void k_eval_mse(const UInstr* __restrict__ code, const ProgDesc* __restrict__ progs,
                                                  const float* __restrict__ params, const int32_t* __restrict__ order,
                                                  const uint8_t* __restrict__ active, int n_progs, int group,
                                                  int n_groups, const float* __restrict__ feat,
                                                  const float* __restrict__ target, int F, int64_t npad,
                                                  int64_t count8, int64_t n_tiles, float* __restrict__ partials,
                                                  int spill_slots) {
    extern __shared__ float smem[];
    const int lane = threadIdx.x & 63;
    const int64_t wave = (int64_t)uni(blockIdx.x) * kWaves + uni(threadIdx.x >> 6);
    float* lds = smem + (size_t)uni(threadIdx.x >> 6) * spill_slots * PPL * 64;
    const int64_t tile = wave / n_groups;
    const int g = (int)(wave - tile * n_groups);
    if (tile >= n_tiles) return;
    const int64_t base = tile * (64 * PPL);

    Files<PPL> R;
    sfor<PPL>([&](auto P) { fget<P(), PPL>(R) = (v32f)(0.0f); });
    load_features<PPL>(R, feat, F, npad, base, lane);
    float y[PPL];
    bool valid[PPL];
#pragma unroll
    for (int p = 0; p < PPL; ++p) {
        const int64_t pt = base + p * 64 + lane;
        y[p] = target[pt];
        valid[p] = pt < count8;
    }

    const int first = g * group;
    const int last = min(first + group, n_progs);
    for (int sp = first; sp < last; ++sp) {
        const int prog = (int)uni((uint32_t)order[sp]);
        float s = 0.0f;
        if (active == nullptr || active[prog]) {
            const ProgDesc pd = progs[prog];
            run_program<PPL>(R, code + uni(pd.code_off), uni(pd.code_len), params + uni(pd.param_off), lds, lane);
            const uint32_t o = uni(pd.out_reg) & 31u;
            sfor<PPL>([&](auto P) {
                const v32f r = fget<P(), PPL>(R);
                const float diff = r[o] - y[P()];
                const float sq = diff * diff;
                s = s + (valid[P()] ? sq : 0.0f);
            });
            s = wave_sum(s);
        }
        if (lane == 0) partials[tile * n_progs + sp] = s;
    }
}

// One thread per scheduling position: Kahan-compensated sum of the tile partials.
__global__ void k_finalize_mse(const float* __restrict__ partials, int64_t n_tiles, int n_progs,
                               const int32_t* __restrict__ order, const uint8_t* __restrict__ active, float n_samples,
                               float* __restrict__ mse_out) {
    const int sp = blockIdx.x * blockDim.x + threadIdx.x;
    if (sp >= n_progs) return;
    const int prog = order[sp];
    if (active != nullptr && !active[prog]) return;
    float s = 0.0f, c = 0.0f;
    for (int64_t t = 0; t < n_tiles; ++t) {
        const float yv = partials[t * n_progs + sp] - c;
        const float tv = s + yv;
        c = (tv - s) - yv;
        s = tv;
    }
    const float mse = s / n_samples;
    mse_out[prog] = __builtin_isfinite(mse) ? mse : FLT_MAX;
}

size_t eval_mse_lds_bytes(const DevPrograms& P, int ppl) {
    return (size_t)kWaves * P.spill_slots * ppl * 64 * sizeof(float);
}

int64_t eval_tiles(const DevDataset& D, int ppl) {
    const int64_t tp = 64 * (int64_t)ppl;
    return (D.count8 + tp - 1) / tp;
}

hipError_t launch_eval_mse(const DevPrograms& P, const DevDataset& D, const uint8_t* active, const EvalTuning& t,
                           float* partials, float* mse_out, hipStream_t s, const KernelTimer* timer) {
    const size_t lds_bytes = eval_mse_lds_bytes(P, t.ppl);
    if (P.n_progs <= 0) return hipSuccess;
    const int64_t n_tiles = eval_tiles(D, t.ppl);
    const int group = t.group > 0 ? t.group : 8;
    const int n_groups = (P.n_progs + group - 1) / group;
    const int64_t waves = n_tiles * n_groups;
    const int64_t blocks = (waves + kWaves - 1) / kWaves;
    if (blocks > INT32_MAX) return hipErrorInvalidValue;
    dim3 grid((unsigned)blocks), block(64 * kWaves);
#define RSR_LAUNCH(PPLV)                                                                                       \
    hipLaunchKernelGGL(k_eval_mse<PPLV>, grid, block, lds_bytes, s, P.code, P.progs, P.params, P.order, active, \
                       P.n_progs, group, n_groups, D.feat, D.target, D.num_features, D.npad, D.count8, n_tiles, \
                       partials, P.spill_slots)
    if (timer) (void)hipEventRecord(timer->begin, s);
    switch (t.ppl) {
        case 1: RSR_LAUNCH(1); break;
        case 4: RSR_LAUNCH(4); break;
        default: RSR_LAUNCH(2); break;
    }
#undef RSR_LAUNCH
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    if (timer) (void)hipEventRecord(timer->end, s);
    const int tb = 256;
    hipLaunchKernelGGL(k_finalize_mse, dim3((P.n_progs + tb - 1) / tb), dim3(tb), 0, s, partials, n_tiles, P.n_progs,
                       P.order, active, (float)D.num_samples, mse_out);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------------
// Reference-order summation (bit-identical to the CPU evaluator's f32 sums)
// ---------------------------------------------------------------------------------
// The reference accumulates diff^2 in 8 lanes, lane l receiving points l, l+8, l+16, ...
// in order (evaluator.rs:159-170), then adds the 8 lanes horizontally.  Here one wave
// owns one program and walks the tiles in order; inside a 64-point tile the eight
// 8-point batches are folded into the 8 accumulators (held by lanes 0..7) one batch
// after the other with shuffles, which reproduces the serial chain exactly.
__device__ __forceinline__ float fold_batches(float acc, float term, int lane) {
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const float v = __shfl(term, (lane & 7) + 8 * k, 64);
        acc = acc + v;
    }
    return acc;
}

// f32x8::reduce_add order: (lo4 + hi4), then ((0 + 2) + (1 + 3)); result in every lane.
__device__ __forceinline__ float reduce_lanes8(float acc) {
    const float a0 = __shfl(acc, 0, 64), a1 = __shfl(acc, 1, 64), a2 = __shfl(acc, 2, 64), a3 = __shfl(acc, 3, 64);
    const float a4 = __shfl(acc, 4, 64), a5 = __shfl(acc, 5, 64), a6 = __shfl(acc, 6, 64), a7 = __shfl(acc, 7, 64);
    const float s0 = a0 + a4, s1 = a1 + a5, s2 = a2 + a6, s3 = a3 + a7;
    return (s0 + s2) + (s1 + s3);
}

__global__ __launch_bounds__(256) void k_eval_mse_exact(const UInstr* __restrict__ code,
                                                        const ProgDesc* __restrict__ progs,
                                                        const float* __restrict__ params,
                                                        const int32_t* __restrict__ order,
                                                        const uint8_t* __restrict__ active, int n_progs,
                                                        const float* __restrict__ feat,
                                                        const float* __restrict__ target, int F, int64_t npad,
                                                        int64_t count8, int64_t n_tiles, float n_samples,
                                                        float* __restrict__ mse_out, int spill_slots) {
    extern __shared__ float smem[];
    const int lane = threadIdx.x & 63;
    const int sp = (int)(uni(blockIdx.x) * kWaves + uni(threadIdx.x >> 6));
    if (sp >= n_progs) return;
    float* lds = smem + (size_t)uni(threadIdx.x >> 6) * spill_slots * 64;
    const int prog = (int)uni((uint32_t)order[sp]);
    if (active != nullptr && !active[prog]) return;
    const ProgDesc pd = progs[prog];
    const uint32_t o = uni(pd.out_reg) & 31u;
    float acc = 0.0f;
    for (int64_t tile = 0; tile < n_tiles; ++tile) {
        const int64_t base = tile * 64;
        Files<1> R;
        R.a = (v32f)(0.0f);
        load_features<1>(R, feat, F, npad, base, lane);
        run_program<1>(R, code + uni(pd.code_off), uni(pd.code_len), params + uni(pd.param_off), lds, lane);
        const int64_t pt = base + lane;
        const v32f rf = R.a;
        const float diff = rf[o] - target[pt];
        const float sq = pt < count8 ? diff * diff : 0.0f;
        acc = fold_batches(acc, sq, lane);
    }
    const float mse = reduce_lanes8(acc) / n_samples;
    if (lane == 0) mse_out[prog] = __builtin_isfinite(mse) ? mse : FLT_MAX;
}

hipError_t launch_eval_mse_exact(const DevPrograms& P, const DevDataset& D, const uint8_t* active, float* mse_out,
                                 hipStream_t s, const KernelTimer* timer) {
    if (P.n_progs <= 0) return hipSuccess;
    const int64_t n_tiles = eval_tiles(D, 1);
    const size_t lds_bytes = eval_mse_lds_bytes(P, 1);
    if (timer) (void)hipEventRecord(timer->begin, s);
    hipLaunchKernelGGL(k_eval_mse_exact, dim3((unsigned)((P.n_progs + kWaves - 1) / kWaves)), dim3(64 * kWaves),
                       lds_bytes, s, P.code, P.progs, P.params, P.order, active, P.n_progs, D.feat, D.target,
                       D.num_features, D.npad, D.count8, n_tiles, (float)D.num_samples, mse_out, P.spill_slots);
    hipError_t e = hipGetLastError();
    if (timer) (void)hipEventRecord(timer->end, s);
    return e;
}

// ---------------------------------------------------------------------------------
// Per-point values
// ---------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_eval_points(const UInstr* __restrict__ code, const ProgDesc* __restrict__ progs,
                                                     const float* __restrict__ params, int first, int count,
                                                     const float* __restrict__ feat, int F, int64_t npad,
                                                     int64_t count8, int64_t n_tiles, float* __restrict__ out,
                                                     int spill_slots) {
    extern __shared__ float smem[];
    const int lane = threadIdx.x & 63;
    float* lds = smem + (size_t)uni(threadIdx.x >> 6) * spill_slots * 64;
    const int64_t wave = (int64_t)uni(blockIdx.x) * kWaves + uni(threadIdx.x >> 6);
    const int64_t tile = wave / count;
    const int i = (int)(wave - tile * count);
    if (tile >= n_tiles) return;
    const int64_t base = tile * 64;
    Files<1> R;
    R.a = (v32f)(0.0f);
    load_features<1>(R, feat, F, npad, base, lane);
    const ProgDesc pd = progs[first + i];
    run_program<1>(R, code + uni(pd.code_off), uni(pd.code_len), params + uni(pd.param_off), lds, lane);
    const int64_t pt = base + lane;
    const v32f rf = R.a;
    if (pt < count8) out[(int64_t)i * count8 + pt] = rf[uni(pd.out_reg) & 31u];
}

hipError_t launch_eval_points(const DevPrograms& P, const DevDataset& D, int first, int count, float* out,
                              hipStream_t s) {
    if (count <= 0) return hipSuccess;
    const int64_t n_tiles = eval_tiles(D, 1);
    const int64_t blocks = (n_tiles * count + kWaves - 1) / kWaves;
    const size_t lds_bytes = (size_t)kWaves * P.spill_slots * 64 * sizeof(float);
    hipLaunchKernelGGL(k_eval_points, dim3((unsigned)blocks), dim3(64 * kWaves), lds_bytes, s, P.code, P.progs,
                       P.params, first, count, D.feat, D.num_features, D.npad, D.count8, n_tiles, out, P.spill_slots);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------------
// Dual (forward-mode) evaluation: MSE and gradient
// ---------------------------------------------------------------------------------
static constexpr int kTan = 4;  // tangent channels per pass

size_t eval_grad_lds_bytes(const DevPrograms& P) {
    return (size_t)kWaves * P.spill_slots * (1 + kTan) * 64 * sizeof(float);
}

__global__ __launch_bounds__(256) void k_eval_grad(const UInstr* __restrict__ code, const ProgDesc* __restrict__ progs,
                                                   const float* __restrict__ params, const int32_t* __restrict__ order,
                                                   const uint8_t* __restrict__ nparams, const uint8_t* __restrict__ active,
                                                   int n_progs, int n_pass, const float* __restrict__ feat,
                                                   const float* __restrict__ target, int F, int64_t npad,
                                                   int64_t count8, int64_t n_tiles, float* __restrict__ partials,
                                                   int spill_slots) {
    extern __shared__ float smem[];
    const int lane = threadIdx.x & 63;
    const int64_t wave = (int64_t)uni(blockIdx.x) * kWaves + uni(threadIdx.x >> 6);
    float* lds = smem + (size_t)uni(threadIdx.x >> 6) * spill_slots * (1 + kTan) * 64;
    // work item = (tile, program, pass); consecutive waves share the tile
    const int64_t per_tile = (int64_t)n_progs * n_pass;
    const int64_t tile = wave / per_tile;
    const int64_t rem = wave - tile * per_tile;
    const int sp = (int)(rem / n_pass);
    const int pass = (int)(rem - (int64_t)sp * n_pass);
    if (tile >= n_tiles) return;
    const int prog = (int)uni((uint32_t)order[sp]);
    const int np = (int)uni(nparams[prog]);
    if ((active != nullptr && !active[prog]) || np == 0 || pass * kTan >= np) return;
    const int64_t base = tile * 64;
    const int64_t pt = base + lane;

    Files<1> R1;
    R1.a = (v32f)(0.0f);
    load_features<1>(R1, feat, F, npad, base, lane);
    DualFiles<kTan> G;
    G.v = R1.a;
    G.a = G.b = G.c = G.d = (v32f)(0.0f);
    const ProgDesc pd = progs[prog];
    const uint32_t k0 = (uint32_t)(pass * kTan);
    run_dual<kTan>(G, code + uni(pd.code_off), uni(pd.code_len), params + uni(pd.param_off), k0, lds, lane);
    const uint32_t o = uni(pd.out_reg) & 31u;
    const bool valid = pt < count8;
    const v32f vfile = G.v;
    const float diff = vfile[o] - target[pt];
    float* dst = partials + (tile * n_progs + sp) * 33;
    if (pass == 0) {
        const float sq = diff * diff;
        const float s = wave_sum(valid ? sq : 0.0f);
        if (lane == 0) dst[0] = s;
    }
    const float two_diff = 2.0f * diff;
    const v32f gf[kTan] = {G.a, G.b, G.c, G.d};
#pragma unroll
    for (int j = 0; j < kTan; ++j) {
        const int k = (int)k0 + j;
        if (k < np) {
            const float gterm = two_diff * gf[j][o];
            const float s = wave_sum(valid ? gterm : 0.0f);
            if (lane == 0) dst[1 + k] = s;
        }
    }
}

// One thread per (scheduling position, slot); slot 0 = squared error, 1 + k = gradient k.
__global__ void k_finalize_grad(const float* __restrict__ partials, int64_t n_tiles, int n_progs,
                                const int32_t* __restrict__ order, const uint8_t* __restrict__ nparams,
                                const uint8_t* __restrict__ active, float n_samples, float* __restrict__ mse_out,
                                float* __restrict__ grad_out) {
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int sp = (int)(idx / 33);
    const int slot = (int)(idx - (int64_t)sp * 33);
    if (sp >= n_progs) return;
    const int prog = order[sp];
    if (active != nullptr && !active[prog]) return;
    const int np = nparams[prog];
    if (np == 0) return;
    float s = 0.0f, c = 0.0f;
    if (slot == 0 || slot - 1 < np) {
        for (int64_t t = 0; t < n_tiles; ++t) {
            const float yv = partials[(t * n_progs + sp) * 33 + slot] - c;
            const float tv = s + yv;
            c = (tv - s) - yv;
            s = tv;
        }
    }
    // the MSE decides whether the gradient is reported (evaluator.rs:214-216); every
    // thread of the program recomputes it from the same partials in the same order
    float m = 0.0f, mc = 0.0f;
    for (int64_t t = 0; t < n_tiles; ++t) {
        const float yv = partials[(t * n_progs + sp) * 33] - mc;
        const float tv = m + yv;
        mc = (tv - m) - yv;
        m = tv;
    }
    const float mse = m / n_samples;
    const bool fin = __builtin_isfinite(mse);
    if (slot == 0) {
        mse_out[prog] = fin ? mse : FLT_MAX;
    } else {
        const int k = slot - 1;
        if (k < 32) grad_out[(int64_t)prog * 32 + k] = (fin && k < np) ? s / n_samples : 0.0f;
    }
}

hipError_t launch_eval_grad(const DevPrograms& P, const DevDataset& D, const uint8_t* nparams_eff,
                            const uint8_t* active, int max_params, float* partials, float* mse_out,
                            float* grad_out, hipStream_t s, const KernelTimer* timer) {
    if (P.n_progs <= 0 || max_params <= 0) return hipSuccess;
    const int64_t n_tiles = eval_tiles(D, 1);
    const int n_pass = (max_params + kTan - 1) / kTan;
    const int64_t waves = n_tiles * P.n_progs * n_pass;
    const int64_t blocks = (waves + kWaves - 1) / kWaves;
    if (blocks > INT32_MAX) return hipErrorInvalidValue;
    const size_t lds_bytes = eval_grad_lds_bytes(P);
    if (timer) (void)hipEventRecord(timer->begin, s);
    hipLaunchKernelGGL(k_eval_grad, dim3((unsigned)blocks), dim3(64 * kWaves), lds_bytes, s, P.code, P.progs, P.params,
                       P.order, nparams_eff, active, P.n_progs, n_pass, D.feat, D.target, D.num_features, D.npad,
                       D.count8, n_tiles, partials, P.spill_slots);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    if (timer) (void)hipEventRecord(timer->end, s);
    const int64_t threads = (int64_t)P.n_progs * 33;
    const int tb = 256;
    hipLaunchKernelGGL(k_finalize_grad, dim3((unsigned)((threads + tb - 1) / tb)), dim3(tb), 0, s, partials, n_tiles,
                       P.n_progs, P.order, nparams_eff, active, (float)D.num_samples, mse_out, grad_out);
    return hipGetLastError();
}

// Reference-order gradient sums: one wave per (program, pass) walks the tiles in order.
__global__ __launch_bounds__(256) void k_eval_grad_exact(const UInstr* __restrict__ code,
                                                         const ProgDesc* __restrict__ progs,
                                                         const float* __restrict__ params,
                                                         const int32_t* __restrict__ order,
                                                         const uint8_t* __restrict__ nparams,
                                                         const uint8_t* __restrict__ active, int n_progs, int n_pass,
                                                         const float* __restrict__ feat,
                                                         const float* __restrict__ target, int F, int64_t npad,
                                                         int64_t count8, int64_t n_tiles, float n_samples,
                                                         float* __restrict__ mse_out, float* __restrict__ grad_out,
                                                         int spill_slots) {
    extern __shared__ float smem[];
    const int lane = threadIdx.x & 63;
    const int64_t wave = (int64_t)uni(blockIdx.x) * kWaves + uni(threadIdx.x >> 6);
    const int sp = (int)(wave / n_pass);
    const int pass = (int)(wave - (int64_t)sp * n_pass);
    if (sp >= n_progs) return;
    float* lds = smem + (size_t)uni(threadIdx.x >> 6) * spill_slots * (1 + kTan) * 64;
    const int prog = (int)uni((uint32_t)order[sp]);
    const int np = (int)uni(nparams[prog]);
    if ((active != nullptr && !active[prog]) || np == 0 || pass * kTan >= np) return;
    const ProgDesc pd = progs[prog];
    const uint32_t o = uni(pd.out_reg) & 31u;
    const uint32_t k0 = (uint32_t)(pass * kTan);
    float acc = 0.0f, gacc[kTan] = {0.0f, 0.0f, 0.0f, 0.0f};
    for (int64_t tile = 0; tile < n_tiles; ++tile) {
        const int64_t base = tile * 64;
        Files<1> R1;
        R1.a = (v32f)(0.0f);
        load_features<1>(R1, feat, F, npad, base, lane);
        DualFiles<kTan> G;
        G.v = R1.a;
        G.a = G.b = G.c = G.d = (v32f)(0.0f);
        run_dual<kTan>(G, code + uni(pd.code_off), uni(pd.code_len), params + uni(pd.param_off), k0, lds, lane);
        const int64_t pt = base + lane;
        const bool valid = pt < count8;
        const v32f vf = G.v;
        const float diff = vf[o] - target[pt];
        acc = fold_batches(acc, valid ? diff * diff : 0.0f, lane);
        const float two_diff = 2.0f * diff;
        const v32f gf[kTan] = {G.a, G.b, G.c, G.d};
#pragma unroll
        for (int j = 0; j < kTan; ++j) gacc[j] = fold_batches(gacc[j], valid ? two_diff * gf[j][o] : 0.0f, lane);
    }
    const float mse = reduce_lanes8(acc) / n_samples;
    const bool fin = __builtin_isfinite(mse);
    if (pass == 0 && lane == 0) mse_out[prog] = fin ? mse : FLT_MAX;
#pragma unroll
    for (int j = 0; j < kTan; ++j) {
        const int k = (int)k0 + j;
        const float g = reduce_lanes8(gacc[j]) / n_samples;
        if (lane == 0 && k < 32) grad_out[(int64_t)prog * 32 + k] = (fin && k < np) ? g : 0.0f;
    }
}

hipError_t launch_eval_grad_exact(const DevPrograms& P, const DevDataset& D, const uint8_t* nparams_eff,
                                  const uint8_t* active, int max_params, float* mse_out, float* grad_out,
                                  hipStream_t s, const KernelTimer* timer) {
    if (P.n_progs <= 0 || max_params <= 0) return hipSuccess;
    const int64_t n_tiles = eval_tiles(D, 1);
    const int n_pass = (max_params + kTan - 1) / kTan;
    const int64_t waves = (int64_t)P.n_progs * n_pass;
    const size_t lds_bytes = eval_grad_lds_bytes(P);
    if (timer) (void)hipEventRecord(timer->begin, s);
    hipLaunchKernelGGL(k_eval_grad_exact, dim3((unsigned)((waves + kWaves - 1) / kWaves)), dim3(64 * kWaves),
                       lds_bytes, s, P.code, P.progs, P.params, P.order, nparams_eff, active, P.n_progs, n_pass,
                       D.feat, D.target, D.num_features, D.npad, D.count8, n_tiles, (float)D.num_samples, mse_out,
                       grad_out, P.spill_slots);
    hipError_t e = hipGetLastError();
    if (timer) (void)hipEventRecord(timer->end, s);
    return e;
}

__global__ __launch_bounds__(64) void k_eval_points_dual(const UInstr* __restrict__ code, const ProgDesc* __restrict__ progs,
                                                         const float* __restrict__ params, int prog, int active_param,
                                                         const float* __restrict__ feat, int F, int64_t npad,
                                                         int64_t count8, float* __restrict__ val_out,
                                                         float* __restrict__ grad_out, int spill_slots) {
    extern __shared__ float smem[];
    const int lane = threadIdx.x & 63;
    float* lds = smem;
    (void)spill_slots;
    const int64_t base = (int64_t)blockIdx.x * 64;
    Files<1> R1;
    R1.a = (v32f)(0.0f);
    load_features<1>(R1, feat, F, npad, base, lane);
    DualFiles<1> G;
    G.v = R1.a;
    G.a = (v32f)(0.0f);
    const ProgDesc pd = progs[prog];
    run_dual<1>(G, code + uni(pd.code_off), uni(pd.code_len), params + uni(pd.param_off), (uint32_t)active_param,
                lds, lane);
    const uint32_t o = uni(pd.out_reg) & 31u;
    const int64_t pt = base + lane;
    const v32f vf = G.v, gfl = G.a;
    if (pt < count8) {
        val_out[pt] = vf[o];
        grad_out[pt] = gfl[o];
    }
}

hipError_t launch_eval_points_dual(const DevPrograms& P, const DevDataset& D, int prog, int active_param,
                                   float* val_out, float* grad_out, hipStream_t s) {
    const int64_t n_tiles = eval_tiles(D, 1);
    const size_t lds_bytes = (size_t)P.spill_slots * 2 * 64 * sizeof(float);
    hipLaunchKernelGGL(k_eval_points_dual, dim3((unsigned)n_tiles), dim3(64), lds_bytes, s, P.code, P.progs, P.params,
                       prog, active_param, D.feat, D.num_features, D.npad, D.count8, val_out, grad_out, P.spill_slots);
    return hipGetLastError();
}

}  // namespace rsr
