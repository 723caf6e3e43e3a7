// kernels.hip -- fitness-evaluation kernels for gfx950.
//
// Every evaluation launch takes a segment table (kernel arguments, up to kMaxSegs
// segments).  A segment is one request's evaluation step: a program batch, a dataset,
// an active mask and output buffers.  One launch therefore evaluates the steps of many
// requests with different datasets, feature counts and population sizes together --
// the continuous-batching path of the server -- and a plain call is a one-segment
// table.
//
// Work decomposition (Kahan summation): a work item is (point tile, program group) of
// one segment.  The four waves of a 256-thread workgroup take four consecutive groups
// of the same tile, so the tile's feature columns are fetched from HBM once and served
// to the other waves from L2/L1.  Each wave loads its 64 * PPL points' features straight
// into the low registers of its register file, where they stay resident while the wave
// interprets every program of its group.  Per (program, tile) the wave writes one
// partial sum; the finalize kernel adds the partials of each program in tile order with
// Kahan-compensated FP32, so results are bitwise reproducible run to run.
//
// Reference-order summation: the same tile-parallel kernels write every per-point term
// (squared error, or gradient term) and k_fold_terms adds them in the reference's order.
#include <float.h>
#include <algorithm>
#include <map>
#include <mutex>
#include <tuple>
#include "interp.h"
#include "kernels.h"

namespace rsr {

static constexpr int kWaves = 4;  // waves per workgroup
static constexpr int kTan = 2;    // tangent channels per dual pass
// Programs per gradient work item (features loaded once for the group): up to 16, fewer
// when a segment would then give the chip fewer than about four waves per SIMD.  On the
// Mat3 workload (313 tiles x 16,384 programs) 16 took the launch from 25.3 to 24.1 ms
// against 4; small requests keep 4 so that they still spread over the CUs.
static int grad_group(int64_t n_tiles, int n_progs) {
    for (int g = 16; g > 4; g /= 2)
        if (n_tiles * ((n_progs + g - 1) / g) >= 4 * 1024) return g;
    return 4;
}
static constexpr uint32_t kXcds = 8;  // accelerator complex dies, each with its own L2

// Workgroups are dealt to the XCDs round-robin (workgroup b runs on XCD b % 8).  This
// logical block index gives each XCD one contiguous range of work items; the evaluation
// kernels order work items tile-major, so each XCD's L2 fetches about an eighth of the
// dataset instead of all of it.  A bijection on [0, gridDim.x).
__device__ __forceinline__ int64_t xcd_block() {
    const uint32_t b = uni(blockIdx.x), G = uni(gridDim.x);
    const uint32_t x = b % kXcds, k = b / kXcds, q = G / kXcds, r = G % kXcds;
    return (int64_t)x * q + (x < r ? x : r) + k;
}

// Work item w of a segment -> (tile, program group).  With xcd_block() each XCD runs a
// contiguous range of work items.  Ordered tile-major, that range is a run of tiles with
// every program group on each: each XCD's L2 then holds about an eighth of the dataset,
// but the code of every program streams through it once per tile.  For segments with at
// least 256 tiles the items are instead ordered in blocks of KB tiles (block, then group,
// then tile), so a group's code serves KB consecutive tiles from the cache (KB = 16 for the
// evaluation, 8 for the gradients: the best of 1, 4, 8, 16, 32, 64 measured on the Lorentz
// and Mat3 workloads; L2 misses of the Lorentz evaluation drop 12x against tile-major).
// Smaller segments keep the tile-major order: there an XCD's range would cut through a
// block's groups, which are sorted by cost, and unbalance the XCDs.  False for w past the
// segment's work.
template <int KB>
__device__ __forceinline__ bool split_work(int64_t w, int64_t n_tiles, int n_groups, int64_t& lt, int& g) {
    if (w >= n_tiles * n_groups) return false;
    const int K = n_tiles >= 256 ? KB : 1;
    const int64_t per = (int64_t)n_groups * K;
    const int64_t blk = w / per, r = w - blk * per;
    const int64_t t0 = blk * K;
    const int kk = (int)((n_tiles - t0) < K ? (n_tiles - t0) : K);  // tiles in this block
    g = (int)(r / kk);
    lt = t0 + (r - (int64_t)g * kk);
    return true;
}

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
    // vmcnt(0): every load has landed before the interpreter runs.  Its indexed moves read
    // the register files through a run-time index that the compiler's wait insertion
    // cannot resolve to one register, so loads still in flight would give every staging
    // read inside the interpreter loop a wait of its own, executed on every micro-op.
    __builtin_amdgcn_s_waitcnt(0x0F70);
}

// Segment owning a work index (first_wave / first_thread are increasing prefix sums).
// A fixed chain of uniform selects over the table, no data-dependent loop.
__device__ __forceinline__ int find_seg_wave(const SegTable& T, int64_t w) {
    int si = 0;
#pragma unroll
    for (int s = 1; s < kMaxSegs; ++s) si = (s < T.n && w >= T.s[s].first_wave) ? s : si;
    return si;
}
__device__ __forceinline__ int find_seg_first_thread(const SegTable& T, int64_t t) {
    int si = 0;
#pragma unroll
    for (int s = 1; s < kMaxSegs; ++s) si = (s < T.n && t >= T.s[s].first_thread) ? s : si;
    return si;
}
// Finalize kernels: every segment's thread range starts on a wave boundary (first_thread
// is a multiple of 64), so the segment follows from the wave's first thread and is
// wave-uniform.
__device__ __forceinline__ int find_seg_thread(const SegTable& T) {
    const int64_t t0 = (int64_t)uni(blockIdx.x) * blockDim.x + ((int64_t)uni(threadIdx.x >> 6) << 6);
    return find_seg_first_thread(T, t0);
}

// One step of Kahan-compensated summation.
__device__ __forceinline__ void kahan_add(float& s, float& c, float v) {
    const float yv = v - c;
    const float tv = s + yv;
    c = (tv - s) - yv;
    s = tv;
}

// f32x8::reduce_add order: (lo4 + hi4), then ((0 + 2) + (1 + 3)); result in every lane.
__device__ __forceinline__ float reduce_lanes8(float acc) {
    const float a0 = lane_f(acc, 0), a1 = lane_f(acc, 1), a2 = lane_f(acc, 2), a3 = lane_f(acc, 3);
    const float a4 = lane_f(acc, 4), a5 = lane_f(acc, 5), a6 = lane_f(acc, 6), a7 = lane_f(acc, 7);
    const float s0 = a0 + a4, s1 = a1 + a5, s2 = a2 + a6, s3 = a3 + a7;
    return (s0 + s2) + (s1 + s3);
}

// ---------------------------------------------------------------------------------
// MSE, Kahan summation
// ---------------------------------------------------------------------------------
// One work item (wave index `wave` of the launch): a tile of points x a group of programs.
// TERMS (the launch's summation mode, the same for every segment of a launch): per-point
// terms for the fold kernel (reference order) instead of per-tile partials (Kahan); each
// mode's kernel holds only its own epilogue.
template <int PPL, bool TERMS>
__device__ __forceinline__ void mse_item(const SegTable& T, int64_t wave, float* lds, int lane) {
    const int si = find_seg_wave(T, wave);
    const SegDesc& S = T.s[si];
    const int64_t w = wave - S.first_wave;
    int64_t lt;  // tile within the launch's range
    int g;
    if (!split_work<16>(w, S.n_tiles, S.n_groups, lt, g)) return;
    if (S.live != nullptr && *S.live == 0) return;
    const int first = g * S.group;
    const int last = min(first + S.group, S.n_small);
    if (S.active != nullptr) {  // nothing to evaluate in this group: skip the feature loads too
        bool any = false;
        for (int sp = first; sp < last; ++sp) any = any || kget(S.active, uni((uint32_t)kget(S.order, sp))) != 0;
        if (!any) return;  // the finalize and fold kernels never read inactive programs' partials
    }
    const int64_t base = (S.tile0 + lt) * (64 * PPL);

    Files<PPL> R;
    sfor<PPL>([&](auto P) { fget<P(), PPL>(R) = (v32f)(0.0f); });
    load_features<PPL>(R, S.feat, S.F, S.npad, base, lane);
    bool valid[PPL];
#pragma unroll
    for (int p = 0; p < PPL; ++p) valid[p] = base + p * 64 + lane < S.count8;
    for (int sp = first; sp < last; ++sp) {
        const int prog = (int)uni((uint32_t)kget(S.order, sp));
        float s = 0.0f;
        if (S.active == nullptr || kget(S.active, prog)) {
            const ProgDesc pd = kget(S.progs, prog);
            run_program<PPL>(R, S.code + uni(pd.code_off), uni(pd.code_len), S.params + uni(pd.param_off), lds,
                             nullptr, lane);
            const uint32_t o = uni(pd.out_reg) & 31u;
            // targets are re-read per program (cache hits) rather than held in registers
            // through the interpreter loop
            float y[PPL];
            sfor<PPL>([&](auto P) { y[P()] = S.target[base + P() * 64 + lane]; });
            if constexpr (TERMS) {
                // reference order: keep every per-point term; the fold kernel adds them.
                // The terms (3.3 GB per launch on the Lorentz workload) are only read back by
                // the fold kernel: non-temporal stores, measured 2% faster per step.
                float* row = S.partials + (int64_t)sp * S.n_tiles * (64 * PPL) + lt * (64 * PPL);
                sfor<PPL>([&](auto P) {
                    const v32f r = fget<P(), PPL>(R);
                    const float diff = r[o] - y[P()];
                    __builtin_nontemporal_store(diff * diff, &row[P() * 64 + lane]);
                });
                continue;
            }
            sfor<PPL>([&](auto P) {
                const v32f r = fget<P(), PPL>(R);
                const float diff = r[o] - y[P()];
                const float sq = diff * diff;
                s = s + (valid[P()] ? sq : 0.0f);
            });
            s = wave_sum(s);
        }
        if (!TERMS && lane == 0) S.partials[lt * S.n_progs + sp] = s;
    }
}

// At four points per lane the register files alone are 128 VGPRs; the occupancy hint
// makes the allocator fit the kernel in 168 (three waves per SIMD).  At three points per
// lane (96 VGPRs of files) it fits in 128: four waves per SIMD.
template <int PPL>
constexpr int kEvalWaves = PPL == 3 ? 4 : 3;
template <int PPL, bool TERMS>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(kEvalWaves<PPL>))) void k_eval_mse(const SegTable T,
                                                                                           int spill_slots) {
    extern __shared__ float smem[];
    float* lds = smem + (size_t)uni(threadIdx.x >> 6) * spill_slots * PPL * 64;
    mse_item<PPL, TERMS>(T, xcd_block() * kWaves + uni(threadIdx.x >> 6), lds, threadIdx.x & 63);
}

// The persistent form: a resident grid drains the launch's work items through a device work
// queue, one counter per XCD.  XCD x owns the same contiguous range of items that
// xcd_block() gives it in the one-shot launch (a run of tiles with their program groups, so
// its L2 keeps that part of the dataset and the groups' code); a wave takes the next item of
// its own XCD's range (workgroup b runs on XCD b % 8) and, once that range is drained, of the
// other XCDs' ranges in turn.  Every take is one atomic increment and the ranges are finite,
// so every wave exits.  The queue is zero at launch and the launch leaves it zero: the last
// wave to finish (queue[kXcds] counts finished waves) clears it.
// (The work loop's extra live state does not fit 128 VGPRs at three points per lane: three
// waves per SIMD for every form.)
template <int PPL, bool TERMS>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void k_eval_mse_q(
    const SegTable T, int spill_slots, int64_t n_waves, int* __restrict__ queue) {
    extern __shared__ float smem[];
    const int64_t G = (n_waves + kWaves - 1) / kWaves;  // blocks of the one-shot launch
    const int64_t q = G / kXcds, r = G % kXcds;
    uint32_t k = 0, x = uni(blockIdx.x) % kXcds;
    for (;;) {
        // the next item: a take from queue x, moving on to the next XCD's range when drained
        const int64_t w0 = ((int64_t)x * q + min<int64_t>(x, r)) * kWaves;
        const int64_t nw = min<int64_t>((q + (x < r ? 1 : 0)) * kWaves, n_waves - w0);
        int t = 0;
        if ((threadIdx.x & 63) == 0) t = atomicAdd(&queue[x], 1);
        t = __builtin_amdgcn_readfirstlane(t);
        if ((int64_t)t >= nw) {
            if (++k == kXcds) break;
            x = (x + 1) % kXcds;
            continue;
        }
        float* lds = smem + (size_t)uni(threadIdx.x >> 6) * spill_slots * PPL * 64;
        mse_item<PPL, TERMS>(T, w0 + t, lds, threadIdx.x & 63);
    }
    if ((threadIdx.x & 63) == 0) {
        const int launched = (int)(uni(gridDim.x) * kWaves);
        if (atomicAdd(&queue[kXcds], 1) == launched - 1) {  // every other wave is past its takes
#pragma unroll
            for (uint32_t y = 0; y <= kXcds; ++y) (void)atomicExch(&queue[y], 0);
        }
    }
}

// One thread per (segment, scheduling position): Kahan sum of the tile partials.
__global__ void k_finalize_mse(const SegTable T) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int si = find_seg_thread(T);
    const SegDesc& S = T.s[si];
    const int64_t sp = t - S.first_thread;
    if (sp >= S.n_progs) return;
    if (S.live != nullptr && *S.live == 0) return;
    const int prog = S.order[sp];
    if (S.active != nullptr && !S.active[prog]) return;
    float s = 0.0f, c = 0.0f;
    if (!S.first) {
        s = S.kstate[2 * sp];
        c = S.kstate[2 * sp + 1];
    }
    const float* src = S.partials + sp;
    const int64_t stride = S.n_progs;
    int64_t lt = 0;
    for (; lt + 16 <= S.n_tiles; lt += 16) {  // loads issued ahead of the serial compensated chain
        float v[16];
#pragma unroll
        for (int u = 0; u < 16; ++u) v[u] = src[(lt + u) * stride];
#pragma unroll
        for (int u = 0; u < 16; ++u) kahan_add(s, c, v[u]);
    }
    for (; lt < S.n_tiles; ++lt) kahan_add(s, c, src[lt * stride]);
    if (!S.last) {
        S.kstate[2 * sp] = s;
        S.kstate[2 * sp + 1] = c;
        return;
    }
    const float mse = s / S.n_samples;
    S.mse_out[prog] = __builtin_isfinite(mse) ? mse : FLT_MAX;
}

// ---------------------------------------------------------------------------------
// MSE, reference-order summation from per-point terms
// ---------------------------------------------------------------------------------
// One wave per position q = sp * fold_rows + slot (one row of terms; MSE: slot 0 only).
// Lanes 0..7 are the reference's eight lane accumulators: lane l adds the terms of points
// l, l+8, l+16, ... of the launch's range in order (evaluator.rs:159-170), starting from the
// state carried from earlier ranges.  The serial chains cannot be split, so the wave keeps
// the memory busy instead: the whole wave loads the row in chunks of kFoldChunk terms with
// coalesced 16-byte loads, transposes each chunk through LDS so that chain l finds its terms
// contiguous, and issues the next chunk's loads before folding the current one.  The last
// range reduces the eight accumulators in f32x8::reduce_add order.  (first_thread counts fold
// waves here, one per position.)
static constexpr int kFoldChunk = 2048;                 // terms per chunk: 8 float4 per lane
static constexpr int kFoldRow = kFoldChunk / 8 + 4;     // LDS row of one chain (+4: banks)
__global__ __launch_bounds__(64) void k_fold_terms(const SegTable T) {
    __shared__ __align__(16) float bt[8 * kFoldRow];
    const int lane = threadIdx.x;
    const int64_t wave = uni(blockIdx.x);
    const int si = find_seg_first_thread(T, wave);
    const SegDesc& S = T.s[si];
    const int64_t n_pos = (int64_t)S.n_progs * S.fold_rows;
    const int64_t q = wave - S.first_thread;
    if (q >= n_pos) return;
    if (S.live != nullptr && *S.live == 0) return;
    const int64_t sp = q / S.fold_rows;
    const int slot = (int)(q - sp * S.fold_rows);
    if (S.active != nullptr && !S.active[S.order[sp]]) return;
    if (S.fold_rows > 1 && slot > 0 && slot - 1 >= (int)S.nparams[S.order[sp]]) return;
    const int64_t row_len = S.n_tiles * S.tile_pts;  // a multiple of 64: whole float4s
    const int64_t p0 = S.tile0 * S.tile_pts;
    const int64_t len = min(p0 + row_len, S.count8) - p0;  // whole 8-point batches
    typedef float f4 __attribute__((ext_vector_type(4)));
    const f4* row = reinterpret_cast<const f4*>(S.partials + q * row_len);
    const int l = lane & 7;
    float acc = (!S.first && lane < 8) ? S.kstate[q * 8 + l] : 0.0f;
    const int64_t n_chunks = (len + kFoldChunk - 1) / kFoldChunk;
    f4 r[8];
    auto load = [&](int64_t c) {
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int64_t i4 = c * (kFoldChunk / 4) + u * 64 + lane;
            r[u] = i4 * 4 < len ? __builtin_nontemporal_load(&row[i4]) : (f4)(0.0f);
        }
    };
    if (n_chunks > 0) load(0);
    // this lane's float4 u holds chunk points 256u + 4 lane + e (e = 0..3): chain
    // 4 (lane & 1) + e, position 32u + (lane >> 1) along it
    float* const wr = bt + 4 * (lane & 1) * kFoldRow + (lane >> 1);
    const float* const rd = bt + l * kFoldRow;
    for (int64_t c = 0; c < n_chunks; ++c) {
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            wr[0 * kFoldRow + 32 * u] = r[u].x;
            wr[1 * kFoldRow + 32 * u] = r[u].y;
            wr[2 * kFoldRow + 32 * u] = r[u].z;
            wr[3 * kFoldRow + 32 * u] = r[u].w;
        }
        __syncthreads();
        if (c + 1 < n_chunks) load(c + 1);  // in flight while this chunk is folded
        const int kmax = (int)(min<int64_t>(kFoldChunk, len - c * kFoldChunk) >> 3);  // terms per chain
        if (lane < 8) {
            int k = 0;
            for (; k + 16 <= kmax; k += 16) {
                f4 v[4];
#pragma unroll
                for (int h = 0; h < 4; ++h) v[h] = *reinterpret_cast<const f4*>(rd + k + 4 * h);
#pragma unroll
                for (int h = 0; h < 4; ++h) acc = (((acc + v[h].x) + v[h].y) + v[h].z) + v[h].w;
            }
            for (; k < kmax; ++k) acc = acc + rd[k];
        }
        __syncthreads();  // the chunk's reads are done before the next chunk is stored
    }
    if (!S.last) {
        if (lane < 8) S.kstate[q * 8 + l] = acc;
        return;
    }
    // f32x8::reduce_add: (lo4 + hi4), then ((0 + 2) + (1 + 3))
    const float sum = reduce_lanes8(acc);
    if (lane != 0) return;
    if (S.fold_rows > 1) {  // gradient rows: the finalize kernel divides and checks
        S.gsum[q] = sum;
        return;
    }
    const float mse = sum / S.n_samples;
    S.mse_out[S.order[sp]] = __builtin_isfinite(mse) ? mse : FLT_MAX;
}

// Gradient terms: one thread per (position, slot); slot 0 = MSE, 1 + k = gradient k, each the
// folded row sum divided by N; the gradient is reported only for a finite MSE
// (evaluator.rs:214-216).
__global__ void k_finalize_grad_terms(const SegTable T) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t t0 = (int64_t)uni(blockIdx.x) * blockDim.x + ((int64_t)uni(threadIdx.x >> 6) << 6);
    int si = 0;
#pragma unroll
    for (int s = 1; s < kMaxSegs; ++s) si = (s < T.n && t0 >= T.s[s].first_fin) ? s : si;
    const SegDesc& S = T.s[si];
    const int64_t q = t - S.first_fin;
    if (q >= (int64_t)S.n_progs * S.slots || !S.last) return;
    if (S.live != nullptr && *S.live == 0) return;
    const int sp = (int)(q / S.slots);
    const int slot = (int)(q - (int64_t)sp * S.slots);
    const int prog = S.order[sp];
    if (S.active != nullptr && !S.active[prog]) return;
    const int np = S.nparams[prog];
    if (np == 0) return;
    const float mse = S.gsum[(int64_t)sp * S.slots] / S.n_samples;
    const bool fin = __builtin_isfinite(mse);
    if (slot == 0) {
        S.mse_out[prog] = fin ? mse : FLT_MAX;
    } else {
        const int k = slot - 1;
        S.grad_out[(int64_t)prog * 32 + k] = (fin && k < np) ? S.gsum[q] / S.n_samples : 0.0f;
    }
}

// ---------------------------------------------------------------------------------
// Gradients (dual), Kahan summation: work item = (tile, group of grad_group() programs),
// every pass of kTan tangent channels of each program inside it
// ---------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_eval_grad(const SegTable T, int spill_slots) {
    extern __shared__ float smem[];
    const int lane = threadIdx.x & 63;
    const int64_t wave = xcd_block() * kWaves + uni(threadIdx.x >> 6);
    const int si = find_seg_wave(T, wave);
    const SegDesc& S = T.s[si];
    const int64_t w = wave - S.first_wave;
    int64_t lt;  // tile within the launch's range
    int g;
    if (!split_work<8>(w, S.n_tiles, S.n_groups, lt, g)) return;
    if (S.live != nullptr && *S.live == 0) return;
    const int first = g * S.group;
    const int last = min(first + S.group, S.n_small);
    float* lds = smem + (size_t)uni(threadIdx.x >> 6) * spill_slots * (1 + kTan) * 64;
    const int64_t base = (S.tile0 + lt) * 64;
    const int64_t pt = base + lane;
    const bool valid = pt < S.count8;

    // The features are loaded once for the group.  Every program and pass reuses the
    // files: the code never writes a feature register, so the feature tangents stay zero,
    // and every temporary is written before it is read.
    Files<1> R1;
    R1.a = (v32f)(0.0f);
    load_features<1>(R1, S.feat, S.F, S.npad, base, lane);
    DualFiles<kTan> G;
    G.v = R1.a;
    zero_tangents<kTan>(G);
    for (int sp = first; sp < last; ++sp) {
        const int prog = (int)uni((uint32_t)kget(S.order, sp));
        const int np = (int)uni(kget(S.nparams, prog));
        if ((S.active != nullptr && !kget(S.active, prog)) || np == 0) continue;
        const ProgDesc pd = kget(S.progs, prog);
        const uint32_t o = uni(pd.out_reg) & 31u;
        float* dst = S.partials + (lt * S.n_progs + sp) * S.slots;
        for (int pass = 0; pass * kTan < np; ++pass) {
            const uint32_t k0 = (uint32_t)(pass * kTan);
            run_dual<kTan>(G, S.code + uni(pd.code_off), uni(pd.code_len), S.params + uni(pd.param_off), k0, lds,
                           nullptr, lane);
            const v32f vfile = G.v;
            const float diff = vfile[o] - S.target[pt];
            const float two_diff = 2.0f * diff;
            v32f gf[kTan];
            tangent_files<kTan>(G, gf);
            if (S.terms) {
                // reference order: per-point terms, row (sp, slot), folded in order later
                float* rows = S.partials + (int64_t)sp * S.slots * S.n_tiles * 64 + lt * 64 + lane;
                const int64_t rstride = S.n_tiles * 64;
                if (pass == 0) __builtin_nontemporal_store(diff * diff, rows);
#pragma unroll
                for (int j = 0; j < kTan; ++j) {
                    const int k = (int)k0 + j;
                    if (k < np) __builtin_nontemporal_store(two_diff * gf[j][o], rows + (1 + k) * rstride);
                }
                continue;
            }
            if (pass == 0) {
                const float sq = diff * diff;
                const float s = wave_sum(valid ? sq : 0.0f);
                if (lane == 0) dst[0] = s;
            }
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
    }
}

// One thread per (segment, scheduling position, slot); slot 0 = squared error, 1 + k =
// gradient k.  Every thread also carries the slot-0 sum, because the MSE decides whether
// the gradient is reported (evaluator.rs:214-216); it is recomputed from the same partials
// in the same order, so all threads of a program agree.  kstate: 4 floats per (sp, slot).
__global__ void k_finalize_grad(const SegTable T) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int si = find_seg_thread(T);
    const SegDesc& S = T.s[si];
    const int64_t idx = t - S.first_thread;
    const int sp = (int)(idx / S.slots);
    const int slot = (int)(idx - (int64_t)sp * S.slots);
    if (sp >= S.n_progs) return;
    if (S.live != nullptr && *S.live == 0) return;
    const int prog = S.order[sp];
    if (S.active != nullptr && !S.active[prog]) return;
    const int np = S.nparams[prog];
    if (np == 0) return;
    float s = 0.0f, c = 0.0f, m = 0.0f, mc = 0.0f;
    float* ks = S.kstate + ((int64_t)sp * S.slots + slot) * 4;
    if (!S.first) {
        s = ks[0];
        c = ks[1];
        m = ks[2];
        mc = ks[3];
    }
    const bool mine = slot == 0 || slot - 1 < np;
    const float* src = S.partials + (int64_t)sp * S.slots;
    const int64_t stride = (int64_t)S.n_progs * S.slots;
    int64_t lt = 0;
    for (; lt + 8 <= S.n_tiles; lt += 8) {  // loads issued ahead of the serial compensated chains
        float v[8], v0[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            v0[u] = src[(lt + u) * stride];
            v[u] = mine ? src[(lt + u) * stride + slot] : 0.0f;
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            if (mine) kahan_add(s, c, v[u]);
            kahan_add(m, mc, v0[u]);
        }
    }
    for (; lt < S.n_tiles; ++lt) {
        const float* row = src + lt * stride;
        if (mine) kahan_add(s, c, row[slot]);
        kahan_add(m, mc, row[0]);
    }
    if (!S.last) {
        ks[0] = s;
        ks[1] = c;
        ks[2] = m;
        ks[3] = mc;
        return;
    }
    const float mse = m / S.n_samples;
    const bool fin = __builtin_isfinite(mse);
    if (slot == 0) {
        S.mse_out[prog] = fin ? mse : FLT_MAX;
    } else {
        const int k = slot - 1;
        S.grad_out[(int64_t)prog * 32 + k] = (fin && k < np) ? s / S.n_samples : 0.0f;
    }
}

// ---------------------------------------------------------------------------------
// Big programs (spill slots in device memory): bounded launches, grid-stride over work
// items; wave w owns the device area gspill + w * gwave_floats for its whole life, so no
// two resident waves share one.  Work items are numbered per segment from first_big.
// ---------------------------------------------------------------------------------
__device__ __forceinline__ int find_seg_big(const SegTable& T, int64_t it) {
    int si = 0;
#pragma unroll
    for (int s = 1; s < kMaxSegs; ++s) si = (s < T.n && it >= T.s[s].first_big) ? s : si;
    return si;
}

// MSE: item = (big position j, tile), program-major so that consecutive items of a wave
// reuse the program's code.  Writes the same per-tile partial (or per-point terms) as the
// main kernel would.
template <int PPL>
__global__ __launch_bounds__(256) void k_eval_mse_big(const SegTable T, int spill_slots, int64_t n_items,
                                                      float* __restrict__ gspill, int64_t gwave_floats) {
    const int lane = threadIdx.x & 63;
    const int64_t wave = (int64_t)uni(blockIdx.x) * kWaves + uni(threadIdx.x >> 6);
    const int64_t n_waves = (int64_t)uni(gridDim.x) * kWaves;
    // every spill slot in device memory: the first kMaxSpillSlots where the LDS ones would be
    float* lds = gspill + wave * gwave_floats;
    float* gsp = lds + kMaxSpillSlots * PPL * 64;
    for (int64_t it = wave; it < n_items; it += n_waves) {
        const int si = find_seg_big(T, it);
        const SegDesc& S = T.s[si];
        const int64_t w = it - S.first_big;
        const int j = (int)(w / S.n_tiles);
        const int64_t lt = w - (int64_t)j * S.n_tiles;
        const int sp = S.n_small + j;
        if (S.live != nullptr && *S.live == 0) continue;
        const int prog = (int)uni((uint32_t)kget(S.order, sp));
        const bool on = S.active == nullptr || kget(S.active, prog);
        const int64_t base = (S.tile0 + lt) * (64 * PPL);
        float s = 0.0f;
        if (on) {
            Files<PPL> R;
            sfor<PPL>([&](auto P) { fget<P(), PPL>(R) = (v32f)(0.0f); });
            load_features<PPL>(R, S.feat, S.F, S.npad, base, lane);
            const ProgDesc pd = kget(S.progs, prog);
            run_program<PPL>(R, S.code + uni(pd.code_off), uni(pd.code_len), S.params + uni(pd.param_off), lds, gsp,
                             lane);
            const uint32_t o = uni(pd.out_reg) & 31u;
            if (S.terms) {
                float* row = S.partials + (int64_t)sp * S.n_tiles * (64 * PPL) + lt * (64 * PPL);
                sfor<PPL>([&](auto P) {
                    const v32f r = fget<P(), PPL>(R);
                    const float diff = r[o] - S.target[base + P() * 64 + lane];
                    row[P() * 64 + lane] = diff * diff;
                });
                continue;
            }
            sfor<PPL>([&](auto P) {
                const v32f r = fget<P(), PPL>(R);
                const float diff = r[o] - S.target[base + P() * 64 + lane];
                const float sq = diff * diff;
                s = s + ((base + P() * 64 + lane < S.count8) ? sq : 0.0f);
            });
            s = wave_sum(s);
        }
        if (!S.terms && lane == 0) S.partials[lt * S.n_progs + sp] = s;
    }
}

// Gradients, Kahan summation: item = (big position j, tile), every pass of the program.
__global__ __launch_bounds__(256) void k_eval_grad_big(const SegTable T, int spill_slots, int64_t n_items,
                                                       float* __restrict__ gspill, int64_t gwave_floats) {
    const int lane = threadIdx.x & 63;
    const int64_t wave = (int64_t)uni(blockIdx.x) * kWaves + uni(threadIdx.x >> 6);
    const int64_t n_waves = (int64_t)uni(gridDim.x) * kWaves;
    float* lds = gspill + wave * gwave_floats;  // every spill slot in device memory
    float* gsp = lds + kMaxSpillSlots * (1 + kTan) * 64;
    for (int64_t it = wave; it < n_items; it += n_waves) {
        const int si = find_seg_big(T, it);
        const SegDesc& S = T.s[si];
        const int64_t w = it - S.first_big;
        const int j = (int)(w / S.n_tiles);
        const int64_t lt = w - (int64_t)j * S.n_tiles;
        const int sp = S.n_small + j;
        if (S.live != nullptr && *S.live == 0) continue;
        const int prog = (int)uni((uint32_t)kget(S.order, sp));
        const int np = (int)uni(kget(S.nparams, prog));
        if ((S.active != nullptr && !kget(S.active, prog)) || np == 0) continue;
        const int64_t base = (S.tile0 + lt) * 64;
        const int64_t pt = base + lane;
        const bool valid = pt < S.count8;
        Files<1> R1;
        R1.a = (v32f)(0.0f);
        load_features<1>(R1, S.feat, S.F, S.npad, base, lane);
        DualFiles<kTan> G;
        G.v = R1.a;
        zero_tangents<kTan>(G);
        const ProgDesc pd = kget(S.progs, prog);
        const uint32_t o = uni(pd.out_reg) & 31u;
        float* dst = S.partials + (lt * S.n_progs + sp) * S.slots;
        for (int pass = 0; pass * kTan < np; ++pass) {
            const uint32_t k0 = (uint32_t)(pass * kTan);
            run_dual<kTan>(G, S.code + uni(pd.code_off), uni(pd.code_len), S.params + uni(pd.param_off), k0, lds,
                           gsp, lane);
            const v32f vfile = G.v;
            const float diff = vfile[o] - S.target[pt];
            const float two_diff = 2.0f * diff;
            v32f gf[kTan];
            tangent_files<kTan>(G, gf);
            if (S.terms) {
                float* rows = S.partials + (int64_t)sp * S.slots * S.n_tiles * 64 + lt * 64 + lane;
                const int64_t rstride = S.n_tiles * 64;
                if (pass == 0) rows[0] = diff * diff;
#pragma unroll
                for (int q = 0; q < kTan; ++q) {
                    const int k = (int)k0 + q;
                    if (k < np) rows[(1 + k) * rstride] = two_diff * gf[q][o];
                }
                continue;
            }
            if (pass == 0) {
                const float sq = diff * diff;
                const float sum = wave_sum(valid ? sq : 0.0f);
                if (lane == 0) dst[0] = sum;
            }
#pragma unroll
            for (int q = 0; q < kTan; ++q) {
                const int k = (int)k0 + q;
                if (k < np) {
                    const float sum = wave_sum(valid ? two_diff * gf[q][o] : 0.0f);
                    if (lane == 0) dst[1 + k] = sum;
                }
            }
        }
    }
}

// ---------------------------------------------------------------------------------
// host side: segment tables and launches
// ---------------------------------------------------------------------------------
int64_t eval_tiles(const DevDataset& D, int ppl) {
    const int64_t tp = 64 * (int64_t)ppl;
    return (D.count8 + tp - 1) / tp;
}

size_t eval_mse_lds_bytes(int spill_slots, int ppl) { return (size_t)kWaves * spill_slots * ppl * 64 * sizeof(float); }
size_t eval_grad_lds_bytes(int spill_slots) { return (size_t)kWaves * spill_slots * (1 + kTan) * 64 * sizeof(float); }

int64_t step_tiles(const EvalStep& st, int ppl) {
    return st.tile_count >= 0 ? st.tile_count : eval_tiles(st.D, ppl) - st.tile_begin;
}
int grad_slots(const EvalStep& st) { return 1 + (st.max_params > 0 ? st.max_params : 0); }
size_t mse_partials_floats(const EvalStep& st, int ppl, bool terms) {
    return (size_t)step_tiles(st, ppl) * st.P.n_progs * (terms ? 64 * ppl : 1);
}
size_t grad_partials_floats(const EvalStep& st) {
    return (size_t)step_tiles(st, 1) * st.P.n_progs * grad_slots(st);
}
size_t mse_state_floats(const EvalStep& st, bool terms) { return (size_t)(terms ? 8 : 2) * st.P.n_progs; }
size_t grad_terms_floats(const EvalStep& st) {
    const size_t rows = (size_t)st.P.n_progs * grad_slots(st);
    return (size_t)step_tiles(st, 1) * 64 * rows + rows;
}
size_t grad_terms_state_floats(const EvalStep& st) { return (size_t)8 * st.P.n_progs * grad_slots(st); }
size_t grad_state_floats(const EvalStep& st) { return (size_t)4 * st.P.n_progs * grad_slots(st); }

static SegDesc make_seg(const EvalStep& st, int ppl, int group) {
    SegDesc d{};
    d.tile_pts = 64 * ppl;
    d.code = st.P.code;
    d.progs = st.P.progs;
    d.params = st.P.params;
    d.order = st.P.order;
    d.active = st.active;
    d.live = st.live;
    d.feat = st.D.feat;
    d.target = st.D.target;
    d.partials = st.partials;
    d.mse_out = st.mse_out;
    d.grad_out = st.grad_out;
    d.nparams = st.nparams;
    d.kstate = st.kstate;
    d.npad = st.D.npad;
    d.count8 = st.D.count8;
    d.tile0 = st.tile_begin;
    d.n_tiles = step_tiles(st, ppl);
    d.slots = grad_slots(st);
    d.first = st.first ? 1 : 0;
    d.last = st.last ? 1 : 0;
    d.n_progs = st.P.n_progs;
    d.n_small = st.P.n_progs - std::min(std::max(st.P.n_big, 0), st.P.n_progs);
    d.group = group;
    d.n_groups = (d.n_small + group - 1) / group;
    d.F = st.D.num_features;
    d.n_samples = (float)st.D.num_samples;
    d.fold_rows = 1;
    return d;
}

static int lds_slots(int spill_slots) { return std::min(spill_slots, kMaxSpillSlots); }

// Floats of one wave's device spill area: every slot, the LDS-range ones included.
static int64_t big_wave_floats(int gslots, int channels) {
    return (int64_t)(kMaxSpillSlots + std::max(gslots, 0)) * channels * 64;
}

// At most 1024 waves (four per CU), fewer when the per-wave area is large: the launch's
// device area stays under 256 MiB.  A multiple of the waves per workgroup.
int64_t big_waves(int gslots, int channels) {
    const int64_t per = big_wave_floats(gslots, channels) * (int64_t)sizeof(float);
    const int64_t w = std::max<int64_t>(64, std::min<int64_t>(1024, ((int64_t)256 << 20) / per));
    return w / kWaves * kWaves;
}

size_t big_spill_floats(const EvalStep* steps, int n, int ppl, bool grad) {
    int g = -1;
    for (int i = 0; i < n; ++i)
        if (steps[i].P.n_big > 0 && steps[i].P.n_progs > 0) g = std::max(g, steps[i].P.gslots);
    if (g < 0) return 0;
    const int ch = grad ? 1 + kTan : ppl;
    return (size_t)(big_waves(g, ch) * big_wave_floats(g, ch));
}

// Launch the big-program kernel K over `items` work items; each launched wave owns one
// device spill area of big.ptr, which must hold every launched wave's area.
template <typename K>
static hipError_t launch_big(K kernel, const SegTable& T, int64_t items, int gslots, int channels, int spill,
                             BigSpill big, hipStream_t s) {
    const int64_t blocks = (std::min<int64_t>(big_waves(gslots, channels), items) + kWaves - 1) / kWaves;
    const int64_t gwave = big_wave_floats(gslots, channels);
    if (big.ptr == nullptr || (size_t)(blocks * kWaves * gwave) > big.floats) return hipErrorInvalidValue;
    hipLaunchKernelGGL(kernel, dim3((unsigned)blocks), dim3(64 * kWaves), 0, s, T, spill, items, big.ptr, gwave);
    return hipGetLastError();
}

// Workgroups of `kernel` that fit the device at once (every CU at its occupancy for this
// block size and dynamic LDS), cached per (kernel, LDS bytes, device).
static int64_t resident_blocks(const void* kernel, size_t lds) {
    static std::mutex mu;
    static std::map<std::tuple<const void*, size_t, int>, int64_t> cache;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 0;
    std::lock_guard<std::mutex> g(mu);
    const auto key = std::make_tuple(kernel, lds, dev);
    auto it = cache.find(key);
    if (it != cache.end()) return it->second;
    int per_cu = 0, cus = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, 64 * kWaves, lds) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        return 0;
    const int64_t n = (int64_t)std::max(per_cu, 1) * std::max(cus, 1);
    cache[key] = n;
    return n;
}

template <int PPL, bool TERMS>
static hipError_t launch_eval_mse_t(const SegTable& T, int64_t waves, int spill, size_t lds, int* queue,
                                    hipStream_t s) {
    const int64_t blocks = (waves + kWaves - 1) / kWaves;
    // the queue pays off only when the work exceeds what the device holds at once: a smaller
    // launch is one wave per item in either form, and its waves would only add the takes
    const int64_t res = queue ? resident_blocks((const void*)k_eval_mse_q<PPL, TERMS>, lds) : 0;
    if (queue == nullptr || res <= 0 || blocks <= res) {
        hipLaunchKernelGGL((k_eval_mse<PPL, TERMS>), dim3((unsigned)blocks), dim3(64 * kWaves), lds, s, T, spill);
        return hipGetLastError();
    }
    if (waves / kXcds >= INT32_MAX) return hipErrorInvalidValue;  // per-XCD counters are 32-bit
    hipLaunchKernelGGL((k_eval_mse_q<PPL, TERMS>), dim3((unsigned)res), dim3(64 * kWaves), lds, s, T, spill, waves,
                       queue);
    return hipGetLastError();
}

template <int PPL>
static hipError_t launch_eval_mse(const SegTable& T, bool terms, int64_t waves, int spill, size_t lds, int* queue,
                                  hipStream_t s) {
    return terms ? launch_eval_mse_t<PPL, true>(T, waves, spill, lds, queue, s)
                 : launch_eval_mse_t<PPL, false>(T, waves, spill, lds, queue, s);
}

hipError_t launch_mse_steps(const EvalStep* steps, int n, const EvalTuning& tune, bool terms, hipStream_t s,
                            const KernelTimer* timer, BigSpill big, int* queue) {
    for (int base = 0; base < n; base += kMaxSegs) {
        const int cnt = min(kMaxSegs, n - base);
        SegTable T{};
        int64_t waves = 0, threads = 0, big_items = 0;
        int spill = 0, gslots = 0;
        for (int i = 0; i < cnt; ++i) {
            const EvalStep& st = steps[base + i];
            if (st.P.n_progs <= 0) continue;
            SegDesc d = make_seg(st, tune.ppl, st.group > 0 ? st.group : tune.group > 0 ? tune.group : 8);
            if (d.n_tiles <= 0) continue;
            d.terms = terms ? 1 : 0;
            d.first_wave = waves;
            d.first_thread = threads;
            d.first_big = big_items;
            waves += d.n_tiles * d.n_groups;
            big_items += d.n_tiles * (d.n_progs - d.n_small);
            if (d.n_progs > d.n_small) gslots = max(gslots, st.P.gslots);
            // finalize threads (whole waves per segment), or fold waves
            threads += terms ? (int64_t)d.n_progs : ((int64_t)d.n_progs + 63) / 64 * 64;
            spill = max(spill, lds_slots(st.P.spill_slots));
            T.s[T.n++] = d;
        }
        if (waves == 0 && big_items == 0) continue;
        const int64_t blocks = (waves + kWaves - 1) / kWaves;
        if (blocks > INT32_MAX) return hipErrorInvalidValue;
        if (timer && base == 0) (void)hipEventRecord(timer->begin, s);
        const size_t lds = eval_mse_lds_bytes(spill, tune.ppl);
        if (waves > 0) {
            hipError_t e;
            switch (tune.ppl) {
                case 1: e = launch_eval_mse<1>(T, terms, waves, spill, lds, queue, s); break;
                case 3: e = launch_eval_mse<3>(T, terms, waves, spill, lds, queue, s); break;
                case 4: e = launch_eval_mse<4>(T, terms, waves, spill, lds, queue, s); break;
                default: e = launch_eval_mse<2>(T, terms, waves, spill, lds, queue, s); break;
            }
            if (e != hipSuccess) return e;
        }
        if (big_items > 0) {
            hipError_t e;
            switch (tune.ppl) {
                case 1: e = launch_big(k_eval_mse_big<1>, T, big_items, gslots, 1, spill, big, s); break;
                case 3: e = launch_big(k_eval_mse_big<3>, T, big_items, gslots, 3, spill, big, s); break;
                case 4: e = launch_big(k_eval_mse_big<4>, T, big_items, gslots, 4, spill, big, s); break;
                default: e = launch_big(k_eval_mse_big<2>, T, big_items, gslots, 2, spill, big, s); break;
            }
            if (e != hipSuccess) return e;
        }
        if (timer && base + cnt >= n) (void)hipEventRecord(timer->end, s);
        const int tb = 256;
        if (terms) {
            if (threads > INT32_MAX) return hipErrorInvalidValue;
            hipLaunchKernelGGL(k_fold_terms, dim3((unsigned)threads), dim3(64), 0, s, T);
        } else
            hipLaunchKernelGGL(k_finalize_mse, dim3((unsigned)((threads + tb - 1) / tb)), dim3(tb), 0, s, T);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

hipError_t launch_grad_steps(const EvalStep* steps, int n, bool terms, hipStream_t s, const KernelTimer* timer,
                             BigSpill big) {
    for (int base = 0; base < n; base += kMaxSegs) {
        const int cnt = min(kMaxSegs, n - base);
        SegTable T{};
        int64_t waves = 0, threads = 0, big_items = 0, fin = 0;
        int spill = 0, gslots = 0;
        for (int i = 0; i < cnt; ++i) {
            const EvalStep& st = steps[base + i];
            if (st.max_params <= 0 || st.P.n_progs <= 0) continue;
            SegDesc d = make_seg(st, 1, grad_group(step_tiles(st, 1), st.P.n_progs));
            if (d.n_tiles <= 0) continue;
            d.terms = terms ? 1 : 0;
            d.first_wave = waves;
            d.first_thread = threads;
            d.first_big = big_items;
            d.first_fin = fin;
            const int nb = d.n_progs - d.n_small;
            waves += d.n_tiles * d.n_groups;
            big_items += d.n_tiles * nb;
            if (nb > 0) gslots = max(gslots, st.P.gslots);
            if (terms) {
                d.fold_rows = d.slots;
                d.gsum = st.partials + (size_t)d.n_tiles * 64 * d.n_progs * d.slots;
                threads += (int64_t)d.n_progs * d.slots;                           // fold waves
                fin += ((int64_t)d.n_progs * d.slots + 63) / 64 * 64;               // finalize threads
            } else {
                threads += ((int64_t)d.n_progs * d.slots + 63) / 64 * 64;  // whole waves per segment
            }
            spill = max(spill, lds_slots(st.P.spill_slots));
            T.s[T.n++] = d;
        }
        if (waves == 0 && big_items == 0) continue;
        const int64_t blocks = (waves + kWaves - 1) / kWaves;
        if (blocks > INT32_MAX) return hipErrorInvalidValue;
        const size_t lds = eval_grad_lds_bytes(spill);
        if (timer && base == 0) (void)hipEventRecord(timer->begin, s);
        if (waves > 0) {
            hipLaunchKernelGGL(k_eval_grad, dim3((unsigned)blocks), dim3(64 * kWaves), lds, s, T, spill);
            hipError_t e = hipGetLastError();
            if (e != hipSuccess) return e;
        }
        if (big_items > 0) {
            const hipError_t e = launch_big(k_eval_grad_big, T, big_items, gslots, 1 + kTan, spill, big, s);
            if (e != hipSuccess) return e;
        }
        if (timer && base + cnt >= n) (void)hipEventRecord(timer->end, s);
        const int tb = 256;
        if (terms) {
            if (threads > INT32_MAX) return hipErrorInvalidValue;
            hipLaunchKernelGGL(k_fold_terms, dim3((unsigned)threads), dim3(64), 0, s, T);
            hipError_t e = hipGetLastError();
            if (e != hipSuccess) return e;
            bool last = false;
            for (int i = 0; i < T.n; ++i) last = last || T.s[i].last;
            if (last) {
                hipLaunchKernelGGL(k_finalize_grad_terms, dim3((unsigned)((fin + tb - 1) / tb)), dim3(tb), 0, s, T);
                if ((e = hipGetLastError()) != hipSuccess) return e;
            }
        } else {
            hipLaunchKernelGGL(k_finalize_grad, dim3((unsigned)((threads + tb - 1) / tb)), dim3(tb), 0, s, T);
            hipError_t e = hipGetLastError();
            if (e != hipSuccess) return e;
        }
    }
    return hipSuccess;
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
    const ProgDesc pd = progs[first + i];
    if (uni(pd.out_reg) & kBigProg) return;  // the big-program launch writes these rows
    const int64_t base = tile * 64;
    Files<1> R;
    R.a = (v32f)(0.0f);
    load_features<1>(R, feat, F, npad, base, lane);
    run_program<1>(R, code + uni(pd.code_off), uni(pd.code_len), params + uni(pd.param_off), lds, nullptr, lane);
    const int64_t pt = base + lane;
    const v32f rf = R.a;
    if (pt < count8) out[(int64_t)i * count8 + pt] = rf[uni(pd.out_reg) & 31u];
}

// The big programs among [first, first + count): grid-stride over (program, tile) items,
// skipping the others (every wave owns its device spill area).
__global__ __launch_bounds__(256) void k_eval_points_big(const UInstr* __restrict__ code,
                                                         const ProgDesc* __restrict__ progs,
                                                         const float* __restrict__ params, int first, int count,
                                                         const float* __restrict__ feat, int F, int64_t npad,
                                                         int64_t count8, int64_t n_tiles, float* __restrict__ out,
                                                         int spill_slots, float* __restrict__ gspill,
                                                         int64_t gwave_floats) {
    const int lane = threadIdx.x & 63;
    const int64_t wave = (int64_t)uni(blockIdx.x) * kWaves + uni(threadIdx.x >> 6);
    const int64_t n_waves = (int64_t)uni(gridDim.x) * kWaves;
    float* lds = gspill + wave * gwave_floats;  // every spill slot in device memory
    float* gsp = lds + kMaxSpillSlots * 64;
    const int64_t n_items = (int64_t)count * n_tiles;
    for (int64_t it = wave; it < n_items; it += n_waves) {
        const int i = (int)(it / n_tiles);
        const int64_t tile = it - (int64_t)i * n_tiles;
        const ProgDesc pd = progs[first + i];
        if (!(uni(pd.out_reg) & kBigProg)) continue;
        const int64_t base = tile * 64;
        Files<1> R;
        R.a = (v32f)(0.0f);
        load_features<1>(R, feat, F, npad, base, lane);
        run_program<1>(R, code + uni(pd.code_off), uni(pd.code_len), params + uni(pd.param_off), lds, gsp, lane);
        const int64_t pt = base + lane;
        const v32f rf = R.a;
        if (pt < count8) out[(int64_t)i * count8 + pt] = rf[uni(pd.out_reg) & 31u];
    }
}

hipError_t launch_eval_points(const DevPrograms& P, const DevDataset& D, int first, int count, bool any_big,
                              float* out, hipStream_t s, BigSpill big) {
    if (count <= 0) return hipSuccess;
    const int64_t n_tiles = eval_tiles(D, 1);
    const int64_t blocks = (n_tiles * count + kWaves - 1) / kWaves;
    const int spill = lds_slots(P.spill_slots);
    const size_t lds_bytes = eval_mse_lds_bytes(spill, 1);
    hipLaunchKernelGGL(k_eval_points, dim3((unsigned)blocks), dim3(64 * kWaves), lds_bytes, s, P.code, P.progs,
                       P.params, first, count, D.feat, D.num_features, D.npad, D.count8, n_tiles, out, spill);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess || !any_big) return e;
    const int64_t bblocks = (std::min<int64_t>(big_waves(P.gslots, 1), n_tiles * count) + kWaves - 1) / kWaves;
    const int64_t gwave = big_wave_floats(P.gslots, 1);
    if (big.ptr == nullptr || (size_t)(bblocks * kWaves * gwave) > big.floats) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_eval_points_big, dim3((unsigned)bblocks), dim3(64 * kWaves), 0, s, P.code, P.progs, P.params,
                       first, count, D.feat, D.num_features, D.npad, D.count8, n_tiles, out, spill, big.ptr, gwave);
    return hipGetLastError();
}

// One program's value and tangent per point; tiles grid-strided over the launch's waves
// (one wave per workgroup), each wave with its own device spill area when gspill is set.
__global__ __launch_bounds__(64) void k_eval_points_dual(const UInstr* __restrict__ code, const ProgDesc* __restrict__ progs,
                                                         const float* __restrict__ params, int prog, int active_param,
                                                         const float* __restrict__ feat, int F, int64_t npad,
                                                         int64_t count8, int64_t n_tiles, float* __restrict__ val_out,
                                                         float* __restrict__ grad_out, float* __restrict__ gspill,
                                                         int64_t gwave_floats) {
    extern __shared__ float smem[];
    const int lane = threadIdx.x & 63;
    // a big program keeps every spill slot in the block's device area
    float* lds = gspill != nullptr ? gspill + (int64_t)blockIdx.x * gwave_floats : smem;
    float* gsp = gspill != nullptr ? lds + kMaxSpillSlots * 2 * 64 : nullptr;
    const ProgDesc pd = progs[prog];
    const uint32_t o = uni(pd.out_reg) & 31u;
    for (int64_t tile = blockIdx.x; tile < n_tiles; tile += gridDim.x) {
        const int64_t base = tile * 64;
        Files<1> R1;
        R1.a = (v32f)(0.0f);
        load_features<1>(R1, feat, F, npad, base, lane);
        DualFiles<1> G;
        G.v = R1.a;
        G.a = (v32f)(0.0f);
        run_dual<1>(G, code + uni(pd.code_off), uni(pd.code_len), params + uni(pd.param_off), (uint32_t)active_param,
                    lds, gsp, lane);
        const int64_t pt = base + lane;
        const v32f vf = G.v, gfl = G.a;
        if (pt < count8) {
            val_out[pt] = vf[o];
            grad_out[pt] = gfl[o];
        }
    }
}

hipError_t launch_eval_points_dual(const DevPrograms& P, const DevDataset& D, int prog, int active_param, bool is_big,
                                   float* val_out, float* grad_out, hipStream_t s, BigSpill big) {
    const int64_t n_tiles = eval_tiles(D, 1);
    const size_t lds_bytes = (size_t)lds_slots(P.spill_slots) * 2 * 64 * sizeof(float);
    int64_t blocks = n_tiles;
    float* g = nullptr;
    int64_t gwave = 0;
    size_t lds = lds_bytes;
    if (is_big) {
        blocks = std::min<int64_t>(big_waves(P.gslots, 2), n_tiles);
        gwave = big_wave_floats(P.gslots, 2);
        if (big.ptr == nullptr || (size_t)(blocks * gwave) > big.floats) return hipErrorInvalidValue;
        g = big.ptr;
        lds = 0;
    }
    hipLaunchKernelGGL(k_eval_points_dual, dim3((unsigned)blocks), dim3(64), lds, s, P.code, P.progs, P.params,
                       prog, active_param, D.feat, D.num_features, D.npad, D.count8, n_tiles, val_out, grad_out, g,
                       gwave);
    return hipGetLastError();
}

}  // namespace rsr
