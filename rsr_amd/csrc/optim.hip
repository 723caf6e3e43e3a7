// optim.hip -- batched constant fitting on the device: L-BFGS and diagonal CMA-ES.
//
// Every individual's optimizer state lives in device arrays laid out field-major
// (value j of individual i at [j * n + i]), one thread per individual.  The expensive
// part -- the MSE / gradient evaluations of all individuals -- runs in the evaluation
// kernels with an active mask, so one launch evaluates every individual still
// working on that step.  Trial points are written straight into the device parameter
// array the evaluation kernels read.
#include <float.h>
#include <math.h>
#include "devmath.h"
#include "optim.h"

namespace rsr {

// ------------------------------------------------------------------------------------
// L-BFGS (src/engine/optimize/lbfgs.rs:9-165)
//   T = float : reference parity (f32 everywhere, second loop starts at slot 0 when the
//               history is not full -- lbfgs.rs:80, including after a reset with head != 0)
//   T = double: the two-loop dots, curvature pairs, gamma, directional derivative and the
//               Armijo test in FP64, second loop from the oldest live pair
// ------------------------------------------------------------------------------------
template <typename T>
__global__ void k_lbfgs_begin(LbfgsDev st, const float* __restrict__ params, const float* __restrict__ d_mse,
                              const float* __restrict__ d_grad) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= st.n) return;
    const int prog = st.prog[i];
    const int np = st.nparam[i];
    const float* pp = params + st.param_off[i];
    for (int j = 0; j < kMaxP; ++j) {
        st.x[j * st.n + i] = j < np ? pp[j] : 0.0f;
        st.g[j * st.n + i] = d_grad[(int64_t)prog * 32 + j];
    }
    st.cur_mse[i] = d_mse[prog];
    st.hist[i] = 0;
    st.head[i] = 0;
    st.flags[i] = np == 0 ? LB_DONE : 0;
}

template <typename T>
__device__ __forceinline__ T& at(T* a, int slot, int j, int n, int i) {
    return a[((int64_t)slot * kMaxP + j) * n + i];
}

template <typename T>
__global__ void k_lbfgs_direction(LbfgsDev st, float* __restrict__ params, uint8_t* __restrict__ active_ls,
                                  int* __restrict__ n_working) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= st.n) return;
    const int prog = st.prog[i];
    active_ls[prog] = 0;
    if (st.flags[i] & LB_DONE) return;
    const int n = st.nparam[i];
    const int N = st.n;
    const int M = st.M;
    T* S = (T*)st.S;
    T* Y = (T*)st.Y;
    T* rho = (T*)st.rho;
    T* alpha = (T*)st.alpha;
    T q[kMaxP];
    T gn = (T)0;
    for (int j = 0; j < n; ++j) {
        const T gj = (T)st.g[j * N + i];
        gn += gj * gj;
    }
    if (sqrt(gn) < (T)1e-5) {
        st.flags[i] |= LB_DONE;
        return;
    }
    for (int j = 0; j < n; ++j) q[j] = (T)st.g[j * N + i];
    int hist = st.hist[i], head = st.head[i];
    int ci = head;
    for (int h = 0; h < hist; ++h) {
        ci = ci == 0 ? M - 1 : ci - 1;
        T d = (T)0;
        for (int j = 0; j < n; ++j) d += at(S, ci, j, N, i) * q[j];
        const T a = rho[ci * N + i] * d;
        alpha[ci * N + i] = a;
        for (int j = 0; j < n; ++j) q[j] -= a * at(Y, ci, j, N, i);
    }
    T gamma = (T)1;
    if (hist > 0) {
        const int li = head == 0 ? M - 1 : head - 1;
        T sy = (T)0, yy = (T)0;
        for (int j = 0; j < n; ++j) {
            sy += at(S, li, j, N, i) * at(Y, li, j, N, i);
            yy += at(Y, li, j, N, i) * at(Y, li, j, N, i);
        }
        if (yy > (T)1e-10) gamma = sy / yy;
    }
    for (int j = 0; j < n; ++j) q[j] *= gamma;
    if constexpr (sizeof(T) == 4) ci = hist < M ? 0 : head;  // reference (lbfgs.rs:80)
    else ci = (head - hist + M) % M;                         // oldest live pair
    for (int h = 0; h < hist; ++h) {
        T d = (T)0;
        for (int j = 0; j < n; ++j) d += at(Y, ci, j, N, i) * q[j];
        const T beta = rho[ci * N + i] * d;
        const T am = alpha[ci * N + i] - beta;
        for (int j = 0; j < n; ++j) q[j] += at(S, ci, j, N, i) * am;
        ci = (ci + 1) % M;
    }
    T ddg = (T)0;
    for (int j = 0; j < n; ++j) {
        const float pj = (float)(-q[j]);
        st.p[j * N + i] = pj;
        ddg += (T)pj * (T)st.g[j * N + i];
    }
    if (ddg > (T)-1e-8) {  // steepest-descent reset (lbfgs.rs:100-106)
        for (int j = 0; j < n; ++j) st.p[j * N + i] = -st.g[j * N + i];
        ddg = -gn;
        hist = 0;
    }
    st.hist[i] = hist;
    ((T*)st.ddg)[i] = ddg;
    st.step[i] = 1.0f;
    st.ls[i] = 0;
    st.flags[i] |= LB_SEARCH;
    float* pp = params + st.param_off[i];
    for (int j = 0; j < n; ++j) pp[j] = st.x[j * N + i] + 1.0f * st.p[j * N + i];
    active_ls[prog] = 1;
    atomicAdd(n_working, 1);
}

template <typename T>
__global__ void k_lbfgs_ls_check(LbfgsDev st, float* __restrict__ params, const float* __restrict__ d_mse,
                                 uint8_t* __restrict__ active_ls, uint8_t* __restrict__ active_grad,
                                 int* __restrict__ n_searching) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= st.n) return;
    if (!(st.flags[i] & LB_SEARCH)) return;
    const int prog = st.prog[i];
    const int n = st.nparam[i];
    const int N = st.n;
    const float m = d_mse[prog];
    const float step = st.step[i];
    bool accept;
    if constexpr (sizeof(T) == 4) {
        const float c1 = 1e-4f;
        accept = m <= st.cur_mse[i] + c1 * step * ((float*)st.ddg)[i];
    } else {
        accept = (double)m <= (double)st.cur_mse[i] + 1e-4 * (double)step * ((double*)st.ddg)[i];
    }
    if (accept || st.ls[i] > 10) {
        st.next_mse[i] = m;
        st.flags[i] &= ~LB_SEARCH;
        st.flags[i] |= LB_MOVED;
        active_ls[prog] = 0;
        active_grad[prog] = 1;
        return;
    }
    const float ns = step * 0.5f;
    st.step[i] = ns;
    st.ls[i] += 1;
    float* pp = params + st.param_off[i];
    for (int j = 0; j < n; ++j) pp[j] = st.x[j * N + i] + ns * st.p[j * N + i];
    atomicAdd(n_searching, 1);
}

template <typename T>
__global__ void k_lbfgs_update(LbfgsDev st, const float* __restrict__ d_grad, uint8_t* __restrict__ active_grad) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= st.n) return;
    const int prog = st.prog[i];
    active_grad[prog] = 0;
    if (!(st.flags[i] & LB_MOVED)) return;
    st.flags[i] &= ~LB_MOVED;
    const int n = st.nparam[i];
    const int N = st.n;
    const int M = st.M;
    T* S = (T*)st.S;
    T* Y = (T*)st.Y;
    T sn[kMaxP], yn[kMaxP];
    T sy = (T)0;
    const float step = st.step[i];
    for (int j = 0; j < n; ++j) {
        const float x = st.x[j * N + i];
        const float nx = x + step * st.p[j * N + i];
        const float g = st.g[j * N + i];
        const float ng = d_grad[(int64_t)prog * 32 + j];
        if constexpr (sizeof(T) == 4) {
            sn[j] = nx - x;
            yn[j] = ng - g;
        } else {
            sn[j] = (T)nx - (T)x;
            yn[j] = (T)ng - (T)g;
        }
        sy += sn[j] * yn[j];
        st.x[j * N + i] = nx;
    }
    for (int j = 0; j < kMaxP; ++j) st.g[j * N + i] = d_grad[(int64_t)prog * 32 + j];
    if (sy > (T)1e-10) {
        const int head = st.head[i];
        for (int j = 0; j < n; ++j) {
            at(S, head, j, N, i) = sn[j];
            at(Y, head, j, N, i) = yn[j];
        }
        ((T*)st.rho)[head * N + i] = (T)1 / sy;
        st.head[i] = (head + 1) % M;
        if (st.hist[i] < M) st.hist[i] += 1;
    }
    st.cur_mse[i] = st.next_mse[i];
}

// One thread per (individual k, round r): candidate k * K + r is x + 2^-r p, the point that
// round r of the sequential search evaluates (the step halves exactly from 1, and the point is
// the same x + step * p expression), active while the individual searches.
__global__ void k_lbfgs_ls_fill(LbfgsDev st) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)st.n * st.K) return;
    const int i = (int)(t / st.K), r = (int)(t - (int64_t)i * st.K);
    const bool on = (st.flags[i] & LB_SEARCH) != 0;
    st.cand_active[t] = on ? 1 : 0;
    if (!on) return;
    float step = 1.0f;
    for (int q = 0; q < r; ++q) step = step * 0.5f;
    const int n = st.nparam[i];
    const int N = st.n;
    float* cp = st.cand_params + st.cand_off[t];
    for (int j = 0; j < n; ++j) cp[j] = st.x[j * N + i] + step * st.p[j * N + i];
}

// One thread per individual: the rounds of lbfgs.rs:113-125 over the candidates' MSEs in
// order; the accepted round's point, step and round count are what the sequential rounds
// leave (k_lbfgs_ls_check), and counter gets the K - 1 evaluations beyond the one the
// working count covers.
template <typename T>
__global__ void k_lbfgs_ls_pick(LbfgsDev st, float* __restrict__ params, const float* __restrict__ cand_mse,
                                uint8_t* __restrict__ active_ls, uint8_t* __restrict__ active_grad,
                                int* __restrict__ n_evals) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= st.n) return;
    if (!(st.flags[i] & LB_SEARCH)) return;
    const int prog = st.prog[i];
    const int n = st.nparam[i];
    const int N = st.n;
    float step = 1.0f;
    for (int r = 0; r < st.K; ++r) {
        const int64_t c = (int64_t)i * st.K + r;
        const float m = cand_mse[c];
        bool accept;
        if constexpr (sizeof(T) == 4) {
            const float c1 = 1e-4f;
            accept = m <= st.cur_mse[i] + c1 * step * ((float*)st.ddg)[i];
        } else {
            accept = (double)m <= (double)st.cur_mse[i] + 1e-4 * (double)step * ((double*)st.ddg)[i];
        }
        if (accept || r == st.K - 1) {  // K = 12: the last round is forced (lbfgs.rs:113-125)
            st.next_mse[i] = m;
            st.step[i] = step;
            st.ls[i] = r;
            st.flags[i] &= ~LB_SEARCH;
            st.flags[i] |= LB_MOVED;
            active_ls[prog] = 0;
            active_grad[prog] = 1;
            const float* cp = st.cand_params + st.cand_off[c];
            float* pp = params + st.param_off[i];
            for (int j = 0; j < n; ++j) pp[j] = cp[j];
            break;
        }
        step = step * 0.5f;
    }
    atomicAdd(n_evals, st.K - 1);
}

__global__ void k_lbfgs_finish(LbfgsDev st, float* __restrict__ params, float* __restrict__ fitness) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= st.n) return;
    const int n = st.nparam[i];
    float* pp = params + st.param_off[i];
    for (int j = 0; j < n; ++j) pp[j] = st.x[j * st.n + i];
    fitness[i] = n > 0 ? st.cur_mse[i] : __builtin_nanf("");
}

#define RSR_BLK(n) dim3((unsigned)(((n) + 127) / 128)), dim3(128)

hipError_t lbfgs_begin(const LbfgsDev& st, const float* params, const float* d_mse, const float* d_grad,
                       hipStream_t s) {
    if (st.fp64) hipLaunchKernelGGL(k_lbfgs_begin<double>, RSR_BLK(st.n), 0, s, st, params, d_mse, d_grad);
    else hipLaunchKernelGGL(k_lbfgs_begin<float>, RSR_BLK(st.n), 0, s, st, params, d_mse, d_grad);
    return hipGetLastError();
}

hipError_t lbfgs_direction(const LbfgsDev& st, float* params, uint8_t* active_ls, int* counter, hipStream_t s) {
    if (st.fp64) hipLaunchKernelGGL(k_lbfgs_direction<double>, RSR_BLK(st.n), 0, s, st, params, active_ls, counter);
    else hipLaunchKernelGGL(k_lbfgs_direction<float>, RSR_BLK(st.n), 0, s, st, params, active_ls, counter);
    return hipGetLastError();
}

hipError_t lbfgs_ls_check(const LbfgsDev& st, float* params, const float* d_mse, uint8_t* active_ls,
                          uint8_t* active_grad, int* counter, hipStream_t s) {
    if (st.fp64)
        hipLaunchKernelGGL(k_lbfgs_ls_check<double>, RSR_BLK(st.n), 0, s, st, params, d_mse, active_ls, active_grad,
                           counter);
    else
        hipLaunchKernelGGL(k_lbfgs_ls_check<float>, RSR_BLK(st.n), 0, s, st, params, d_mse, active_ls, active_grad,
                           counter);
    return hipGetLastError();
}

hipError_t lbfgs_ls_fill(const LbfgsDev& st, hipStream_t s) {
    hipLaunchKernelGGL(k_lbfgs_ls_fill, RSR_BLK((int64_t)st.n * st.K), 0, s, st);
    return hipGetLastError();
}

hipError_t lbfgs_ls_pick(const LbfgsDev& st, float* params, const float* cand_mse, uint8_t* active_ls,
                         uint8_t* active_grad, int* counter, hipStream_t s) {
    if (st.fp64)
        hipLaunchKernelGGL(k_lbfgs_ls_pick<double>, RSR_BLK(st.n), 0, s, st, params, cand_mse, active_ls, active_grad,
                           counter);
    else
        hipLaunchKernelGGL(k_lbfgs_ls_pick<float>, RSR_BLK(st.n), 0, s, st, params, cand_mse, active_ls, active_grad,
                           counter);
    return hipGetLastError();
}

hipError_t lbfgs_update(const LbfgsDev& st, const float* d_grad, uint8_t* active_grad, hipStream_t s) {
    if (st.fp64) hipLaunchKernelGGL(k_lbfgs_update<double>, RSR_BLK(st.n), 0, s, st, d_grad, active_grad);
    else hipLaunchKernelGGL(k_lbfgs_update<float>, RSR_BLK(st.n), 0, s, st, d_grad, active_grad);
    return hipGetLastError();
}

hipError_t lbfgs_finish(const LbfgsDev& st, float* params, float* fitness, hipStream_t s) {
    hipLaunchKernelGGL(k_lbfgs_finish, RSR_BLK(st.n), 0, s, st, params, fitness);
    return hipGetLastError();
}

// ------------------------------------------------------------------------------------
// Diagonal CMA-ES (src/engine/optimize/cmaes.rs:19-192)
// ------------------------------------------------------------------------------------
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy as
// 1, 2, 3", SC'11): counter-based, so candidate (i, it, c, j) draws the same numbers
// regardless of launch geometry.
__host__ __device__ __forceinline__ void philox4x32(uint32_t c[4], uint32_t k0, uint32_t k1) {
    for (int r = 0; r < 10; ++r) {
        const uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
        const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        const uint32_t n0 = hi1 ^ c[1] ^ k0, n2 = hi0 ^ c[3] ^ k1;
        c[0] = n0;
        c[1] = lo1;
        c[2] = n2;
        c[3] = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
}

// Standard normal for (individual, iteration, candidate, dimension): two uniforms in
// [0, 1) from the top 24 bits, then Box-Muller exactly as cmaes.rs:13-17.
__host__ __device__ __forceinline__ float cma_normal(uint64_t seed, uint32_t ind, uint32_t it, uint32_t cand,
                                                     uint32_t j) {
    uint32_t c[4] = {cand * 32u + j, it, ind, 0x43414D45u};
    philox4x32(c, (uint32_t)seed, (uint32_t)(seed >> 32));
    float u1 = (float)(c[0] >> 8) * (1.0f / 16777216.0f);
    const float u2 = (float)(c[1] >> 8) * (1.0f / 16777216.0f);
    u1 = u1 > 1e-8f ? u1 : 1e-8f;  // .max(1e-8)
    return __builtin_sqrtf(-2.0f * dm_ln(u1)) * dm_cos(2.0f * 3.14159274101257324f * u2);
}

// Starting point of every individual (cmaes.rs:104-110): best = the doubled mean, scored
// by the parent's MSE plus 0.01 * its L1 norm (summed in dimension order).
__global__ void k_cmaes_begin(CmaDev st, const float* __restrict__ d_mse) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= st.n) return;
    const int N = st.n;
    float l1 = 0.0f;
    for (int j = 0; j < st.ndim[i]; ++j) l1 += __builtin_fabsf(st.mean[j * N + i]);
    const float m = d_mse[st.prog[i]];
    st.best_mse[i] = m;
    st.best_fit[i] = m + 0.01f * l1;
}

__global__ void k_cmaes_sample(CmaDev st, float* __restrict__ cand_params, int it) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = t / st.lambda, c = t - i * st.lambda;
    if (i >= st.n) return;
    const int N = st.n;
    const int64_t ic = (int64_t)i * st.lambda + c;
    const bool done = (st.flags[i] & CMA_DONE) != 0;
    st.cand_active[ic] = done ? 0 : 1;
    if (done) return;
    const int nd = st.ndim[i];
    float l1 = 0.0f;
    float* pp = cand_params + st.cand_off[ic];
    const float sigma = st.sigma[i];
    for (int j = 0; j < nd; ++j) {
        const float z = st.normals ? st.normals[(((int64_t)it * N + i) * st.lambda + c) * kMaxP + j]
                                   : cma_normal(st.seed, (uint32_t)st.prog[i], (uint32_t)it, (uint32_t)c, (uint32_t)j);
        const float step = z * __builtin_sqrtf(st.cdiag[j * N + i]);
        const float val = st.mean[j * N + i] + sigma * step;
        st.steps[ic * kMaxP + j] = step;
        st.pop[ic * kMaxP + j] = val;
        l1 += __builtin_fabsf(val);
        if (j < st.nparam[i]) pp[j] = val;  // unflatten_params of the 32-slot buffer
    }
    st.l1[ic] = l1;
}

__global__ void k_cmaes_update(CmaDev st, const float* __restrict__ cand_mse, int* __restrict__ n_working) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= st.n) return;
    if (st.flags[i] & CMA_DONE) return;
    const int N = st.n, L = st.lambda, mu = st.mu;
    const int nd = st.ndim[i];
    const int64_t base = (int64_t)i * L;
    float fit[kMaxLambda];
    int idx[kMaxLambda];
    for (int c = 0; c < L; ++c) {
        fit[c] = cand_mse[base + c] + 0.01f * st.l1[base + c];
        idx[c] = c;
    }
    // ascending fitness; ties by candidate index (deterministic)
    for (int a = 1; a < L; ++a) {
        const int k = idx[a];
        int b = a - 1;
        while (b >= 0 && (fit[idx[b]] > fit[k] || (fit[idx[b]] == fit[k] && idx[b] > k))) {
            idx[b + 1] = idx[b];
            --b;
        }
        idx[b + 1] = k;
    }
    const int bi = idx[0];
    if (fit[bi] < st.best_fit[i]) {
        st.best_fit[i] = fit[bi];
        st.best_mse[i] = cand_mse[base + bi];
        for (int j = 0; j < nd; ++j) st.best[j * N + i] = st.pop[(base + bi) * kMaxP + j];
    }
    if (st.best_mse[i] < 1e-8f) {
        st.flags[i] |= CMA_DONE;
        return;
    }
    float sm[kMaxP], old[kMaxP];
    for (int j = 0; j < nd; ++j) {
        sm[j] = 0.0f;
        old[j] = st.mean[j * N + i];
    }
    for (int r = 0; r < mu; ++r) {
        const int k = idx[r];
        const float w = st.weights[r];
        for (int j = 0; j < nd; ++j) {
            sm[j] += w * st.steps[(base + k) * kMaxP + j];
            st.mean[j * N + i] += w * (st.pop[(base + k) * kMaxP + j] - old[j]);
        }
    }
    const float cc = st.c_c[i], ccov = st.c_cov[i], me = st.mu_eff;
    const float sq = __builtin_sqrtf(cc * (2.0f - cc) * me);
    for (int j = 0; j < nd; ++j) {
        const float pc = (1.0f - cc) * st.pc[j * N + i] + sq * sm[j];
        st.pc[j * N + i] = pc;
        float cu = 0.0f;
        for (int r = 0; r < mu; ++r) {
            const float sv = st.steps[(base + idx[r]) * kMaxP + j];
            cu += st.weights[r] * (sv * sv);
        }
        float cd = (1.0f - ccov) * st.cdiag[j * N + i] + (ccov / me) * (pc * pc) + ccov * (1.0f - 1.0f / me) * cu;
        cd = cd < 1e-6f ? 1e-6f : (cd > 1e6f ? 1e6f : cd);
        st.cdiag[j * N + i] = cd;
    }
    st.sigma[i] *= st.sigma_decay;
    st.iters[i] += 1;
    atomicAdd(n_working, 1);
}

__global__ void k_cmaes_finish(CmaDev st, float* __restrict__ params, float* __restrict__ fitness,
                               int* __restrict__ iters_out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= st.n) return;
    const int np = st.nparam[i];
    float* pp = params + st.param_off[i];
    for (int j = 0; j < np; ++j) pp[j] = st.best[j * st.n + i];
    fitness[i] = st.ndim[i] > 0 ? st.best_mse[i] : __builtin_nanf("");
    if (iters_out) iters_out[i] = st.iters[i] + ((st.flags[i] & CMA_DONE) ? 1 : 0);
}

hipError_t cmaes_begin(const CmaDev& st, const float* d_mse, hipStream_t s) {
    hipLaunchKernelGGL(k_cmaes_begin, dim3((unsigned)((st.n + 127) / 128)), dim3(128), 0, s, st, d_mse);
    return hipGetLastError();
}

hipError_t cmaes_sample(const CmaDev& st, float* cand_params, int it, hipStream_t s) {
    const int64_t t = (int64_t)st.n * st.lambda;
    hipLaunchKernelGGL(k_cmaes_sample, dim3((unsigned)((t + 127) / 128)), dim3(128), 0, s, st, cand_params, it);
    return hipGetLastError();
}

hipError_t cmaes_update(const CmaDev& st, const float* cand_mse, int* counter, hipStream_t s) {
    hipLaunchKernelGGL(k_cmaes_update, RSR_BLK(st.n), 0, s, st, cand_mse, counter);
    return hipGetLastError();
}

hipError_t cmaes_finish(const CmaDev& st, float* params, float* fitness, int* iters_out, hipStream_t s) {
    hipLaunchKernelGGL(k_cmaes_finish, RSR_BLK(st.n), 0, s, st, params, fitness, iters_out);
    return hipGetLastError();
}

float cma_normal_host(uint64_t seed, uint32_t ind, uint32_t it, uint32_t cand, uint32_t j) {
    return cma_normal(seed, ind, it, cand, j);
}

// ------------------------------------------------------------------------------------
// Nelder-Mead (nelder_mead.rs:9-165), one thread per individual.  Fitness of a point =
// mse + 0.01 * sum|v| (the L1 summed in dimension order); alpha 1, gamma 2, rho 0.5,
// sigma 0.5; stop when |worst - best| < 1e-7 or best mse < 1e-8.  The simplex is sorted
// by a stable insertion sort of record indices (the reference's sort_unstable_by leaves
// the order of equal fitnesses unspecified).
// ------------------------------------------------------------------------------------
namespace {

struct NmView {
    const NmDev& st;
    int i;
    __device__ float& fit(int r) const { return st.fit[(int64_t)r * st.n + i]; }
    __device__ float& mse(int r) const { return st.mse[(int64_t)r * st.n + i]; }
    __device__ float& pt(int r, int j) const { return st.pt[((int64_t)r * kMaxP + j) * st.n + i]; }
    __device__ float& cen(int j) const { return st.cen[(int64_t)j * st.n + i]; }
    __device__ float& refl(int j) const { return st.refl[(int64_t)j * st.n + i]; }
    __device__ float& trial(int j) const { return st.trial[(int64_t)j * st.n + i]; }
    __device__ int& idx(int s) const { return st.idx[(int64_t)s * st.n + i]; }
    __device__ int off(int j) const { return st.opt_off[(int64_t)j * st.n + i]; }
};

__device__ float nm_fitness(float m, const float* v, int64_t stride, int nd) {
    float l1 = 0.0f;
    for (int j = 0; j < nd; ++j) l1 += __builtin_fabsf(v[(int64_t)j * stride]);
    return m + 0.01f * l1;
}

__device__ void nm_sort(const NmView& V, int nd) {
    for (int a = 1; a <= nd; ++a) {
        const int k = V.idx(a);
        const float kf = V.fit(k);
        int b = a - 1;
        while (b >= 0 && V.fit(V.idx(b)) > kf) {
            V.idx(b + 1) = V.idx(b);
            --b;
        }
        V.idx(b + 1) = k;
    }
}

}  // namespace

__global__ void k_nm_begin(NmDev st, const float* __restrict__ params, float* __restrict__ cand_params) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= st.n) return;
    const NmView V{st, i};
    const int nd = st.ndim[i];
    const float* pp = params + st.param_off[i];
    for (int r = 0; r < kNmSlots; ++r) {
        const int64_t c = (int64_t)i * kNmSlots + r;
        st.cand_active[c] = r <= nd ? 1 : 0;
        if (r > nd) continue;
        float* cp = cand_params + st.cand_off[c];
        for (int j = 0; j < nd; ++j) {
            float v = pp[V.off(j)];
            if (r == j + 1) v += __builtin_fabsf(v) < 1e-4f ? 0.01f : v * 0.10f;
            V.pt(r, j) = v;
            cp[V.off(j)] = v;
        }
        V.idx(r) = r;
    }
    st.phase[i] = NM_INIT;
    st.iter[i] = 0;
}

__global__ void k_nm_step(NmDev st, float* __restrict__ cand_params, const float* __restrict__ cand_mse,
                          int* __restrict__ n_working) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= st.n) return;
    const NmView V{st, i};
    const int nd = st.ndim[i];
    const int64_t c0 = (int64_t)i * kNmSlots;
    int ph = st.phase[i];
    for (int r = 0; r < kNmSlots; ++r) st.cand_active[c0 + r] = 0;
    if (ph == NM_DONE) return;
    const int64_t S = st.n;  // field stride
    bool iterate = false;
    switch (ph) {
        case NM_INIT:
            for (int r = 0; r <= nd; ++r) {
                const float m = cand_mse[c0 + r];
                V.mse(r) = m;
                V.fit(r) = nm_fitness(m, &V.pt(r, 0), S, nd);
            }
            iterate = true;
            break;
        case NM_REFLECT: {
            const float m = cand_mse[c0];
            const float rfv = nm_fitness(m, &V.refl(0), S, nd);
            const float bf = V.fit(V.idx(0)), swf = V.fit(V.idx(nd - 1)), wf = V.fit(V.idx(nd));
            const int w = V.idx(nd);
            if (rfv >= bf && rfv < swf) {
                V.fit(w) = rfv;
                V.mse(w) = m;
                for (int j = 0; j < nd; ++j) V.pt(w, j) = V.refl(j);
                st.iter[i] += 1;
                iterate = true;
                break;
            }
            st.rf[i] = rfv;
            st.rm[i] = m;
            float* cp = cand_params + st.cand_off[c0];
            if (rfv < bf) {
                for (int j = 0; j < nd; ++j) {
                    const float e = V.cen(j) + 2.0f * (V.refl(j) - V.cen(j));
                    V.trial(j) = e;
                    cp[V.off(j)] = e;
                }
                st.phase[i] = NM_EXPAND;
            } else {
                const bool inside = rfv < wf;
                for (int j = 0; j < nd; ++j) {
                    const float src = inside ? V.refl(j) : V.pt(w, j);
                    const float cv = V.cen(j) + 0.5f * (src - V.cen(j));
                    V.trial(j) = cv;
                    cp[V.off(j)] = cv;
                }
                st.lim[i] = inside ? rfv : wf;
                st.phase[i] = NM_CONTRACT;
            }
            st.cand_active[c0] = 1;
            atomicAdd(n_working, 1);
            return;
        }
        case NM_EXPAND: {
            const float m = cand_mse[c0];
            const float ef = nm_fitness(m, &V.trial(0), S, nd);
            const int w = V.idx(nd);
            const bool take = ef < st.rf[i];
            V.fit(w) = take ? ef : st.rf[i];
            V.mse(w) = take ? m : st.rm[i];
            for (int j = 0; j < nd; ++j) V.pt(w, j) = take ? V.trial(j) : V.refl(j);
            st.iter[i] += 1;
            iterate = true;
            break;
        }
        case NM_CONTRACT: {
            const float m = cand_mse[c0];
            const float cf = nm_fitness(m, &V.trial(0), S, nd);
            if (cf < st.lim[i]) {
                const int w = V.idx(nd);
                V.fit(w) = cf;
                V.mse(w) = m;
                for (int j = 0; j < nd; ++j) V.pt(w, j) = V.trial(j);
                st.iter[i] += 1;
                iterate = true;
                break;
            }
            // shrink every record towards the best one
            const int b = V.idx(0);
            for (int s = 1; s <= nd; ++s) {
                const int r = V.idx(s);
                float* cp = cand_params + st.cand_off[c0 + s];
                for (int j = 0; j < nd; ++j) {
                    const float bp = V.pt(b, j);
                    const float v = bp + 0.5f * (V.pt(r, j) - bp);
                    V.pt(r, j) = v;
                    cp[V.off(j)] = v;
                }
                st.cand_active[c0 + s] = 1;
            }
            st.phase[i] = NM_SHRINK;
            atomicAdd(n_working, 1);
            return;
        }
        case NM_SHRINK:
            for (int s = 1; s <= nd; ++s) {
                const int r = V.idx(s);
                const float m = cand_mse[c0 + s];
                V.mse(r) = m;
                V.fit(r) = nm_fitness(m, &V.pt(r, 0), S, nd);
            }
            st.iter[i] += 1;
            iterate = true;
            break;
        default: break;
    }
    if (!iterate) return;
    // the top of the reference's iteration loop
    if (st.iter[i] >= st.max_iter) {
        st.phase[i] = NM_DONE;
        return;
    }
    nm_sort(V, nd);
    const float bf = V.fit(V.idx(0)), bm = V.mse(V.idx(0)), wf = V.fit(V.idx(nd));
    if (__builtin_fabsf(wf - bf) < 1e-7f || bm < 1e-8f) {
        st.phase[i] = NM_DONE;
        return;
    }
    for (int j = 0; j < nd; ++j) V.cen(j) = 0.0f;
    for (int s = 0; s < nd; ++s) {
        const int r = V.idx(s);
        for (int j = 0; j < nd; ++j) V.cen(j) += V.pt(r, j);
    }
    for (int j = 0; j < nd; ++j) V.cen(j) /= (float)nd;
    const int w = V.idx(nd);
    float* cp = cand_params + st.cand_off[c0];
    for (int j = 0; j < nd; ++j) {
        const float v = V.cen(j) + 1.0f * (V.cen(j) - V.pt(w, j));
        V.refl(j) = v;
        cp[V.off(j)] = v;
    }
    st.cand_active[c0] = 1;
    st.phase[i] = NM_REFLECT;
    atomicAdd(n_working, 1);
}

__global__ void k_nm_finish(NmDev st, float* __restrict__ params, float* __restrict__ fitness,
                            int* __restrict__ iters_out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= st.n) return;
    const NmView V{st, i};
    const int nd = st.ndim[i];
    nm_sort(V, nd);
    const int b = V.idx(0);
    float* pp = params + st.param_off[i];
    for (int j = 0; j < nd; ++j) pp[V.off(j)] = V.pt(b, j);
    fitness[i] = V.mse(b);
    if (iters_out) iters_out[i] = st.iter[i];
}

hipError_t nm_begin(const NmDev& st, const float* params, float* cand_params, hipStream_t s) {
    hipLaunchKernelGGL(k_nm_begin, RSR_BLK(st.n), 0, s, st, params, cand_params);
    return hipGetLastError();
}

hipError_t nm_step(const NmDev& st, float* cand_params, const float* cand_mse, int* counter, hipStream_t s) {
    hipLaunchKernelGGL(k_nm_step, RSR_BLK(st.n), 0, s, st, cand_params, cand_mse, counter);
    return hipGetLastError();
}

hipError_t nm_finish(const NmDev& st, float* params, float* fitness, int* iters_out, hipStream_t s) {
    hipLaunchKernelGGL(k_nm_finish, RSR_BLK(st.n), 0, s, st, params, fitness, iters_out);
    return hipGetLastError();
}

}  // namespace rsr
