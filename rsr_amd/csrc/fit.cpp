// fit.cpp -- C ABI of the batched constant fitters (L-BFGS, CMA-ES) and the
// optimize_constants dispatch.
#include <hip/hip_runtime.h>
#include <math.h>
#include <string.h>

#include <algorithm>
#include <memory>
#include <numeric>
#include <vector>

#include "../../include/rsr_amd.h"
#include "kernels.h"
#include "optim.h"
#include "runtime_internal.h"

using namespace rsr;

#define HIP_TRY(expr, what)                               \
    do {                                                  \
        hipError_t e_ = (expr);                           \
        if (e_ != hipSuccess) return hip_error(e_, what); \
    } while (0)

namespace {

// Device buffers of one fit call (released on return).
struct Scratch {
    rsr_context* ctx;
    std::vector<std::unique_ptr<DevBuf>> bufs;
    explicit Scratch(rsr_context* c) : ctx(c) {}
    template <typename T>
    int get(size_t count, T** out) {
        bufs.emplace_back(new DevBuf());
        int rc = bufs.back()->alloc(ctx, std::max<size_t>(count, 1) * sizeof(T));
        if (rc) return rc;
        HIP_TRY(hipMemsetAsync(bufs.back()->ptr, 0, std::max<size_t>(count, 1) * sizeof(T), ctx->stream), "memset");
        *out = (T*)bufs.back()->ptr;
        return RSR_OK;
    }
    template <typename T>
    int put(const std::vector<T>& v, T** out) {
        int rc = get<T>(v.size(), out);
        if (rc) return rc;
        if (!v.empty())
            HIP_TRY(hipMemcpyAsync(*out, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, ctx->stream), "upload");
        return RSR_OK;
    }
};

// Copy the device parameters back into the host constants (unflatten, program.rs:147-192).
int download_params(rsr_programs* p) {
    rsr_context* ctx = p->ctx;
    HIP_TRY(hipMemcpyAsync(p->params.data(), p->dparams.ptr, p->params.size() * sizeof(float), hipMemcpyDeviceToHost,
                           ctx->stream),
            "params download");
    HIP_TRY(hipStreamSynchronize(ctx->stream), "params download");
    for (int i = 0; i < p->n; ++i) {
        int64_t v = p->value_off[i];
        uint32_t f = p->param_off[i];
        for (int64_t c = p->const_off[i]; c < p->const_off[i + 1]; ++c) {
            const int s = vt_size(p->ctypes[c]);
            for (int k = 0; k < s; ++k) p->values[v + k] = p->params[f + k];
            v += s ? s : 1;
            f += s;
        }
    }
    return RSR_OK;
}

// MSE + gradient of the programs flagged in d_active (device mask), into d_mse / d_grad.
int grad_masked(rsr_context* ctx, rsr_programs* p, const rsr_dataset* ds, const uint8_t* d_active, float* d_mse,
                float* d_grad) {
    Lowering* Ld;
    int rc = get_lowering(p, ds->num_features, true, &Ld);
    if (rc) return rc;
    const DevDataset D = dev_view(ds);
    const DevPrograms DP = dev_view(p, Ld);
    if (eval_grad_lds_bytes(DP) > ctx->max_lds) return set_error(RSR_E_UNSUPPORTED, "dual spill area too large");
    HIP_TRY(hipMemsetAsync(d_grad, 0, (size_t)p->n * 32 * sizeof(float), ctx->stream), "grad clear");
    if (ctx->summation == RSR_SUM_REFERENCE) {
        HIP_TRY(launch_eval_grad_exact(DP, D, (const uint8_t*)Ld->dnparams.ptr, d_active, Ld->max_params, d_mse,
                                       d_grad, ctx->stream, ctx->next_timer()),
                "grad launch");
        return RSR_OK;
    }
    const size_t nparts = (size_t)eval_tiles(D, 1) * p->n * 33;
    if ((rc = ensure_scratch(ctx, ctx->scratch_partials, nparts * sizeof(float)))) return rc;
    HIP_TRY(launch_eval_grad(DP, D, (const uint8_t*)Ld->dnparams.ptr, d_active, Ld->max_params,
                             (float*)ctx->scratch_partials.ptr, d_mse, d_grad, ctx->stream, ctx->next_timer()),
            "grad launch");
    return RSR_OK;
}

int read_counter(rsr_context* ctx, int* d_counter, int* value) {
    HIP_TRY(hipMemcpyAsync(value, d_counter, sizeof(int), hipMemcpyDeviceToHost, ctx->stream), "counter");
    HIP_TRY(hipStreamSynchronize(ctx->stream), "counter");
    return RSR_OK;
}

}  // namespace

// run_lbfgs (src/engine/optimize/lbfgs.rs:9-165) for every program with mask[i] != 0.
int rsr::fit_lbfgs_impl(rsr_context* ctx, rsr_programs* p, const rsr_dataset* ds, int max_iterations, int history,
                        int mode, const uint8_t* mask, float* fitness_out) {
    if (history < 1 || history > 8 || max_iterations < 0)
        return set_error(RSR_E_INVALID_ARG, "history must be 1..8, iterations >= 0");
    const bool fp64 = mode == RSR_LBFGS_FP64;
    Lowering *Lp, *Ld;
    int rc = get_lowering(p, ds->num_features, false, &Lp);
    if (rc) return rc;
    if ((rc = get_lowering(p, ds->num_features, true, &Ld))) return rc;
    std::vector<int> prog;
    std::vector<int> nparam;
    std::vector<uint32_t> poff;
    for (int i = 0; i < p->n; ++i) {
        if (mask && !mask[i]) continue;
        if (Ld->status[i] != PROG_OK || Lp->status[i] != PROG_OK) continue;
        prog.push_back(i);
        nparam.push_back(Ld->nparams[i]);
        poff.push_back(p->param_off[i]);
    }
    for (int i = 0; i < p->n; ++i) fitness_out[i] = NAN;
    const int n = (int)prog.size();
    if (n == 0) return RSR_OK;
    Scratch sc(ctx);
    LbfgsDev st;
    st.n = n;
    st.M = history;
    st.fp64 = fp64 ? 1 : 0;
    const size_t tb = fp64 ? sizeof(double) : sizeof(float);
    int *d_prog, *d_np;
    uint32_t* d_poff;
    if ((rc = sc.put(prog, &d_prog)) || (rc = sc.put(nparam, &d_np)) || (rc = sc.put(poff, &d_poff))) return rc;
    st.prog = d_prog;
    st.nparam = d_np;
    st.param_off = d_poff;
    if ((rc = sc.get((size_t)kMaxP * n, &st.x)) || (rc = sc.get((size_t)kMaxP * n, &st.g)) ||
        (rc = sc.get((size_t)kMaxP * n, &st.p)) || (rc = sc.get((size_t)n, &st.step)) ||
        (rc = sc.get((size_t)n, &st.cur_mse)) || (rc = sc.get((size_t)n, &st.next_mse)))
        return rc;
    char *S, *Y, *rho, *alpha, *ddg;
    if ((rc = sc.get((size_t)8 * kMaxP * n * tb, &S)) || (rc = sc.get((size_t)8 * kMaxP * n * tb, &Y)) ||
        (rc = sc.get((size_t)8 * n * tb, &rho)) || (rc = sc.get((size_t)8 * n * tb, &alpha)) ||
        (rc = sc.get((size_t)n * tb, &ddg)))
        return rc;
    st.S = S;
    st.Y = Y;
    st.rho = rho;
    st.alpha = alpha;
    st.ddg = ddg;
    if ((rc = sc.get((size_t)n, &st.hist)) || (rc = sc.get((size_t)n, &st.head)) || (rc = sc.get((size_t)n, &st.ls)) ||
        (rc = sc.get((size_t)n, &st.flags)))
        return rc;
    uint8_t *act_fit, *act_ls, *act_grad;
    float *d_mse, *d_grad, *d_fit;
    int* d_counter;
    std::vector<uint8_t> fitmask(p->n, 0);
    for (int i : prog) fitmask[i] = 1;
    if ((rc = sc.put(fitmask, &act_fit)) || (rc = sc.get((size_t)p->n, &act_ls)) ||
        (rc = sc.get((size_t)p->n, &act_grad)) || (rc = sc.get((size_t)p->n, &d_mse)) ||
        (rc = sc.get((size_t)p->n * 32, &d_grad)) || (rc = sc.get((size_t)n, &d_fit)) ||
        (rc = sc.get((size_t)1, &d_counter)))
        return rc;
    float* params = (float*)p->dparams.ptr;

    // (current_mse, current_grad) = compute_mse_with_gradient (lbfgs.rs:29-30)
    if ((rc = grad_masked(ctx, p, ds, act_fit, d_mse, d_grad))) return rc;
    HIP_TRY(lbfgs_begin(st, params, d_mse, d_grad, ctx->stream), "lbfgs_begin");
    for (int it = 0; it < max_iterations; ++it) {
        HIP_TRY(hipMemsetAsync(d_counter, 0, sizeof(int), ctx->stream), "counter");
        HIP_TRY(lbfgs_direction(st, params, act_ls, d_counter, ctx->stream), "lbfgs_direction");
        int working = 0;
        if ((rc = read_counter(ctx, d_counter, &working))) return rc;
        if (working == 0) break;
        for (int ls = 0; ls < 12; ++ls) {  // Armijo backtracking, at most 12 evaluations
            if ((rc = eval_mse_device(ctx, p, ds, act_ls, d_mse, nullptr))) return rc;
            HIP_TRY(hipMemsetAsync(d_counter, 0, sizeof(int), ctx->stream), "counter");
            HIP_TRY(lbfgs_ls_check(st, params, d_mse, act_ls, act_grad, d_counter, ctx->stream), "ls_check");
            int searching = 0;
            if ((rc = read_counter(ctx, d_counter, &searching))) return rc;
            if (searching == 0) break;
        }
        if ((rc = grad_masked(ctx, p, ds, act_grad, d_mse, d_grad))) return rc;
        HIP_TRY(lbfgs_update(st, d_grad, act_grad, ctx->stream), "lbfgs_update");
    }
    HIP_TRY(lbfgs_finish(st, params, d_fit, ctx->stream), "lbfgs_finish");
    std::vector<float> fit(n);
    HIP_TRY(hipMemcpyAsync(fit.data(), d_fit, (size_t)n * sizeof(float), hipMemcpyDeviceToHost, ctx->stream),
            "fitness");
    if ((rc = download_params(p))) return rc;
    for (int k = 0; k < n; ++k) fitness_out[prog[k]] = fit[k];
    return RSR_OK;
}

extern "C" int rsr_fit_lbfgs(rsr_context* ctx, rsr_programs* p, const rsr_dataset* ds, int max_iterations,
                             int history, int mode, const uint8_t* mask, float* fitness_out) {
    if (!ctx || !p || !ds || (p->n && !fitness_out)) return set_error(RSR_E_INVALID_ARG, "null argument");
    if (mode != RSR_LBFGS_REFERENCE && mode != RSR_LBFGS_FP64) return set_error(RSR_E_INVALID_ARG, "bad mode");
    std::lock_guard<std::mutex> g(ctx->mu);
    HIP_TRY(hipSetDevice(ctx->device), "hipSetDevice");
    return fit_lbfgs_impl(ctx, p, ds, max_iterations, history, mode, mask, fitness_out);
}

// run_cma_es (src/engine/optimize/cmaes.rs:19-192) for every program with mask[i] != 0.
int rsr::fit_cmaes_impl(rsr_context* ctx, rsr_programs* p, const rsr_dataset* ds, int max_iterations, int lambda,
                        int mu, uint64_t seed, const float* normals, const uint8_t* mask, float* fitness_out,
                        int32_t* iters_out) {
    if (lambda < 2 || lambda > kMaxLambda || mu < 1 || mu > lambda || max_iterations < 0)
        return set_error(RSR_E_INVALID_ARG, "lambda must be 2..%d and 1 <= mu <= lambda", kMaxLambda);
    Lowering* Lp;
    int rc = get_lowering(p, ds->num_features, false, &Lp);
    if (rc) return rc;
    std::vector<int> prog, nparam, ndim;
    std::vector<uint32_t> poff;
    for (int i = 0; i < p->n; ++i) {
        fitness_out[i] = NAN;
        if (iters_out) iters_out[i] = 0;
        if (mask && !mask[i]) continue;
        if (Lp->status[i] != PROG_OK) continue;
        const int pc = (int)(p->param_off[i + 1] - p->param_off[i]);
        if (pc == 0) continue;  // cmaes.rs:25-28
        prog.push_back(i);
        nparam.push_back(std::min(pc, kMaxP));
        ndim.push_back(std::min(std::min(pc, kMaxP) + pc, kMaxP));
        poff.push_back(p->param_off[i]);
    }
    const int n = (int)prog.size();
    if (n == 0) return RSR_OK;
    // recombination weights and strategy constants in f32 as cmaes.rs:85-102
    std::vector<float> w(mu);
    float sumw = 0.0f;
    const float lmu = logf((float)mu + 0.5f);
    for (int i = 0; i < mu; ++i) {
        const float v = lmu - logf((float)(i + 1));
        w[i] = v > 0.0f ? v : 0.0f;
        sumw += w[i];
    }
    float me = 0.0f;
    for (int i = 0; i < mu; ++i) {
        w[i] = w[i] / sumw;
        me += w[i] * w[i];
    }
    me = 1.0f / me;
    std::vector<float> cc(n), ccov(n), mean((size_t)kMaxP * n, 0.0f), cdiag((size_t)kMaxP * n, 1.0f),
        best((size_t)kMaxP * n, 0.0f), sigma(n, 0.5f), bfit(n), bmse(n);
    std::vector<int64_t> cand_off((size_t)n * lambda);
    std::vector<float> cand_params;
    for (int k = 0; k < n; ++k) {
        const int i = prog[k];
        const float nf = (float)ndim[k];
        cc[k] = (4.0f + me / nf) / (nf + 4.0f + 2.0f * me / nf);
        const float t1 = nf + 1.414f, t2 = nf + 2.0f;
        ccov[k] = (1.0f / me) * (2.0f / (t1 * t1)) + (1.0f - 1.0f / me) * ((2.0f * me - 1.0f) / (t2 * t2 + me));
        const float* pp = &p->params[p->param_off[i]];
        const int pc = (int)(p->param_off[i + 1] - p->param_off[i]);
        for (int j = 0; j < nparam[k]; ++j) mean[(size_t)j * n + k] = pp[j];
        for (int j = nparam[k]; j < ndim[k]; ++j) mean[(size_t)j * n + k] = pp[j - nparam[k]];
        for (int j = 0; j < ndim[k]; ++j) best[(size_t)j * n + k] = mean[(size_t)j * n + k];
        for (int c = 0; c < lambda; ++c) {
            cand_off[(size_t)k * lambda + c] = (int64_t)cand_params.size();
            cand_params.insert(cand_params.end(), pp, pp + pc);
        }
    }
    cand_params.push_back(0.0f);
    Scratch sc(ctx);
    // initial MSE of every individual (cmaes.rs:104) and the L1 of the doubled mean
    std::vector<uint8_t> fitmask(p->n, 0);
    for (int i : prog) fitmask[i] = 1;
    uint8_t* act_fit;
    float* d_mse0;
    if ((rc = sc.put(fitmask, &act_fit)) || (rc = sc.get((size_t)p->n, &d_mse0))) return rc;
    if ((rc = eval_mse_device(ctx, p, ds, act_fit, d_mse0, nullptr))) return rc;
    std::vector<float> mse0(p->n);
    HIP_TRY(hipMemcpyAsync(mse0.data(), d_mse0, (size_t)p->n * sizeof(float), hipMemcpyDeviceToHost, ctx->stream),
            "mse0");
    HIP_TRY(hipStreamSynchronize(ctx->stream), "mse0");
    for (int k = 0; k < n; ++k) {
        float l1 = 0.0f;
        for (int j = 0; j < ndim[k]; ++j) l1 += fabsf(mean[(size_t)j * n + k]);
        bmse[k] = mse0[prog[k]];
        bfit[k] = bmse[k] + 0.01f * l1;
    }
    CmaDev st;
    st.n = n;
    st.lambda = lambda;
    st.mu = mu;
    st.mu_eff = me;
    st.sigma_decay = expf(-0.01f);  // (-0.01_f32).exp() (cmaes.rs:178)
    st.seed = seed;
    float *d_w, *d_normals = nullptr, *d_cc, *d_ccov;
    int *d_prog, *d_np, *d_nd;
    uint32_t* d_poff;
    int64_t* d_coff;
    if ((rc = sc.put(w, &d_w)) || (rc = sc.put(prog, &d_prog)) || (rc = sc.put(nparam, &d_np)) ||
        (rc = sc.put(ndim, &d_nd)) || (rc = sc.put(poff, &d_poff)) || (rc = sc.put(cand_off, &d_coff)) ||
        (rc = sc.put(cc, &d_cc)) || (rc = sc.put(ccov, &d_ccov)))
        return rc;
    if (normals) {
        std::vector<float> nv(normals, normals + (size_t)max_iterations * n * lambda * kMaxP);
        if ((rc = sc.put(nv, &d_normals))) return rc;
    }
    st.weights = d_w;
    st.normals = d_normals;
    st.prog = d_prog;
    st.nparam = d_np;
    st.ndim = d_nd;
    st.param_off = d_poff;
    st.cand_off = d_coff;
    st.c_c = d_cc;
    st.c_cov = d_ccov;
    if ((rc = sc.put(mean, &st.mean)) || (rc = sc.put(cdiag, &st.cdiag)) || (rc = sc.get((size_t)kMaxP * n, &st.pc)) ||
        (rc = sc.put(best, &st.best)) || (rc = sc.put(sigma, &st.sigma)) || (rc = sc.put(bmse, &st.best_mse)) ||
        (rc = sc.put(bfit, &st.best_fit)) || (rc = sc.get((size_t)n * lambda * kMaxP, &st.steps)) ||
        (rc = sc.get((size_t)n * lambda * kMaxP, &st.pop)) || (rc = sc.get((size_t)n * lambda, &st.l1)) ||
        (rc = sc.get((size_t)n * lambda, &st.cand_active)) || (rc = sc.get((size_t)n, &st.flags)) ||
        (rc = sc.get((size_t)n, &st.iters)))
        return rc;
    // candidate batch: every individual's program repeated lambda times, own parameter blocks
    float* d_cparams;
    if ((rc = sc.put(cand_params, &d_cparams))) return rc;
    std::vector<ProgDesc> hdesc((size_t)n * lambda);
    std::vector<ProgDesc> base_desc(p->n);
    HIP_TRY(hipMemcpyAsync(base_desc.data(), Lp->ddesc.ptr, (size_t)p->n * sizeof(ProgDesc), hipMemcpyDeviceToHost,
                           ctx->stream),
            "desc download");
    HIP_TRY(hipStreamSynchronize(ctx->stream), "desc download");
    std::vector<int32_t> corder((size_t)n * lambda);
    for (int k = 0; k < n; ++k)
        for (int c = 0; c < lambda; ++c) {
            ProgDesc d = base_desc[prog[k]];
            d.param_off = (uint32_t)cand_off[(size_t)k * lambda + c];
            hdesc[(size_t)k * lambda + c] = d;
        }
    // schedule candidates longest-first like the parent batch
    std::vector<int> rank(p->n);
    for (int r = 0; r < p->n; ++r) rank[Lp->order[r]] = r;
    std::iota(corder.begin(), corder.end(), 0);
    std::stable_sort(corder.begin(), corder.end(),
                     [&](int a, int b) { return rank[prog[a / lambda]] < rank[prog[b / lambda]]; });
    ProgDesc* d_desc;
    int32_t* d_order;
    float *d_cmse, *d_fit;
    int *d_counter, *d_iters;
    if ((rc = sc.put(hdesc, &d_desc)) || (rc = sc.put(corder, &d_order)) ||
        (rc = sc.get((size_t)n * lambda, &d_cmse)) || (rc = sc.get((size_t)n, &d_fit)) ||
        (rc = sc.get((size_t)1, &d_counter)) || (rc = sc.get((size_t)n, &d_iters)))
        return rc;
    DevPrograms CP;
    CP.code = (const UInstr*)Lp->dcode.ptr;
    CP.progs = d_desc;
    CP.params = d_cparams;
    CP.order = d_order;
    CP.n_progs = n * lambda;
    CP.spill_slots = Lp->spill_slots;
    const DevDataset D = dev_view(ds);
    EvalTuning t = ctx->tuning;
    while (t.ppl > 1 && eval_mse_lds_bytes(CP, t.ppl) > ctx->max_lds) t.ppl /= 2;
    const size_t nparts = (size_t)eval_tiles(D, t.ppl) * CP.n_progs;
    if (ctx->summation != RSR_SUM_REFERENCE &&
        (rc = ensure_scratch(ctx, ctx->scratch_partials, nparts * sizeof(float))))
        return rc;
    for (int it = 0; it < max_iterations; ++it) {
        HIP_TRY(cmaes_sample(st, d_cparams, it, ctx->stream), "cmaes_sample");
        if (ctx->summation == RSR_SUM_REFERENCE)
            HIP_TRY(launch_eval_mse_exact(CP, D, st.cand_active, d_cmse, ctx->stream, ctx->next_timer()), "cand eval");
        else
            HIP_TRY(launch_eval_mse(CP, D, st.cand_active, t, (float*)ctx->scratch_partials.ptr, d_cmse, ctx->stream,
                                    ctx->next_timer()),
                    "cand eval");
        HIP_TRY(hipMemsetAsync(d_counter, 0, sizeof(int), ctx->stream), "counter");
        HIP_TRY(cmaes_update(st, d_cmse, d_counter, ctx->stream), "cmaes_update");
        if ((it & 3) == 3 || it + 1 == max_iterations) {  // stop early once every individual has converged
            int working = 0;
            if ((rc = read_counter(ctx, d_counter, &working))) return rc;
            if (working == 0) break;
        }
    }
    float* params = (float*)p->dparams.ptr;
    HIP_TRY(cmaes_finish(st, params, d_fit, d_iters, ctx->stream), "cmaes_finish");
    std::vector<float> fit(n);
    std::vector<int> its(n);
    HIP_TRY(hipMemcpyAsync(fit.data(), d_fit, (size_t)n * sizeof(float), hipMemcpyDeviceToHost, ctx->stream), "fit");
    HIP_TRY(hipMemcpyAsync(its.data(), d_iters, (size_t)n * sizeof(int), hipMemcpyDeviceToHost, ctx->stream), "its");
    if ((rc = download_params(p))) return rc;
    for (int k = 0; k < n; ++k) {
        fitness_out[prog[k]] = fit[k];
        if (iters_out) iters_out[prog[k]] = its[k];
    }
    return RSR_OK;
}

extern "C" int rsr_fit_cmaes(rsr_context* ctx, rsr_programs* p, const rsr_dataset* ds, int max_iterations,
                             int lambda, int mu, uint64_t seed, const float* normals, const uint8_t* mask,
                             float* fitness_out, int32_t* iters_out) {
    if (!ctx || !p || !ds || (p->n && !fitness_out)) return set_error(RSR_E_INVALID_ARG, "null argument");
    std::lock_guard<std::mutex> g(ctx->mu);
    HIP_TRY(hipSetDevice(ctx->device), "hipSetDevice");
    return fit_cmaes_impl(ctx, p, ds, max_iterations, lambda, mu, seed, normals, mask, fitness_out, iters_out);
}

extern "C" float rsr_cmaes_normal(uint64_t seed, uint32_t individual, uint32_t iteration, uint32_t candidate,
                                  uint32_t dim) {
    return cma_normal_host(seed, individual, iteration, candidate, dim);
}
