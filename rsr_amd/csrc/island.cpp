// island.cpp -- the evaluation step of the search loop's islands, batched over islands
// (Island::evaluate_buffer, src/engine/search/island.rs:215-261; the islands' rayon loop,
// src/engine/search/engine.rs:52).
//
// The reference walks each island's buffer in order, and whether an individual is
// constant-fitted depends on the island's best fitness so far, which the individuals
// before it may have lowered: a loop-carried dependency.  Here the GPU work of every
// island is done in four batched rounds and the reference's loop is then replayed on the
// host with the results:
//   1. the mini-batch MSE of every buffer individual (the caller scores each island's best,
//      best_mini_mse);
//   2. the random draws, which depend only on those MSEs, decide who is promising and
//      who draws a random fit; the full-dataset MSE of the promising individuals;
//   3. the best fitness only ever decreases along the loop, so an individual can be fitted
//      only if its MSE is finite and below the island's best at the start, or it drew a
//      random fit: that superset is fitted, snapped, simplified (rsr_optimize_constants,
//      rsr_simplify_programs) and rescored on the full dataset in one batch;
//   4. the replay takes a fit only where the reference's loop would have made it.
// Every fit is independent of the others in its batch, so the replay's results are the
// reference loop's.
#include <math.h>
#include <float.h>
#include <stdint.h>

#include <vector>

#include "../../include/rsr_amd.h"
#include "isa.h"
#include "runtime_internal.h"
#include "backend.h"
#include "wire.h"

using namespace rsr;

namespace {

struct Batch {  // an rsr_programs handle destroyed with the scope
    rsr_programs* p = nullptr;
    ~Batch() {
        if (p) rsr_programs_destroy(p);
    }
};

// compute_mse that tolerates rejected programs (their MSE is NaN)
int eval(const Backend& B, const Wire& w, const rsr_dataset* ds, std::vector<float>& out) {
    out.assign((size_t)w.n(), NAN);
    if (w.n() == 0) return RSR_OK;
    Batch b;
    int rc = w.create(B.ctx, &b.p);
    if (rc) return rc;
    rc = B.eval(b.p, ds, out.data());
    return rc == RSR_E_PROGRAM ? RSR_OK : rc;
}

// A request through the server, waited for.
int served(rsr_server* srv, const rsr_request& q) {
    uint64_t ticket = 0;
    int rc = rsr_server_submit(srv, &q, &ticket);
    if (rc) return rc;
    int status = RSR_OK;
    rc = rsr_server_wait(srv, ticket, -1.0, &status);
    (void)rsr_server_release(srv, ticket);
    return rc ? rc : status;
}

}  // namespace

int rsr::Backend::eval(rsr_programs* p, const rsr_dataset* ds, float* out) const {
    if (!srv) return rsr_eval_mse(ctx, p, ds, out);
    rsr_request q{};
    q.kind = RSR_REQ_EVAL;
    q.programs = p;
    q.dataset = ds;
    q.out = out;
    return served(srv, q);
}

int rsr::Backend::optimize(rsr_programs* p, const rsr_dataset* ds, int iterations, uint64_t seed, float threshold,
                           float* mse_out) const {
    if (!srv || threshold != 0.05f)
        return rsr_optimize_constants(ctx, p, ds, iterations, seed, threshold, nullptr, mse_out, nullptr);
    rsr_request q{};
    q.kind = RSR_REQ_OPTIMIZE;
    q.programs = p;
    q.dataset = ds;
    q.iterations = iterations;
    q.seed = seed;
    q.out = mse_out;
    return served(srv, q);
}

int rsr::islands_evaluate(const Backend& B, const rsr_dataset* dataset, const rsr_dataset* mini_batch,
                                    int n_islands, const int64_t* island_offsets, const uint8_t* ops,
                                    const uint16_t* args, const int64_t* code_offsets, const uint8_t* const_types,
                                    const int64_t* const_offsets, const float* const_values,
                                    const int64_t* value_offsets, const float* best_fitness,
                                    const float* best_mini_mse, const float* draws, const int64_t* draw_offsets,
                                    float opt_prob, int opt_iterations, float parsimony, float threshold,
                                    uint64_t seed, int64_t code_capacity, int64_t const_capacity,
                                    int64_t value_capacity, uint8_t* ops_out, uint16_t* args_out,
                                    int64_t* code_offsets_out, uint8_t* const_types_out, int64_t* const_offsets_out,
                                    float* const_values_out, int64_t* value_offsets_out, float* fitness_out,
                                    float* mse_out, uint8_t* optimized_out, int32_t* best_index_out,
                                    int64_t* draws_used_out) {
    rsr_context* ctx = B.ctx;
    if (!ctx || !dataset || !mini_batch || n_islands < 0 || !island_offsets || !code_offsets || !const_offsets ||
        !value_offsets || !best_fitness || !best_mini_mse || !draw_offsets || !code_offsets_out ||
        !const_offsets_out || !value_offsets_out || !best_index_out || !draws_used_out)
        return set_error(RSR_E_INVALID_ARG, "null argument");
    const int64_t n = island_offsets[n_islands];
    if (island_offsets[0] != 0 || n < 0 || (n > 0 && (!fitness_out || !mse_out || !optimized_out)))
        return set_error(RSR_E_INVALID_ARG, "invalid island offsets or null outputs");
    for (int g = 0; g < n_islands; ++g)
        if (island_offsets[g + 1] < island_offsets[g] || draw_offsets[g + 1] < draw_offsets[g])
            return set_error(RSR_E_INVALID_ARG, "island %d: decreasing offsets", g);
    if ((n > 0 && (code_offsets[n] > code_offsets[0] && (!ops || !args))) ||
        (n > 0 && const_offsets[n] > const_offsets[0] && (!const_types || !const_values)) ||
        (draw_offsets[n_islands] > draw_offsets[0] && !draws))
        return set_error(RSR_E_INVALID_ARG, "null program or draw arrays");
    for (int64_t i = 0; i < n; ++i)
        if (code_offsets[i + 1] < code_offsets[i] || const_offsets[i + 1] < const_offsets[i])
            return set_error(RSR_E_INVALID_ARG, "individual %lld: decreasing offsets", (long long)i);
    Wire buf;
    for (int64_t i = 0; i < n; ++i)
        buf.add(ops, args, code_offsets, const_types, const_offsets, const_values, value_offsets, i);

    // 1. mini-batch MSEs
    std::vector<float> mini;
    int rc = eval(B, buf, mini_batch, mini);
    if (rc) return rc;

    // 2. draws (in the reference's order), promising individuals, full MSEs
    std::vector<uint8_t> promising((size_t)n, 0), random_opt((size_t)n, 0);
    for (int g = 0; g < n_islands; ++g) {
        int64_t d = draw_offsets[g];
        auto next = [&](float& u) {
            if (d >= draw_offsets[g + 1]) return false;
            u = draws[d++];
            return true;
        };
        const float baseline = best_mini_mse[g];
        for (int64_t i = island_offsets[g]; i < island_offsets[g + 1]; ++i) {
            bool p = mini[i] < baseline * 2.0f;
            float u;
            if (!p) {
                if (!next(u)) return set_error(RSR_E_INVALID_ARG, "island %d: draw stream exhausted", g);
                p = u < 0.05f;
            }
            if (p) {
                if (!next(u)) return set_error(RSR_E_INVALID_ARG, "island %d: draw stream exhausted", g);
                random_opt[i] = u < opt_prob;
            }
            promising[i] = p;
        }
        draws_used_out[g] = d - draw_offsets[g];
    }
    Wire prom;
    std::vector<int64_t> prom_idx;
    for (int64_t i = 0; i < n; ++i)
        if (promising[i]) {
            prom.add(buf, (int)i);
            prom_idx.push_back(i);
        }
    std::vector<float> full_p;
    if ((rc = eval(B, prom, dataset, full_p))) return rc;
    std::vector<float> full((size_t)n, NAN);
    for (size_t k = 0; k < prom_idx.size(); ++k) full[prom_idx[k]] = full_p[k];

    // 3. fit, snap, simplify and rescore every individual the loop could fit
    Wire cand;
    std::vector<int64_t> cand_idx;
    for (int g = 0; g < n_islands; ++g)
        for (int64_t i = island_offsets[g]; i < island_offsets[g + 1]; ++i)
            if (promising[i] && isfinite(full[i]) && (full[i] < best_fitness[g] || random_opt[i])) {
                cand.add(buf, (int)i);
                cand_idx.push_back(i);
            }
    Wire fitted;  // the candidates with snapped constants, then simplified
    std::vector<float> fitted_mse;
    if (cand.n() > 0) {
        Batch b;
        if ((rc = cand.create(ctx, &b.p))) return rc;
        std::vector<float> unused((size_t)cand.n());
        rc = B.optimize(b.p, dataset, opt_iterations, seed, threshold, unused.data());
        if (rc && rc != RSR_E_PROGRAM) return rc;
        Wire snapped = cand;
        if ((rc = rsr_programs_get_constants(b.p, snapped.cv.data()))) return rc;
        const int64_t cap = snapped.co.back() > 0 ? snapped.co.back() : 1;
        fitted.ops.resize((size_t)cap);
        fitted.args.resize((size_t)cap);
        fitted.ct.resize((size_t)cap);
        fitted.cv.resize((size_t)(9 * cap));
        fitted.co.assign((size_t)cand.n() + 1, 0);
        fitted.cof.assign((size_t)cand.n() + 1, 0);
        fitted.vo.assign((size_t)cand.n() + 1, 0);
        const uint8_t u8 = 0;
        const uint16_t u16 = 0;
        const float f0 = 0.0f;
        rc = rsr_simplify_programs(cand.n(), snapped.ops.empty() ? &u8 : snapped.ops.data(),
                                   snapped.args.empty() ? &u16 : snapped.args.data(), snapped.co.data(),
                                   snapped.ct.empty() ? &u8 : snapped.ct.data(), snapped.cof.data(),
                                   snapped.cv.empty() ? &f0 : snapped.cv.data(), snapped.vo.data(), cap, cap, 9 * cap,
                                   fitted.ops.data(), fitted.args.data(), fitted.co.data(), fitted.ct.data(),
                                   fitted.cof.data(), fitted.cv.data(), fitted.vo.data());
        if (rc) return rc;
        fitted.ops.resize((size_t)fitted.co.back());
        fitted.args.resize((size_t)fitted.co.back());
        fitted.ct.resize((size_t)fitted.cof.back());
        fitted.cv.resize((size_t)fitted.vo.back());
        if ((rc = eval(B, fitted, dataset, fitted_mse))) return rc;
    }
    std::vector<int> cand_of((size_t)n, -1);
    for (size_t k = 0; k < cand_idx.size(); ++k) cand_of[cand_idx[k]] = (int)k;

    // 4. the reference loop, replayed
    Wire out;
    for (int64_t i = 0; i < n; ++i) out.add(buf, (int)i);  // replaced below where a fit is taken
    std::vector<uint8_t> take((size_t)n, 0);
    std::vector<float> final_mse((size_t)n);
    for (int g = 0; g < n_islands; ++g) {
        float best = best_fitness[g];
        int32_t best_i = -1;
        for (int64_t i = island_offsets[g]; i < island_offsets[g + 1]; ++i) {
            float m = mini[i];
            if (promising[i]) {
                m = full[i];
                const bool elite = m < best;
                if ((elite || random_opt[i]) && isfinite(m)) {
                    take[i] = 1;
                    m = fitted_mse[cand_of[i]];
                }
            }
            final_mse[i] = m;
            float fit = FLT_MAX;
            if (isfinite(m)) {
                // complexity of the program the individual now has
                const Wire& src = take[i] ? fitted : buf;
                const int j = take[i] ? cand_of[i] : (int)i;
                const int64_t offs[2] = {0, src.co[j + 1] - src.co[j]};
                const uint8_t none = 0;
                int32_t comp = 0;
                if ((rc = rsr_complexity(1, offs[1] ? src.ops.data() + src.co[j] : &none, offs, &comp))) return rc;
                fit = m + (float)comp * parsimony;  // island.rs:250-251
            }
            fitness_out[i] = fit;
            mse_out[i] = m;
            optimized_out[i] = take[i];
            if (fit < best) {
                best = fit;
                best_i = (int32_t)(i - island_offsets[g]);
            }
        }
        best_index_out[g] = best_i;
    }

    // the buffer after the step: fitted individuals carry their simplified programs
    int64_t nc = 0, nk = 0, nv = 0;
    code_offsets_out[0] = const_offsets_out[0] = value_offsets_out[0] = 0;
    for (int64_t i = 0; i < n; ++i) {
        const Wire& src = take[i] ? fitted : buf;
        const int j = take[i] ? cand_of[i] : (int)i;
        const int64_t c0 = src.co[j], c1 = src.co[j + 1], k0 = src.cof[j], k1 = src.cof[j + 1];
        const int64_t v0 = src.vo[j], v1 = src.vo[j + 1];
        if (nc + (c1 - c0) > code_capacity || nk + (k1 - k0) > const_capacity || nv + (v1 - v0) > value_capacity)
            return set_error(RSR_E_INVALID_ARG, "individual %lld: output capacity exceeded", (long long)i);
        if ((c1 > c0 && (!ops_out || !args_out)) || (k1 > k0 && (!const_types_out || !const_values_out)))
            return set_error(RSR_E_INVALID_ARG, "null output program arrays");
        for (int64_t k = c0; k < c1; ++k, ++nc) {
            ops_out[nc] = src.ops[k];
            args_out[nc] = src.args[k];
        }
        for (int64_t k = k0; k < k1; ++k) const_types_out[nk++] = src.ct[k];
        for (int64_t k = v0; k < v1; ++k) const_values_out[nv++] = src.cv[k];
        code_offsets_out[i + 1] = nc;
        const_offsets_out[i + 1] = nk;
        value_offsets_out[i + 1] = nv;
    }
    return RSR_OK;
}

extern "C" int rsr_islands_evaluate(rsr_context* ctx, const rsr_dataset* dataset, const rsr_dataset* mini_batch,
                                    int n_islands, const int64_t* island_offsets, const uint8_t* ops,
                                    const uint16_t* args, const int64_t* code_offsets, const uint8_t* const_types,
                                    const int64_t* const_offsets, const float* const_values,
                                    const int64_t* value_offsets, const float* best_fitness,
                                    const float* best_mini_mse, const float* draws, const int64_t* draw_offsets,
                                    float opt_prob, int opt_iterations, float parsimony, float threshold,
                                    uint64_t seed, int64_t code_capacity, int64_t const_capacity,
                                    int64_t value_capacity, uint8_t* ops_out, uint16_t* args_out,
                                    int64_t* code_offsets_out, uint8_t* const_types_out, int64_t* const_offsets_out,
                                    float* const_values_out, int64_t* value_offsets_out, float* fitness_out,
                                    float* mse_out, uint8_t* optimized_out, int32_t* best_index_out,
                                    int64_t* draws_used_out) {
    Backend B;
    B.ctx = ctx;
    return islands_evaluate(B, dataset, mini_batch, n_islands, island_offsets, ops, args, code_offsets, const_types, const_offsets, const_values, value_offsets, best_fitness, best_mini_mse, draws, draw_offsets, opt_prob, opt_iterations, parsimony, threshold, seed, code_capacity, const_capacity, value_capacity, ops_out, args_out, code_offsets_out, const_types_out, const_offsets_out, const_values_out, value_offsets_out, fitness_out, mse_out, optimized_out, best_index_out, draws_used_out);
}
