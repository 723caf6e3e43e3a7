// engine.cpp -- the search loop as a client of the evaluator: Island::{new,
// step_generation, fill_next_generation, nuke} (src/engine/search/island.rs:33-213) and
// Engine::{run, migrate_individuals, get_global_best, get_pareto_front}
// (src/engine/search/engine.rs:18-156).
//
// The reference steps its islands on rayon threads, each island calling the CPU
// evaluator once per individual.  Here one generation of every island is a handful of
// merged device calls: the random injections of all islands in one variation launch,
// all crossovers in another, each mutation cycle of all mutation offspring in one launch
// followed by one mini-batch evaluation of all of them, the evaluation step of all
// islands in one rsr_islands_evaluate, and the nukes' generation and scoring batched
// likewise.  The host keeps the populations in the wire format between those calls and
// runs what is per-island bookkeeping: NSGA-II ranks, tournaments, the halls of fame,
// stagnation and migration.
//
// Random streams.  Each island owns a Xoshiro256++ stream seeded with seed + island
// (the reference: 42 + i, island.rs:43, engine.rs:28) that makes the island's sequential
// decisions in the reference's order: crossover or mutation (island.rs:123), the
// tournaments' draws (selection.rs:13-26) and the evaluation step's draws
// (island.rs:220-230).  The variation operators draw from per-offspring streams keyed by
// (seed, island, generation, slot, cycle), so that every offspring of a generation is
// produced in one launch; the reference draws those from the island stream in between.
#include <float.h>
#include <math.h>
#include <stdint.h>
#include <string.h>

#include <map>
#include <memory>
#include <utility>
#include <vector>

#include "../../include/rsr_amd.h"
#include "isa.h"
#include "backend.h"
#include "runtime_internal.h"
#include "variation.h"
#include "variation_host.h"
#include "wire.h"

using namespace rsr;

namespace {

// Individual::has_forbidden_patterns (individual.rs:112-197)
bool forbidden_patterns(const uint8_t* ops, int64_t len) {
    enum : uint16_t { TRIG = 1, EXP = 2, LN = 4, POWER = 8, TRANSPOSE = 16, INVERSE = 32, DET = 64 };
    std::vector<uint16_t> st;
    st.reserve(32);
    for (int64_t i = 0; i < len; ++i) {
        const int op = ops[i];
        if (v_is_load(op)) {
            st.push_back(0);
            continue;
        }
        const int n = v_arity(op);
        if ((int)st.size() < n) return true;
        uint16_t f = 0;
        for (int k = 0; k < n; ++k) {
            f |= st.back();
            st.pop_back();
        }
        uint16_t add = 0, bad = 0;
        switch (op) {
            case SinF: case CosF: bad = TRIG | EXP | LN; add = TRIG; break;
            case ExpF: bad = TRIG | EXP | LN | POWER; add = EXP; break;
            case LnF: bad = TRIG | EXP | LN; add = LN; break;
            case SqrtF: case SqrF: bad = POWER | TRIG | LN | EXP; add = POWER; break;
            case TransposeM2: case TransposeM3: bad = TRANSPOSE; add = TRANSPOSE; break;
            case InverseM2: case InverseM3: bad = INVERSE; add = INVERSE; break;
            case DetM2: case DetM3: bad = DET; add = DET; break;
            default: break;
        }
        if (f & bad) return true;
        st.push_back(f | add);
    }
    return false;
}

uint64_t mix(uint64_t x) {
    uint64_t s = x;
    return splitmix64(s);
}
// The stream of one variation task.
uint64_t task_seed(uint64_t seed, int64_t island, int64_t generation, int64_t slot, int64_t cycle) {
    uint64_t x = mix(seed ^ 0x5851F42D4C957F2Dull);
    x = mix(x ^ (uint64_t)island);
    x = mix(x ^ (uint64_t)generation);
    x = mix(x ^ (uint64_t)slot);
    return mix(x ^ (uint64_t)cycle);
}

struct Batch {
    rsr_programs* p = nullptr;
    ~Batch() {
        if (p) rsr_programs_destroy(p);
    }
};

// compute_mse of every program (NaN for programs the lowering rejects)
int eval(const Backend& B, const Wire& w, const rsr_dataset* ds, std::vector<float>& out) {
    out.assign((size_t)w.n(), NAN);
    if (w.n() == 0) return RSR_OK;
    Batch b;
    int rc = w.create(B.ctx, &b.p);
    if (rc) return rc;
    rc = B.eval(b.p, ds, out.data());
    return rc == RSR_E_PROGRAM ? RSR_OK : rc;
}

// Individual::simplify of every program (rsr_simplify_programs)
int simplify(const Wire& in, Wire& out) {
    const int n = in.n();
    out.clear();
    if (n == 0) return RSR_OK;
    const int64_t code = std::max<int64_t>((int64_t)in.ops.size(), 1);
    std::vector<uint8_t> ops(code), ct(code);
    std::vector<uint16_t> args(code);
    std::vector<float> cv(9 * code);
    std::vector<int64_t> co(n + 1), cof(n + 1), vo(n + 1);
    const uint8_t u8 = 0;
    const uint16_t u16 = 0;
    const float f = 0.0f;
    int rc = rsr_simplify_programs(n, in.ops.empty() ? &u8 : in.ops.data(), in.args.empty() ? &u16 : in.args.data(),
                                   in.co.data(), in.ct.empty() ? &u8 : in.ct.data(), in.cof.data(),
                                   in.cv.empty() ? &f : in.cv.data(), in.vo.data(), code, code, 9 * code, ops.data(),
                                   args.data(), co.data(), ct.data(), cof.data(), cv.data(), vo.data());
    if (rc) return rc;
    for (int i = 0; i < n; ++i) out.add(ops.data(), args.data(), co.data(), ct.data(), cof.data(), cv.data(), vo.data(), i);
    return RSR_OK;
}

int complexity_of(const Wire& w, int i) {
    int32_t c = 0;
    const int64_t co[2] = {0, w.len(i)};
    (void)rsr_complexity(1, w.ops.data() + w.co[i], co, &c);
    return c;
}

float penalised(float mse, int complexity, float parsimony) {
    const float penalty = (float)complexity * parsimony;
    return mse + penalty;
}

struct Hof {  // complexity -> (MSE, program); the reference's HashMap<usize, (f32, Individual)>
    std::map<int, std::pair<float, Wire>> m;
    // insert when the complexity is new or the MSE beats the entry (island.rs:242-250)
    void offer(int c, float mse, const Wire& w, int i) {
        auto it = m.find(c);
        if (it != m.end() && !(mse < it->second.first)) return;
        Wire one;
        one.add(w, i);
        m[c] = {mse, std::move(one)};
    }
};

struct Island {
    Wire pop;
    std::vector<float> fit;
    std::vector<uint64_t> age;
    Wire best;  // one program
    float best_fit = FLT_MAX;
    uint64_t best_age = 0;
    VRng rng;
    int stagnation = 0;
    Hof hof;
};

}  // namespace

struct rsr_engine {
    rsr_context* ctx = nullptr;
    rsr::Backend B;  // evaluations and fits: direct on ctx, or through a server
    rsr_engine_config cfg{};
    const rsr_dataset* ds = nullptr;
    rsr_dataset* mini = nullptr;
    VaryEnv env;
    std::vector<Island> isl;
    Hof ghof;
    int64_t generation = 0;  // generations run
    int64_t nukes = 0;
    float pure_mse = FLT_MAX;
    bool target_reached = false;
    bool finished = false;
    ~rsr_engine() {
        if (mini) rsr_dataset_destroy(mini);
    }
};

namespace {

// generate_random_ast(Float, 5) + simplify for `count` new individuals of each listed
// island (slots give the streams).
int generate(rsr_engine* E, const std::vector<std::pair<int, int64_t>>& who, int64_t gen, Wire& out) {
    std::vector<rsr_vary_task> tasks;
    for (const auto& w : who) {
        rsr_vary_task t{};
        t.kind = RSR_VARY_GENERATE;
        t.type = VtFloat;
        t.depth = 5;
        t.seed = task_seed(E->cfg.seed, w.first, gen, w.second, 0);
        tasks.push_back(t);
    }
    Wire raw;
    int rc = vary(E->ctx, tasks, Wire(), E->env, raw, nullptr);
    if (rc) return rc;
    return simplify(raw, out);
}

int island_of_global_best(const rsr_engine* E) {  // get_global_best: the first minimum
    int g = 0;
    for (int i = 1; i < (int)E->isl.size(); ++i)
        if (E->isl[i].best_fit < E->isl[g].best_fit) g = i;
    return g;
}

// One generation of every island (Island::step_generation, island.rs:71-98).
int step_islands(rsr_engine* E) {
    rsr_context* ctx = E->ctx;
    const rsr_engine_config& C = E->cfg;
    const int G = (int)E->isl.size();
    const int pop = C.island_size;
    const int64_t gen = E->generation;
    int rc;
    std::vector<float> old_best(G);
    // ages, ranks and crowding distances (island.rs:73-77)
    std::vector<std::vector<uint32_t>> rank(G);
    std::vector<std::vector<float>> crowd(G);
    for (int g = 0; g < G; ++g) {
        Island& I = E->isl[g];
        old_best[g] = I.best_fit;
        for (auto& a : I.age) a++;
        rank[g].assign(I.pop.n(), 0);
        crowd[g].assign(I.pop.n(), 0.0f);
        if ((rc = rsr_assign_rank_and_crowding(I.pop.n(), I.fit.data(), I.age.data(), rank[g].data(),
                                               crowd[g].data())))
            return rc;
    }
    // fill_next_generation (island.rs:100-185): the plan of every island from its stream
    enum Kind { BEST, RANDOM, CROSS, MUTATE };
    struct Slot {
        Kind kind;
        int a = -1, b = -1;  // parents in the island's population
    };
    std::vector<std::vector<Slot>> plan(G);
    const int n_rand = (int)std::max((float)pop * C.random_injection_rate, (float)C.min_random_injection);
    auto tournament = [&](int g) {
        Island& I = E->isl[g];
        std::vector<uint32_t> d(C.tournament_size);
        for (auto& x : d) x = vrng_below(I.rng, (uint32_t)I.pop.n());
        uint32_t w = 0;
        (void)rsr_tournament(I.pop.n(), rank[g].data(), crowd[g].data(), C.tournament_size, d.data(), &w);
        return (int)w;
    };
    for (int g = 0; g < G; ++g) {
        auto& P = plan[g];
        P.push_back({BEST});
        for (int r = 0; r < n_rand && (int)P.size() < pop; ++r) P.push_back({RANDOM});
        while ((int)P.size() < pop) {
            const float p = vrng_f32(E->isl[g].rng);
            Slot s;
            if (p < C.crossover_rate) {
                s.kind = CROSS;
                s.a = tournament(g);
                s.b = tournament(g);
            } else {
                s.kind = MUTATE;
                s.a = tournament(g);
            }
            P.push_back(s);
        }
    }
    // the buffers: per island, per slot, one program (filled below)
    struct Child {
        Wire w;  // one program
        float fit = FLT_MAX;
        uint64_t age = 0;
    };
    std::vector<std::vector<Child>> buf(G);
    for (int g = 0; g < G; ++g) buf[g].resize(plan[g].size());
    // random injections
    {
        std::vector<std::pair<int, int64_t>> who;
        std::vector<std::pair<int, int>> at;
        for (int g = 0; g < G; ++g)
            for (int s = 0; s < (int)plan[g].size(); ++s)
                if (plan[g][s].kind == RANDOM) {
                    who.push_back({g, s});
                    at.push_back({g, s});
                }
        Wire made;
        if ((rc = generate(E, who, gen, made))) return rc;
        for (size_t k = 0; k < at.size(); ++k) buf[at[k].first][at[k].second].w.add(made, (int)k);
    }
    for (int g = 0; g < G; ++g) {
        Island& I = E->isl[g];
        buf[g][0].w.add(I.best, 0);
        buf[g][0].fit = I.best_fit;
        buf[g][0].age = I.best_age;
    }
    // crossovers of every island in one launch (island.rs:126-148)
    {
        Wire parents;
        std::vector<rsr_vary_task> tasks;
        std::vector<std::pair<int, int>> at;
        for (int g = 0; g < G; ++g)
            for (int s = 0; s < (int)plan[g].size(); ++s)
                if (plan[g][s].kind == CROSS) {
                    rsr_vary_task t{};
                    t.kind = RSR_VARY_CROSSOVER;
                    t.a = parents.n();
                    parents.add(E->isl[g].pop, plan[g][s].a);
                    t.b = parents.n();
                    parents.add(E->isl[g].pop, plan[g][s].b);
                    t.max_size = C.max_tree_size;
                    t.seed = task_seed(C.seed, g, gen, s, 0);
                    tasks.push_back(t);
                    at.push_back({g, s});
                }
        Wire raw, kept, done;
        if ((rc = vary(ctx, tasks, parents, E->env, raw, nullptr))) return rc;
        for (size_t k = 0; k < at.size(); ++k) {  // a child with a forbidden pattern is parent 1
            if (forbidden_patterns(raw.ops.data() + raw.co[k], raw.len((int)k)))
                kept.add(parents, tasks[k].a);
            else
                kept.add(raw, (int)k);
        }
        if ((rc = simplify(kept, done))) return rc;
        for (size_t k = 0; k < at.size(); ++k) {
            const Island& I = E->isl[at[k].first];
            const Slot& s = plan[at[k].first][at[k].second];
            Child& c = buf[at[k].first][at[k].second];
            c.w.add(done, (int)k);
            c.age = std::max(I.age[s.a], I.age[s.b]);
        }
    }
    // mutation: the candidates, scored on the mini-batch when never scored
    // (island.rs:150-155), then mutation_cycles rounds of mutate, simplify, score on the
    // mini-batch and keep when better (island.rs:156-181)
    {
        std::vector<std::pair<int, int>> at;
        Wire cand;
        std::vector<float> cur;
        for (int g = 0; g < G; ++g)
            for (int s = 0; s < (int)plan[g].size(); ++s)
                if (plan[g][s].kind == MUTATE) {
                    const Island& I = E->isl[g];
                    at.push_back({g, s});
                    cand.add(I.pop, plan[g][s].a);
                    cur.push_back(I.fit[plan[g][s].a]);
                }
        std::vector<uint64_t> ages(at.size());
        for (size_t k = 0; k < at.size(); ++k) ages[k] = E->isl[at[k].first].age[plan[at[k].first][at[k].second].a];
        {
            Wire unscored;
            std::vector<size_t> idx;
            for (size_t k = 0; k < at.size(); ++k)
                if (cur[k] == FLT_MAX) {
                    unscored.add(cand, (int)k);
                    idx.push_back(k);
                }
            std::vector<float> m;
            if ((rc = eval(E->B, unscored, E->mini, m))) return rc;
            for (size_t q = 0; q < idx.size(); ++q) cur[idx[q]] = m[q];
        }
        for (int cyc = 0; cyc < C.mutation_cycles && !at.empty(); ++cyc) {
            std::vector<rsr_vary_task> tasks(at.size());
            for (size_t k = 0; k < at.size(); ++k) {
                rsr_vary_task& t = tasks[k];
                t = rsr_vary_task{};
                t.kind = RSR_VARY_MUTATE;
                t.a = (int)k;
                t.depth = (uint8_t)C.mutation_max_depth;
                t.max_size = C.max_tree_size;
                t.seed = task_seed(C.seed, at[k].first, gen, at[k].second, 1 + cyc);
            }
            Wire raw, mut;
            if ((rc = vary(ctx, tasks, cand, E->env, raw, nullptr))) return rc;
            if ((rc = simplify(raw, mut))) return rc;
            Wire ok;
            std::vector<size_t> idx;
            for (size_t k = 0; k < at.size(); ++k)
                if (!forbidden_patterns(mut.ops.data() + mut.co[k], mut.len((int)k))) {
                    ok.add(mut, (int)k);
                    idx.push_back(k);
                }
            std::vector<float> m;
            if ((rc = eval(E->B, ok, E->mini, m))) return rc;
            std::vector<uint8_t> take(at.size(), 0);
            for (size_t q = 0; q < idx.size(); ++q)
                if (m[q] < cur[idx[q]]) {
                    take[idx[q]] = 1;
                    cur[idx[q]] = m[q];
                }
            Wire next;
            size_t q = 0;
            for (size_t k = 0; k < at.size(); ++k) {
                while (q < idx.size() && idx[q] < k) q++;
                if (take[k])
                    next.add(ok, (int)q);
                else
                    next.add(cand, (int)k);
            }
            cand = std::move(next);
        }
        for (size_t k = 0; k < at.size(); ++k) {
            Child& c = buf[at[k].first][at[k].second];
            c.w.add(cand, (int)k);
            c.fit = cur[k];
            c.age = ages[k];
        }
    }
    // evaluate_buffer of every island in one call (island.rs:215-261)
    {
        Wire all, bests;
        std::vector<int64_t> off{0};
        for (int g = 0; g < G; ++g) {
            for (auto& c : buf[g]) all.add(c.w, 0);
            off.push_back(all.n());
            bests.add(E->isl[g].best, 0);
        }
        std::vector<float> best_mini;
        if ((rc = eval(E->B, bests, E->mini, best_mini))) return rc;
        std::vector<float> best_fit(G), draws;
        std::vector<int64_t> doff{0};
        std::vector<VRng> saved(G);
        for (int g = 0; g < G; ++g) {
            best_fit[g] = E->isl[g].best_fit;
            saved[g] = E->isl[g].rng;
            VRng r = E->isl[g].rng;  // at most two draws per individual
            for (int k = 0; k < 2 * (int)buf[g].size(); ++k) draws.push_back(vrng_f32(r));
            doff.push_back((int64_t)draws.size());
        }
        const int n = all.n();
        const int64_t code = std::max<int64_t>((int64_t)all.ops.size(), 1);
        std::vector<uint8_t> ops(code), ct(code), opt(std::max(n, 1));
        std::vector<uint16_t> args(code);
        std::vector<float> cv(9 * code), fit(std::max(n, 1)), mse(std::max(n, 1));
        std::vector<int64_t> co(n + 1), cof(n + 1), vo(n + 1), used(G);
        std::vector<int32_t> best_idx(G);
        const uint8_t u8 = 0;
        const uint16_t u16 = 0;
        const float f0 = 0.0f;
        rc = islands_evaluate(E->B, E->ds, E->mini, G, off.data(), all.ops.empty() ? &u8 : all.ops.data(),
                                  all.args.empty() ? &u16 : all.args.data(), all.co.data(),
                                  all.ct.empty() ? &u8 : all.ct.data(), all.cof.data(),
                                  all.cv.empty() ? &f0 : all.cv.data(), all.vo.data(), best_fit.data(),
                                  best_mini.data(), draws.data(), doff.data(), C.opt_prob, C.opt_iterations,
                                  C.parsimony_penalty, 0.05f, C.seed + (uint64_t)gen, code, code, 9 * code, ops.data(),
                                  args.data(), co.data(), ct.data(), cof.data(), cv.data(), vo.data(), fit.data(),
                                  mse.data(), opt.data(), best_idx.data(), used.data());
        if (rc) return rc;
        Wire after;
        for (int i = 0; i < n; ++i) after.add(ops.data(), args.data(), co.data(), ct.data(), cof.data(), cv.data(), vo.data(), i);
        for (int g = 0; g < G; ++g) {
            Island& I = E->isl[g];
            I.rng = saved[g];
            for (int64_t k = 0; k < used[g]; ++k) (void)vrng_f32(I.rng);
            Wire np;
            std::vector<float> nf;
            std::vector<uint64_t> na;
            for (int64_t i = off[g]; i < off[g + 1]; ++i) {
                np.add(after, (int)i);
                nf.push_back(fit[i]);
                na.push_back(buf[g][i - off[g]].age);
                if (isfinite(mse[i])) I.hof.offer(complexity_of(after, (int)i), mse[i], after, (int)i);
            }
            if (best_idx[g] >= 0) {
                const int i = (int)off[g] + best_idx[g];
                I.best.clear();
                I.best.add(after, i);
                I.best_fit = fit[i];
                I.best_age = buf[g][best_idx[g]].age;
            }
            I.pop = std::move(np);  // mem::swap(individuals, next_gen_buffer)
            I.fit = std::move(nf);
            I.age = std::move(na);
        }
    }
    // stagnation and nuke (island.rs:84-97, 187-213)
    std::vector<int> nuke;
    for (int g = 0; g < G; ++g) {
        Island& I = E->isl[g];
        const float improvement = old_best[g] - I.best_fit;
        if (improvement > C.min_improvement)
            I.stagnation = 0;
        else
            I.stagnation++;
        if (I.stagnation >= C.stagnation_threshold * 4) nuke.push_back(g);
    }
    if (!nuke.empty()) {
        std::vector<std::pair<int, int64_t>> who;
        for (int g : nuke)
            for (int s = 1; s < pop; ++s) who.push_back({g, (int64_t)pop + s});  // slots apart from the plan's
        Wire made;
        if ((rc = generate(E, who, gen, made))) return rc;
        std::vector<float> m;
        if ((rc = eval(E->B, made, E->ds, m))) return rc;
        int k = 0;
        for (int g : nuke) {
            Island& I = E->isl[g];
            Wire np;
            std::vector<float> nf;
            std::vector<uint64_t> na;
            np.add(I.best, 0);
            nf.push_back(I.best_fit);
            na.push_back(I.best_age);
            for (int s = 1; s < pop; ++s, ++k) {
                np.add(made, k);
                const float mk = m[k];
                nf.push_back(isfinite(mk) ? penalised(mk, complexity_of(made, k), C.parsimony_penalty) : FLT_MAX);
                na.push_back(0);
            }
            I.pop = std::move(np);
            I.fit = std::move(nf);
            I.age = std::move(na);
            I.stagnation = 0;
            E->nukes++;
        }
    }
    return RSR_OK;
}

// One iteration of Engine::run's loop (engine.rs:51-91).
int engine_step(rsr_engine* E) {
    int rc = step_islands(E);
    if (rc) return rc;
    for (auto& I : E->isl)  // the islands' halls of fame into the global one
        for (auto& kv : I.hof.m) E->ghof.offer(kv.first, kv.second.first, kv.second.second, 0);
    const Island& Bst = E->isl[island_of_global_best(E)];
    std::vector<float> m;
    if ((rc = eval(E->B, Bst.best, E->ds, m))) return rc;
    E->pure_mse = m[0];
    const int64_t generation = E->generation++;
    if (E->pure_mse <= E->cfg.target_mse) {
        E->target_reached = true;
        return RSR_OK;
    }
    if (generation > 0 && generation % E->cfg.migration_interval == 0 && E->isl.size() >= 2) {
        // migrate_individuals (engine.rs:97-114): island i receives island i-1's best
        const int G = (int)E->isl.size();
        std::vector<Wire> mig(G);
        std::vector<float> mf(G);
        std::vector<uint64_t> ma(G);
        for (int g = 0; g < G; ++g) {
            mig[g].add(E->isl[g].best, 0);
            mf[g] = E->isl[g].best_fit;
            ma[g] = E->isl[g].best_age;
        }
        for (int g = 0; g < G; ++g) {
            const int src = (g + G - 1) % G;
            Island& I = E->isl[g];
            const int last = I.pop.n() - 1;
            Wire np;
            for (int i = 0; i < last; ++i) np.add(I.pop, i);
            np.add(mig[src], 0);
            I.pop = std::move(np);
            I.fit[last] = mf[src];
            I.age[last] = ma[src];
        }
    }
    return RSR_OK;
}

// The end of Engine::run (engine.rs:78-90): the global best fitted with
// final_opt_iterations, snapped and simplified, rescored, and made island 0's best.
int engine_finish(rsr_engine* E) {
    const int g = island_of_global_best(E);
    Wire best;
    best.add(E->isl[g].best, 0);
    Batch b;
    int rc = best.create(E->ctx, &b.p);
    if (rc) return rc;
    float mse = NAN;
    rc = E->B.optimize(b.p, E->ds, E->cfg.final_opt_iterations, E->cfg.seed, 0.05f, &mse);
    if (rc && rc != RSR_E_PROGRAM) return rc;
    int64_t nv = 0;
    int nprog = 0;
    (void)rsr_programs_count(b.p, &nprog, &nv);
    std::vector<float> vals(std::max<int64_t>(nv, 1));
    if ((rc = rsr_programs_get_constants(b.p, vals.data()))) return rc;
    std::copy(vals.begin(), vals.begin() + nv, best.cv.begin());
    Wire simp;
    if ((rc = simplify(best, simp))) return rc;
    std::vector<float> m;
    if ((rc = eval(E->B, simp, E->ds, m))) return rc;
    Island& I0 = E->isl[0];
    I0.best = std::move(simp);
    I0.best_fit = penalised(m[0], complexity_of(I0.best, 0), E->cfg.parsimony_penalty);
    E->finished = true;
    return RSR_OK;
}

int put_programs(const std::vector<const Wire*>& ws, int64_t cap_p, int64_t cap_c, int64_t cap_k, int64_t cap_v,
                 int* n_out, uint8_t* ops, uint16_t* args, int64_t* co, uint8_t* ct, int64_t* cof, float* cv,
                 int64_t* vo) {
    Wire all;
    for (const Wire* w : ws)
        for (int i = 0; i < w->n(); ++i) all.add(*w, i);
    *n_out = all.n();
    if (all.n() > cap_p || (int64_t)all.ops.size() > cap_c || (int64_t)all.ct.size() > cap_k ||
        (int64_t)all.cv.size() > cap_v)
        return set_error(RSR_E_INVALID_ARG, "%d programs need %zu code, %zu constant and %zu value entries", all.n(),
                         all.ops.size(), all.ct.size(), all.cv.size());
    if (!co || !cof || !vo || (!all.ops.empty() && (!ops || !args)) || (!all.ct.empty() && !ct) ||
        (!all.cv.empty() && !cv))
        return set_error(RSR_E_INVALID_ARG, "null output");
    if (!all.ops.empty()) {
        memcpy(ops, all.ops.data(), all.ops.size());
        memcpy(args, all.args.data(), 2 * all.args.size());
    }
    if (!all.ct.empty()) memcpy(ct, all.ct.data(), all.ct.size());
    if (!all.cv.empty()) memcpy(cv, all.cv.data(), 4 * all.cv.size());
    memcpy(co, all.co.data(), 8 * all.co.size());
    memcpy(cof, all.cof.data(), 8 * all.cof.size());
    memcpy(vo, all.vo.data(), 8 * all.vo.size());
    return RSR_OK;
}

}  // namespace

extern "C" void rsr_engine_default_config(rsr_engine_config* c) {
    if (!c) return;
    *c = rsr_engine_config{};
    c->num_islands = 24;
    c->island_size = 25;
    c->max_generations = 3000;
    c->crossover_rate = 0.10f;
    c->tournament_size = 2;
    c->migration_interval = 25;
    c->parsimony_penalty = 0.000005f;
    c->opt_prob = 0.01f;
    c->opt_iterations = 100;
    c->final_opt_iterations = 4000;
    c->stagnation_threshold = 1000;
    c->target_mse = 1e-7f;
    c->min_improvement = 1e-6f;
    c->random_injection_rate = 0.10f;
    c->min_random_injection = 2;
    c->max_tree_size = 32;
    c->mutation_max_depth = 4;
    c->mutation_cycles = 2;
    c->mini_batch_size = 64;
    c->seed = 42;
}

extern "C" int rsr_engine_create(rsr_context* ctx, const rsr_engine_config* cfg, const rsr_dataset* dataset,
                                 int n_vars, const uint8_t* var_types, const uint8_t* var_indices, int n_allowed,
                                 const uint8_t* allowed_ops, rsr_engine** out) {
    if (!ctx || !cfg || !dataset || !out || n_vars < 0 || n_allowed < 0 || (n_vars && (!var_types || !var_indices)) ||
        (n_allowed && !allowed_ops))
        return set_error(RSR_E_INVALID_ARG, "null argument");
    *out = nullptr;
    const rsr_engine_config& c = *cfg;
    if (c.num_islands < 1 || c.island_size < 2 || c.tournament_size < 1 || c.migration_interval < 1 ||
        c.max_tree_size < 1 || c.mutation_max_depth < 0 || c.mutation_max_depth > 5 || c.mutation_cycles < 0 ||
        c.mini_batch_size < 1 || c.opt_iterations < 0 || c.final_opt_iterations < 0 || c.stagnation_threshold < 0)
        return set_error(RSR_E_INVALID_ARG, "invalid engine configuration");
    int nf = 0;
    int64_t ns = 0, nb = 0;
    int rc = rsr_dataset_info(dataset, &nf, &ns, &nb);
    if (rc) return rc;
    for (int i = 0; i < n_vars; ++i)
        if (var_types[i] > VtMat3 || var_types[i] == VtBool || var_types[i] == VtInt)
            return set_error(RSR_E_INVALID_ARG, "variable %d: type %d", i, var_types[i]);
    bool float_root = false;  // a Float program must be constructible
    for (int i = 0; i < n_allowed; ++i) {
        if (allowed_ops[i] < AddF || allowed_ops[i] >= NegF)
            return set_error(RSR_E_INVALID_ARG, "allowed operator %d is not an operator", allowed_ops[i]);
        float_root = float_root || v_ret(allowed_ops[i]) == VtFloat;
    }
    (void)float_root;
    auto E = std::make_unique<rsr_engine>();
    E->ctx = ctx;
    E->B.ctx = ctx;
    E->cfg = c;
    E->ds = dataset;
    E->env.var_type.assign(var_types, var_types + n_vars);
    E->env.var_idx.assign(var_indices, var_indices + n_vars);
    E->env.ops.assign(allowed_ops, allowed_ops + n_allowed);
    // the mini-batch: Dataset::subset(mini_batch_size) (engine.rs:48-49)
    if ((rc = rsr_dataset_subset(ctx, dataset, c.mini_batch_size, &E->mini))) return rc;
    // Island::new (island.rs:33-68): island_size random Float trees of depth 5, simplified;
    // the first is the initial best (its fitness unset: FLT_MAX)
    std::vector<std::pair<int, int64_t>> who;
    for (int g = 0; g < c.num_islands; ++g)
        for (int s = 0; s < c.island_size; ++s) who.push_back({g, s});
    Wire made;
    if ((rc = generate(E.get(), who, -1, made))) return rc;
    E->isl.resize(c.num_islands);
    int k = 0;
    for (int g = 0; g < c.num_islands; ++g) {
        Island& I = E->isl[g];
        I.rng = vrng_seed(c.seed + (uint64_t)g);
        for (int s = 0; s < c.island_size; ++s, ++k) {
            I.pop.add(made, k);
            I.fit.push_back(FLT_MAX);
            I.age.push_back(0);
        }
        I.best.add(I.pop, 0);
    }
    *out = E.release();
    return RSR_OK;
}

extern "C" int rsr_engine_set_server(rsr_engine* e, rsr_server* server) {
    if (!e) return set_error(RSR_E_INVALID_ARG, "null engine");
    e->B.srv = server;
    return RSR_OK;
}

extern "C" int rsr_engine_step(rsr_engine* e, int* target_reached) {
    if (!e) return set_error(RSR_E_INVALID_ARG, "null engine");
    if (e->finished) return set_error(RSR_E_INVALID_ARG, "the run has finished");
    int rc = engine_step(e);
    if (target_reached) *target_reached = e->target_reached ? 1 : 0;
    return rc;
}

extern "C" int rsr_engine_run(rsr_engine* e, int* generations_out) {
    if (!e) return set_error(RSR_E_INVALID_ARG, "null engine");
    if (e->finished) return set_error(RSR_E_INVALID_ARG, "the run has finished");
    int rc = RSR_OK;
    while (rc == RSR_OK && e->generation < e->cfg.max_generations && !e->target_reached) rc = engine_step(e);
    if (rc == RSR_OK) rc = engine_finish(e);
    if (generations_out) *generations_out = (int)e->generation;
    return rc;
}

extern "C" int rsr_engine_get_info(const rsr_engine* e, rsr_engine_info* info) {
    if (!e || !info) return set_error(RSR_E_INVALID_ARG, "null argument");
    info->generation = e->generation;
    info->nukes = e->nukes;
    info->best_island = island_of_global_best(e);
    info->best_fitness = e->isl[info->best_island].best_fit;
    info->best_mse = e->pure_mse;
    info->target_reached = e->target_reached ? 1 : 0;
    info->finished = e->finished ? 1 : 0;
    info->hof_size = (int32_t)e->ghof.m.size();
    return RSR_OK;
}

extern "C" int rsr_engine_programs(const rsr_engine* e, int which, int island, int64_t program_capacity,
                                   int64_t code_capacity, int64_t const_capacity, int64_t value_capacity, int* n_out,
                                   uint8_t* ops_out, uint16_t* args_out, int64_t* code_offsets_out,
                                   uint8_t* const_types_out, int64_t* const_offsets_out, float* const_values_out,
                                   int64_t* value_offsets_out, float* value_out, int64_t* count_out) {
    if (!e || !n_out) return set_error(RSR_E_INVALID_ARG, "null argument");
    const int G = (int)e->isl.size();
    std::vector<const Wire*> ws;
    std::vector<float> val;
    std::vector<int64_t> cnt;
    std::vector<Wire> front;
    if (which == RSR_ENGINE_POPULATION || which == RSR_ENGINE_ISLAND_BEST) {
        if (island < 0 || island >= G) return set_error(RSR_E_INVALID_ARG, "island %d of %d", island, G);
        const Island& I = e->isl[island];
        if (which == RSR_ENGINE_POPULATION) {
            ws.push_back(&I.pop);
            val = I.fit;
            for (auto a : I.age) cnt.push_back((int64_t)a);
        } else {
            ws.push_back(&I.best);
            val.push_back(I.best_fit);
            cnt.push_back((int64_t)I.best_age);
        }
    } else if (which == RSR_ENGINE_GLOBAL_BEST) {
        const Island& I = e->isl[island_of_global_best(e)];
        ws.push_back(&I.best);
        val.push_back(I.best_fit);
        cnt.push_back((int64_t)I.best_age);
    } else if (which == RSR_ENGINE_PARETO || which == RSR_ENGINE_HALL_OF_FAME || which == RSR_ENGINE_ISLAND_HOF) {
        // get_pareto_front (engine.rs:131-155): the hall of fame by complexity, keeping the
        // entries that lower the best MSE so far
        const Hof* h = &e->ghof;
        if (which == RSR_ENGINE_ISLAND_HOF) {
            if (island < 0 || island >= G) return set_error(RSR_E_INVALID_ARG, "island %d of %d", island, G);
            h = &e->isl[island].hof;
        }
        float best = FLT_MAX;
        for (const auto& kv : h->m) {
            if (which == RSR_ENGINE_PARETO && !(kv.second.first < best)) continue;
            best = std::min(best, kv.second.first);
            ws.push_back(&kv.second.second);
            val.push_back(kv.second.first);
            cnt.push_back(kv.first);
        }
    } else {
        return set_error(RSR_E_INVALID_ARG, "unknown selection %d", which);
    }
    int rc = put_programs(ws, program_capacity, code_capacity, const_capacity, value_capacity, n_out, ops_out, args_out,
                          code_offsets_out, const_types_out, const_offsets_out, const_values_out, value_offsets_out);
    if (rc) return rc;
    if (value_out) memcpy(value_out, val.data(), 4 * val.size());
    if (count_out) memcpy(count_out, cnt.data(), 8 * cnt.size());
    return RSR_OK;
}

extern "C" int rsr_engine_destroy(rsr_engine* e) {
    delete e;
    return RSR_OK;
}
