// select.cpp -- complexity, fitness and NSGA-II selection of the search loop that consumes
// the evaluator's MSEs (host code: O(n^2) over an island of a few dozen individuals).
#include <float.h>
#include <math.h>
#include <stdint.h>

#include <algorithm>
#include <numeric>
#include <vector>

#include "../../include/rsr_amd.h"
#include "isa.h"
#include "runtime_internal.h"

using namespace rsr;

namespace {

// Node::weight (src/engine/expr/node.rs:21-33) with the operator weights of
// src/domains/basic/mod.rs:156-163 and src/domains/linalg/mod.rs:434-478.  A constant
// weighs its type's size (the LoadConst opcode carries the type).
int node_weight(int op) {
    switch (op) {
        case LoadVarF: case LoadVarB: case LoadVarI: case LoadVarV2: case LoadVarV3: case LoadVarM2: case LoadVarM3:
            return 1;  // a variable of any type
        case LoadConstF: case LoadConstB: case LoadConstI: return 1;
        case LoadConstV2: return 2;
        case LoadConstV3: return 3;
        case LoadConstM2: return 4;
        case LoadConstM3: return 9;
        case AddF: case SubF: case MulF: return 1;
        case DivF: case SqrF: case SqrtF: return 2;
        case SinF: case CosF: return 3;
        case ExpF: case LnF: return 4;
        case NegF: return 1;  // the extension opcode, weighted like a subtraction
        case AddV2: case SubV2: case AddV3: case SubV3: case AddM2: case SubM2: case AddM3: case SubM3: return 1;
        case MakeVec2: case GetXV2: case GetYV2: case ScaleV2: case DotV2: case MakeVec3: case GetXV3: case GetYV3:
        case GetZV3: case ScaleV3: case DotV3: case ScaleM2: case ScaleM3: return 2;
        case NormV2: case NormV3: case MakeMat2: case MulM2: case MulM2V2: case TraceM2: case TransposeM2:
        case MulM3V3: case TraceM3: case TransposeM3: return 3;
        case CrossV3: case DetM2: case InverseM2: case MulM3: return 4;
        case MakeMat3: case DetM3: return 5;
        case InverseM3: return 6;
        default: return 0;
    }
}

// dominates (selection.rs:5-11) on (fitness, age), both minimised.
bool dominates(float fa, uint64_t aa, float fb, uint64_t ab) {
    return (fa <= fb && aa <= ab) && (fa < fb || aa < ab);
}

}  // namespace

extern "C" int rsr_complexity(int n, const uint8_t* ops, const int64_t* code_offsets, int32_t* complexity_out) {
    if (n < 0 || (n && (!ops || !code_offsets || !complexity_out))) return set_error(RSR_E_INVALID_ARG, "null argument");
    for (int i = 0; i < n; ++i) {
        int32_t w = 0;
        for (int64_t k = code_offsets[i]; k < code_offsets[i + 1]; ++k) w += node_weight(ops[k]);
        complexity_out[i] = w;
    }
    return RSR_OK;
}

extern "C" int rsr_fitness(int n, const float* mse, const int32_t* complexity, float parsimony, float* out) {
    if (n < 0 || (n && (!mse || !complexity || !out))) return set_error(RSR_E_INVALID_ARG, "null argument");
    for (int i = 0; i < n; ++i) {
        if (!isfinite(mse[i])) {
            out[i] = FLT_MAX;
        } else {
            const float penalty = (float)complexity[i] * parsimony;
            out[i] = mse[i] + penalty;
        }
    }
    return RSR_OK;
}

extern "C" int rsr_assign_rank_and_crowding(int n, const float* fitness, const uint64_t* age, uint32_t* rank,
                                            float* crowding) {
    if (n < 0 || (n && (!fitness || !age || !rank || !crowding))) return set_error(RSR_E_INVALID_ARG, "null argument");
    if (n == 0) return RSR_OK;
    // non-dominated fronts (selection.rs:39-78)
    std::vector<int> dom_count(n, 0);
    std::vector<std::vector<int>> dominated(n);
    std::vector<std::vector<int>> fronts(1);
    for (int i = 0; i < n; ++i) {
        for (int j = 0; j < n; ++j) {
            if (i == j) continue;
            if (dominates(fitness[i], age[i], fitness[j], age[j]))
                dominated[i].push_back(j);
            else if (dominates(fitness[j], age[j], fitness[i], age[i]))
                dom_count[i]++;
        }
        if (dom_count[i] == 0) {
            rank[i] = 0;
            fronts[0].push_back(i);
        }
    }
    for (size_t f = 0; f < fronts.size(); ++f) {
        std::vector<int> next;
        for (int p : fronts[f])
            for (int q : dominated[p])
                if (--dom_count[q] == 0) {
                    rank[q] = (uint32_t)(f + 1);
                    next.push_back(q);
                }
        if (next.empty()) break;
        fronts.push_back(std::move(next));
    }
    // crowding distance over fitness, then over age (selection.rs:80-134); stable sorts
    for (const auto& front : fronts) {
        if (front.empty()) continue;
        for (int i : front) crowding[i] = 0.0f;
        if (front.size() < 3) {
            for (int i : front) crowding[i] = FLT_MAX;
            continue;
        }
        std::vector<int> s = front;
        std::stable_sort(s.begin(), s.end(), [&](int a, int b) { return fitness[a] < fitness[b]; });
        const size_t m = s.size();
        crowding[s[0]] = FLT_MAX;
        crowding[s[m - 1]] = FLT_MAX;
        const float fr = fitness[s[m - 1]] - fitness[s[0]];
        if (fr > 1e-9f)
            for (size_t k = 1; k + 1 < m; ++k) {
                const float dist = (fitness[s[k + 1]] - fitness[s[k - 1]]) / fr;
                if (crowding[s[k]] != FLT_MAX) crowding[s[k]] += dist;
            }
        std::stable_sort(s.begin(), s.end(), [&](int a, int b) { return age[a] < age[b]; });
        crowding[s[0]] = FLT_MAX;
        crowding[s[m - 1]] = FLT_MAX;
        const float amin = (float)age[s[0]], amax = (float)age[s[m - 1]];
        const float ar = amax - amin;
        if (ar > 1e-9f)
            for (size_t k = 1; k + 1 < m; ++k) {
                const float dist = ((float)age[s[k + 1]] - (float)age[s[k - 1]]) / ar;
                if (crowding[s[k]] != FLT_MAX) crowding[s[k]] += dist;
            }
    }
    return RSR_OK;
}

extern "C" int rsr_tournament(int n, const uint32_t* rank, const float* crowding, int k, const uint32_t* draws,
                              uint32_t* winner) {
    if (n <= 0 || k < 1 || !rank || !crowding || !draws || !winner)
        return set_error(RSR_E_INVALID_ARG, "invalid tournament arguments");
    uint32_t best = draws[0];
    if (best >= (uint32_t)n) return set_error(RSR_E_INVALID_ARG, "draw out of range");
    for (int t = 1; t < k; ++t) {
        const uint32_t c = draws[t];
        if (c >= (uint32_t)n) return set_error(RSR_E_INVALID_ARG, "draw out of range");
        // nsga2_compare(challenger, best) == Less (selection.rs:28-37)
        const bool less = rank[c] < rank[best] || (rank[c] == rank[best] && crowding[best] < crowding[c]);
        if (less) best = c;
    }
    *winner = best;
    return RSR_OK;
}
