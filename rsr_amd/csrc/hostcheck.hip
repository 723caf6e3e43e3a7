// hostcheck.hip -- executes lowered micro-code on the host, one point at a time.
//
// A test hook for the lowering only: the CPU test-suite uses it to check the
// postfix -> micro-code translation (register assignment, aliasing, op sequences)
// against the oracle without a GPU.  No evaluation entry point calls it; they all
// launch the HIP kernels.  The arithmetic helpers are the device ones (devmath.h is
// __host__ __device__), so this also pins the device transcendental code on the CPU.
#include <math.h>
#include <string.h>
#include <algorithm>
#include <vector>
#include "../../include/rsr_amd.h"
#include "devmath.h"
#include "lower.h"

using namespace rsr;

// One micro-op on the host register files (values R, tangents G with respect to flat
// parameter k).
static void run_one(uint32_t op, uint32_t x, uint32_t d, uint32_t a, uint32_t b, uint32_t c, uint32_t e,
                    const float* params, float* R, float* G, int k, float* spill) {
    {
        float v = 0.0f, g = 0.0f;
        switch (op) {
            case U_IMM: v = params[x]; g = ((int)x == k) ? 1.0f : 0.0f; break;
            case U_ADD: v = R[a] + R[b]; g = G[a] + G[b]; break;
            case U_SUB: v = R[a] - R[b]; g = G[a] - G[b]; break;
            // operand-swapped forms: value code only (no tangents)
            case U_RSUB: v = R[b] - R[a]; break;
            case U_RDIV: v = R[b] / R[a]; break;
            case U_MULM3V3R: {
                float rv[3];
                for (int i = 0; i < 3; ++i) rv[i] = ((R[b + i] * R[a]) + (R[b + 3 + i] * R[a + 1])) + (R[b + 6 + i] * R[a + 2]);
                for (int i = 0; i < 3; ++i) R[d + i] = rv[i];
                return;
            }
            case U_MUL: v = R[a] * R[b]; g = (R[a] * G[b]) + (G[a] * R[b]); break;
            case U_DIV: v = R[a] / R[b]; g = ((G[a] * R[b]) - (R[a] * G[b])) / (R[b] * R[b]); break;
            case U_SQR: v = R[a] * R[a]; g = (2.0f * R[a]) * G[a]; break;
            case U_SQRT: v = sqrtf(R[a]); g = G[a] / (2.0f * v); break;
            case U_SINK: v = dm_sin(params[x]); break;  // constant operand: value code only
            case U_COSK: v = dm_cos(params[x]); break;
            case U_EXPK: v = dm_exp(params[x]); break;
            case U_LNK: v = dm_ln(params[x]); break;
            case U_SQRTK: v = sqrtf(params[x]); break;
            case U_NEGK: v = -params[x]; break;
            case U_SQRK: v = params[x] * params[x]; break;
            case U_SINU: v = dm_sin(R[a]); break;  // point-invariant forms: value code only
            case U_COSU: v = dm_cos(R[a]); break;
            case U_EXPU: v = dm_exp(R[a]); break;
            case U_LNU: v = dm_ln(R[a]); break;
            case U_SQRTU: v = sqrtf(R[a]); break;
            case U_SIN: v = dm_sin(R[a]); g = G[a] * dm_cos(R[a]); break;
            case U_COS: v = dm_cos(R[a]); g = G[a] * (-dm_sin(R[a])); break;
            case U_EXP: v = dm_exp(R[a]); g = G[a] * v; break;
            case U_LN: v = dm_ln(R[a]); g = G[a] / R[a]; break;
            case U_NEG: v = -R[a]; g = -G[a]; break;
            case U_MULADD: v = (R[a] * R[b]) + (R[c] * R[e]);
                g = ((R[a] * G[b]) + (G[a] * R[b])) + ((R[c] * G[e]) + (G[c] * R[e])); break;
            case U_MULSUB: v = (R[a] * R[b]) - (R[c] * R[e]);
                g = ((R[a] * G[b]) + (G[a] * R[b])) - ((R[c] * G[e]) + (G[c] * R[e])); break;
            case U_ADDMUL: v = R[a] + (R[b] * R[c]); g = G[a] + ((R[b] * G[c]) + (G[b] * R[c])); break;
            case U_SQRTACC: v = sqrtf(R[a] + (R[b] * R[b])); g = (G[a] + ((R[b] * G[b]) + (G[b] * R[b]))) / (2.0f * v); break;
            case U_ADD3: v = (R[a] + R[b]) + R[c]; g = (G[a] + G[b]) + G[c]; break;
            case U_RCPSAFE: v = 1.0f / ((fabsf(R[a]) < 1e-9f) ? 1.0f : R[a]); break;
            case U_INVSEL: {
                const float xa = (ux_flags(x) & kInvNeg) ? -R[a] : R[a];
                v = (fabsf(R[c]) < 1e-9f) ? ((ux_flags(x) & kInvOne) ? 1.0f : 0.0f) : xa * R[b];
                break;
            }
            case U_SPILL: v = R[a]; g = G[a]; spill[2 * x] = v; spill[2 * x + 1] = g; break;
            case U_FILL: v = spill[2 * x]; g = spill[2 * x + 1]; break;
            case U_DOT3:
                v = ((R[a] * R[b]) + (R[a + 1] * R[b + 1])) + (R[a + 2] * R[b + 2]);
                g = (((R[a] * G[b]) + (G[a] * R[b])) + ((R[a + 1] * G[b + 1]) + (G[a + 1] * R[b + 1]))) +
                    ((R[a + 2] * G[b + 2]) + (G[a + 2] * R[b + 2]));
                break;
            case U_NORM3: {
                v = sqrtf(((R[a] * R[a]) + (R[a + 1] * R[a + 1])) + (R[a + 2] * R[a + 2]));
                const float t0 = ((R[a] * G[a]) + (G[a] * R[a])) + ((R[a + 1] * G[a + 1]) + (G[a + 1] * R[a + 1]));
                g = (t0 + ((R[a + 2] * G[a + 2]) + (G[a + 2] * R[a + 2]))) / (2.0f * v);
                break;
            }
            case U_ADDK: case U_SUBK: case U_KSUB: case U_MULK: case U_DIVK: case U_KDIV: {
                const float p = params[x], gp = ((int)x == k) ? 1.0f : 0.0f;
                switch (op) {
                    case U_ADDK: v = R[a] + p; g = G[a] + gp; break;
                    case U_SUBK: v = R[a] - p; g = G[a] - gp; break;
                    case U_KSUB: v = p - R[a]; g = gp - G[a]; break;
                    case U_MULK: v = R[a] * p; g = (R[a] * gp) + (G[a] * p); break;
                    case U_DIVK: v = R[a] / p; g = ((G[a] * p) - (R[a] * gp)) / (p * p); break;
                    default: v = p / R[a]; g = ((gp * R[a]) - (p * G[a])) / (R[a] * R[a]); break;
                }
                break;
            }
            case U_MULM3V3: {
                float rv[3], rg[3];
                for (int i = 0; i < 3; ++i) {
                    const float t = (R[a + i] * R[b]) + (R[a + 3 + i] * R[b + 1]);
                    const float tg = ((R[a + i] * G[b]) + (G[a + i] * R[b])) +
                                     ((R[a + 3 + i] * G[b + 1]) + (G[a + 3 + i] * R[b + 1]));
                    rv[i] = t + (R[a + 6 + i] * R[b + 2]);
                    rg[i] = tg + ((R[a + 6 + i] * G[b + 2]) + (G[a + 6 + i] * R[b + 2]));
                }
                for (int i = 0; i < 3; ++i) {
                    R[d + i] = rv[i];
                    G[d + i] = rg[i];
                }
                return;
            }
            case U_SCALEM3: case U_SCALEM3K: {
                const bool kk = op == U_SCALEM3K;
                const float sv = kk ? params[x] : R[b], gs = kk ? (((int)x == k) ? 1.0f : 0.0f) : G[b];
                float rv[9], rg[9];
                for (int q = 0; q < 9; ++q) {
                    rv[q] = R[a + q] * sv;
                    rg[q] = (R[a + q] * gs) + (G[a + q] * sv);
                }
                for (int q = 0; q < 9; ++q) {
                    R[d + q] = rv[q];
                    G[d + q] = rg[q];
                }
                return;
            }
            case U_ADDV4: case U_SUBV4: case U_SCALEV4: case U_SCALEV4K: case U_MULM2: case U_INVM2: {
                float rv[4], rg[4];
                const float p = params[x], gp = ((int)x == k) ? 1.0f : 0.0f;
                for (int q = 0; q < 4; ++q) {
                    switch (op) {
                        case U_ADDV4: rv[q] = R[a + q] + R[b + q]; rg[q] = G[a + q] + G[b + q]; break;
                        case U_SUBV4: rv[q] = R[a + q] - R[b + q]; rg[q] = G[a + q] - G[b + q]; break;
                        case U_SCALEV4: rv[q] = R[a + q] * R[b]; rg[q] = (R[a + q] * G[b]) + (G[a + q] * R[b]); break;
                        case U_SCALEV4K: rv[q] = R[a + q] * p; rg[q] = (R[a + q] * gp) + (G[a + q] * p); break;
                        case U_MULM2: {
                            const int r = q & 1, c = q >> 1;
                            rv[q] = (R[a + r] * R[b + 2 * c]) + (R[a + r + 2] * R[b + 2 * c + 1]);
                            rg[q] = ((R[a + r] * G[b + 2 * c]) + (G[a + r] * R[b + 2 * c])) +
                                    ((R[a + r + 2] * G[b + 2 * c + 1]) + (G[a + r + 2] * R[b + 2 * c + 1]));
                            break;
                        }
                        default: {  // inverse (linalg/eval.rs:316-329); tangents unused (never lowered dual)
                            const float det = (R[a] * R[a + 3]) - (R[a + 1] * R[a + 2]);
                            const bool sing = fabsf(det) < 1e-9f;
                            const float inv = 1.0f / (sing ? 1.0f : det);
                            static const int src[4] = {3, 1, 2, 0};
                            const float m = (q == 1 || q == 2) ? -R[a + src[q]] : R[a + src[q]];
                            rv[q] = sing ? ((q == 0 || q == 3) ? 1.0f : 0.0f) : m * inv;
                            rg[q] = 0.0f;
                        }
                    }
                }
                for (int q = 0; q < 4; ++q) {
                    R[d + q] = rv[q];
                    G[d + q] = rg[q];
                }
                return;
            }
            case U_RSUBV3: case U_RCROSS3: case U_SCALEV3S: {
                float rv[3];
                for (int q = 0; q < 3; ++q) {
                    const int i = (q + 1) % 3, j = (q + 2) % 3;
                    rv[q] = op == U_RSUBV3 ? R[b + q] - R[a + q]
                          : op == U_SCALEV3S ? R[b + q] * R[a]
                                             : (R[b + i] * R[a + j]) - (R[b + j] * R[a + i]);
                }
                for (int q = 0; q < 3; ++q) R[d + q] = rv[q];
                return;
            }
            case U_ADDV3: case U_SUBV3: case U_SCALEV3: case U_SCALEV3K: case U_IMMV3: case U_CROSS3: {
                float rv[3], rg[3];
                for (int q = 0; q < 3; ++q) {
                    switch (op) {
                        case U_ADDV3: rv[q] = R[a + q] + R[b + q]; rg[q] = G[a + q] + G[b + q]; break;
                        case U_SUBV3: rv[q] = R[a + q] - R[b + q]; rg[q] = G[a + q] - G[b + q]; break;
                        case U_SCALEV3: rv[q] = R[a + q] * R[b]; rg[q] = (R[a + q] * G[b]) + (G[a + q] * R[b]); break;
                        case U_SCALEV3K: {
                            const float p = params[x], gp = ((int)x == k) ? 1.0f : 0.0f;
                            rv[q] = R[a + q] * p;
                            rg[q] = (R[a + q] * gp) + (G[a + q] * p);
                            break;
                        }
                        case U_IMMV3: rv[q] = params[x + q]; rg[q] = ((int)(x + q) == k) ? 1.0f : 0.0f; break;
                        default: {  // cross: component q = (a[i]*b[j]) - (a[j]*b[i])
                            const int i = (q + 1) % 3, j = (q + 2) % 3;
                            rv[q] = (R[a + i] * R[b + j]) - (R[a + j] * R[b + i]);
                            rg[q] = ((R[a + i] * G[b + j]) + (G[a + i] * R[b + j])) -
                                    ((R[a + j] * G[b + i]) + (G[a + j] * R[b + i]));
                        }
                    }
                }
                for (int q = 0; q < 3; ++q) {
                    R[d + q] = rv[q];
                    G[d + q] = rg[q];
                }
                return;
            }
            default: break;
        }
        R[d] = v;
        G[d] = g;
    }
}

// R and G hold kRegs + 3 entries: registers [kRegs, kRegs + 3) stand for the
// interpreter's accumulator, so a kFwd operand is read from there.  A kNoWb result is copied there and
// its file registers are overwritten with NaN, so any read of the file that the
// forwarding pass wrongly declared dead poisons the program's result.
static void run_host(const std::vector<UInstr>& code, const float* params, float* R, float* G, int k,
                     float* spill /* [slot][2] */) {
    float Fv[3] = {0.0f, 0.0f, 0.0f}, Fg[3] = {0.0f, 0.0f, 0.0f};
    for (const UInstr& u : code) {
        const uint32_t op = u.w0 & 255u, x = u.w0 >> 8, d = u.w1 & 31u;
        for (int q = 0; q < 3; ++q) {
            R[kRegs + q] = Fv[q];
            G[kRegs + q] = Fg[q];
        }
        const uint32_t a = (u.w2 & kFwd) ? (uint32_t)kRegs : (u.w2 & 31u);
        const uint32_t b = u.w3 & 31u;
        const uint32_t c = ux_c(x), e = ux_e(x);
        const int nd = ugroups(op).nd;
        run_one(op, x, d, a, b, c, e, params, R, G, k, spill);
        for (int q = 0; q < nd && q < 3; ++q) {
            Fv[q] = R[d + q];
            Fg[q] = G[d + q];
        }
        if (u.w1 & kNoWb)
            for (int q = 0; q < nd; ++q) {
                R[d + q] = __builtin_nanf("");
                G[d + q] = __builtin_nanf("");
            }
    }
}

// Lower one program (reference wire format) and run it on n points given as
// feature-major columns feat[f * n + i].  vals_out[i] = value; when dual != 0 the
// tangent w.r.t. flat parameter `active_param` goes to tan_out[i].  Returns the
// lowering status; n_uops / max_reg describe the emitted code.
extern "C" RSR_API int rsr_lower_check(const uint8_t* ops, const uint16_t* args, int n_code, const uint8_t* ctypes,
                                       const float* cvals, int n_consts, int num_features, int dual,
                                       int active_param, const float* feat, int64_t n, float* vals_out,
                                       float* tan_out, int32_t* n_uops, int32_t* max_reg) {
    // n_uops / max_reg report the code size and register use; spill slots are folded
    // into max_reg as 32 + slots when the program spills
    RefProgram p{ops, args, n_code, ctypes, cvals, n_consts};
    Lowered lw = lower_program(p, num_features, dual != 0);
    if (n_uops) *n_uops = (int32_t)lw.code.size();
    if (max_reg) *max_reg = (int32_t)(lw.spill_slots ? kRegs + lw.spill_slots : lw.max_reg);
    if (lw.status != PROG_OK) return lw.status;
    std::vector<float> params;
    flatten_params(p, params);
    params.push_back(0.0f);
    float R[kRegs + 3], G[kRegs + 3], spill[2 * kMaxSlots];
    for (int64_t i = 0; i < n; ++i) {
        // temporaries and spill slots start as NaN: a read before a write poisons the result
        for (int r = 0; r < kRegs + 3; ++r) { R[r] = __builtin_nanf(""); G[r] = __builtin_nanf(""); }
        for (int s = 0; s < 2 * kMaxSlots; ++s) spill[s] = __builtin_nanf("");
        for (int f = 0; f < num_features; ++f) { R[f] = feat[(int64_t)f * n + i]; G[f] = 0.0f; }
        run_host(lw.code, params.data(), R, G, dual ? active_param : -1, spill);
        if (vals_out) vals_out[i] = R[lw.out_reg];
        if (tan_out) tan_out[i] = G[lw.out_reg];
    }
    return 0;
}

// Lowered micro-code of one program (value code when dual == 0) as four words per
// micro-op, for tests that check which micro-ops and flags the lowering emits.
extern "C" RSR_API int rsr_lower_code(const uint8_t* ops, const uint16_t* args, int n_code, const uint8_t* ctypes,
                                      const float* cvals, int n_consts, int num_features, int dual,
                                      uint32_t* words_out, int32_t capacity, int32_t* n_uops) {
    RefProgram p{ops, args, n_code, ctypes, cvals, n_consts};
    Lowered lw = lower_program(p, num_features, dual != 0);
    if (n_uops) *n_uops = (int32_t)lw.code.size();
    if (lw.status != PROG_OK) return lw.status;
    const int32_t n = std::min<int32_t>(capacity, (int32_t)lw.code.size());
    for (int32_t i = 0; i < n; ++i) {
        words_out[4 * i] = lw.code[i].w0;
        words_out[4 * i + 1] = lw.code[i].w1;
        words_out[4 * i + 2] = lw.code[i].w2;
        words_out[4 * i + 3] = lw.code[i].w3;
    }
    return 0;
}
