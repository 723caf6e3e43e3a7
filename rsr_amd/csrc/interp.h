// interp.h -- the wave-uniform micro-op interpreter (device side).
//
// One wavefront runs one program at a time over 64 * PPL data points (PPL points per
// lane).  The micro-op stream is read through the scalar data cache (the program
// pointer and pc are wave-uniform, so the loads are s_load), decoded in SGPRs and
// dispatched by a uniform branch tree: no lane ever diverges.  Register operands are
// wave-uniform indices into a per-lane register file held in VGPRs
// (ext_vector_type(32) indexed by an SGPR compiles to s_set_gpr_idx_on + v_mov, no
// scratch).  Every iteration reads its operands inside the switch and writes exactly
// one register component at a single point after it, which keeps the register file in
// place (no vector copies).
//
// The arithmetic of every micro-op is the reference formula step by step (see isa.h);
// the dual variant carries T tangent channels and applies the forward-mode rules of
// src/engine/eval/autodiff.rs:32-133 to each channel.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <utility>
#include "isa.h"
#include "devmath.h"

namespace rsr {

typedef float v32f __attribute__((ext_vector_type(32)));

__device__ __forceinline__ uint32_t uni(uint32_t x) { return __builtin_amdgcn_readfirstlane(x); }

// N register files as distinct members (an array of vectors would become a stack
// allocation); element I is reached through a compile-time index only.
template <int N> struct Files;
template <> struct Files<1> { v32f a; };
template <> struct Files<2> { v32f a, b; };
template <> struct Files<4> { v32f a, b, c, d; };

// Value file plus T tangent files in one aggregate (two separate aggregates passed by
// reference are not promoted to registers by the optimizer).
template <int T> struct DualFiles;
template <> struct DualFiles<1> { v32f v, a; };
template <> struct DualFiles<4> { v32f v, a, b, c, d; };

template <int I, int N>
__device__ __forceinline__ v32f& fget(Files<N>& f) {
    if constexpr (I == 0) return f.a;
    else if constexpr (I == 1) return f.b;
    else if constexpr (I == 2) return f.c;
    else return f.d;
}

// compile-time loop: fn(std::integral_constant<int, 0>) ... fn(<N-1>)
template <int N, class Fn>
__device__ __forceinline__ void sfor(Fn&& fn) {
    [&]<int... I>(std::integer_sequence<int, I...>) {
        (fn(std::integral_constant<int, I>{}), ...);
    }(std::make_integer_sequence<int, N>{});
}

// Primal interpreter over PPL register files (one per point of the lane).
// `lds` is this wave's spill area: slot s, point p at lds[(s * PPL + p) * 64 + lane].
template <int PPL>
__device__ __forceinline__ void run_program(Files<PPL>& R, const UInstr* __restrict__ code, uint32_t len,
                                            const float* __restrict__ params, float* __restrict__ lds, int lane) {
    uint32_t nw0 = len ? uni(code[0].w0) : 0u, nw1 = len ? uni(code[0].w1) : 0u;
    for (uint32_t pc = 0; pc < len; ++pc) {
        const uint32_t w0 = nw0, w1 = nw1;
        // prefetch the next micro-op (the code array carries one padding entry)
        nw0 = uni(code[pc + 1].w0);
        nw1 = uni(code[pc + 1].w1);
        const uint32_t op = w0 & 255u;
        const uint32_t d = (w0 >> 8) & 31u, a = (w0 >> 16) & 31u, b = (w0 >> 24) & 31u;
        const uint32_t c = w1 & 31u, e = (w1 >> 8) & 31u;
        float res[PPL];
        switch (op) {
            case U_IMM: {
                const float v = params[w1];
                sfor<PPL>([&](auto P) { res[P()] = v; });
                break;
            }
#define RSR_CASE(OPC, EXPR)                               \
    case OPC: {                                           \
        sfor<PPL>([&](auto P) {                           \
            const v32f r = fget<P(), PPL>(R);             \
            res[P()] = (EXPR);                            \
        });                                               \
        break;                                            \
    }
            RSR_CASE(U_ADD, r[a] + r[b])
            RSR_CASE(U_SUB, r[a] - r[b])
            RSR_CASE(U_MUL, r[a] * r[b])
            RSR_CASE(U_DIV, r[a] / r[b])
            RSR_CASE(U_SQR, r[a] * r[a])
            RSR_CASE(U_SQRT, __builtin_sqrtf(r[a]))
            RSR_CASE(U_SIN, dm_sin(r[a]))
            RSR_CASE(U_COS, dm_cos(r[a]))
            RSR_CASE(U_EXP, dm_exp(r[a]))
            RSR_CASE(U_LN, dm_ln(r[a]))
            RSR_CASE(U_NEG, -r[a])
            RSR_CASE(U_MULADD, (r[a] * r[b]) + (r[c] * r[e]))
            RSR_CASE(U_MULSUB, (r[a] * r[b]) - (r[c] * r[e]))
            RSR_CASE(U_ADDMUL, r[a] + (r[b] * r[c]))
            RSR_CASE(U_SQRTACC, __builtin_sqrtf(r[a] + (r[b] * r[b])))
            RSR_CASE(U_ADD3, (r[a] + r[b]) + r[c])
            RSR_CASE(U_RCPSAFE, 1.0f / ((__builtin_fabsf(r[a]) < 1e-9f) ? 1.0f : r[a]))
            case U_INVSEL: {
                const float k = (w1 & (2u << 16)) ? 1.0f : 0.0f;
                const bool ng = (w1 & (1u << 16)) != 0u;
                sfor<PPL>([&](auto P) {
                    const v32f r = fget<P(), PPL>(R);
                    const float x = ng ? -r[a] : r[a];
                    res[P()] = (__builtin_fabsf(r[c]) < 1e-9f) ? k : x * r[b];
                });
                break;
            }
#undef RSR_CASE
            case U_SPILL:
                sfor<PPL>([&](auto P) {
                    const v32f r = fget<P(), PPL>(R);
                    lds[(w1 * PPL + P()) * 64u + lane] = r[a];
                    res[P()] = r[a];
                });
                break;
            case U_FILL:
                sfor<PPL>([&](auto P) { res[P()] = lds[(w1 * PPL + P()) * 64u + lane]; });
                break;
            default:
                sfor<PPL>([&](auto P) { res[P()] = 0.0f; });
                break;
        }
        sfor<PPL>([&](auto P) {
            v32f t = fget<P(), PPL>(R);
            t[d] = res[P()];
            fget<P(), PPL>(R) = t;
        });
    }
}

// Dual interpreter: value file V plus T tangent files; channel j is the tangent with
// respect to flat parameter k0 + j.
// Expand STMT once per tangent channel j with Gj = that channel's register file (by value).
#define RSR_TAN1(STMT, J, M) { constexpr int j = J; const v32f Gj = G.M; (void)Gj; STMT; }
#define RSR_TAN(STMT)                                                  \
    do {                                                               \
        RSR_TAN1(STMT, 0, a)                                           \
        if constexpr (T >= 2) { RSR_TAN1(STMT, 1, b) }                 \
        if constexpr (T >= 4) { RSR_TAN1(STMT, 2, c) RSR_TAN1(STMT, 3, d) } \
    } while (0)
// Spill slot s, channel ch (0 = value, 1 + j = tangent j) at lds[(s * (1 + T) + ch) * 64 + lane].
template <int T>
__device__ __forceinline__ void run_dual(DualFiles<T>& G, const UInstr* __restrict__ code, uint32_t len,
                                         const float* __restrict__ params, uint32_t k0, float* __restrict__ lds,
                                         int lane) {
    uint32_t nw0 = len ? uni(code[0].w0) : 0u, nw1 = len ? uni(code[0].w1) : 0u;
    for (uint32_t pc = 0; pc < len; ++pc) {
        const uint32_t w0 = nw0, w1 = nw1;
        nw0 = uni(code[pc + 1].w0);
        nw1 = uni(code[pc + 1].w1);
        const uint32_t op = w0 & 255u;
        const uint32_t d = (w0 >> 8) & 31u, a = (w0 >> 16) & 31u, b = (w0 >> 24) & 31u;
        const uint32_t c = w1 & 31u, e = (w1 >> 8) & 31u;
        float rv;
        float rg[T];
        const v32f Vv = G.v;
#define V Vv
        switch (op) {
            case U_IMM:
                rv = params[w1];
                RSR_TAN(rg[j] = (w1 == k0 + (uint32_t)j) ? 1.0f : 0.0f);
                break;
            case U_ADD:
                rv = V[a] + V[b];
                RSR_TAN(rg[j] = Gj[a] + Gj[b]);
                break;
            case U_SUB:
                rv = V[a] - V[b];
                RSR_TAN(rg[j] = Gj[a] - Gj[b]);
                break;
            case U_MUL: {
                const float x = V[a], y = V[b];
                rv = x * y;
                RSR_TAN(rg[j] = (x * Gj[b]) + (Gj[a] * y));
                break;
            }
            case U_DIV: {
                const float x = V[a], y = V[b];
                rv = x / y;
                const float yy = y * y;
                RSR_TAN(rg[j] = ((Gj[a] * y) - (x * Gj[b])) / yy);
                break;
            }
            case U_SQR: {
                const float x = V[a];
                rv = x * x;
                const float x2 = 2.0f * x;
                RSR_TAN(rg[j] = x2 * Gj[a]);
                break;
            }
            case U_SQRT: {
                const float s = __builtin_sqrtf(V[a]);
                rv = s;
                const float s2 = 2.0f * s;
                RSR_TAN(rg[j] = Gj[a] / s2);
                break;
            }
            case U_SIN: {
                const float x = V[a];
                rv = dm_sin(x);
                const float cs = dm_cos(x);
                RSR_TAN(rg[j] = Gj[a] * cs);
                break;
            }
            case U_COS: {
                const float x = V[a];
                rv = dm_cos(x);
                const float ns = -dm_sin(x);
                RSR_TAN(rg[j] = Gj[a] * ns);
                break;
            }
            case U_EXP: {
                const float ex = dm_exp(V[a]);
                rv = ex;
                RSR_TAN(rg[j] = Gj[a] * ex);
                break;
            }
            case U_LN: {
                const float x = V[a];
                rv = dm_ln(x);
                RSR_TAN(rg[j] = Gj[a] / x);
                break;
            }
            case U_NEG:
                rv = -V[a];
                RSR_TAN(rg[j] = -Gj[a]);
                break;
            case U_MULADD: case U_MULSUB: {
                const float xa = V[a], xb = V[b], xc = V[c], xe = V[e];
                const bool sub = op == U_MULSUB;
                const float p1 = xa * xb, p2 = xc * xe;
                rv = sub ? p1 - p2 : p1 + p2;
                RSR_TAN(const float t1 = (xa * Gj[b]) + (Gj[a] * xb); const float t2 = (xc * Gj[e]) + (Gj[c] * xe);
                        rg[j] = sub ? t1 - t2 : t1 + t2);
                break;
            }
            case U_ADDMUL: {
                const float xb = V[b], xc = V[c];
                rv = V[a] + (xb * xc);
                RSR_TAN(rg[j] = Gj[a] + ((xb * Gj[c]) + (Gj[b] * xc)));
                break;
            }
            case U_SQRTACC: {
                const float xb = V[b];
                const float s = __builtin_sqrtf(V[a] + (xb * xb));
                rv = s;
                const float s2 = 2.0f * s;
                RSR_TAN(rg[j] = (Gj[a] + ((xb * Gj[b]) + (Gj[b] * xb))) / s2);
                break;
            }
            case U_ADD3:
                rv = (V[a] + V[b]) + V[c];
                RSR_TAN(rg[j] = (Gj[a] + Gj[b]) + Gj[c]);
                break;
            case U_SPILL:
                rv = V[a];
                lds[(w1 * (1 + T)) * 64u + lane] = rv;
                RSR_TAN(rg[j] = Gj[a]; lds[(w1 * (1 + T) + 1 + j) * 64u + lane] = rg[j]);
                break;
            case U_FILL:
                rv = lds[(w1 * (1 + T)) * 64u + lane];
                RSR_TAN(rg[j] = lds[(w1 * (1 + T) + 1 + j) * 64u + lane]);
                break;
            default:
                rv = 0.0f;
                RSR_TAN(rg[j] = 0.0f);
                break;
        }
#undef V
        {
            v32f t = G.v;
            t[d] = rv;
            G.v = t;
        }
        {
            v32f t = G.a; t[d] = rg[0]; G.a = t;
        }
        if constexpr (T >= 2) { v32f t = G.b; t[d] = rg[1]; G.b = t; }
        if constexpr (T >= 4) {
            { v32f t = G.c; t[d] = rg[2]; G.c = t; }
            { v32f t = G.d; t[d] = rg[3]; G.d = t; }
        }
    }
}
#undef RSR_TAN
#undef RSR_TAN1

// Sum over the 64 lanes of a wavefront in a fixed butterfly order (deterministic).
__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
    return v;
}

}  // namespace rsr
