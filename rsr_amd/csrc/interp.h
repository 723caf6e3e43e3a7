// interp.h -- the wave-uniform micro-op interpreter (device side).
//
// One wavefront runs one program at a time over 64 * PPL data points (PPL points per
// lane).  The micro-op stream is read through the scalar data cache (the program
// pointer and pc are wave-uniform, so the loads are s_load), decoded in SGPRs and
// dispatched by a uniform branch tree: no lane ever diverges.  Register operands are
// wave-uniform indices into a per-lane register file held in VGPRs
// (ext_vector_type(32) indexed by an SGPR compiles to s_set_gpr_idx_on + v_mov, no
// scratch).  Every iteration reads operands a and b before the switch, computes inside
// it, and writes result component 0 after it (block ops write their further components
// inside the switch), which keeps the register file in place (no vector copies).
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

// Program constants are read through the constant address space: the pointer and the
// index are wave-uniform, so the read is one s_load (waited on with lgkmcnt) instead of a
// vector load whose vmcnt(0) wait would also drain the lane's outstanding dataset loads.
// The buffer is only written between launches, never during one.
typedef const __attribute__((address_space(4))) float* kptr;
__device__ __forceinline__ float kld(const float* p, uint32_t i) { return ((kptr)p)[i]; }
// The same for the other launch-constant tables a wave reads at uniform addresses (micro-
// code, program descriptors, scheduling order, active flags): s_loads by construction,
// not by the compiler's proof that nothing in the kernel writes them, which a wait or a
// fence placed before the loads would defeat.
template <class T>
__device__ __forceinline__ T kget(const T* p, int64_t i) {
    if constexpr (sizeof(T) == 16) {  // four-dword records (16-byte aligned arrays)
        typedef uint32_t u4 __attribute__((ext_vector_type(4)));
        return __builtin_bit_cast(T, ((const __attribute__((address_space(4))) u4*)p)[i]);
    } else {
        return ((const __attribute__((address_space(4))) T*)p)[i];
    }
}

typedef float v32f __attribute__((ext_vector_type(32)));

// The lane index for spill-slot addressing, recomputed where it is used (two VALU ops).
// It depends on the micro-op's slot field (x >> 24 is zero for every valid slot), so the
// compiler cannot hoist it out of the interpreter loop and keep it in a VGPR for the
// whole program, which at four points per lane pushed the kernel into scratch spills.
__device__ __forceinline__ uint32_t spill_lane(uint32_t x) {
    return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, x >> 24));
}

__device__ __forceinline__ uint32_t uni(uint32_t x) { return __builtin_amdgcn_readfirstlane(x); }

// N register files as distinct members (an array of vectors would become a stack
// allocation); element I is reached through a compile-time index only.
template <int N> struct Files;
template <> struct Files<1> { v32f a; };
template <> struct Files<2> { v32f a, b; };
template <> struct Files<3> { v32f a, b, c; };
template <> struct Files<4> { v32f a, b, c, d; };

// Value file plus T tangent files in one aggregate (two separate aggregates passed by
// reference are not promoted to registers by the optimizer).
template <int T> struct DualFiles;
template <> struct DualFiles<1> { v32f v, a; };
template <> struct DualFiles<2> { v32f v, a, b; };
template <> struct DualFiles<4> { v32f v, a, b, c, d; };

template <int T>
__device__ __forceinline__ void zero_tangents(DualFiles<T>& G) {
    G.a = (v32f)(0.0f);
    if constexpr (T >= 2) G.b = (v32f)(0.0f);
    if constexpr (T >= 4) {
        G.c = (v32f)(0.0f);
        G.d = (v32f)(0.0f);
    }
}

// The T tangent files of a DualFiles<T>, in channel order.
template <int T>
__device__ __forceinline__ void tangent_files(const DualFiles<T>& G, v32f (&gf)[T]) {
    gf[0] = G.a;
    if constexpr (T >= 2) gf[1] = G.b;
    if constexpr (T >= 4) {
        gf[2] = G.c;
        gf[3] = G.d;
    }
}

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

// Correctly rounded f32 square root: the same bits as __builtin_sqrtf.  The compiler's
// expansion of that (correctly rounded mode) is the hardware estimate followed by the
// choice between it and its neighbours one ulp either side by the signs of their exact
// residuals x - s' * s; around it, it scales inputs below 2^-96 (where the residuals would
// underflow) and returns +-0 and +inf unchanged, six more VALU operations on every input.
// sqrt_core is the choice alone.  It gives the expansion's bits for every x with
// |x| >= 2^-96: +inf, NaN and negative inputs come out of the estimate as they do from the
// full sequence (a negative input's NaN passes the scaling multiply unchanged).  sqrt_n
// takes it when no value of the wave is below 2^-96 in magnitude (one compare per value
// and one uniform branch), and the full sequence otherwise.
__device__ __forceinline__ float sqrt_core(float x) {
    const float s = __builtin_amdgcn_sqrtf(x);
    const uint32_t sb = __builtin_bit_cast(uint32_t, s);
    const float sd = __builtin_bit_cast(float, sb - 1u), su = __builtin_bit_cast(float, sb + 1u);
    const float rd = __builtin_fmaf(-sd, s, x), ru = __builtin_fmaf(-su, s, x);
    const float r = rd <= 0.0f ? sd : s;
    return ru > 0.0f ? su : r;
}

template <int N>
__device__ __forceinline__ void sqrt_n(float (&v)[N]) {
    uint64_t slow = 0;
    sfor<N>([&](auto P) { slow |= __builtin_amdgcn_ballot_w64(__builtin_fabsf(v[P()]) < 0x1p-96f); });
    if (slow == 0) {
        sfor<N>([&](auto P) { v[P()] = sqrt_core(v[P()]); });
    } else {
        sfor<N>([&](auto P) { v[P()] = __builtin_sqrtf(v[P()]); });
    }
}

__device__ __forceinline__ float sqrt1(float x) {
    float v[1] = {x};
    sqrt_n<1>(v);
    return v[0];
}

// The micro-op at code[i], in SGPRs (one 16-byte scalar load).
struct Uop {
    uint32_t w0, d, a, b;
};
__device__ __forceinline__ Uop load_uop(const UInstr* __restrict__ code, uint32_t i) {
    const UInstr c = kget(code, i);
    Uop u;
    u.w0 = uni(c.w0);
    u.d = uni(c.w1);
    u.a = uni(c.w2);
    u.b = uni(c.w3);
    return u;
}

// Primal interpreter at one point per lane (run_program below dispatches to it).
// `lds` is this wave's spill area: slot s, point p at lds[(s * PPL + p) * 64 + lane];
// slots from kMaxSpillSlots up live in the wave's device-memory area `gsp` (slot
// kMaxSpillSlots + g, point p at gsp[(g * PPL + p) * 64 + lane]), which only the launches
// of big programs provide (null elsewhere: such slots then read as zero, never fault).
// The same micro-op semantics as run_program, with the accumulator machinery off (the
// code is written for both and `kAcc` selects): at one point per lane a staging move or
// a write-back costs one VALU op, less than the branch that would skip it.  Operands a
// and b are always read from the file (one component; three-component operands read
// their others in their case), and results are always written to it (three-component
// results write components 1 and 2 in their case).  kFwd, kVec3, kNone and kNoWb are
// ignored, which is exact since the file then holds every value.
template <int PPL>
__device__ __forceinline__ void run_program_p1(Files<PPL>& R, const UInstr* __restrict__ code, uint32_t len,
                                               const float* __restrict__ params, float* __restrict__ lds,
                                               float* __restrict__ gsp, int lane) {
    Uop nx = load_uop(code, 0);  // the code array ends with one padding entry
    constexpr bool kAcc = false;
    float A0[PPL], A1[PPL], A2[PPL];
    sfor<PPL>([&](auto P) { A0[P()] = 0.0f; A1[P()] = 0.0f; A2[P()] = 0.0f; });
    for (uint32_t pc = 0; pc < len; ++pc) {
        const Uop u = nx;
        nx = load_uop(code, pc + 1);  // prefetch
        const uint32_t op = u.w0 & 255u, x = u.w0 >> 8;
        // The masks (no-ops on valid code) give instruction selection the known-zero high
        // bits it needs to fold the +1, +2, ... of block components into the base register
        // of one indexed-move region instead of computing a new index per component.
        const uint32_t d = u.d & 31u, a = u.a & 31u, b = u.b & 31u;
        float B0[PPL], B1[PPL], B2[PPL];
        if (kAcc && (u.b & kVec3)) {
            sfor<PPL>([&](auto P) {
                const v32f r = fget<P(), PPL>(R);
                B0[P()] = r[b];
                B1[P()] = r[b + 1];
                B2[P()] = r[b + 2];
            });
        } else if (kAcc ? !(u.b & kNone) : true) {
            sfor<PPL>([&](auto P) { B0[P()] = fget<P(), PPL>(R)[b]; });
        }
        // components 1 and 2 of a three-component operand, and the three-component result
        auto a1 = [&](auto P) {
            if constexpr (kAcc) return A1[P()];
            else return fget<P(), PPL>(R)[a + 1];
        };
        auto a2 = [&](auto P) {
            if constexpr (kAcc) return A2[P()];
            else return fget<P(), PPL>(R)[a + 2];
        };
        auto b1 = [&](auto P) {
            if constexpr (kAcc) return B1[P()];
            else return fget<P(), PPL>(R)[b + 1];
        };
        auto b2 = [&](auto P) {
            if constexpr (kAcc) return B2[P()];
            else return fget<P(), PPL>(R)[b + 2];
        };
        auto put3 = [&](auto P, float t0, float t1, float t2) {
            A0[P()] = t0;
            if constexpr (kAcc) {
                A1[P()] = t1;
                A2[P()] = t2;
            } else {
                v32f t = fget<P(), PPL>(R);
                t[d + 1] = t1;
                t[d + 2] = t2;
                fget<P(), PPL>(R) = t;
            }
        };
        // A one-component result leaves components 1 and 2 of the accumulator as they were:
        // no consumer reads them, and keeping them costs no instruction (declaring them
        // undefined made the compiler clear them with moves at every join, 5% of the
        // evaluation kernel's time on the Lorentz workload).
        auto scalar_result = [&]() {};
        if (kAcc ? !(u.a & (kFwd | kNone)) : true) {
            if (kAcc && (u.a & kVec3)) {
                sfor<PPL>([&](auto P) {
                    const v32f r = fget<P(), PPL>(R);
                    A0[P()] = r[a];
                    A1[P()] = r[a + 1];
                    A2[P()] = r[a + 2];
                });
            } else {
                sfor<PPL>([&](auto P) { A0[P()] = fget<P(), PPL>(R)[a]; });
            }
        }
        switch (op) {
            case U_IMM: {
                const float v = kld(params, x);
                sfor<PPL>([&](auto P) { A0[P()] = v; });
                scalar_result();
                break;
            }
#define RSR_CASE(OPC, EXPR)                                   \
    case OPC:                                                 \
        sfor<PPL>([&](auto P) {                               \
            const float A = A0[P()], B = B0[P()];             \
            (void)B;                                          \
            A0[P()] = (EXPR);                                 \
        });                                                   \
        scalar_result();                                      \
        break;
            RSR_CASE(U_ADD, A + B)
            RSR_CASE(U_SUB, A - B)
            RSR_CASE(U_MUL, A * B)
            RSR_CASE(U_DIV, A / B)
            RSR_CASE(U_SQR, A * A)
            RSR_CASE(U_SIN, dm_sin(A))
            RSR_CASE(U_COS, dm_cos(A))
            RSR_CASE(U_EXP, dm_exp(A))
            RSR_CASE(U_LN, dm_ln(A))
            RSR_CASE(U_NEG, -A)
            RSR_CASE(U_RCPSAFE, 1.0f / ((__builtin_fabsf(A) < 1e-9f) ? 1.0f : A))
            RSR_CASE(U_RSUB, B - A)
            RSR_CASE(U_RDIV, B / A)
#undef RSR_CASE
            case U_SQRT:
                sqrt_n<PPL>(A0);
                scalar_result();
                break;
            case U_SQRTACC:
                sfor<PPL>([&](auto P) { A0[P()] = A0[P()] + (B0[P()] * B0[P()]); });
                sqrt_n<PPL>(A0);
                scalar_result();
                break;
            // point-invariant operand: one evaluation, copied to the lane's other points
#define RSR_UCASE(OPC, EXPR)                                  \
    case OPC: {                                               \
        const float A = A0[0];                                \
        const float r0 = (EXPR);                              \
        sfor<PPL>([&](auto P) { A0[P()] = r0; });             \
        scalar_result();                                      \
        break;                                                \
    }
            RSR_UCASE(U_SINU, dm_sin(A))
            RSR_UCASE(U_COSU, dm_cos(A))
            RSR_UCASE(U_EXPU, dm_exp(A))
            RSR_UCASE(U_LNU, dm_ln(A))
            RSR_UCASE(U_SQRTU, sqrt1(A))
#undef RSR_UCASE
            // unary op of a constant: one evaluation, copied to every point of the lane
#define RSR_UKCASE(OPC, EXPR)                                 \
    case OPC: {                                               \
        const float A = kld(params, x);                       \
        const float r0 = (EXPR);                              \
        sfor<PPL>([&](auto P) { A0[P()] = r0; });             \
        scalar_result();                                      \
        break;                                                \
    }
            RSR_UKCASE(U_SINK, dm_sin(A))
            RSR_UKCASE(U_COSK, dm_cos(A))
            RSR_UKCASE(U_EXPK, dm_exp(A))
            RSR_UKCASE(U_LNK, dm_ln(A))
            RSR_UKCASE(U_SQRTK, sqrt1(A))
            RSR_UKCASE(U_NEGK, -A)
            RSR_UKCASE(U_SQRK, A * A)
#undef RSR_UKCASE
            case U_MULADD:
            case U_MULSUB: {
                const uint32_t c = x & 31u, e = (x >> 5) & 31u;
                const bool sub = op == U_MULSUB;
                sfor<PPL>([&](auto P) {
                    const v32f r = fget<P(), PPL>(R);
                    const float p1 = A0[P()] * B0[P()], p2 = r[c] * r[e];
                    A0[P()] = sub ? p1 - p2 : p1 + p2;
                });
                scalar_result();
                break;
            }
            case U_ADDMUL: {
                const uint32_t c = x & 31u;
                sfor<PPL>([&](auto P) {
                    const v32f r = fget<P(), PPL>(R);
                    A0[P()] = A0[P()] + (B0[P()] * r[c]);
                });
                scalar_result();
                break;
            }
            case U_ADD3: {
                const uint32_t c = x & 31u;
                sfor<PPL>([&](auto P) {
                    const v32f r = fget<P(), PPL>(R);
                    A0[P()] = (A0[P()] + B0[P()]) + r[c];
                });
                scalar_result();
                break;
            }
            case U_INVSEL: {
                const uint32_t c = x & 31u, fl = (x >> 10) & 3u;
                const float k = (fl & kInvOne) ? 1.0f : 0.0f;
                const bool ng = (fl & kInvNeg) != 0u;
                sfor<PPL>([&](auto P) {
                    const v32f r = fget<P(), PPL>(R);
                    const float xv = ng ? -A0[P()] : A0[P()];
                    A0[P()] = (__builtin_fabsf(r[c]) < 1e-9f) ? k : xv * B0[P()];
                });
                scalar_result();
                break;
            }
            case U_SPILL:
                if (x < (uint32_t)kMaxSpillSlots) {
                    sfor<PPL>([&](auto P) { lds[(x * PPL + P()) * 64u + spill_lane(x)] = A0[P()]; });
                } else if (gsp != nullptr) {
                    const uint32_t g = x - (uint32_t)kMaxSpillSlots;
                    sfor<PPL>([&](auto P) { gsp[(g * PPL + P()) * 64u + spill_lane(x)] = A0[P()]; });
                }
                scalar_result();
                break;
            case U_FILL:
                if (x < (uint32_t)kMaxSpillSlots) {
                    sfor<PPL>([&](auto P) { A0[P()] = lds[(x * PPL + P()) * 64u + spill_lane(x)]; });
                } else {
                    const uint32_t g = x - (uint32_t)kMaxSpillSlots;
                    if (gsp != nullptr) {
                        sfor<PPL>([&](auto P) { A0[P()] = gsp[(g * PPL + P()) * 64u + spill_lane(x)]; });
                    } else {
                        sfor<PPL>([&](auto P) { A0[P()] = 0.0f; });
                    }
                }
                scalar_result();
                break;
            // Vec3 block ops: every operand component is staged, the result replaces A
#define RSR_V3CASE(OPC, E0, E1, E2)                                             \
    case OPC:                                                                   \
        sfor<PPL>([&](auto P) {                                                 \
            const float x0 = A0[P()], x1 = a1(P), x2 = a2(P);                   \
            const float y0 = B0[P()], y1 = b1(P), y2 = b2(P);                   \
            (void)y1; (void)y2;                                                 \
            put3(P, (E0), (E1), (E2));                                          \
        });                                                                     \
        break;
            RSR_V3CASE(U_ADDV3, x0 + y0, x1 + y1, x2 + y2)
            RSR_V3CASE(U_SUBV3, x0 - y0, x1 - y1, x2 - y2)
            RSR_V3CASE(U_SCALEV3, x0 * y0, x1 * y0, x2 * y0)
            RSR_V3CASE(U_CROSS3, (x1 * y2) - (x2 * y1), (x2 * y0) - (x0 * y2), (x0 * y1) - (x1 * y0))
            RSR_V3CASE(U_RSUBV3, y0 - x0, y1 - x1, y2 - x2)
            RSR_V3CASE(U_RCROSS3, (y1 * x2) - (y2 * x1), (y2 * x0) - (y0 * x2), (y0 * x1) - (y1 * x0))
            RSR_V3CASE(U_SCALEV3S, y0 * x0, y1 * x0, y2 * x0)
#undef RSR_V3CASE
#define RSR_KCASE(OPC, EXPR)                               \
    case OPC: {                                            \
        const float K = kld(params, x);                    \
        sfor<PPL>([&](auto P) {                            \
            const float A = A0[P()];                       \
            A0[P()] = (EXPR);                              \
        });                                                \
        scalar_result();                                   \
        break;                                             \
    }
            RSR_KCASE(U_ADDK, A + K)
            RSR_KCASE(U_SUBK, A - K)
            RSR_KCASE(U_KSUB, K - A)
            RSR_KCASE(U_MULK, A * K)
            RSR_KCASE(U_DIVK, A / K)
            RSR_KCASE(U_KDIV, K / A)
#undef RSR_KCASE
            case U_SCALEV3K: {
                const float K = kld(params, x);
                sfor<PPL>([&](auto P) { put3(P, A0[P()] * K, a1(P) * K, a2(P) * K); });
                break;
            }
            case U_MULM3V3: {
                __builtin_assume(a < 24u);
                sfor<PPL>([&](auto P) {
                    const v32f r = fget<P(), PPL>(R);
                    const float v0 = B0[P()], v1 = b1(P), v2 = b2(P);
                    const float o0 = ((A0[P()] * v0) + (r[a + 3] * v1)) + (r[a + 6] * v2);
                    const float o1 = ((r[a + 1] * v0) + (r[a + 4] * v1)) +
                                     (r[a + 7] * v2);
                    const float o2 = ((r[a + 2] * v0) + (r[a + 5] * v1)) +
                                     (r[a + 8] * v2);
                    put3(P, o0, o1, o2);
                });
                break;
            }
            case U_MULM3V3R: {
                __builtin_assume(b < 24u);
                sfor<PPL>([&](auto P) {
                    const v32f r = fget<P(), PPL>(R);
                    const float v0 = A0[P()], v1 = a1(P), v2 = a2(P);
                    const float o0 = ((B0[P()] * v0) + (r[b + 3] * v1)) + (r[b + 6] * v2);
                    const float o1 = ((r[b + 1] * v0) + (r[b + 4] * v1)) +
                                     (r[b + 7] * v2);
                    const float o2 = ((r[b + 2] * v0) + (r[b + 5] * v1)) +
                                     (r[b + 8] * v2);
                    put3(P, o0, o1, o2);
                });
                break;
            }
            case U_SCALEM3:
            case U_SCALEM3K: {
                __builtin_assume(a < 24u);
                __builtin_assume(d < 24u);
                const float K = kld(params, x);
                const bool kk = op == U_SCALEM3K;
                // per point: all nine reads, then the writes (files of different points are disjoint)
                sfor<PPL>([&](auto P) {
                    v32f t = fget<P(), PPL>(R);
                    const float sv = kk ? K : B0[P()];
                    float rr[8];
#pragma unroll
                    for (int q = 1; q < 9; ++q) rr[q - 1] = t[a + q] * sv;
                    A0[P()] = A0[P()] * sv;
#pragma unroll
                    for (int q = 1; q < 9; ++q) t[d + q] = rr[q - 1];
                    fget<P(), PPL>(R) = t;
                });
                scalar_result();
                break;
            }
            // Mat2 / four-component block ops: components 1..3 are read and written here,
            // component 0 goes through the accumulator
#define RSR_V4CASE(OPC, E0, E1, E2, E3, READB)                                  \
    case OPC: {                                                                 \
        __builtin_assume(a < 29u);                                              \
        __builtin_assume(b < 29u);                                              \
        __builtin_assume(d < 29u);                                              \
        const float K = kld(params, x);                                         \
        (void)K;                                                                \
        sfor<PPL>([&](auto P) {                                                 \
            v32f t = fget<P(), PPL>(R);                                         \
            const float a0 = A0[P()], a1 = t[a + 1], a2 = t[a + 2], a3 = t[a + 3]; \
            const float b0 = B0[P()];                                           \
            float b1 = 0.0f, b2 = 0.0f, b3 = 0.0f;                              \
            if (READB) { b1 = t[b + 1]; b2 = t[b + 2]; b3 = t[b + 3]; }         \
            (void)b0; (void)b1; (void)b2; (void)b3;                             \
            const float o0 = (E0), o1 = (E1), o2 = (E2), o3 = (E3);             \
            A0[P()] = o0;                                                       \
            t[d + 1] = o1;                                                      \
            t[d + 2] = o2;                                                      \
            t[d + 3] = o3;                                                      \
            fget<P(), PPL>(R) = t;                                              \
        });                                                                     \
        scalar_result();                                                        \
        break;                                                                  \
    }
            RSR_V4CASE(U_ADDV4, a0 + b0, a1 + b1, a2 + b2, a3 + b3, true)
            RSR_V4CASE(U_SUBV4, a0 - b0, a1 - b1, a2 - b2, a3 - b3, true)
            RSR_V4CASE(U_SCALEV4, a0 * b0, a1 * b0, a2 * b0, a3 * b0, false)
            RSR_V4CASE(U_SCALEV4K, a0 * K, a1 * K, a2 * K, a3 * K, false)
            RSR_V4CASE(U_MULM2, (a0 * b0) + (a2 * b1), (a1 * b0) + (a3 * b1), (a0 * b2) + (a2 * b3),
                       (a1 * b2) + (a3 * b3), true)
#undef RSR_V4CASE
            case U_INVM2: {
                __builtin_assume(a < 29u);
                __builtin_assume(d < 29u);
                sfor<PPL>([&](auto P) {
                    v32f t = fget<P(), PPL>(R);
                    const float a0 = A0[P()], a1 = t[a + 1], a2 = t[a + 2], a3 = t[a + 3];
                    const float det = (a0 * a3) - (a1 * a2);
                    const bool sing = __builtin_fabsf(det) < 1e-9f;
                    const float inv = 1.0f / (sing ? 1.0f : det);
                    A0[P()] = sing ? 1.0f : a3 * inv;
                    t[d + 1] = sing ? 0.0f : (-a1) * inv;
                    t[d + 2] = sing ? 0.0f : (-a2) * inv;
                    t[d + 3] = sing ? 1.0f : a0 * inv;
                    fget<P(), PPL>(R) = t;
                });
                scalar_result();
                break;
            }
            case U_IMMV3: {
                const float v0 = kld(params, x), v1 = kld(params, x + 1), v2 = kld(params, x + 2);
                sfor<PPL>([&](auto P) { put3(P, v0, v1, v2); });
                break;
            }
            case U_DOT3:
                sfor<PPL>([&](auto P) {
                    A0[P()] = ((A0[P()] * B0[P()]) + (a1(P) * b1(P))) + (a2(P) * b2(P));
                });
                scalar_result();
                break;
            case U_NORM3:
                sfor<PPL>([&](auto P) {
                    const float x0 = A0[P()], x1 = a1(P), x2 = a2(P);
                    A0[P()] = ((x0 * x0) + (x1 * x1)) + (x2 * x2);
                });
                sqrt_n<PPL>(A0);
                scalar_result();
                break;
            default:
                sfor<PPL>([&](auto P) { A0[P()] = 0.0f; });
                scalar_result();
                break;
        }
        // The write-back index also depends on the a and b words (their bits 16 and up
        // are zero), which keeps them live to the end of the iteration: otherwise the
        // compiler moves the prefetched next words into their registers right after the
        // staging, which waits for the prefetch and exposes its latency on every op.
        const uint32_t dw = (d + ((u.a | u.b) >> 16)) & 31u;
        if (kAcc ? !(u.d & kNoWb) : true) {
            if (kAcc && (u.d & kVec3)) {
                sfor<PPL>([&](auto P) {
                    v32f t = fget<P(), PPL>(R);
                    t[dw] = A0[P()];
                    t[dw + 1] = A1[P()];
                    t[dw + 2] = A2[P()];
                    fget<P(), PPL>(R) = t;
                });
            } else {
                sfor<PPL>([&](auto P) {
                    v32f t = fget<P(), PPL>(R);
                    t[dw] = A0[P()];
                    fget<P(), PPL>(R) = t;
                });
            }
        }
    }
}

// Primal interpreter over PPL register files (one per point of the lane).
// `lds` is this wave's spill area: slot s, point p at lds[(s * PPL + p) * 64 + lane].
// The interpreter is an accumulator machine over the register file: A0..A2 hold the
// last result (one or three components) across iterations.  Operand a is staged into
// them unless it is that result (kFwd), operand b into B0..B2, both before the dispatch
// and one indexed-read group per operand; the switch computes the new result into A;
// the write-back after it stores A into the file unless kNoWb.  Results with four or
// nine components write their components 1.. to the file inside the switch.  After a
// one-component result A1 and A2 are dead (no consumer reads them).
template <int PPL>
__device__ __forceinline__ void run_program(Files<PPL>& R, const UInstr* __restrict__ code, uint32_t len,
                                            const float* __restrict__ params, float* __restrict__ lds,
                                            float* __restrict__ gsp, int lane) {
    if constexpr (PPL == 1) {
        run_program_p1<PPL>(R, code, len, params, lds, gsp, lane);
        return;
    }
    Uop nx = load_uop(code, 0);  // the code array ends with one padding entry
    float A0[PPL], A1[PPL], A2[PPL];
    sfor<PPL>([&](auto P) { A0[P()] = 0.0f; A1[P()] = 0.0f; A2[P()] = 0.0f; });
    for (uint32_t pc = 0; pc < len; ++pc) {
        const Uop u = nx;
        nx = load_uop(code, pc + 1);  // prefetch
        const uint32_t op = u.w0 & 255u, x = u.w0 >> 8;
        // The masks (no-ops on valid code) give instruction selection the known-zero high
        // bits it needs to fold the +1, +2, ... of block components into the base register
        // of one indexed-move region instead of computing a new index per component.
        const uint32_t d = u.d & 31u, a = u.a & 31u, b = u.b & 31u;
        float B0[PPL], B1[PPL], B2[PPL];
        if (u.b & kVec3) {
            sfor<PPL>([&](auto P) {
                const v32f r = fget<P(), PPL>(R);
                B0[P()] = r[b];
                B1[P()] = r[b + 1];
                B2[P()] = r[b + 2];
            });
        } else if (!(u.b & kNone)) {
            sfor<PPL>([&](auto P) { B0[P()] = fget<P(), PPL>(R)[b]; });
        }
        // A one-component result leaves components 1 and 2 of the accumulator as they were:
        // no consumer reads them, and keeping them costs no instruction (declaring them
        // undefined made the compiler clear them with moves at every join, 5% of the
        // evaluation kernel's time on the Lorentz workload).
        auto scalar_result = [&]() {};
        if (!(u.a & (kFwd | kNone))) {
            if (u.a & kVec3) {
                sfor<PPL>([&](auto P) {
                    const v32f r = fget<P(), PPL>(R);
                    A0[P()] = r[a];
                    A1[P()] = r[a + 1];
                    A2[P()] = r[a + 2];
                });
            } else {
                sfor<PPL>([&](auto P) { A0[P()] = fget<P(), PPL>(R)[a]; });
            }
        }
        switch (op) {
            case U_IMM: {
                const float v = kld(params, x);
                sfor<PPL>([&](auto P) { A0[P()] = v; });
                scalar_result();
                break;
            }
#define RSR_CASE(OPC, EXPR)                                   \
    case OPC:                                                 \
        sfor<PPL>([&](auto P) {                               \
            const float A = A0[P()], B = B0[P()];             \
            (void)B;                                          \
            A0[P()] = (EXPR);                                 \
        });                                                   \
        scalar_result();                                      \
        break;
            RSR_CASE(U_ADD, A + B)
            RSR_CASE(U_SUB, A - B)
            RSR_CASE(U_MUL, A * B)
            RSR_CASE(U_DIV, A / B)
            RSR_CASE(U_SQR, A * A)
            RSR_CASE(U_SIN, dm_sin(A))
            RSR_CASE(U_COS, dm_cos(A))
            RSR_CASE(U_EXP, dm_exp(A))
            RSR_CASE(U_LN, dm_ln(A))
            RSR_CASE(U_NEG, -A)
            RSR_CASE(U_RCPSAFE, 1.0f / ((__builtin_fabsf(A) < 1e-9f) ? 1.0f : A))
            RSR_CASE(U_RSUB, B - A)
            RSR_CASE(U_RDIV, B / A)
#undef RSR_CASE
            case U_SQRT:
                sqrt_n<PPL>(A0);
                scalar_result();
                break;
            case U_SQRTACC:
                sfor<PPL>([&](auto P) { A0[P()] = A0[P()] + (B0[P()] * B0[P()]); });
                sqrt_n<PPL>(A0);
                scalar_result();
                break;
            // point-invariant operand: one evaluation, copied to the lane's other points
#define RSR_UCASE(OPC, EXPR)                                  \
    case OPC: {                                               \
        const float A = A0[0];                                \
        const float r0 = (EXPR);                              \
        sfor<PPL>([&](auto P) { A0[P()] = r0; });             \
        scalar_result();                                      \
        break;                                                \
    }
            RSR_UCASE(U_SINU, dm_sin(A))
            RSR_UCASE(U_COSU, dm_cos(A))
            RSR_UCASE(U_EXPU, dm_exp(A))
            RSR_UCASE(U_LNU, dm_ln(A))
            RSR_UCASE(U_SQRTU, sqrt1(A))
#undef RSR_UCASE
            // unary op of a constant: one evaluation, copied to every point of the lane
#define RSR_UKCASE(OPC, EXPR)                                 \
    case OPC: {                                               \
        const float A = kld(params, x);                       \
        const float r0 = (EXPR);                              \
        sfor<PPL>([&](auto P) { A0[P()] = r0; });             \
        scalar_result();                                      \
        break;                                                \
    }
            RSR_UKCASE(U_SINK, dm_sin(A))
            RSR_UKCASE(U_COSK, dm_cos(A))
            RSR_UKCASE(U_EXPK, dm_exp(A))
            RSR_UKCASE(U_LNK, dm_ln(A))
            RSR_UKCASE(U_SQRTK, sqrt1(A))
            RSR_UKCASE(U_NEGK, -A)
            RSR_UKCASE(U_SQRK, A * A)
#undef RSR_UKCASE
            case U_MULADD:
            case U_MULSUB: {
                const uint32_t c = x & 31u, e = (x >> 5) & 31u;
                const bool sub = op == U_MULSUB;
                sfor<PPL>([&](auto P) {
                    const v32f r = fget<P(), PPL>(R);
                    const float p1 = A0[P()] * B0[P()], p2 = r[c] * r[e];
                    A0[P()] = sub ? p1 - p2 : p1 + p2;
                });
                scalar_result();
                break;
            }
            case U_ADDMUL: {
                const uint32_t c = x & 31u;
                sfor<PPL>([&](auto P) {
                    const v32f r = fget<P(), PPL>(R);
                    A0[P()] = A0[P()] + (B0[P()] * r[c]);
                });
                scalar_result();
                break;
            }
            case U_ADD3: {
                const uint32_t c = x & 31u;
                sfor<PPL>([&](auto P) {
                    const v32f r = fget<P(), PPL>(R);
                    A0[P()] = (A0[P()] + B0[P()]) + r[c];
                });
                scalar_result();
                break;
            }
            case U_INVSEL: {
                const uint32_t c = x & 31u, fl = (x >> 10) & 3u;
                const float k = (fl & kInvOne) ? 1.0f : 0.0f;
                const bool ng = (fl & kInvNeg) != 0u;
                sfor<PPL>([&](auto P) {
                    const v32f r = fget<P(), PPL>(R);
                    const float xv = ng ? -A0[P()] : A0[P()];
                    A0[P()] = (__builtin_fabsf(r[c]) < 1e-9f) ? k : xv * B0[P()];
                });
                scalar_result();
                break;
            }
            case U_SPILL:
                if (x < (uint32_t)kMaxSpillSlots) {
                    sfor<PPL>([&](auto P) { lds[(x * PPL + P()) * 64u + spill_lane(x)] = A0[P()]; });
                } else if (gsp != nullptr) {
                    const uint32_t g = x - (uint32_t)kMaxSpillSlots;
                    sfor<PPL>([&](auto P) { gsp[(g * PPL + P()) * 64u + spill_lane(x)] = A0[P()]; });
                }
                scalar_result();
                break;
            case U_FILL:
                if (x < (uint32_t)kMaxSpillSlots) {
                    sfor<PPL>([&](auto P) { A0[P()] = lds[(x * PPL + P()) * 64u + spill_lane(x)]; });
                } else {
                    const uint32_t g = x - (uint32_t)kMaxSpillSlots;
                    if (gsp != nullptr) {
                        sfor<PPL>([&](auto P) { A0[P()] = gsp[(g * PPL + P()) * 64u + spill_lane(x)]; });
                    } else {
                        sfor<PPL>([&](auto P) { A0[P()] = 0.0f; });
                    }
                }
                scalar_result();
                break;
            // Vec3 block ops: every operand component is staged, the result replaces A
#define RSR_V3CASE(OPC, E0, E1, E2)                                             \
    case OPC:                                                                   \
        sfor<PPL>([&](auto P) {                                                 \
            const float a0 = A0[P()], a1 = A1[P()], a2 = A2[P()];               \
            const float b0 = B0[P()], b1 = B1[P()], b2 = B2[P()];               \
            (void)b1; (void)b2;                                                 \
            const float t0 = (E0), t1 = (E1), t2 = (E2);                        \
            A0[P()] = t0;                                                       \
            A1[P()] = t1;                                                       \
            A2[P()] = t2;                                                       \
        });                                                                     \
        break;
            RSR_V3CASE(U_ADDV3, a0 + b0, a1 + b1, a2 + b2)
            RSR_V3CASE(U_SUBV3, a0 - b0, a1 - b1, a2 - b2)
            RSR_V3CASE(U_SCALEV3, a0 * b0, a1 * b0, a2 * b0)
            RSR_V3CASE(U_CROSS3, (a1 * b2) - (a2 * b1), (a2 * b0) - (a0 * b2), (a0 * b1) - (a1 * b0))
            RSR_V3CASE(U_RSUBV3, b0 - a0, b1 - a1, b2 - a2)
            RSR_V3CASE(U_RCROSS3, (b1 * a2) - (b2 * a1), (b2 * a0) - (b0 * a2), (b0 * a1) - (b1 * a0))
            RSR_V3CASE(U_SCALEV3S, b0 * a0, b1 * a0, b2 * a0)
#undef RSR_V3CASE
#define RSR_KCASE(OPC, EXPR)                               \
    case OPC: {                                            \
        const float K = kld(params, x);                    \
        sfor<PPL>([&](auto P) {                            \
            const float A = A0[P()];                       \
            A0[P()] = (EXPR);                              \
        });                                                \
        scalar_result();                                   \
        break;                                             \
    }
            RSR_KCASE(U_ADDK, A + K)
            RSR_KCASE(U_SUBK, A - K)
            RSR_KCASE(U_KSUB, K - A)
            RSR_KCASE(U_MULK, A * K)
            RSR_KCASE(U_DIVK, A / K)
            RSR_KCASE(U_KDIV, K / A)
#undef RSR_KCASE
            case U_SCALEV3K: {
                const float K = kld(params, x);
                sfor<PPL>([&](auto P) {
                    A0[P()] = A0[P()] * K;
                    A1[P()] = A1[P()] * K;
                    A2[P()] = A2[P()] * K;
                });
                break;
            }
            case U_MULM3V3: {
                __builtin_assume(a < 24u);
                sfor<PPL>([&](auto P) {
                    const v32f r = fget<P(), PPL>(R);
                    const float v0 = B0[P()], v1 = B1[P()], v2 = B2[P()];
                    const float o0 = ((A0[P()] * v0) + (r[a + 3] * v1)) + (r[a + 6] * v2);
                    const float o1 = ((r[a + 1] * v0) + (r[a + 4] * v1)) +
                                     (r[a + 7] * v2);
                    const float o2 = ((r[a + 2] * v0) + (r[a + 5] * v1)) +
                                     (r[a + 8] * v2);
                    A0[P()] = o0;
                    A1[P()] = o1;
                    A2[P()] = o2;
                });
                break;
            }
            case U_MULM3V3R: {
                __builtin_assume(b < 24u);
                sfor<PPL>([&](auto P) {
                    const v32f r = fget<P(), PPL>(R);
                    const float v0 = A0[P()], v1 = A1[P()], v2 = A2[P()];
                    const float o0 = ((B0[P()] * v0) + (r[b + 3] * v1)) + (r[b + 6] * v2);
                    const float o1 = ((r[b + 1] * v0) + (r[b + 4] * v1)) +
                                     (r[b + 7] * v2);
                    const float o2 = ((r[b + 2] * v0) + (r[b + 5] * v1)) +
                                     (r[b + 8] * v2);
                    A0[P()] = o0;
                    A1[P()] = o1;
                    A2[P()] = o2;
                });
                break;
            }
            case U_SCALEM3:
            case U_SCALEM3K: {
                __builtin_assume(a < 24u);
                __builtin_assume(d < 24u);
                const float K = kld(params, x);
                const bool kk = op == U_SCALEM3K;
                // per point: all nine reads, then the writes (files of different points are disjoint)
                sfor<PPL>([&](auto P) {
                    v32f t = fget<P(), PPL>(R);
                    const float sv = kk ? K : B0[P()];
                    float rr[8];
#pragma unroll
                    for (int q = 1; q < 9; ++q) rr[q - 1] = t[a + q] * sv;
                    A0[P()] = A0[P()] * sv;
#pragma unroll
                    for (int q = 1; q < 9; ++q) t[d + q] = rr[q - 1];
                    fget<P(), PPL>(R) = t;
                });
                scalar_result();
                break;
            }
            // Mat2 / four-component block ops: components 1..3 are read and written here,
            // component 0 goes through the accumulator
#define RSR_V4CASE(OPC, E0, E1, E2, E3, READB)                                  \
    case OPC: {                                                                 \
        __builtin_assume(a < 29u);                                              \
        __builtin_assume(b < 29u);                                              \
        __builtin_assume(d < 29u);                                              \
        const float K = kld(params, x);                                         \
        (void)K;                                                                \
        sfor<PPL>([&](auto P) {                                                 \
            v32f t = fget<P(), PPL>(R);                                         \
            const float a0 = A0[P()], a1 = t[a + 1], a2 = t[a + 2], a3 = t[a + 3]; \
            const float b0 = B0[P()];                                           \
            float b1 = 0.0f, b2 = 0.0f, b3 = 0.0f;                              \
            if (READB) { b1 = t[b + 1]; b2 = t[b + 2]; b3 = t[b + 3]; }         \
            (void)b0; (void)b1; (void)b2; (void)b3;                             \
            const float o0 = (E0), o1 = (E1), o2 = (E2), o3 = (E3);             \
            A0[P()] = o0;                                                       \
            t[d + 1] = o1;                                                      \
            t[d + 2] = o2;                                                      \
            t[d + 3] = o3;                                                      \
            fget<P(), PPL>(R) = t;                                              \
        });                                                                     \
        scalar_result();                                                        \
        break;                                                                  \
    }
            RSR_V4CASE(U_ADDV4, a0 + b0, a1 + b1, a2 + b2, a3 + b3, true)
            RSR_V4CASE(U_SUBV4, a0 - b0, a1 - b1, a2 - b2, a3 - b3, true)
            RSR_V4CASE(U_SCALEV4, a0 * b0, a1 * b0, a2 * b0, a3 * b0, false)
            RSR_V4CASE(U_SCALEV4K, a0 * K, a1 * K, a2 * K, a3 * K, false)
            RSR_V4CASE(U_MULM2, (a0 * b0) + (a2 * b1), (a1 * b0) + (a3 * b1), (a0 * b2) + (a2 * b3),
                       (a1 * b2) + (a3 * b3), true)
#undef RSR_V4CASE
            case U_INVM2: {
                __builtin_assume(a < 29u);
                __builtin_assume(d < 29u);
                sfor<PPL>([&](auto P) {
                    v32f t = fget<P(), PPL>(R);
                    const float a0 = A0[P()], a1 = t[a + 1], a2 = t[a + 2], a3 = t[a + 3];
                    const float det = (a0 * a3) - (a1 * a2);
                    const bool sing = __builtin_fabsf(det) < 1e-9f;
                    const float inv = 1.0f / (sing ? 1.0f : det);
                    A0[P()] = sing ? 1.0f : a3 * inv;
                    t[d + 1] = sing ? 0.0f : (-a1) * inv;
                    t[d + 2] = sing ? 0.0f : (-a2) * inv;
                    t[d + 3] = sing ? 1.0f : a0 * inv;
                    fget<P(), PPL>(R) = t;
                });
                scalar_result();
                break;
            }
            case U_IMMV3: {
                const float v0 = kld(params, x), v1 = kld(params, x + 1), v2 = kld(params, x + 2);
                sfor<PPL>([&](auto P) {
                    A0[P()] = v0;
                    A1[P()] = v1;
                    A2[P()] = v2;
                });
                break;
            }
            case U_DOT3:
                sfor<PPL>([&](auto P) {
                    A0[P()] = ((A0[P()] * B0[P()]) + (A1[P()] * B1[P()])) + (A2[P()] * B2[P()]);
                });
                scalar_result();
                break;
            case U_NORM3:
                sfor<PPL>([&](auto P) {
                    const float a0 = A0[P()], a1 = A1[P()], a2 = A2[P()];
                    A0[P()] = ((a0 * a0) + (a1 * a1)) + (a2 * a2);
                });
                sqrt_n<PPL>(A0);
                scalar_result();
                break;
            default:
                sfor<PPL>([&](auto P) { A0[P()] = 0.0f; });
                scalar_result();
                break;
        }
        // The write-back index also depends on the a and b words (their bits 16 and up
        // are zero), which keeps them live to the end of the iteration: otherwise the
        // compiler moves the prefetched next words into their registers right after the
        // staging, which waits for the prefetch and exposes its latency on every op.
        const uint32_t dw = (d + ((u.a | u.b) >> 16)) & 31u;
        if (!(u.d & kNoWb)) {
            if (u.d & kVec3) {
                sfor<PPL>([&](auto P) {
                    v32f t = fget<P(), PPL>(R);
                    t[dw] = A0[P()];
                    t[dw + 1] = A1[P()];
                    t[dw + 2] = A2[P()];
                    fget<P(), PPL>(R) = t;
                });
            } else {
                sfor<PPL>([&](auto P) {
                    v32f t = fget<P(), PPL>(R);
                    t[dw] = A0[P()];
                    fget<P(), PPL>(R) = t;
                });
            }
        }
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
// Spill slot s, channel ch (0 = value, 1 + j = tangent j) at lds[(s * (1 + T) + ch) * 64 + lane]
// for s < kMaxSpillSlots, else at gsp[((s - kMaxSpillSlots) * (1 + T) + ch) * 64 + lane].
// As in the primal interpreter, operands a and b of every channel are read before the
// dispatch and the results written after it.
template <int T>
__device__ __forceinline__ void run_dual(DualFiles<T>& G, const UInstr* __restrict__ code, uint32_t len,
                                         const float* __restrict__ params, uint32_t k0, float* __restrict__ lds,
                                         float* __restrict__ gsp, int lane) {
    Uop nx = load_uop(code, 0);
    for (uint32_t pc = 0; pc < len; ++pc) {
        const Uop u = nx;
        nx = load_uop(code, pc + 1);
        const uint32_t op = u.w0 & 255u, x = u.w0 >> 8;
        // The masks (no-ops on valid code) give instruction selection the known-zero high
        // bits it needs to fold the +1, +2, ... of block components into the base register
        // of one indexed-move region instead of computing a new index per component.
        const uint32_t d = u.d & 31u, a = u.a & 31u, b = u.b & 31u;
        float va, vb, ga[T], gb[T];
        {
            const v32f Vv = G.v;
            va = Vv[a];
            RSR_TAN(ga[j] = Gj[a]);
            vb = Vv[b];
            RSR_TAN(gb[j] = Gj[b]);
        }
        float rv;
        float rg[T];
        // components 1.. of a Vec3 / Mat2 result, in the value file and every tangent file
        auto put3 = [&](float v1, float v2, const float (&g1)[T], const float (&g2)[T]) {
            { v32f t = G.v; t[d + 1] = v1; t[d + 2] = v2; G.v = t; }
            { v32f t = G.a; t[d + 1] = g1[0]; t[d + 2] = g2[0]; G.a = t; }
            if constexpr (T >= 2) { v32f t = G.b; t[d + 1] = g1[1]; t[d + 2] = g2[1]; G.b = t; }
            if constexpr (T >= 4) {
                { v32f t = G.c; t[d + 1] = g1[2]; t[d + 2] = g2[2]; G.c = t; }
                { v32f t = G.d; t[d + 1] = g1[3]; t[d + 2] = g2[3]; G.d = t; }
            }
        };
        auto put4 = [&](const float (&out)[4], const float (&og)[4][T]) {
            { v32f t = G.v; t[d + 1] = out[1]; t[d + 2] = out[2]; t[d + 3] = out[3]; G.v = t; }
            { v32f t = G.a; t[d + 1] = og[1][0]; t[d + 2] = og[2][0]; t[d + 3] = og[3][0]; G.a = t; }
            if constexpr (T >= 2) { v32f t = G.b; t[d + 1] = og[1][1]; t[d + 2] = og[2][1]; t[d + 3] = og[3][1]; G.b = t; }
            if constexpr (T >= 4) {
                { v32f t = G.c; t[d + 1] = og[1][2]; t[d + 2] = og[2][2]; t[d + 3] = og[3][2]; G.c = t; }
                { v32f t = G.d; t[d + 1] = og[1][3]; t[d + 2] = og[2][3]; t[d + 3] = og[3][3]; G.d = t; }
            }
        };
        switch (op) {
            case U_IMM:
                rv = kld(params, x);
                RSR_TAN(rg[j] = (x == k0 + (uint32_t)j) ? 1.0f : 0.0f);
                break;
            case U_ADD:
                rv = va + vb;
                RSR_TAN(rg[j] = ga[j] + gb[j]);
                break;
            case U_SUB:
                rv = va - vb;
                RSR_TAN(rg[j] = ga[j] - gb[j]);
                break;
            case U_MUL:
                rv = va * vb;
                RSR_TAN(rg[j] = (va * gb[j]) + (ga[j] * vb));
                break;
            case U_DIV: {
                rv = va / vb;
                const float yy = vb * vb;
                RSR_TAN(rg[j] = ((ga[j] * vb) - (va * gb[j])) / yy);
                break;
            }
            case U_SQR: {
                rv = va * va;
                const float x2 = 2.0f * va;
                RSR_TAN(rg[j] = x2 * ga[j]);
                break;
            }
            case U_SQRT: {
                const float sq = sqrt1(va);
                rv = sq;
                const float s2 = 2.0f * sq;
                RSR_TAN(rg[j] = ga[j] / s2);
                break;
            }
            case U_SIN: {
                float cs;
                dm_sincos(va, rv, cs);
                RSR_TAN(rg[j] = ga[j] * cs);
                break;
            }
            case U_COS: {
                float sn;
                dm_sincos(va, sn, rv);
                const float ns = -sn;
                RSR_TAN(rg[j] = ga[j] * ns);
                break;
            }
            case U_EXP: {
                const float ex = dm_exp(va);
                rv = ex;
                RSR_TAN(rg[j] = ga[j] * ex);
                break;
            }
            case U_LN:
                rv = dm_ln(va);
                RSR_TAN(rg[j] = ga[j] / va);
                break;
            case U_NEG:
                rv = -va;
                RSR_TAN(rg[j] = -ga[j]);
                break;
            case U_MULADD:
            case U_MULSUB: {
                const uint32_t c = x & 31u, e = (x >> 5) & 31u;
                const v32f Vv = G.v;
                const float xc = Vv[c], xe = Vv[e];
                const bool sub = op == U_MULSUB;
                const float p1 = va * vb, p2 = xc * xe;
                rv = sub ? p1 - p2 : p1 + p2;
                RSR_TAN(const float t1 = (va * gb[j]) + (ga[j] * vb); const float t2 = (xc * Gj[e]) + (Gj[c] * xe);
                        rg[j] = sub ? t1 - t2 : t1 + t2);
                break;
            }
            case U_ADDMUL: {
                const uint32_t c = x & 31u;
                const v32f Vv = G.v;
                const float xc = Vv[c];
                rv = va + (vb * xc);
                RSR_TAN(rg[j] = ga[j] + ((vb * Gj[c]) + (gb[j] * xc)));
                break;
            }
            case U_SQRTACC: {
                const float sq = sqrt1(va + (vb * vb));
                rv = sq;
                const float s2 = 2.0f * sq;
                RSR_TAN(rg[j] = (ga[j] + ((vb * gb[j]) + (gb[j] * vb))) / s2);
                break;
            }
            case U_ADD3: {
                const uint32_t c = x & 31u;
                const v32f Vv = G.v;
                rv = (va + vb) + Vv[c];
                RSR_TAN(rg[j] = (ga[j] + gb[j]) + Gj[c]);
                break;
            }
            case U_SPILL: {
                // slots below kMaxSpillSlots in LDS, the others in the wave's device area
                const bool in_lds = x < (uint32_t)kMaxSpillSlots;
                float* base = in_lds ? lds : gsp;
                const uint32_t sx = in_lds ? x : x - (uint32_t)kMaxSpillSlots;
                rv = va;
                RSR_TAN(rg[j] = ga[j]);
                if (base != nullptr) {
                    base[(sx * (1 + T)) * 64u + lane] = rv;
                    RSR_TAN(base[(sx * (1 + T) + 1 + j) * 64u + lane] = rg[j]);
                }
                break;
            }
            case U_FILL: {
                const bool in_lds = x < (uint32_t)kMaxSpillSlots;
                const float* base = in_lds ? lds : gsp;
                const uint32_t sx = in_lds ? x : x - (uint32_t)kMaxSpillSlots;
                if (base != nullptr) {
                    rv = base[(sx * (1 + T)) * 64u + lane];
                    RSR_TAN(rg[j] = base[(sx * (1 + T) + 1 + j) * 64u + lane]);
                } else {
                    rv = 0.0f;
                    RSR_TAN(rg[j] = 0.0f);
                }
                break;
            }
            case U_DOT3: {
                __builtin_assume(a < 30u);
                __builtin_assume(b < 30u);
                const v32f Vv = G.v;
                const float a1 = Vv[a + 1], a2 = Vv[a + 2], b1 = Vv[b + 1], b2 = Vv[b + 2];
                rv = ((va * vb) + (a1 * b1)) + (a2 * b2);
                RSR_TAN(rg[j] = (((va * gb[j]) + (ga[j] * vb)) + ((a1 * Gj[b + 1]) + (Gj[a + 1] * b1))) +
                                ((a2 * Gj[b + 2]) + (Gj[a + 2] * b2)));
                break;
            }
            case U_NORM3: {
                __builtin_assume(a < 30u);
                const v32f Vv = G.v;
                const float a1 = Vv[a + 1], a2 = Vv[a + 2];
                const float sq = sqrt1(((va * va) + (a1 * a1)) + (a2 * a2));
                rv = sq;
                const float s2 = 2.0f * sq;
                RSR_TAN(rg[j] = ((((va * ga[j]) + (ga[j] * va)) + ((a1 * Gj[a + 1]) + (Gj[a + 1] * a1))) +
                                 ((a2 * Gj[a + 2]) + (Gj[a + 2] * a2))) / s2);
                break;
            }
            case U_ADDK: case U_SUBK: case U_KSUB: case U_MULK: case U_DIVK: case U_KDIV: {
                const float K = kld(params, x);
                float gk[T];
                RSR_TAN(gk[j] = (x == k0 + (uint32_t)j) ? 1.0f : 0.0f);
                if (op == U_ADDK) {
                    rv = va + K;
                    RSR_TAN(rg[j] = ga[j] + gk[j]);
                } else if (op == U_SUBK) {
                    rv = va - K;
                    RSR_TAN(rg[j] = ga[j] - gk[j]);
                } else if (op == U_KSUB) {
                    rv = K - va;
                    RSR_TAN(rg[j] = gk[j] - ga[j]);
                } else if (op == U_MULK) {
                    rv = va * K;
                    RSR_TAN(rg[j] = (va * gk[j]) + (ga[j] * K));
                } else if (op == U_DIVK) {
                    rv = va / K;
                    const float kk = K * K;
                    RSR_TAN(rg[j] = ((ga[j] * K) - (va * gk[j])) / kk);
                } else {
                    rv = K / va;
                    const float aa = va * va;
                    RSR_TAN(rg[j] = ((gk[j] * va) - (K * ga[j])) / aa);
                }
                break;
            }
            case U_MULM3V3: {
                __builtin_assume(a < 24u);
                __builtin_assume(b < 30u);
                __builtin_assume(d < 30u);
                const v32f Vv = G.v;
                float out[3], og[3][T];
#pragma unroll
                for (int i = 0; i < 3; ++i) {
                    const float m0 = Vv[a + i], m1 = Vv[a + 3 + i], m2 = Vv[a + 6 + i];
                    const float v0 = Vv[b], v1 = Vv[b + 1], v2 = Vv[b + 2];
                    const float t = (m0 * v0) + (m1 * v1);
                    out[i] = t + (m2 * v2);
                    RSR_TAN(const float tg = ((m0 * Gj[b]) + (Gj[a + i] * v0)) + ((m1 * Gj[b + 1]) + (Gj[a + 3 + i] * v1));
                            og[i][j] = tg + ((m2 * Gj[b + 2]) + (Gj[a + 6 + i] * v2)));
                }
                rv = out[0];
                RSR_TAN(rg[j] = og[0][j]);
                {
                    v32f t = G.v;
                    t[d + 1] = out[1];
                    t[d + 2] = out[2];
                    G.v = t;
                }
                { v32f t = G.a; t[d + 1] = og[1][0]; t[d + 2] = og[2][0]; G.a = t; }
                if constexpr (T >= 2) { v32f t = G.b; t[d + 1] = og[1][1]; t[d + 2] = og[2][1]; G.b = t; }
                if constexpr (T >= 4) {
                    { v32f t = G.c; t[d + 1] = og[1][2]; t[d + 2] = og[2][2]; G.c = t; }
                    { v32f t = G.d; t[d + 1] = og[1][3]; t[d + 2] = og[2][3]; G.d = t; }
                }
                break;
            }
            case U_SCALEM3:
            case U_SCALEM3K: {
                __builtin_assume(a < 24u);
                __builtin_assume(d < 24u);
                const v32f Vv = G.v;
                const bool kk = op == U_SCALEM3K;
                const float sv = kk ? kld(params, x) : vb;
                float gs[T];
                RSR_TAN(gs[j] = kk ? ((x == k0 + (uint32_t)j) ? 1.0f : 0.0f) : gb[j]);
                float out[9], og[9][T];
#pragma unroll
                for (int q = 0; q < 9; ++q) {
                    const float m = Vv[a + q];
                    out[q] = m * sv;
                    RSR_TAN(og[q][j] = (m * gs[j]) + (Gj[a + q] * sv));
                }
                rv = out[0];
                RSR_TAN(rg[j] = og[0][j]);
                {
                    v32f t = G.v;
#pragma unroll
                    for (int q = 1; q < 9; ++q) t[d + q] = out[q];
                    G.v = t;
                }
                {
                    v32f t = G.a;
#pragma unroll
                    for (int q = 1; q < 9; ++q) t[d + q] = og[q][0];
                    G.a = t;
                }
                if constexpr (T >= 2) {
                    v32f t = G.b;
#pragma unroll
                    for (int q = 1; q < 9; ++q) t[d + q] = og[q][1];
                    G.b = t;
                }
                if constexpr (T >= 4) {
                    {
                        v32f t = G.c;
#pragma unroll
                        for (int q = 1; q < 9; ++q) t[d + q] = og[q][2];
                        G.c = t;
                    }
                    {
                        v32f t = G.d;
#pragma unroll
                        for (int q = 1; q < 9; ++q) t[d + q] = og[q][3];
                        G.d = t;
                    }
                }
                break;
            }
            // Block ops write their components 1.. inside the case.  Ops whose operand b is
            // a block (b + k inside the file) and ops whose b is a scalar (any register,
            // 30 and 31 included) are separate cases, so that no component index is ever
            // formed past the end of the file.
            case U_ADDV4: case U_SUBV4: case U_MULM2: {
                __builtin_assume(a < 29u);
                __builtin_assume(b < 29u);
                __builtin_assume(d < 29u);
                const v32f Vv = G.v;
                float A[4], B[4], out[4], og[4][T];
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    A[q] = Vv[a + q];
                    B[q] = Vv[b + q];
                }
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    if (op == U_ADDV4) {
                        out[q] = A[q] + B[q];
                        RSR_TAN(og[q][j] = Gj[a + q] + Gj[b + q]);
                    } else if (op == U_SUBV4) {
                        out[q] = A[q] - B[q];
                        RSR_TAN(og[q][j] = Gj[a + q] - Gj[b + q]);
                    } else {
                        const int r = q & 1, c = q >> 1;
                        out[q] = (A[r] * B[2 * c]) + (A[r + 2] * B[2 * c + 1]);
                        RSR_TAN(og[q][j] = ((A[r] * Gj[b + 2 * c]) + (Gj[a + r] * B[2 * c])) +
                                           ((A[r + 2] * Gj[b + 2 * c + 1]) + (Gj[a + r + 2] * B[2 * c + 1])));
                    }
                }
                rv = out[0];
                RSR_TAN(rg[j] = og[0][j]);
                put4(out, og);
                break;
            }
            case U_SCALEV4: case U_SCALEV4K: {
                __builtin_assume(a < 29u);
                __builtin_assume(d < 29u);
                const v32f Vv = G.v;
                const bool kk = op == U_SCALEV4K;
                // the scalar: register b, or the parameter x (tangent 1 on its own channel)
                const float sv = kk ? kld(params, x) : vb;
                float gs[T];
                RSR_TAN(gs[j] = kk ? ((x == k0 + (uint32_t)j) ? 1.0f : 0.0f) : gb[j]);
                float out[4], og[4][T];
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const float aq = Vv[a + q];
                    out[q] = aq * sv;
                    RSR_TAN(og[q][j] = (aq * gs[j]) + (Gj[a + q] * sv));
                }
                rv = out[0];
                RSR_TAN(rg[j] = og[0][j]);
                put4(out, og);
                break;
            }
            case U_ADDV3: case U_SUBV3: case U_CROSS3: {
                __builtin_assume(a < 30u);
                __builtin_assume(b < 30u);
                __builtin_assume(d < 30u);
                // components 1 and 2 of the value and of every tangent channel
                const v32f Vv = G.v;
                const float a1 = Vv[a + 1], a2 = Vv[a + 2], b1 = Vv[b + 1], b2 = Vv[b + 2];
                float v1, v2, g1[T], g2[T];
                if (op == U_ADDV3) {
                    rv = va + vb; v1 = a1 + b1; v2 = a2 + b2;
                    RSR_TAN(rg[j] = ga[j] + gb[j]; g1[j] = Gj[a + 1] + Gj[b + 1]; g2[j] = Gj[a + 2] + Gj[b + 2]);
                } else if (op == U_SUBV3) {
                    rv = va - vb; v1 = a1 - b1; v2 = a2 - b2;
                    RSR_TAN(rg[j] = ga[j] - gb[j]; g1[j] = Gj[a + 1] - Gj[b + 1]; g2[j] = Gj[a + 2] - Gj[b + 2]);
                } else {
                    rv = (a1 * b2) - (a2 * b1); v1 = (a2 * vb) - (va * b2); v2 = (va * b1) - (a1 * vb);
                    RSR_TAN(rg[j] = ((a1 * Gj[b + 2]) + (Gj[a + 1] * b2)) - ((a2 * Gj[b + 1]) + (Gj[a + 2] * b1));
                            g1[j] = ((a2 * gb[j]) + (Gj[a + 2] * vb)) - ((va * Gj[b + 2]) + (ga[j] * b2));
                            g2[j] = ((va * Gj[b + 1]) + (ga[j] * b1)) - ((a1 * gb[j]) + (Gj[a + 1] * vb)));
                }
                put3(v1, v2, g1, g2);
                break;
            }
            case U_SCALEV3: case U_SCALEV3K: case U_IMMV3: {
                __builtin_assume(a < 30u);
                __builtin_assume(d < 30u);
                float v1, v2, g1[T], g2[T];
                if (op == U_IMMV3) {
                    rv = kld(params, x); v1 = kld(params, x + 1); v2 = kld(params, x + 2);
                    RSR_TAN(rg[j] = (x == k0 + (uint32_t)j) ? 1.0f : 0.0f;
                            g1[j] = (x + 1u == k0 + (uint32_t)j) ? 1.0f : 0.0f;
                            g2[j] = (x + 2u == k0 + (uint32_t)j) ? 1.0f : 0.0f);
                } else {
                    const v32f Vv = G.v;
                    const float a1 = Vv[a + 1], a2 = Vv[a + 2];
                    // the scalar: register b, or the parameter x (tangent 1 on its own channel)
                    const bool kk = op == U_SCALEV3K;
                    const float sv = kk ? kld(params, x) : vb;
                    float gs[T];
                    RSR_TAN(gs[j] = kk ? ((x == k0 + (uint32_t)j) ? 1.0f : 0.0f) : gb[j]);
                    rv = va * sv; v1 = a1 * sv; v2 = a2 * sv;
                    RSR_TAN(rg[j] = (va * gs[j]) + (ga[j] * sv); g1[j] = (a1 * gs[j]) + (Gj[a + 1] * sv);
                            g2[j] = (a2 * gs[j]) + (Gj[a + 2] * sv));
                }
                put3(v1, v2, g1, g2);
                break;
            }
            default:
                rv = 0.0f;
                RSR_TAN(rg[j] = 0.0f);
                break;
        }
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

// v from the lane that DPP control Ctrl names (rows of 16 lanes; lanes of the rows that
// RowMask leaves out, and lanes whose source is out of range, read 0).  Every lane of the
// wave is active.
template <int Ctrl, int RowMask = 0xf>
__device__ __forceinline__ float dpp_f(float v) {
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), Ctrl, RowMask, 0xf,
                                                                 false));
}

__device__ __forceinline__ float lane_f(float v, int l) {
    return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), l));
}

// Sum over the 64 lanes of a wavefront in a fixed order (deterministic), the same value in
// every lane.  Four DPP steps leave every lane of a row of 16 holding the row's sum: the
// partner lane of quad_perm [1,0,3,2], of quad_perm [2,3,0,1], of row_half_mirror (lane i
// of 8 reads 7 - i) and of row_mirror (lane i of 16 reads 15 - i); a + b and b + a are the
// same float, so the partners agree.  row_bcast:15 then adds row 0's sum into row 1 and
// row 2's into row 3, row_bcast:31 adds row 1's (r0 + r1) into row 3: lane 63 holds
// (r2 + r3) + (r0 + r1).  Called with the whole wave active.
__device__ __forceinline__ float wave_sum(float v) {
    v = v + dpp_f<0xB1>(v);        // quad_perm [1,0,3,2]
    v = v + dpp_f<0x4E>(v);        // quad_perm [2,3,0,1]
    v = v + dpp_f<0x141>(v);       // row_half_mirror
    v = v + dpp_f<0x140>(v);       // row_mirror
    v = v + dpp_f<0x142, 0xa>(v);  // row_bcast:15 into rows 1 and 3
    v = v + dpp_f<0x143, 0xc>(v);  // row_bcast:31 into rows 2 and 3
    return lane_f(v, 63);
}

}  // namespace rsr
