/*
 * rsr_oracle.c -- CPU oracle for the RSR fitness-evaluation path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is a plain-C restatement of the
 * reference evaluator (daruandris/rsr, Rust) and is used solely as the
 * checker by tests/, __graft_entry__.smoke() and the cpu_baseline leg of
 * bench.py.  The product library (rsr_amd/) never links, loads or calls it.
 *
 * Build: gcc -O2 -std=c11 -ffp-contract=off -fno-fast-math (x86-64 SSE2, so
 * every float operation below is one IEEE-754 binary32 operation, no FMA
 * contraction, no excess precision), exactly like `wide::f32x8` arithmetic.
 *
 * What each part follows (reference file:line):
 *   - Instruction / opcode numbering ........ src/engine/domain.rs:46-63, src/lib.rs:44-48,
 *                                             src/domains/basic/mod.rs:9-20, src/domains/linalg/mod.rs:9-51
 *   - VmState (32-deep zeroed typed stacks) .. src/engine/eval/state.rs:12-79
 *   - eval_simd ............................. src/engine/eval/evaluator.rs:20-150
 *   - Basic ops ............................. src/domains/basic/eval.rs:5-63
 *   - Linalg ops ............................ src/domains/linalg/eval.rs:1-506
 *   - compute_mse ........................... src/engine/eval/evaluator.rs:153-174
 *   - compute_mse_with_gradient ............. src/engine/eval/evaluator.rs:177-224
 *   - eval_simd_dual ........................ src/engine/eval/evaluator.rs:228-388
 *   - DualSimd + dual ops ................... src/engine/eval/autodiff.rs:11-133, 219-873;
 *                                             linalg dual dispatch src/domains/linalg/mod.rs:502-690
 *   - Dataset::new (AoSoA, padding, z-score)  src/engine/data/dataset.rs:37-169
 *   - Dataset::subset ....................... src/engine/data/dataset.rs:360-447
 *   - Parameterized flatten/unflatten ....... src/engine/expr/program.rs:84-193
 *   - run_lbfgs ............................. src/engine/optimize/lbfgs.rs:9-165
 *   - run_cma_es ............................ src/engine/optimize/cmaes.rs:19-192
 *   - run_nelder_mead ....................... src/engine/optimize/nelder_mead.rs:9-165
 *   - NSGA-II rank / crowding / tournament .. src/engine/search/operators/selection.rs:5-134
 *
 * The transcendental functions sin/cos/exp/ln live in the third-party crate
 * `wide 1.1.1` (Cargo.lock:506-513) whose source is absent.  Here they are
 * Cody-Waite reductions (with fused multiply-adds for sin/cos/exp) and the Cephes
 * single-precision minimax polynomials (S. L. Moshier, Cephes Math Library:
 * sinf.c, cosf.c, expf.c, logf.c), the device library's own sequences restated
 * operation for operation.  They are pinned against FP64 libm by ulp-error tests, not
 * against wide (parity unpinned for these four functions).  The final
 * horizontal add of the 8 MSE lanes (`f32x8::reduce_add`) is likewise inside
 * wide; it is restated as (lo4 + hi4) followed by ((0+2)+(1+3)) (unpinned).
 */
#include <math.h>
#include <stdint.h>
#include <string.h>
#include <float.h>
#include <stdlib.h>

#define ORC_API __attribute__((visibility("default")))

/* ---- opcodes: reference Instruction enum order (domain.rs:46-63 + lib.rs:44-48) ---- */
enum {
    OP_LOADVARF = 0, OP_LOADCONSTF, OP_LOADVARB, OP_LOADCONSTB, OP_LOADVARI, OP_LOADCONSTI,
    OP_LOADVARV2, OP_LOADCONSTV2, OP_LOADVARV3, OP_LOADCONSTV3, OP_LOADVARM2, OP_LOADCONSTM2,
    OP_LOADVARM3, OP_LOADCONSTM3,
    /* Basic (basic/mod.rs:9-20) */
    OP_ADDF = 14, OP_SUBF, OP_MULF, OP_DIVF, OP_SINF, OP_COSF, OP_EXPF, OP_SQRF, OP_SQRTF, OP_LNF,
    /* Linalg (linalg/mod.rs:9-51) */
    OP_MAKEVEC2 = 24, OP_MAKEVEC3, OP_GETXV2, OP_GETYV2, OP_GETXV3, OP_GETYV3, OP_GETZV3,
    OP_ADDV2, OP_SUBV2, OP_SCALEV2, OP_DOTV2, OP_NORMV2,
    OP_ADDV3, OP_SUBV3, OP_SCALEV3, OP_DOTV3, OP_NORMV3, OP_CROSSV3,
    OP_MAKEMAT2, OP_ADDM2, OP_SUBM2, OP_SCALEM2, OP_MULM2, OP_MULM2V2, OP_DETM2, OP_TRACEM2,
    OP_TRANSPOSEM2, OP_INVERSEM2,
    OP_MAKEMAT3, OP_ADDM3, OP_SUBM3, OP_SCALEM3, OP_MULM3, OP_MULM3V3, OP_DETM3, OP_TRACEM3,
    OP_TRANSPOSEM3, OP_INVERSEM3,
    /* Build extension (not a reference opcode): exact sign flip, used by the scalar workload. */
    OP_NEGF = 62,
    OP_COUNT = 63
};

/* ValueType (types.rs:1-10) */
enum { VT_FLOAT = 0, VT_INT = 1, VT_BOOL = 2, VT_VEC2 = 3, VT_VEC3 = 4, VT_MAT2 = 5, VT_MAT3 = 6 };

/* error codes returned by the oracle */
#define ORC_OK 0
#define ORC_E_STACK (-1)    /* stack underflow / overflow (UB in the reference) */
#define ORC_E_OPCODE (-2)   /* unknown opcode */
#define ORC_E_ARG (-3)      /* bad argument (feature / constant index out of range) */

static int vt_size(int vt) {
    switch (vt) {
        case VT_FLOAT: return 1;
        case VT_VEC2: return 2;
        case VT_VEC3: return 3;
        case VT_MAT2: return 4;
        case VT_MAT3: return 9;
        default: return 0;
    }
}

/* =====================================================================
 * Transcendentals.  The reference's come from the crate `wide 1.1.1`, whose source is not
 * here, so the bits are the device library's own definition, restated operation for
 * operation: sin / cos / exp reduce with fused multiply-adds (Cody-Waite: pi/2 in three
 * float parts, ln 2 in two) and evaluate the Cephes single-precision minimax polynomials
 * (S. L. Moshier, Cephes Math Library, sinf.c / cosf.c / expf.c) in Horner form with fused
 * multiply-adds; exp scales by ldexpf.  ln is Cephes logf with plain mul/add.  fmaf,
 * rintf and ldexpf are correctly rounded, as the device's v_fma_f32 / v_rndne_f32 /
 * v_ldexp_f32 are.  Accuracy (at most 2 ulp against FP64) is pinned by the test-suite.
 * ===================================================================== */
ORC_API float orc_sinf(float xx);
ORC_API float orc_cosf(float xx);
ORC_API float orc_expf(float xx);
ORC_API float orc_logf(float xx);

static uint32_t f2u(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }
static float u2f(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }

/* |x| = k*pi/2 + r, k = rint(|x|*2/pi); s ~ sin(r), c ~ cos(r); n = k (quadrant bits). */
static void sincos_reduced(float ax, float *s, float *c, uint32_t *n) {
    float k = rintf(ax * 0.636619746685028076171875f);
    float r = fmaf(-k, 1.57079637050628662109375f, ax);
    r = fmaf(-k, -4.37113882867379300296e-08f, r);
    r = fmaf(-k, -1.71512451000588190000e-15f, r);
    float z = r * r;
    float p = fmaf(z, -1.9515295891E-4f, 8.3321608736E-3f);
    p = fmaf(z, p, -1.6666654611E-1f);
    *s = fmaf(r * z, p, r);
    float q = fmaf(z, 2.443315711809948E-5f, -1.388731625493765E-3f);
    q = fmaf(z, q, 4.166664568298827E-2f);
    *c = fmaf(z * z, q, fmaf(-0.5f, z, 1.0f));
    *n = (uint32_t)(int)fminf(k, 16777216.0f);
}

ORC_API float orc_sinf(float xx) {
    float ax = fabsf(xx);
    if (!(ax <= 16777215.0f)) return 0.0f * xx; /* NaN for inf/NaN; +-0 past 2^24 - 1 (Cephes T24M1) */
    float s, c; uint32_t n;
    sincos_reduced(ax, &s, &c, &n);
    float y = (n & 1u) ? c : s;
    return u2f(f2u(y) ^ (((n << 30) ^ f2u(xx)) & 0x80000000u));
}

ORC_API float orc_cosf(float xx) {
    float ax = fabsf(xx);
    if (isnan(ax) || isinf(ax)) return NAN;
    if (ax > 16777215.0f) return 1.0f;
    float s, c; uint32_t n;
    sincos_reduced(ax, &s, &c, &n);
    float y = (n & 1u) ? s : c;
    return u2f(f2u(y) ^ (((n + 1u) << 30) & 0x80000000u));
}

ORC_API float orc_expf(float xx) {
    if (isnan(xx)) return xx;
    float x = fminf(fmaxf(xx, -104.0f), 89.0f); /* past these, ldexpf gives 0 or inf */
    float k = rintf(x * 1.44269502162933349609375f);
    float r = fmaf(-k, 0.693147182464599609375f, x);
    r = fmaf(-k, -1.904654212125933566e-09f, r);
    float p = fmaf(1.9875691500E-4f, r, 1.3981999507E-3f);
    p = fmaf(p, r, 8.3334519073E-3f);
    p = fmaf(p, r, 4.1665795894E-2f);
    p = fmaf(p, r, 1.6666665459E-1f);
    p = fmaf(p, r, 5.0000001201E-1f);
    float y = fmaf(r * r, p, r) + 1.0f;
    return ldexpf(y, (int)k);
}

ORC_API float orc_logf(float xx) {
    if (isnan(xx)) return xx;
    if (xx < 0.0f) return NAN;
    if (xx == 0.0f) return -INFINITY;
    if (isinf(xx)) return xx;
    int e;
    float x = frexpf(xx, &e); /* x in [0.5, 1) */
    if (x < 0.707106781186547524f) { e -= 1; x = x + x - 1.0f; }
    else { x = x - 1.0f; }
    float z = x * x;
    float y = ((((((((7.0376836292E-2f * x - 1.1514610310E-1f) * x + 1.1676998740E-1f) * x
                    - 1.2420140846E-1f) * x + 1.4249322787E-1f) * x - 1.6668057665E-1f) * x
                 + 2.0000714765E-1f) * x - 2.4999993993E-1f) * x + 3.3333331174E-1f) * x * z;
    float fe = (float)e;
    if (e) y = y + -2.12194440e-4f * fe;
    y = y + -0.5f * z;
    z = x + y;
    if (e) z = z + 0.693359375f * fe;
    return z;
}

/* Accuracy sweep (test support): every stride-th float of [lo, hi] (finite, lo <= hi, both of
 * one sign or lo = -hi mirrored) through orc_sinf (0) / orc_cosf (1) / orc_expf (2) /
 * orc_logf (3) against the FP64 libm value.  out[0] = largest error in ulps of the float
 * nearest the exact value where that is a normal float (sin/cos: where |f(x)| >= 1e-3;
 * ln: where |ln x| >= 1e-6), out[1] = largest absolute error elsewhere, out[2] = the
 * argument of out[0], out[3] = points swept.  Floats are walked by their bit patterns over
 * [0, hi] and mirrored to -x when lo < 0. */
static double ulp_at(double t) {
    float tf = (float)fabs(t);
    if (tf < FLT_MIN) return (double)1.40129846e-45f;
    if (isinf(tf)) tf = FLT_MAX;
    return (double)(nextafterf(tf, INFINITY) - tf);
}
ORC_API int orc_ulp_sweep(int which, float lo, float hi, uint32_t stride, double *out) {
    if (!(lo <= hi) || stride == 0 || which < 0 || which > 3) return -1;
    out[0] = out[1] = out[2] = out[3] = 0.0;
    const float amax = fmaxf(fabsf(lo), fabsf(hi));
    const float amin = (lo <= 0.0f && hi >= 0.0f) ? 0.0f : fminf(fabsf(lo), fabsf(hi));
    uint32_t u0, u1;
    memcpy(&u0, &amin, 4);
    memcpy(&u1, &amax, 4);
    for (uint64_t u = u0; u <= u1; u += stride) {
        uint32_t b = (uint32_t)u;
        float a;
        memcpy(&a, &b, 4);
        for (int sg = 0; sg < 2; ++sg) {
            const float x = sg ? -a : a;
            if (x < lo || x > hi || (sg && a == 0.0f)) continue;
            float got;
            double t;
            switch (which) {
                case 0: got = orc_sinf(x); t = sin((double)x); break;
                case 1: got = orc_cosf(x); t = cos((double)x); break;
                case 2: got = orc_expf(x); t = exp((double)x); break;
                default: got = orc_logf(x); t = log((double)x); break;
            }
            out[3] += 1.0;
            int rel;
            if (which <= 1) rel = fabs(t) >= 1e-3;
            else if (which == 2) rel = t >= (double)FLT_MIN && t <= (double)FLT_MAX;
            else rel = fabs(t) >= 1e-6 && isfinite(t);
            if (which == 2 && t > (double)FLT_MAX) { if (!isinf(got)) out[1] = INFINITY; continue; }
            const double e = fabs((double)got - t);
            if (rel) {
                const double eu = e / ulp_at(t);
                if (eu > out[0]) { out[0] = eu; out[2] = x; }
            } else if (e > out[1]) {
                out[1] = e;
            }
        }
    }
    return 0;
}

/* =====================================================================
 * 8-lane f32 VM: VmState (state.rs:12-49) and eval_simd (evaluator.rs:20-150)
 * ===================================================================== */
#define LANES 8
typedef struct { float l[LANES]; } F8;

typedef struct {
    F8 f[32]; int sp_f;
    F8 v2[32][2]; int sp_v2;
    F8 v3[32][3]; int sp_v3;
    F8 m2[32][4]; int sp_m2;
    F8 m3[32][9]; int sp_m3;
} Vm;

/* A program as the C ABI carries it: reference Instruction stream + Scalar constants. */
typedef struct {
    const uint8_t *ops;
    const uint16_t *args;
    int n_code;
    const uint8_t *ctypes;  /* ValueType per constant */
    const float *cvals;     /* flat constant payloads (Float 1, Vec2 2, Vec3 3, Mat2 4, Mat3 9, Int/Bool 1) */
    int n_consts;
} OrcProgram;

/* offset of constant c's payload inside cvals (Int/Bool take one slot) */
static int cval_offset(const OrcProgram *p, int c) {
    int off = 0;
    for (int i = 0; i < c; ++i) { int s = vt_size(p->ctypes[i]); off += s ? s : 1; }
    return off;
}

#define CHK(cond) do { if (!(cond)) return ORC_E_STACK; } while (0)
#define FORL for (int l = 0; l < LANES; ++l)

static F8 splat(float v) { F8 r; FORL r.l[l] = v; return r; }

/* one batch of 8 points: features = F8[num_features] (feature_flat slice) */
static int eval_simd8(const OrcProgram *p, const F8 *feat, int nfeat, F8 *out) {
    static __thread Vm vm;
    memset(&vm, 0, sizeof(vm)); /* VmState::new() zeroes (state.rs:46-48) */
    Vm *c = &vm;
    for (int pc = 0; pc < p->n_code; ++pc) {
        int op = p->ops[pc];
        int a = p->args[pc];
        switch (op) {
        case OP_LOADVARF: CHK(c->sp_f < 32); if (a >= nfeat) return ORC_E_ARG; c->f[c->sp_f++] = feat[a]; break;
        case OP_LOADCONSTF: CHK(c->sp_f < 32); if (a >= p->n_consts) return ORC_E_ARG;
            if (p->ctypes[a] == VT_FLOAT) c->f[c->sp_f] = splat(p->cvals[cval_offset(p, a)]);
            c->sp_f++; break;
        case OP_LOADVARV2: CHK(c->sp_v2 < 32); if (a + 1 >= nfeat) return ORC_E_ARG;
            for (int k = 0; k < 2; ++k) c->v2[c->sp_v2][k] = feat[a + k]; c->sp_v2++; break;
        case OP_LOADCONSTV2: CHK(c->sp_v2 < 32); if (a >= p->n_consts) return ORC_E_ARG;
            if (p->ctypes[a] == VT_VEC2) { int o = cval_offset(p, a); for (int k = 0; k < 2; ++k) c->v2[c->sp_v2][k] = splat(p->cvals[o + k]); }
            c->sp_v2++; break;
        case OP_LOADVARV3: CHK(c->sp_v3 < 32); if (a + 2 >= nfeat) return ORC_E_ARG;
            for (int k = 0; k < 3; ++k) c->v3[c->sp_v3][k] = feat[a + k]; c->sp_v3++; break;
        case OP_LOADCONSTV3: CHK(c->sp_v3 < 32); if (a >= p->n_consts) return ORC_E_ARG;
            if (p->ctypes[a] == VT_VEC3) { int o = cval_offset(p, a); for (int k = 0; k < 3; ++k) c->v3[c->sp_v3][k] = splat(p->cvals[o + k]); }
            c->sp_v3++; break;
        case OP_LOADVARM2: CHK(c->sp_m2 < 32); if (a + 3 >= nfeat) return ORC_E_ARG;
            for (int k = 0; k < 4; ++k) c->m2[c->sp_m2][k] = feat[a + k]; c->sp_m2++; break;
        case OP_LOADCONSTM2: CHK(c->sp_m2 < 32); if (a >= p->n_consts) return ORC_E_ARG;
            if (p->ctypes[a] == VT_MAT2) { int o = cval_offset(p, a); for (int k = 0; k < 4; ++k) c->m2[c->sp_m2][k] = splat(p->cvals[o + k]); }
            c->sp_m2++; break;
        case OP_LOADVARM3: CHK(c->sp_m3 < 32); if (a + 8 >= nfeat) return ORC_E_ARG;
            for (int k = 0; k < 9; ++k) c->m3[c->sp_m3][k] = feat[a + k]; c->sp_m3++; break;
        case OP_LOADCONSTM3: CHK(c->sp_m3 < 32); if (a >= p->n_consts) return ORC_E_ARG;
            if (p->ctypes[a] == VT_MAT3) { int o = cval_offset(p, a); for (int k = 0; k < 9; ++k) c->m3[c->sp_m3][k] = splat(p->cvals[o + k]); }
            c->sp_m3++; break;
        case OP_LOADVARB: case OP_LOADCONSTB: case OP_LOADVARI: case OP_LOADCONSTI:
            break; /* silent no-op (evaluator.rs:144 -> domain.rs:81-88 matches nothing) */

        /* ---- Basic (basic/eval.rs) ---- */
        case OP_ADDF: CHK(c->sp_f >= 2); c->sp_f -= 2; FORL c->f[c->sp_f].l[l] = c->f[c->sp_f].l[l] + c->f[c->sp_f + 1].l[l]; c->sp_f++; break;
        case OP_SUBF: CHK(c->sp_f >= 2); c->sp_f -= 2; FORL c->f[c->sp_f].l[l] = c->f[c->sp_f].l[l] - c->f[c->sp_f + 1].l[l]; c->sp_f++; break;
        case OP_MULF: CHK(c->sp_f >= 2); c->sp_f -= 2; FORL c->f[c->sp_f].l[l] = c->f[c->sp_f].l[l] * c->f[c->sp_f + 1].l[l]; c->sp_f++; break;
        case OP_DIVF: CHK(c->sp_f >= 2); c->sp_f -= 2; FORL c->f[c->sp_f].l[l] = c->f[c->sp_f].l[l] / c->f[c->sp_f + 1].l[l]; c->sp_f++; break;
        case OP_SINF: CHK(c->sp_f >= 1); FORL c->f[c->sp_f - 1].l[l] = orc_sinf(c->f[c->sp_f - 1].l[l]); break;
        case OP_COSF: CHK(c->sp_f >= 1); FORL c->f[c->sp_f - 1].l[l] = orc_cosf(c->f[c->sp_f - 1].l[l]); break;
        case OP_EXPF: CHK(c->sp_f >= 1); FORL c->f[c->sp_f - 1].l[l] = orc_expf(c->f[c->sp_f - 1].l[l]); break;
        case OP_SQRF: CHK(c->sp_f >= 1); FORL { float v = c->f[c->sp_f - 1].l[l]; c->f[c->sp_f - 1].l[l] = v * v; } break;
        case OP_SQRTF: CHK(c->sp_f >= 1); FORL c->f[c->sp_f - 1].l[l] = sqrtf(c->f[c->sp_f - 1].l[l]); break;
        case OP_LNF: CHK(c->sp_f >= 1); FORL c->f[c->sp_f - 1].l[l] = orc_logf(c->f[c->sp_f - 1].l[l]); break;
        case OP_NEGF: CHK(c->sp_f >= 1); FORL c->f[c->sp_f - 1].l[l] = -c->f[c->sp_f - 1].l[l]; break;

        /* ---- Linalg (linalg/eval.rs) ---- */
        case OP_MAKEVEC2: CHK(c->sp_f >= 2 && c->sp_v2 < 32); c->sp_f -= 2;
            c->v2[c->sp_v2][0] = c->f[c->sp_f]; c->v2[c->sp_v2][1] = c->f[c->sp_f + 1]; c->sp_v2++; break;
        case OP_MAKEVEC3: CHK(c->sp_f >= 3 && c->sp_v3 < 32); c->sp_f -= 3;
            for (int k = 0; k < 3; ++k) c->v3[c->sp_v3][k] = c->f[c->sp_f + k]; c->sp_v3++; break;
        case OP_GETXV2: case OP_GETYV2: CHK(c->sp_v2 >= 1 && c->sp_f < 32); c->sp_v2--;
            c->f[c->sp_f++] = c->v2[c->sp_v2][op == OP_GETXV2 ? 0 : 1]; break;
        case OP_GETXV3: case OP_GETYV3: case OP_GETZV3: CHK(c->sp_v3 >= 1 && c->sp_f < 32); c->sp_v3--;
            c->f[c->sp_f++] = c->v3[c->sp_v3][op - OP_GETXV3]; break;
        case OP_ADDV2: case OP_SUBV2: CHK(c->sp_v2 >= 2); c->sp_v2 -= 2;
            for (int k = 0; k < 2; ++k) FORL { float x = c->v2[c->sp_v2][k].l[l], y = c->v2[c->sp_v2 + 1][k].l[l]; c->v2[c->sp_v2][k].l[l] = op == OP_ADDV2 ? x + y : x - y; }
            c->sp_v2++; break;
        case OP_SCALEV2: CHK(c->sp_f >= 1 && c->sp_v2 >= 1); c->sp_f--; c->sp_v2--;
            for (int k = 0; k < 2; ++k) FORL c->v2[c->sp_v2][k].l[l] = c->v2[c->sp_v2][k].l[l] * c->f[c->sp_f].l[l];
            c->sp_v2++; break;
        case OP_DOTV2: CHK(c->sp_v2 >= 2 && c->sp_f < 32); c->sp_v2 -= 2;
            FORL { F8 *A = c->v2[c->sp_v2], *B = c->v2[c->sp_v2 + 1]; c->f[c->sp_f].l[l] = (A[0].l[l] * B[0].l[l]) + (A[1].l[l] * B[1].l[l]); }
            c->sp_f++; break;
        case OP_NORMV2: CHK(c->sp_v2 >= 1 && c->sp_f < 32);
            FORL { F8 *V = c->v2[c->sp_v2 - 1]; c->f[c->sp_f].l[l] = sqrtf((V[0].l[l] * V[0].l[l]) + (V[1].l[l] * V[1].l[l])); }
            c->sp_f++; c->sp_v2--; break;
        case OP_ADDV3: case OP_SUBV3: CHK(c->sp_v3 >= 2); c->sp_v3 -= 2;
            for (int k = 0; k < 3; ++k) FORL { float x = c->v3[c->sp_v3][k].l[l], y = c->v3[c->sp_v3 + 1][k].l[l]; c->v3[c->sp_v3][k].l[l] = op == OP_ADDV3 ? x + y : x - y; }
            c->sp_v3++; break;
        case OP_SCALEV3: CHK(c->sp_f >= 1 && c->sp_v3 >= 1); c->sp_f--; c->sp_v3--;
            for (int k = 0; k < 3; ++k) FORL c->v3[c->sp_v3][k].l[l] = c->v3[c->sp_v3][k].l[l] * c->f[c->sp_f].l[l];
            c->sp_v3++; break;
        case OP_DOTV3: CHK(c->sp_v3 >= 2 && c->sp_f < 32); c->sp_v3 -= 2;
            FORL { F8 *A = c->v3[c->sp_v3], *B = c->v3[c->sp_v3 + 1];
                   c->f[c->sp_f].l[l] = (A[0].l[l] * B[0].l[l]) + (A[1].l[l] * B[1].l[l]) + (A[2].l[l] * B[2].l[l]); }
            c->sp_f++; break;
        case OP_NORMV3: CHK(c->sp_v3 >= 1 && c->sp_f < 32);
            FORL { F8 *V = c->v3[c->sp_v3 - 1]; c->f[c->sp_f].l[l] = sqrtf((V[0].l[l] * V[0].l[l]) + (V[1].l[l] * V[1].l[l]) + (V[2].l[l] * V[2].l[l])); }
            c->sp_f++; c->sp_v3--; break;
        case OP_CROSSV3: CHK(c->sp_v3 >= 2); c->sp_v3 -= 2;
            FORL { float a0 = c->v3[c->sp_v3][0].l[l], a1 = c->v3[c->sp_v3][1].l[l], a2 = c->v3[c->sp_v3][2].l[l];
                   float b0 = c->v3[c->sp_v3 + 1][0].l[l], b1 = c->v3[c->sp_v3 + 1][1].l[l], b2 = c->v3[c->sp_v3 + 1][2].l[l];
                   c->v3[c->sp_v3][0].l[l] = a1 * b2 - a2 * b1;
                   c->v3[c->sp_v3][1].l[l] = a2 * b0 - a0 * b2;
                   c->v3[c->sp_v3][2].l[l] = a0 * b1 - a1 * b0; }
            c->sp_v3++; break;
        case OP_MAKEMAT2: CHK(c->sp_v2 >= 2 && c->sp_m2 < 32); c->sp_v2 -= 2;
            c->m2[c->sp_m2][0] = c->v2[c->sp_v2][0]; c->m2[c->sp_m2][1] = c->v2[c->sp_v2][1];
            c->m2[c->sp_m2][2] = c->v2[c->sp_v2 + 1][0]; c->m2[c->sp_m2][3] = c->v2[c->sp_v2 + 1][1];
            c->sp_m2++; break;
        case OP_ADDM2: case OP_SUBM2: CHK(c->sp_m2 >= 2); c->sp_m2 -= 2;
            for (int k = 0; k < 4; ++k) FORL { float x = c->m2[c->sp_m2][k].l[l], y = c->m2[c->sp_m2 + 1][k].l[l]; c->m2[c->sp_m2][k].l[l] = op == OP_ADDM2 ? x + y : x - y; }
            c->sp_m2++; break;
        case OP_SCALEM2: CHK(c->sp_f >= 1 && c->sp_m2 >= 1); c->sp_f--; c->sp_m2--;
            for (int k = 0; k < 4; ++k) FORL c->m2[c->sp_m2][k].l[l] = c->m2[c->sp_m2][k].l[l] * c->f[c->sp_f].l[l];
            c->sp_m2++; break;
        case OP_MULM2: CHK(c->sp_m2 >= 2); c->sp_m2 -= 2;
            FORL { float a[4], b[4]; for (int k = 0; k < 4; ++k) { a[k] = c->m2[c->sp_m2][k].l[l]; b[k] = c->m2[c->sp_m2 + 1][k].l[l]; }
                   c->m2[c->sp_m2][0].l[l] = a[0] * b[0] + a[2] * b[1];
                   c->m2[c->sp_m2][1].l[l] = a[1] * b[0] + a[3] * b[1];
                   c->m2[c->sp_m2][2].l[l] = a[0] * b[2] + a[2] * b[3];
                   c->m2[c->sp_m2][3].l[l] = a[1] * b[2] + a[3] * b[3]; }
            c->sp_m2++; break;
        case OP_MULM2V2: CHK(c->sp_m2 >= 1 && c->sp_v2 >= 1); c->sp_m2--; c->sp_v2--;
            FORL { float m[4], v[2]; for (int k = 0; k < 4; ++k) m[k] = c->m2[c->sp_m2][k].l[l]; v[0] = c->v2[c->sp_v2][0].l[l]; v[1] = c->v2[c->sp_v2][1].l[l];
                   c->v2[c->sp_v2][0].l[l] = m[0] * v[0] + m[2] * v[1];
                   c->v2[c->sp_v2][1].l[l] = m[1] * v[0] + m[3] * v[1]; }
            c->sp_v2++; break;
        case OP_DETM2: CHK(c->sp_m2 >= 1 && c->sp_f < 32); c->sp_m2--;
            FORL { F8 *M = c->m2[c->sp_m2]; c->f[c->sp_f].l[l] = (M[0].l[l] * M[3].l[l]) - (M[1].l[l] * M[2].l[l]); }
            c->sp_f++; break;
        case OP_TRACEM2: CHK(c->sp_m2 >= 1 && c->sp_f < 32); c->sp_m2--;
            FORL { F8 *M = c->m2[c->sp_m2]; c->f[c->sp_f].l[l] = M[0].l[l] + M[3].l[l]; }
            c->sp_f++; break;
        case OP_TRANSPOSEM2: CHK(c->sp_m2 >= 1); { F8 *M = c->m2[c->sp_m2 - 1]; F8 t = M[1]; M[1] = M[2]; M[2] = t; } break;
        case OP_INVERSEM2: CHK(c->sp_m2 >= 1);
            FORL { F8 *M = c->m2[c->sp_m2 - 1]; float m0 = M[0].l[l], m1 = M[1].l[l], m2 = M[2].l[l], m3 = M[3].l[l];
                   float det = (m0 * m3) - (m1 * m2);
                   int sing = fabsf(det) < 1e-9f;
                   float safe = sing ? 1.0f : det;
                   float inv = 1.0f / safe;
                   M[0].l[l] = sing ? 1.0f : m3 * inv;
                   M[1].l[l] = sing ? 0.0f : (-m1) * inv;
                   M[2].l[l] = sing ? 0.0f : (-m2) * inv;
                   M[3].l[l] = sing ? 1.0f : m0 * inv; }
            break;
        case OP_MAKEMAT3: CHK(c->sp_v3 >= 3 && c->sp_m3 < 32); c->sp_v3 -= 3;
            for (int col = 0; col < 3; ++col) for (int k = 0; k < 3; ++k) c->m3[c->sp_m3][col * 3 + k] = c->v3[c->sp_v3 + col][k];
            c->sp_m3++; break;
        case OP_ADDM3: case OP_SUBM3: CHK(c->sp_m3 >= 2); c->sp_m3 -= 2;
            for (int k = 0; k < 9; ++k) FORL { float x = c->m3[c->sp_m3][k].l[l], y = c->m3[c->sp_m3 + 1][k].l[l]; c->m3[c->sp_m3][k].l[l] = op == OP_ADDM3 ? x + y : x - y; }
            c->sp_m3++; break;
        case OP_SCALEM3: CHK(c->sp_f >= 1 && c->sp_m3 >= 1); c->sp_f--; c->sp_m3--;
            for (int k = 0; k < 9; ++k) FORL c->m3[c->sp_m3][k].l[l] = c->m3[c->sp_m3][k].l[l] * c->f[c->sp_f].l[l];
            c->sp_m3++; break;
        case OP_MULM3: CHK(c->sp_m3 >= 2); c->sp_m3 -= 2;
            FORL { float a[9], b[9], o[9]; for (int k = 0; k < 9; ++k) { a[k] = c->m3[c->sp_m3][k].l[l]; b[k] = c->m3[c->sp_m3 + 1][k].l[l]; }
                   for (int col = 0; col < 3; ++col) for (int r = 0; r < 3; ++r)
                       o[col * 3 + r] = a[r] * b[col * 3] + a[3 + r] * b[col * 3 + 1] + a[6 + r] * b[col * 3 + 2];
                   for (int k = 0; k < 9; ++k) c->m3[c->sp_m3][k].l[l] = o[k]; }
            c->sp_m3++; break;
        case OP_MULM3V3: CHK(c->sp_m3 >= 1 && c->sp_v3 >= 1); c->sp_m3--; c->sp_v3--;
            FORL { float m[9], v[3]; for (int k = 0; k < 9; ++k) m[k] = c->m3[c->sp_m3][k].l[l]; for (int k = 0; k < 3; ++k) v[k] = c->v3[c->sp_v3][k].l[l];
                   for (int r = 0; r < 3; ++r) c->v3[c->sp_v3][r].l[l] = m[r] * v[0] + m[3 + r] * v[1] + m[6 + r] * v[2]; }
            c->sp_v3++; break;
        case OP_DETM3: CHK(c->sp_m3 >= 1 && c->sp_f < 32); c->sp_m3--;
            FORL { float m[9]; for (int k = 0; k < 9; ++k) m[k] = c->m3[c->sp_m3][k].l[l];
                   c->f[c->sp_f].l[l] = m[0] * (m[4] * m[8] - m[5] * m[7]) - m[3] * (m[1] * m[8] - m[2] * m[7]) + m[6] * (m[1] * m[5] - m[2] * m[4]); }
            c->sp_f++; break;
        case OP_TRACEM3: CHK(c->sp_m3 >= 1 && c->sp_f < 32); c->sp_m3--;
            FORL { F8 *M = c->m3[c->sp_m3]; c->f[c->sp_f].l[l] = M[0].l[l] + M[4].l[l] + M[8].l[l]; }
            c->sp_f++; break;
        case OP_TRANSPOSEM3: CHK(c->sp_m3 >= 1); { F8 *M = c->m3[c->sp_m3 - 1]; F8 t[9]; memcpy(t, M, sizeof(t));
            static const int perm[9] = {0, 3, 6, 1, 4, 7, 2, 5, 8}; for (int k = 0; k < 9; ++k) M[k] = t[perm[k]]; } break;
        case OP_INVERSEM3: CHK(c->sp_m3 >= 1);
            FORL { F8 *M = c->m3[c->sp_m3 - 1]; float m[9]; for (int k = 0; k < 9; ++k) m[k] = M[k].l[l];
                   float det = m[0] * (m[4] * m[8] - m[5] * m[7]) - m[3] * (m[1] * m[8] - m[2] * m[7]) + m[6] * (m[1] * m[5] - m[2] * m[4]);
                   int sing = fabsf(det) < 1e-9f;
                   float inv = 1.0f / (sing ? 1.0f : det);
                   float adj[9] = {
                       (m[4] * m[8] - m[5] * m[7]) * inv, -(m[1] * m[8] - m[2] * m[7]) * inv, (m[1] * m[5] - m[2] * m[4]) * inv,
                       -(m[3] * m[8] - m[5] * m[6]) * inv, (m[0] * m[8] - m[2] * m[6]) * inv, -(m[0] * m[5] - m[2] * m[3]) * inv,
                       (m[3] * m[7] - m[4] * m[6]) * inv, -(m[0] * m[7] - m[1] * m[6]) * inv, (m[0] * m[4] - m[1] * m[3]) * inv };
                   for (int k = 0; k < 9; ++k) M[k].l[l] = sing ? ((k == 0 || k == 4 || k == 8) ? 1.0f : 0.0f) : adj[k]; }
            break;
        default: return ORC_E_OPCODE;
        }
    }
    *out = c->f[0]; /* stack_f[0] (evaluator.rs:149) */
    return ORC_OK;
}

/* =====================================================================
 * Dual VM (state.rs:56-79, evaluator.rs:228-388, autodiff.rs)
 * ===================================================================== */
typedef struct { F8 v, g; } D8;
typedef struct {
    D8 f[32]; int sp_f;
    D8 v2[32][2]; int sp_v2;
    D8 v3[32][3]; int sp_v3;
    D8 m2[32][4]; int sp_m2;
    D8 m3[32][9]; int sp_m3;
} DVm;

static D8 d_const(F8 v) { D8 r; r.v = v; r.g = splat(0.0f); return r; }
static D8 d_add(D8 a, D8 b) { D8 r; FORL { r.v.l[l] = a.v.l[l] + b.v.l[l]; r.g.l[l] = a.g.l[l] + b.g.l[l]; } return r; }
static D8 d_sub(D8 a, D8 b) { D8 r; FORL { r.v.l[l] = a.v.l[l] - b.v.l[l]; r.g.l[l] = a.g.l[l] - b.g.l[l]; } return r; }
static D8 d_mul(D8 a, D8 b) { D8 r; FORL { r.v.l[l] = a.v.l[l] * b.v.l[l]; r.g.l[l] = (a.v.l[l] * b.g.l[l]) + (a.g.l[l] * b.v.l[l]); } return r; }
static D8 d_div(D8 a, D8 b) { D8 r; FORL { r.v.l[l] = a.v.l[l] / b.v.l[l]; r.g.l[l] = ((a.g.l[l] * b.v.l[l]) - (a.v.l[l] * b.g.l[l])) / (b.v.l[l] * b.v.l[l]); } return r; }
static D8 d_sin(D8 a) { D8 r; FORL { r.v.l[l] = orc_sinf(a.v.l[l]); r.g.l[l] = a.g.l[l] * orc_cosf(a.v.l[l]); } return r; }
static D8 d_cos(D8 a) { D8 r; FORL { r.v.l[l] = orc_cosf(a.v.l[l]); r.g.l[l] = a.g.l[l] * (-orc_sinf(a.v.l[l])); } return r; }
static D8 d_exp(D8 a) { D8 r; FORL { float e = orc_expf(a.v.l[l]); r.v.l[l] = e; r.g.l[l] = a.g.l[l] * e; } return r; }
static D8 d_sqr(D8 a) { D8 r; FORL { r.v.l[l] = a.v.l[l] * a.v.l[l]; r.g.l[l] = (2.0f * a.v.l[l]) * a.g.l[l]; } return r; }
static D8 d_sqrt(D8 a) { D8 r; FORL { float s = sqrtf(a.v.l[l]); r.v.l[l] = s; r.g.l[l] = a.g.l[l] / (2.0f * s); } return r; }
static D8 d_ln(D8 a) { D8 r; FORL { r.v.l[l] = orc_logf(a.v.l[l]); r.g.l[l] = a.g.l[l] / a.v.l[l]; } return r; }
static D8 d_neg(D8 a) { D8 r; FORL { r.v.l[l] = -a.v.l[l]; r.g.l[l] = -a.g.l[l]; } return r; }

/* flat parameter index of constant c (evaluator.rs:232-245: get_flat_start_idx) */
static int flat_start(const OrcProgram *p, int c) {
    int f = 0;
    for (int i = 0; i < c; ++i) f += vt_size(p->ctypes[i]);
    return f;
}

static int eval_simd_dual8(const OrcProgram *p, const F8 *feat, int nfeat, int active, D8 *out) {
    static __thread DVm vm;
    memset(&vm, 0, sizeof(vm)); /* DualVmState::new() zeroes (state.rs:76-78) */
    DVm *c = &vm;
    for (int pc = 0; pc < p->n_code; ++pc) {
        int op = p->ops[pc];
        int a = p->args[pc];
        switch (op) {
        case OP_LOADVARF: CHK(c->sp_f < 32); if (a >= nfeat) return ORC_E_ARG; c->f[c->sp_f++] = d_const(feat[a]); break;
        case OP_LOADVARV2: CHK(c->sp_v2 < 32); if (a + 1 >= nfeat) return ORC_E_ARG; for (int k = 0; k < 2; ++k) c->v2[c->sp_v2][k] = d_const(feat[a + k]); c->sp_v2++; break;
        case OP_LOADVARV3: CHK(c->sp_v3 < 32); if (a + 2 >= nfeat) return ORC_E_ARG; for (int k = 0; k < 3; ++k) c->v3[c->sp_v3][k] = d_const(feat[a + k]); c->sp_v3++; break;
        case OP_LOADVARM2: CHK(c->sp_m2 < 32); if (a + 3 >= nfeat) return ORC_E_ARG; for (int k = 0; k < 4; ++k) c->m2[c->sp_m2][k] = d_const(feat[a + k]); c->sp_m2++; break;
        case OP_LOADVARM3: CHK(c->sp_m3 < 32); if (a + 8 >= nfeat) return ORC_E_ARG; for (int k = 0; k < 9; ++k) c->m3[c->sp_m3][k] = d_const(feat[a + k]); c->sp_m3++; break;
        case OP_LOADCONSTF: case OP_LOADCONSTV2: case OP_LOADCONSTV3: case OP_LOADCONSTM2: case OP_LOADCONSTM3: {
            if (a >= p->n_consts) return ORC_E_ARG;
            int want = op == OP_LOADCONSTF ? VT_FLOAT : op == OP_LOADCONSTV2 ? VT_VEC2 : op == OP_LOADCONSTV3 ? VT_VEC3 : op == OP_LOADCONSTM2 ? VT_MAT2 : VT_MAT3;
            int n = vt_size(want);
            D8 *slot; int *sp;
            switch (want) {
                case VT_FLOAT: slot = &c->f[c->sp_f]; sp = &c->sp_f; break;
                case VT_VEC2: slot = c->v2[c->sp_v2]; sp = &c->sp_v2; break;
                case VT_VEC3: slot = c->v3[c->sp_v3]; sp = &c->sp_v3; break;
                case VT_MAT2: slot = c->m2[c->sp_m2]; sp = &c->sp_m2; break;
                default: slot = c->m3[c->sp_m3]; sp = &c->sp_m3; break;
            }
            CHK(*sp < 32);
            if (p->ctypes[a] == want) {
                int o = cval_offset(p, a), fl = flat_start(p, a);
                for (int k = 0; k < n; ++k) { slot[k].v = splat(p->cvals[o + k]); slot[k].g = splat(fl + k == active ? 1.0f : 0.0f); }
            }
            (*sp)++;
            break; }
        case OP_LOADVARB: case OP_LOADCONSTB: case OP_LOADVARI: case OP_LOADCONSTI: break;

        case OP_ADDF: CHK(c->sp_f >= 2); c->sp_f -= 2; c->f[c->sp_f] = d_add(c->f[c->sp_f], c->f[c->sp_f + 1]); c->sp_f++; break;
        case OP_SUBF: CHK(c->sp_f >= 2); c->sp_f -= 2; c->f[c->sp_f] = d_sub(c->f[c->sp_f], c->f[c->sp_f + 1]); c->sp_f++; break;
        case OP_MULF: CHK(c->sp_f >= 2); c->sp_f -= 2; c->f[c->sp_f] = d_mul(c->f[c->sp_f], c->f[c->sp_f + 1]); c->sp_f++; break;
        case OP_DIVF: CHK(c->sp_f >= 2); c->sp_f -= 2; c->f[c->sp_f] = d_div(c->f[c->sp_f], c->f[c->sp_f + 1]); c->sp_f++; break;
        case OP_SINF: CHK(c->sp_f >= 1); c->f[c->sp_f - 1] = d_sin(c->f[c->sp_f - 1]); break;
        case OP_COSF: CHK(c->sp_f >= 1); c->f[c->sp_f - 1] = d_cos(c->f[c->sp_f - 1]); break;
        case OP_EXPF: CHK(c->sp_f >= 1); c->f[c->sp_f - 1] = d_exp(c->f[c->sp_f - 1]); break;
        case OP_SQRF: CHK(c->sp_f >= 1); c->f[c->sp_f - 1] = d_sqr(c->f[c->sp_f - 1]); break;
        case OP_SQRTF: CHK(c->sp_f >= 1); c->f[c->sp_f - 1] = d_sqrt(c->f[c->sp_f - 1]); break;
        case OP_LNF: CHK(c->sp_f >= 1); c->f[c->sp_f - 1] = d_ln(c->f[c->sp_f - 1]); break;
        case OP_NEGF: CHK(c->sp_f >= 1); c->f[c->sp_f - 1] = d_neg(c->f[c->sp_f - 1]); break;

        case OP_MAKEVEC2: CHK(c->sp_f >= 2 && c->sp_v2 < 32); c->sp_f -= 2; c->v2[c->sp_v2][0] = c->f[c->sp_f]; c->v2[c->sp_v2][1] = c->f[c->sp_f + 1]; c->sp_v2++; break;
        case OP_MAKEVEC3: CHK(c->sp_f >= 3 && c->sp_v3 < 32); c->sp_f -= 3; for (int k = 0; k < 3; ++k) c->v3[c->sp_v3][k] = c->f[c->sp_f + k]; c->sp_v3++; break;
        case OP_GETXV2: case OP_GETYV2: CHK(c->sp_v2 >= 1 && c->sp_f < 32); c->sp_v2--; c->f[c->sp_f++] = c->v2[c->sp_v2][op == OP_GETXV2 ? 0 : 1]; break;
        case OP_GETXV3: case OP_GETYV3: case OP_GETZV3: CHK(c->sp_v3 >= 1 && c->sp_f < 32); c->sp_v3--; c->f[c->sp_f++] = c->v3[c->sp_v3][op - OP_GETXV3]; break;
        case OP_ADDV2: case OP_SUBV2: CHK(c->sp_v2 >= 2); c->sp_v2 -= 2;
            for (int k = 0; k < 2; ++k) c->v2[c->sp_v2][k] = op == OP_ADDV2 ? d_add(c->v2[c->sp_v2][k], c->v2[c->sp_v2 + 1][k]) : d_sub(c->v2[c->sp_v2][k], c->v2[c->sp_v2 + 1][k]);
            c->sp_v2++; break;
        case OP_SCALEV2: CHK(c->sp_f >= 1 && c->sp_v2 >= 1); c->sp_f--; c->sp_v2--;
            for (int k = 0; k < 2; ++k) c->v2[c->sp_v2][k] = d_mul(c->v2[c->sp_v2][k], c->f[c->sp_f]); c->sp_v2++; break;
        case OP_DOTV2: CHK(c->sp_v2 >= 2 && c->sp_f < 32); c->sp_v2 -= 2;
            { D8 *A = c->v2[c->sp_v2], *B = c->v2[c->sp_v2 + 1]; c->f[c->sp_f] = d_add(d_mul(A[0], B[0]), d_mul(A[1], B[1])); }
            c->sp_f++; break;
        case OP_NORMV2: CHK(c->sp_v2 >= 1 && c->sp_f < 32);
            { D8 *V = c->v2[c->sp_v2 - 1]; c->f[c->sp_f] = d_sqrt(d_add(d_mul(V[0], V[0]), d_mul(V[1], V[1]))); }
            c->sp_f++; c->sp_v2--; break;
        case OP_ADDV3: case OP_SUBV3: CHK(c->sp_v3 >= 2); c->sp_v3 -= 2;
            for (int k = 0; k < 3; ++k) c->v3[c->sp_v3][k] = op == OP_ADDV3 ? d_add(c->v3[c->sp_v3][k], c->v3[c->sp_v3 + 1][k]) : d_sub(c->v3[c->sp_v3][k], c->v3[c->sp_v3 + 1][k]);
            c->sp_v3++; break;
        case OP_SCALEV3: CHK(c->sp_f >= 1 && c->sp_v3 >= 1); c->sp_f--; c->sp_v3--;
            for (int k = 0; k < 3; ++k) c->v3[c->sp_v3][k] = d_mul(c->v3[c->sp_v3][k], c->f[c->sp_f]); c->sp_v3++; break;
        case OP_DOTV3: CHK(c->sp_v3 >= 2 && c->sp_f < 32); c->sp_v3 -= 2;
            { /* dual_dot_v3 (autodiff.rs:127-133) */
              D8 *A = c->v3[c->sp_v3], *B = c->v3[c->sp_v3 + 1]; D8 r;
              FORL { r.v.l[l] = (A[0].v.l[l] * B[0].v.l[l]) + (A[1].v.l[l] * B[1].v.l[l]) + (A[2].v.l[l] * B[2].v.l[l]);
                     r.g.l[l] = (A[0].v.l[l] * B[0].g.l[l] + A[0].g.l[l] * B[0].v.l[l])
                              + (A[1].v.l[l] * B[1].g.l[l] + A[1].g.l[l] * B[1].v.l[l])
                              + (A[2].v.l[l] * B[2].g.l[l] + A[2].g.l[l] * B[2].v.l[l]); }
              c->f[c->sp_f] = r; }
            c->sp_f++; break;
        case OP_NORMV3: CHK(c->sp_v3 >= 1 && c->sp_f < 32);
            { D8 *V = c->v3[c->sp_v3 - 1]; c->f[c->sp_f] = d_sqrt(d_add(d_add(d_mul(V[0], V[0]), d_mul(V[1], V[1])), d_mul(V[2], V[2]))); }
            c->sp_f++; c->sp_v3--; break;
        case OP_CROSSV3: CHK(c->sp_v3 >= 2); c->sp_v3 -= 2;
            { D8 A[3], B[3]; memcpy(A, c->v3[c->sp_v3], sizeof(A)); memcpy(B, c->v3[c->sp_v3 + 1], sizeof(B));
              c->v3[c->sp_v3][0] = d_sub(d_mul(A[1], B[2]), d_mul(A[2], B[1]));
              c->v3[c->sp_v3][1] = d_sub(d_mul(A[2], B[0]), d_mul(A[0], B[2]));
              c->v3[c->sp_v3][2] = d_sub(d_mul(A[0], B[1]), d_mul(A[1], B[0])); }
            c->sp_v3++; break;
        case OP_MAKEMAT2: CHK(c->sp_v2 >= 2 && c->sp_m2 < 32); c->sp_v2 -= 2;
            c->m2[c->sp_m2][0] = c->v2[c->sp_v2][0]; c->m2[c->sp_m2][1] = c->v2[c->sp_v2][1];
            c->m2[c->sp_m2][2] = c->v2[c->sp_v2 + 1][0]; c->m2[c->sp_m2][3] = c->v2[c->sp_v2 + 1][1]; c->sp_m2++; break;
        case OP_ADDM2: case OP_SUBM2: CHK(c->sp_m2 >= 2); c->sp_m2 -= 2;
            for (int k = 0; k < 4; ++k) c->m2[c->sp_m2][k] = op == OP_ADDM2 ? d_add(c->m2[c->sp_m2][k], c->m2[c->sp_m2 + 1][k]) : d_sub(c->m2[c->sp_m2][k], c->m2[c->sp_m2 + 1][k]);
            c->sp_m2++; break;
        case OP_SCALEM2: CHK(c->sp_f >= 1 && c->sp_m2 >= 1); c->sp_f--; c->sp_m2--;
            for (int k = 0; k < 4; ++k) c->m2[c->sp_m2][k] = d_mul(c->m2[c->sp_m2][k], c->f[c->sp_f]); c->sp_m2++; break;
        case OP_MULM2: CHK(c->sp_m2 >= 2); c->sp_m2 -= 2;
            { D8 A[4], B[4]; memcpy(A, c->m2[c->sp_m2], sizeof(A)); memcpy(B, c->m2[c->sp_m2 + 1], sizeof(B));
              c->m2[c->sp_m2][0] = d_add(d_mul(A[0], B[0]), d_mul(A[2], B[1]));
              c->m2[c->sp_m2][1] = d_add(d_mul(A[1], B[0]), d_mul(A[3], B[1]));
              c->m2[c->sp_m2][2] = d_add(d_mul(A[0], B[2]), d_mul(A[2], B[3]));
              c->m2[c->sp_m2][3] = d_add(d_mul(A[1], B[2]), d_mul(A[3], B[3])); }
            c->sp_m2++; break;
        case OP_MULM2V2: CHK(c->sp_m2 >= 1 && c->sp_v2 >= 1); c->sp_m2--; c->sp_v2--;
            { D8 M[4], V[2]; memcpy(M, c->m2[c->sp_m2], sizeof(M)); memcpy(V, c->v2[c->sp_v2], sizeof(V));
              c->v2[c->sp_v2][0] = d_add(d_mul(M[0], V[0]), d_mul(M[2], V[1]));
              c->v2[c->sp_v2][1] = d_add(d_mul(M[1], V[0]), d_mul(M[3], V[1])); }
            c->sp_v2++; break;
        case OP_DETM2: CHK(c->sp_m2 >= 1 && c->sp_f < 32); c->sp_m2--;
            { D8 *M = c->m2[c->sp_m2]; c->f[c->sp_f] = d_sub(d_mul(M[0], M[3]), d_mul(M[1], M[2])); } c->sp_f++; break;
        case OP_TRACEM2: CHK(c->sp_m2 >= 1 && c->sp_f < 32); c->sp_m2--;
            { D8 *M = c->m2[c->sp_m2]; c->f[c->sp_f] = d_add(M[0], M[3]); } c->sp_f++; break;
        case OP_TRANSPOSEM2: CHK(c->sp_m2 >= 1); { D8 *M = c->m2[c->sp_m2 - 1]; D8 t = M[1]; M[1] = M[2]; M[2] = t; } break;
        case OP_INVERSEM2: case OP_INVERSEM3: break; /* dual inverse is a no-op (linalg/mod.rs:643, 687) */
        case OP_MAKEMAT3: CHK(c->sp_v3 >= 3 && c->sp_m3 < 32); c->sp_v3 -= 3;
            for (int col = 0; col < 3; ++col) for (int k = 0; k < 3; ++k) c->m3[c->sp_m3][col * 3 + k] = c->v3[c->sp_v3 + col][k];
            c->sp_m3++; break;
        case OP_ADDM3: case OP_SUBM3: CHK(c->sp_m3 >= 2); c->sp_m3 -= 2;
            for (int k = 0; k < 9; ++k) c->m3[c->sp_m3][k] = op == OP_ADDM3 ? d_add(c->m3[c->sp_m3][k], c->m3[c->sp_m3 + 1][k]) : d_sub(c->m3[c->sp_m3][k], c->m3[c->sp_m3 + 1][k]);
            c->sp_m3++; break;
        case OP_SCALEM3: CHK(c->sp_f >= 1 && c->sp_m3 >= 1); c->sp_f--; c->sp_m3--;
            for (int k = 0; k < 9; ++k) c->m3[c->sp_m3][k] = d_mul(c->m3[c->sp_m3][k], c->f[c->sp_f]); c->sp_m3++; break;
        case OP_MULM3: CHK(c->sp_m3 >= 2); c->sp_m3 -= 2;
            { D8 A[9], B[9]; memcpy(A, c->m3[c->sp_m3], sizeof(A)); memcpy(B, c->m3[c->sp_m3 + 1], sizeof(B));
              for (int col = 0; col < 3; ++col) for (int r = 0; r < 3; ++r)
                  c->m3[c->sp_m3][col * 3 + r] = d_add(d_add(d_mul(A[r], B[col * 3]), d_mul(A[3 + r], B[col * 3 + 1])), d_mul(A[6 + r], B[col * 3 + 2])); }
            c->sp_m3++; break;
        case OP_MULM3V3: CHK(c->sp_m3 >= 1 && c->sp_v3 >= 1); c->sp_m3--; c->sp_v3--;
            { D8 M[9], V[3]; memcpy(M, c->m3[c->sp_m3], sizeof(M)); memcpy(V, c->v3[c->sp_v3], sizeof(V));
              for (int r = 0; r < 3; ++r) c->v3[c->sp_v3][r] = d_add(d_add(d_mul(M[r], V[0]), d_mul(M[3 + r], V[1])), d_mul(M[6 + r], V[2])); }
            c->sp_v3++; break;
        case OP_DETM3: CHK(c->sp_m3 >= 1 && c->sp_f < 32); c->sp_m3--;
            { D8 *m = c->m3[c->sp_m3];
              D8 t1 = d_sub(d_mul(m[4], m[8]), d_mul(m[5], m[7]));
              D8 t2 = d_sub(d_mul(m[1], m[8]), d_mul(m[2], m[7]));
              D8 t3 = d_sub(d_mul(m[1], m[5]), d_mul(m[2], m[4]));
              c->f[c->sp_f] = d_add(d_sub(d_mul(m[0], t1), d_mul(m[3], t2)), d_mul(m[6], t3)); }
            c->sp_f++; break;
        case OP_TRACEM3: CHK(c->sp_m3 >= 1 && c->sp_f < 32); c->sp_m3--;
            { D8 *M = c->m3[c->sp_m3]; c->f[c->sp_f] = d_add(d_add(M[0], M[4]), M[8]); } c->sp_f++; break;
        case OP_TRANSPOSEM3: CHK(c->sp_m3 >= 1); { D8 *M = c->m3[c->sp_m3 - 1]; D8 t[9]; memcpy(t, M, sizeof(t));
            static const int perm[9] = {0, 3, 6, 1, 4, 7, 2, 5, 8}; for (int k = 0; k < 9; ++k) M[k] = t[perm[k]]; } break;
        default: return ORC_E_OPCODE;
        }
    }
    *out = c->f[0];
    return ORC_OK;
}

/* ---- reduce_add of an f32x8 (order inside `wide`, unpinned; see header) ---- */
static float reduce_add8(const F8 *a) {
    float s0 = a->l[0] + a->l[4], s1 = a->l[1] + a->l[5], s2 = a->l[2] + a->l[6], s3 = a->l[3] + a->l[7];
    return (s0 + s2) + (s1 + s3);
}

static OrcProgram mkprog(const uint8_t *ops, const uint16_t *args, int n_code,
                         const uint8_t *ctypes, const float *cvals, int n_consts) {
    OrcProgram p = {ops, args, n_code, ctypes, cvals, n_consts};
    return p;
}

/* eval_simd over every batch: per-point values (num_batches*8 floats, padding lanes included). */
ORC_API int orc_eval_points(const uint8_t *ops, const uint16_t *args, int n_code,
                            const uint8_t *ctypes, const float *cvals, int n_consts,
                            const float *feature_flat, int num_features, int64_t num_batches,
                            float *out) {
    OrcProgram p = mkprog(ops, args, n_code, ctypes, cvals, n_consts);
    for (int64_t b = 0; b < num_batches; ++b) {
        F8 r;
        int rc = eval_simd8(&p, (const F8 *)(feature_flat + b * num_features * LANES), num_features, &r);
        if (rc) return rc;
        memcpy(out + b * LANES, r.l, sizeof(r.l));
    }
    return ORC_OK;
}

/* compute_mse (evaluator.rs:153-174) */
ORC_API int orc_compute_mse(const uint8_t *ops, const uint16_t *args, int n_code,
                            const uint8_t *ctypes, const float *cvals, int n_consts,
                            const float *feature_flat, const float *target_batches, int num_features,
                            int64_t num_batches, int64_t num_samples, float *mse_out) {
    OrcProgram p = mkprog(ops, args, n_code, ctypes, cvals, n_consts);
    F8 acc = splat(0.0f);
    for (int64_t b = 0; b < num_batches; ++b) {
        F8 r;
        int rc = eval_simd8(&p, (const F8 *)(feature_flat + b * num_features * LANES), num_features, &r);
        if (rc) return rc;
        const float *t = target_batches + b * LANES;
        FORL { float d = r.l[l] - t[l]; acc.l[l] = acc.l[l] + d * d; }
    }
    float mse = reduce_add8(&acc) / (float)num_samples;
    *mse_out = isfinite(mse) ? mse : FLT_MAX;
    return ORC_OK;
}

static int param_count(const OrcProgram *p) {
    int n = 0;
    for (int i = 0; i < p->n_consts; ++i) n += vt_size(p->ctypes[i]);
    return n;
}

/* compute_mse_with_gradient (evaluator.rs:177-224): one dual pass per parameter. */
ORC_API int orc_compute_mse_with_gradient(const uint8_t *ops, const uint16_t *args, int n_code,
                                          const uint8_t *ctypes, const float *cvals, int n_consts,
                                          const float *feature_flat, const float *target_batches,
                                          int num_features, int64_t num_batches, int64_t num_samples,
                                          float *mse_out, float *grad_out /* [32] */) {
    OrcProgram p = mkprog(ops, args, n_code, ctypes, cvals, n_consts);
    int na = param_count(&p);
    if (na > 32) na = 32;
    for (int k = 0; k < 32; ++k) grad_out[k] = 0.0f;
    if (na == 0)
        return orc_compute_mse(ops, args, n_code, ctypes, cvals, n_consts, feature_flat, target_batches,
                               num_features, num_batches, num_samples, mse_out);
    F8 acc = splat(0.0f);
    F8 gs[32];
    for (int k = 0; k < 32; ++k) gs[k] = splat(0.0f);
    for (int64_t b = 0; b < num_batches; ++b) {
        const F8 *feat = (const F8 *)(feature_flat + b * num_features * LANES);
        const float *t = target_batches + b * LANES;
        F8 diff = splat(0.0f);
        for (int k = 0; k < na; ++k) {
            D8 r;
            int rc = eval_simd_dual8(&p, feat, num_features, k, &r);
            if (rc) return rc;
            if (k == 0) {
                FORL { diff.l[l] = r.v.l[l] - t[l]; acc.l[l] = acc.l[l] + diff.l[l] * diff.l[l]; }
            }
            FORL gs[k].l[l] = gs[k].l[l] + (2.0f * diff.l[l]) * r.g.l[l];
        }
    }
    float ns = (float)num_samples;
    float mse = reduce_add8(&acc) / ns;
    if (!isfinite(mse)) { *mse_out = FLT_MAX; return ORC_OK; }
    for (int k = 0; k < na; ++k) grad_out[k] = reduce_add8(&gs[k]) / ns;
    *mse_out = mse;
    return ORC_OK;
}

/* eval_simd_dual over all batches for one active parameter: values and tangents per point. */
ORC_API int orc_eval_points_dual(const uint8_t *ops, const uint16_t *args, int n_code,
                                 const uint8_t *ctypes, const float *cvals, int n_consts,
                                 const float *feature_flat, int num_features, int64_t num_batches,
                                 int active, float *val_out, float *grad_out) {
    OrcProgram p = mkprog(ops, args, n_code, ctypes, cvals, n_consts);
    for (int64_t b = 0; b < num_batches; ++b) {
        D8 r;
        int rc = eval_simd_dual8(&p, (const F8 *)(feature_flat + b * num_features * LANES), num_features, active, &r);
        if (rc) return rc;
        memcpy(val_out + b * LANES, r.v.l, sizeof(r.v.l));
        memcpy(grad_out + b * LANES, r.g.l, sizeof(r.g.l));
    }
    return ORC_OK;
}

/* =====================================================================
 * FP64 truth: the same formulas in double with libm transcendentals,
 * evaluated per point (padding lanes included, divisor = real N).
 * ===================================================================== */
typedef struct {
    double f[32]; int sp_f;
    double v2[32][2]; int sp_v2;
    double v3[32][3]; int sp_v3;
    double m2[32][4]; int sp_m2;
    double m3[32][9]; int sp_m3;
} VmD;

static int eval_point_f64(const OrcProgram *p, const double *x, int nfeat, double *out) {
    static __thread VmD vm;
    memset(&vm, 0, sizeof(vm));
    VmD *c = &vm;
    for (int pc = 0; pc < p->n_code; ++pc) {
        int op = p->ops[pc], a = p->args[pc];
        switch (op) {
        case OP_LOADVARF: CHK(c->sp_f < 32); if (a >= nfeat) return ORC_E_ARG; c->f[c->sp_f++] = x[a]; break;
        case OP_LOADVARV2: CHK(c->sp_v2 < 32); if (a + 1 >= nfeat) return ORC_E_ARG; for (int k = 0; k < 2; ++k) c->v2[c->sp_v2][k] = x[a + k]; c->sp_v2++; break;
        case OP_LOADVARV3: CHK(c->sp_v3 < 32); if (a + 2 >= nfeat) return ORC_E_ARG; for (int k = 0; k < 3; ++k) c->v3[c->sp_v3][k] = x[a + k]; c->sp_v3++; break;
        case OP_LOADVARM2: CHK(c->sp_m2 < 32); if (a + 3 >= nfeat) return ORC_E_ARG; for (int k = 0; k < 4; ++k) c->m2[c->sp_m2][k] = x[a + k]; c->sp_m2++; break;
        case OP_LOADVARM3: CHK(c->sp_m3 < 32); if (a + 8 >= nfeat) return ORC_E_ARG; for (int k = 0; k < 9; ++k) c->m3[c->sp_m3][k] = x[a + k]; c->sp_m3++; break;
        case OP_LOADCONSTF: CHK(c->sp_f < 32); if (a >= p->n_consts) return ORC_E_ARG; if (p->ctypes[a] == VT_FLOAT) c->f[c->sp_f] = p->cvals[cval_offset(p, a)]; c->sp_f++; break;
        case OP_LOADCONSTV2: CHK(c->sp_v2 < 32); if (a >= p->n_consts) return ORC_E_ARG; if (p->ctypes[a] == VT_VEC2) { int o = cval_offset(p, a); for (int k = 0; k < 2; ++k) c->v2[c->sp_v2][k] = p->cvals[o + k]; } c->sp_v2++; break;
        case OP_LOADCONSTV3: CHK(c->sp_v3 < 32); if (a >= p->n_consts) return ORC_E_ARG; if (p->ctypes[a] == VT_VEC3) { int o = cval_offset(p, a); for (int k = 0; k < 3; ++k) c->v3[c->sp_v3][k] = p->cvals[o + k]; } c->sp_v3++; break;
        case OP_LOADCONSTM2: CHK(c->sp_m2 < 32); if (a >= p->n_consts) return ORC_E_ARG; if (p->ctypes[a] == VT_MAT2) { int o = cval_offset(p, a); for (int k = 0; k < 4; ++k) c->m2[c->sp_m2][k] = p->cvals[o + k]; } c->sp_m2++; break;
        case OP_LOADCONSTM3: CHK(c->sp_m3 < 32); if (a >= p->n_consts) return ORC_E_ARG; if (p->ctypes[a] == VT_MAT3) { int o = cval_offset(p, a); for (int k = 0; k < 9; ++k) c->m3[c->sp_m3][k] = p->cvals[o + k]; } c->sp_m3++; break;
        case OP_LOADVARB: case OP_LOADCONSTB: case OP_LOADVARI: case OP_LOADCONSTI: break;
        case OP_ADDF: CHK(c->sp_f >= 2); c->sp_f -= 2; c->f[c->sp_f] = c->f[c->sp_f] + c->f[c->sp_f + 1]; c->sp_f++; break;
        case OP_SUBF: CHK(c->sp_f >= 2); c->sp_f -= 2; c->f[c->sp_f] = c->f[c->sp_f] - c->f[c->sp_f + 1]; c->sp_f++; break;
        case OP_MULF: CHK(c->sp_f >= 2); c->sp_f -= 2; c->f[c->sp_f] = c->f[c->sp_f] * c->f[c->sp_f + 1]; c->sp_f++; break;
        case OP_DIVF: CHK(c->sp_f >= 2); c->sp_f -= 2; c->f[c->sp_f] = c->f[c->sp_f] / c->f[c->sp_f + 1]; c->sp_f++; break;
        case OP_SINF: CHK(c->sp_f >= 1); c->f[c->sp_f - 1] = sin(c->f[c->sp_f - 1]); break;
        case OP_COSF: CHK(c->sp_f >= 1); c->f[c->sp_f - 1] = cos(c->f[c->sp_f - 1]); break;
        case OP_EXPF: CHK(c->sp_f >= 1); c->f[c->sp_f - 1] = exp(c->f[c->sp_f - 1]); break;
        case OP_SQRF: CHK(c->sp_f >= 1); c->f[c->sp_f - 1] = c->f[c->sp_f - 1] * c->f[c->sp_f - 1]; break;
        case OP_SQRTF: CHK(c->sp_f >= 1); c->f[c->sp_f - 1] = sqrt(c->f[c->sp_f - 1]); break;
        case OP_LNF: CHK(c->sp_f >= 1); c->f[c->sp_f - 1] = log(c->f[c->sp_f - 1]); break;
        case OP_NEGF: CHK(c->sp_f >= 1); c->f[c->sp_f - 1] = -c->f[c->sp_f - 1]; break;
        case OP_MAKEVEC2: CHK(c->sp_f >= 2 && c->sp_v2 < 32); c->sp_f -= 2; c->v2[c->sp_v2][0] = c->f[c->sp_f]; c->v2[c->sp_v2][1] = c->f[c->sp_f + 1]; c->sp_v2++; break;
        case OP_MAKEVEC3: CHK(c->sp_f >= 3 && c->sp_v3 < 32); c->sp_f -= 3; for (int k = 0; k < 3; ++k) c->v3[c->sp_v3][k] = c->f[c->sp_f + k]; c->sp_v3++; break;
        case OP_GETXV2: case OP_GETYV2: CHK(c->sp_v2 >= 1 && c->sp_f < 32); c->sp_v2--; c->f[c->sp_f++] = c->v2[c->sp_v2][op == OP_GETXV2 ? 0 : 1]; break;
        case OP_GETXV3: case OP_GETYV3: case OP_GETZV3: CHK(c->sp_v3 >= 1 && c->sp_f < 32); c->sp_v3--; c->f[c->sp_f++] = c->v3[c->sp_v3][op - OP_GETXV3]; break;
        case OP_ADDV2: case OP_SUBV2: CHK(c->sp_v2 >= 2); c->sp_v2 -= 2; for (int k = 0; k < 2; ++k) c->v2[c->sp_v2][k] = op == OP_ADDV2 ? c->v2[c->sp_v2][k] + c->v2[c->sp_v2 + 1][k] : c->v2[c->sp_v2][k] - c->v2[c->sp_v2 + 1][k]; c->sp_v2++; break;
        case OP_SCALEV2: CHK(c->sp_f >= 1 && c->sp_v2 >= 1); c->sp_f--; c->sp_v2--; for (int k = 0; k < 2; ++k) c->v2[c->sp_v2][k] *= c->f[c->sp_f]; c->sp_v2++; break;
        case OP_DOTV2: CHK(c->sp_v2 >= 2 && c->sp_f < 32); c->sp_v2 -= 2; c->f[c->sp_f++] = c->v2[c->sp_v2][0] * c->v2[c->sp_v2 + 1][0] + c->v2[c->sp_v2][1] * c->v2[c->sp_v2 + 1][1]; break;
        case OP_NORMV2: CHK(c->sp_v2 >= 1 && c->sp_f < 32); { double *v = c->v2[c->sp_v2 - 1]; c->f[c->sp_f++] = sqrt(v[0] * v[0] + v[1] * v[1]); } c->sp_v2--; break;
        case OP_ADDV3: case OP_SUBV3: CHK(c->sp_v3 >= 2); c->sp_v3 -= 2; for (int k = 0; k < 3; ++k) c->v3[c->sp_v3][k] = op == OP_ADDV3 ? c->v3[c->sp_v3][k] + c->v3[c->sp_v3 + 1][k] : c->v3[c->sp_v3][k] - c->v3[c->sp_v3 + 1][k]; c->sp_v3++; break;
        case OP_SCALEV3: CHK(c->sp_f >= 1 && c->sp_v3 >= 1); c->sp_f--; c->sp_v3--; for (int k = 0; k < 3; ++k) c->v3[c->sp_v3][k] *= c->f[c->sp_f]; c->sp_v3++; break;
        case OP_DOTV3: CHK(c->sp_v3 >= 2 && c->sp_f < 32); c->sp_v3 -= 2; { double *A = c->v3[c->sp_v3], *B = c->v3[c->sp_v3 + 1]; c->f[c->sp_f++] = A[0] * B[0] + A[1] * B[1] + A[2] * B[2]; } break;
        case OP_NORMV3: CHK(c->sp_v3 >= 1 && c->sp_f < 32); { double *v = c->v3[c->sp_v3 - 1]; c->f[c->sp_f++] = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]); } c->sp_v3--; break;
        case OP_CROSSV3: CHK(c->sp_v3 >= 2); c->sp_v3 -= 2; { double A[3], B[3]; memcpy(A, c->v3[c->sp_v3], sizeof(A)); memcpy(B, c->v3[c->sp_v3 + 1], sizeof(B));
            c->v3[c->sp_v3][0] = A[1] * B[2] - A[2] * B[1]; c->v3[c->sp_v3][1] = A[2] * B[0] - A[0] * B[2]; c->v3[c->sp_v3][2] = A[0] * B[1] - A[1] * B[0]; } c->sp_v3++; break;
        case OP_MAKEMAT2: CHK(c->sp_v2 >= 2 && c->sp_m2 < 32); c->sp_v2 -= 2; c->m2[c->sp_m2][0] = c->v2[c->sp_v2][0]; c->m2[c->sp_m2][1] = c->v2[c->sp_v2][1]; c->m2[c->sp_m2][2] = c->v2[c->sp_v2 + 1][0]; c->m2[c->sp_m2][3] = c->v2[c->sp_v2 + 1][1]; c->sp_m2++; break;
        case OP_ADDM2: case OP_SUBM2: CHK(c->sp_m2 >= 2); c->sp_m2 -= 2; for (int k = 0; k < 4; ++k) c->m2[c->sp_m2][k] = op == OP_ADDM2 ? c->m2[c->sp_m2][k] + c->m2[c->sp_m2 + 1][k] : c->m2[c->sp_m2][k] - c->m2[c->sp_m2 + 1][k]; c->sp_m2++; break;
        case OP_SCALEM2: CHK(c->sp_f >= 1 && c->sp_m2 >= 1); c->sp_f--; c->sp_m2--; for (int k = 0; k < 4; ++k) c->m2[c->sp_m2][k] *= c->f[c->sp_f]; c->sp_m2++; break;
        case OP_MULM2: CHK(c->sp_m2 >= 2); c->sp_m2 -= 2; { double A[4], B[4]; memcpy(A, c->m2[c->sp_m2], sizeof(A)); memcpy(B, c->m2[c->sp_m2 + 1], sizeof(B));
            c->m2[c->sp_m2][0] = A[0] * B[0] + A[2] * B[1]; c->m2[c->sp_m2][1] = A[1] * B[0] + A[3] * B[1]; c->m2[c->sp_m2][2] = A[0] * B[2] + A[2] * B[3]; c->m2[c->sp_m2][3] = A[1] * B[2] + A[3] * B[3]; } c->sp_m2++; break;
        case OP_MULM2V2: CHK(c->sp_m2 >= 1 && c->sp_v2 >= 1); c->sp_m2--; c->sp_v2--; { double *m = c->m2[c->sp_m2], v0 = c->v2[c->sp_v2][0], v1 = c->v2[c->sp_v2][1];
            c->v2[c->sp_v2][0] = m[0] * v0 + m[2] * v1; c->v2[c->sp_v2][1] = m[1] * v0 + m[3] * v1; } c->sp_v2++; break;
        case OP_DETM2: CHK(c->sp_m2 >= 1 && c->sp_f < 32); c->sp_m2--; { double *m = c->m2[c->sp_m2]; c->f[c->sp_f++] = m[0] * m[3] - m[1] * m[2]; } break;
        case OP_TRACEM2: CHK(c->sp_m2 >= 1 && c->sp_f < 32); c->sp_m2--; { double *m = c->m2[c->sp_m2]; c->f[c->sp_f++] = m[0] + m[3]; } break;
        case OP_TRANSPOSEM2: CHK(c->sp_m2 >= 1); { double *m = c->m2[c->sp_m2 - 1]; double t = m[1]; m[1] = m[2]; m[2] = t; } break;
        case OP_INVERSEM2: CHK(c->sp_m2 >= 1); { double *m = c->m2[c->sp_m2 - 1]; double det = m[0] * m[3] - m[1] * m[2];
            if (fabs(det) < 1e-9) { m[0] = 1; m[1] = 0; m[2] = 0; m[3] = 1; }
            else { double inv = 1.0 / det, m0 = m[0], m1 = m[1], m2v = m[2], m3 = m[3]; m[0] = m3 * inv; m[1] = -m1 * inv; m[2] = -m2v * inv; m[3] = m0 * inv; } } break;
        case OP_MAKEMAT3: CHK(c->sp_v3 >= 3 && c->sp_m3 < 32); c->sp_v3 -= 3; for (int col = 0; col < 3; ++col) for (int k = 0; k < 3; ++k) c->m3[c->sp_m3][col * 3 + k] = c->v3[c->sp_v3 + col][k]; c->sp_m3++; break;
        case OP_ADDM3: case OP_SUBM3: CHK(c->sp_m3 >= 2); c->sp_m3 -= 2; for (int k = 0; k < 9; ++k) c->m3[c->sp_m3][k] = op == OP_ADDM3 ? c->m3[c->sp_m3][k] + c->m3[c->sp_m3 + 1][k] : c->m3[c->sp_m3][k] - c->m3[c->sp_m3 + 1][k]; c->sp_m3++; break;
        case OP_SCALEM3: CHK(c->sp_f >= 1 && c->sp_m3 >= 1); c->sp_f--; c->sp_m3--; for (int k = 0; k < 9; ++k) c->m3[c->sp_m3][k] *= c->f[c->sp_f]; c->sp_m3++; break;
        case OP_MULM3: CHK(c->sp_m3 >= 2); c->sp_m3 -= 2; { double A[9], B[9]; memcpy(A, c->m3[c->sp_m3], sizeof(A)); memcpy(B, c->m3[c->sp_m3 + 1], sizeof(B));
            for (int col = 0; col < 3; ++col) for (int r = 0; r < 3; ++r) c->m3[c->sp_m3][col * 3 + r] = A[r] * B[col * 3] + A[3 + r] * B[col * 3 + 1] + A[6 + r] * B[col * 3 + 2]; } c->sp_m3++; break;
        case OP_MULM3V3: CHK(c->sp_m3 >= 1 && c->sp_v3 >= 1); c->sp_m3--; c->sp_v3--; { double *m = c->m3[c->sp_m3], v[3]; memcpy(v, c->v3[c->sp_v3], sizeof(v));
            for (int r = 0; r < 3; ++r) c->v3[c->sp_v3][r] = m[r] * v[0] + m[3 + r] * v[1] + m[6 + r] * v[2]; } c->sp_v3++; break;
        case OP_DETM3: CHK(c->sp_m3 >= 1 && c->sp_f < 32); c->sp_m3--; { double *m = c->m3[c->sp_m3];
            c->f[c->sp_f++] = m[0] * (m[4] * m[8] - m[5] * m[7]) - m[3] * (m[1] * m[8] - m[2] * m[7]) + m[6] * (m[1] * m[5] - m[2] * m[4]); } break;
        case OP_TRACEM3: CHK(c->sp_m3 >= 1 && c->sp_f < 32); c->sp_m3--; { double *m = c->m3[c->sp_m3]; c->f[c->sp_f++] = m[0] + m[4] + m[8]; } break;
        case OP_TRANSPOSEM3: CHK(c->sp_m3 >= 1); { double *M = c->m3[c->sp_m3 - 1], t[9]; memcpy(t, M, sizeof(t));
            static const int perm[9] = {0, 3, 6, 1, 4, 7, 2, 5, 8}; for (int k = 0; k < 9; ++k) M[k] = t[perm[k]]; } break;
        case OP_INVERSEM3: CHK(c->sp_m3 >= 1); { double *M = c->m3[c->sp_m3 - 1], m[9]; memcpy(m, M, sizeof(m));
            double det = m[0] * (m[4] * m[8] - m[5] * m[7]) - m[3] * (m[1] * m[8] - m[2] * m[7]) + m[6] * (m[1] * m[5] - m[2] * m[4]);
            if (fabs(det) < 1e-9) { for (int k = 0; k < 9; ++k) M[k] = (k == 0 || k == 4 || k == 8) ? 1.0 : 0.0; }
            else { double inv = 1.0 / det;
                M[0] = (m[4] * m[8] - m[5] * m[7]) * inv; M[1] = -(m[1] * m[8] - m[2] * m[7]) * inv; M[2] = (m[1] * m[5] - m[2] * m[4]) * inv;
                M[3] = -(m[3] * m[8] - m[5] * m[6]) * inv; M[4] = (m[0] * m[8] - m[2] * m[6]) * inv; M[5] = -(m[0] * m[5] - m[2] * m[3]) * inv;
                M[6] = (m[3] * m[7] - m[4] * m[6]) * inv; M[7] = -(m[0] * m[7] - m[1] * m[6]) * inv; M[8] = (m[0] * m[4] - m[1] * m[3]) * inv; } } break;
        default: return ORC_E_OPCODE;
        }
    }
    *out = c->f[0];
    return ORC_OK;
}

/* FP64 per-point values and MSE over the same AoSoA dataset (padding lanes included). */
ORC_API int orc_eval_points_f64(const uint8_t *ops, const uint16_t *args, int n_code,
                                const uint8_t *ctypes, const float *cvals, int n_consts,
                                const float *feature_flat, int num_features, int64_t num_batches,
                                double *out) {
    OrcProgram p = mkprog(ops, args, n_code, ctypes, cvals, n_consts);
    double x[256];
    if (num_features > 256) return ORC_E_ARG;
    for (int64_t b = 0; b < num_batches; ++b)
        for (int l = 0; l < LANES; ++l) {
            for (int f = 0; f < num_features; ++f) x[f] = feature_flat[(b * num_features + f) * LANES + l];
            int rc = eval_point_f64(&p, x, num_features, &out[b * LANES + l]);
            if (rc) return rc;
        }
    return ORC_OK;
}

ORC_API int orc_compute_mse_f64(const uint8_t *ops, const uint16_t *args, int n_code,
                                const uint8_t *ctypes, const float *cvals, int n_consts,
                                const float *feature_flat, const float *target_batches, int num_features,
                                int64_t num_batches, int64_t num_samples, double *mse_out) {
    OrcProgram p = mkprog(ops, args, n_code, ctypes, cvals, n_consts);
    double x[256], s = 0.0, comp = 0.0;
    if (num_features > 256) return ORC_E_ARG;
    for (int64_t b = 0; b < num_batches; ++b)
        for (int l = 0; l < LANES; ++l) {
            for (int f = 0; f < num_features; ++f) x[f] = feature_flat[(b * num_features + f) * LANES + l];
            double v;
            int rc = eval_point_f64(&p, x, num_features, &v);
            if (rc) return rc;
            double d = v - (double)target_batches[b * LANES + l];
            double y = d * d - comp, t = s + y;  /* Kahan in f64 */
            comp = (t - s) - y; s = t;
        }
    *mse_out = s / (double)num_samples;
    return ORC_OK;
}

/* =====================================================================
 * Dataset::new (dataset.rs:37-169) and subset (dataset.rs:360-447)
 * ===================================================================== */
ORC_API int64_t orc_dataset_num_batches(int64_t num_samples) { return (num_samples + 7) / 8; }

/* rows: [num_samples][num_features] row-major; outputs sized num_batches*num_features*8 and num_batches*8.
 * stats_out (optional): feature means[F], std devs[F], target mean, target std. */
ORC_API int orc_dataset_new(const float *rows, const float *y, int64_t num_samples, int num_features,
                            int normalize, float *feature_flat, float *target_batches, float *stats_out) {
    int F = num_features;
    float *mean = (float *)calloc((size_t)F + 1, sizeof(float));
    float *sd = (float *)malloc(((size_t)F + 1) * sizeof(float));
    for (int f = 0; f < F; ++f) sd[f] = 1.0f;
    float tmean = 0.0f, tsd = 1.0f;
    float ns = (float)num_samples;
    if (normalize) {
        for (int f = 0; f < F; ++f) {
            float sum = 0.0f;
            for (int64_t i = 0; i < num_samples; ++i) sum += rows[i * F + f];
            float m = sum / ns;
            mean[f] = m;
            float var = 0.0f;
            for (int64_t i = 0; i < num_samples; ++i) { float d = rows[i * F + f] - m; var += d * d; }
            var = var / ns;
            sd[f] = var < 1e-9f ? 1.0f : sqrtf(var);
        }
        float ts = 0.0f;
        for (int64_t i = 0; i < num_samples; ++i) ts += y[i];
        tmean = ts / ns;
        float tv = 0.0f;
        for (int64_t i = 0; i < num_samples; ++i) { float d = y[i] - tmean; tv += d * d; }
        tv = tv / ns;
        tsd = tv < 1e-9f ? 1.0f : sqrtf(tv);
    }
    int64_t nb = orc_dataset_num_batches(num_samples);
    for (int64_t b = 0; b < nb; ++b) {
        for (int f = 0; f < F; ++f)
            for (int l = 0; l < LANES; ++l) {
                int64_t i = b * LANES + l;
                feature_flat[(b * F + f) * LANES + l] = i < num_samples ? (rows[i * F + f] - mean[f]) / sd[f] : 0.0f;
            }
        for (int l = 0; l < LANES; ++l) {
            int64_t i = b * LANES + l;
            target_batches[b * LANES + l] = i < num_samples ? (y[i] - tmean) / tsd : 0.0f;
        }
    }
    if (stats_out) {
        for (int f = 0; f < F; ++f) { stats_out[f] = mean[f]; stats_out[F + f] = sd[f]; }
        stats_out[2 * F] = tmean; stats_out[2 * F + 1] = tsd;
    }
    free(mean); free(sd);
    return ORC_OK;
}

/* subset: strided sample idx = round(i * step), step = (N-1)/max(n-1, 1) in f64; padding lanes 0. */
ORC_API int orc_dataset_subset(const float *feature_flat, const float *target_batches, int num_features,
                               int64_t num_samples, int64_t target_samples,
                               float *out_flat, float *out_target, int64_t *out_samples) {
    int64_t samples = target_samples < num_samples ? target_samples : num_samples;
    int F = num_features;
    int64_t nb_out = orc_dataset_num_batches(samples);
    if (samples == num_samples) {
        memcpy(out_flat, feature_flat, (size_t)(nb_out * F * LANES) * sizeof(float));
        memcpy(out_target, target_batches, (size_t)(nb_out * LANES) * sizeof(float));
        *out_samples = samples;
        return ORC_OK;
    }
    double denom = (double)samples - 1.0;
    if (denom < 1.0) denom = 1.0;
    double step = ((double)num_samples - 1.0) / denom;
    for (int64_t b = 0; b < nb_out; ++b)
        for (int l = 0; l < LANES; ++l) {
            int64_t i = b * LANES + l;
            int64_t orig = (int64_t)round((double)i * step);
            for (int f = 0; f < F; ++f)
                out_flat[(b * F + f) * LANES + l] = orig >= num_samples ? 0.0f
                    : feature_flat[((orig / 8) * F + f) * LANES + (orig % 8)];
            out_target[b * LANES + l] = orig >= num_samples ? 0.0f : target_batches[(orig / 8) * LANES + (orig % 8)];
        }
    *out_samples = samples;
    return ORC_OK;
}

/* =====================================================================
 * Optimizers.  The program's constants are mutated in place through
 * flatten/unflatten over a 32-slot buffer (program.rs:100-192).
 * ===================================================================== */
typedef struct {
    OrcProgram p;
    float *cvals_mut;
    const float *feat, *tgt;
    int nfeat;
    int64_t nb, ns;
} OptCtx;

static void flatten(const OptCtx *o, float *buf, int cap) {
    int ptr = 0, off = 0;
    for (int i = 0; i < o->p.n_consts; ++i) {
        int s = vt_size(o->p.ctypes[i]);
        for (int k = 0; k < s; ++k) if (ptr < cap) buf[ptr++] = o->cvals_mut[off + k];
        off += s ? s : 1;
    }
}
static void unflatten(OptCtx *o, const float *buf, int cap) {
    int ptr = 0, off = 0;
    for (int i = 0; i < o->p.n_consts; ++i) {
        int s = vt_size(o->p.ctypes[i]);
        for (int k = 0; k < s; ++k) if (ptr < cap) o->cvals_mut[off + k] = buf[ptr++];
        off += s ? s : 1;
    }
}
static float mse_of(OptCtx *o) {
    float m;
    orc_compute_mse(o->p.ops, o->p.args, o->p.n_code, o->p.ctypes, o->cvals_mut, o->p.n_consts,
                    o->feat, o->tgt, o->nfeat, o->nb, o->ns, &m);
    return m;
}
static float mse_grad_of(OptCtx *o, float *g) {
    float m;
    orc_compute_mse_with_gradient(o->p.ops, o->p.args, o->p.n_code, o->p.ctypes, o->cvals_mut, o->p.n_consts,
                                  o->feat, o->tgt, o->nfeat, o->nb, o->ns, &m, g);
    return m;
}

#define MAXP 32

/* run_lbfgs (lbfgs.rs:9-165), reference-parity mode: f32, history M, fixed iterations,
 * including the second-loop start-at-slot-0 quirk after a reset (lbfgs.rs:80).
 * Returns the final fitness (current_mse) and writes the constants back into cvals_mut. */
ORC_API int orc_run_lbfgs(const uint8_t *ops, const uint16_t *args, int n_code,
                          const uint8_t *ctypes, float *cvals_mut, int n_consts,
                          const float *feature_flat, const float *target_batches, int num_features,
                          int64_t num_batches, int64_t num_samples, int max_iterations, int M,
                          float *fitness_out) {
    OptCtx o = {mkprog(ops, args, n_code, ctypes, cvals_mut, n_consts), cvals_mut,
                feature_flat, target_batches, num_features, num_batches, num_samples};
    int n = param_count(&o.p);
    if (n > MAXP) n = MAXP;
    if (n == 0) { *fitness_out = NAN; return 1; } /* returns without touching fitness */
    if (M > 8) M = 8;
    float x[MAXP] = {0}, s[8][MAXP] = {{0}}, yv[8][MAXP] = {{0}}, rho[8] = {0}, alpha[8] = {0}, q[MAXP] = {0};
    flatten(&o, x, MAXP);
    float cur_grad[MAXP];
    float cur_mse = mse_grad_of(&o, cur_grad);
    int hist = 0, head = 0;
    for (int it = 0; it < max_iterations; ++it) {
        float gn = 0.0f;
        for (int j = 0; j < n; ++j) gn += cur_grad[j] * cur_grad[j];
        if (sqrtf(gn) < 1e-5f) break;
        for (int j = 0; j < n; ++j) q[j] = cur_grad[j];
        int ci = head;
        for (int h = 0; h < hist; ++h) {
            ci = ci == 0 ? M - 1 : ci - 1;
            float d = 0.0f;
            for (int j = 0; j < n; ++j) d += s[ci][j] * q[j];
            alpha[ci] = rho[ci] * d;
            for (int j = 0; j < n; ++j) q[j] -= alpha[ci] * yv[ci][j];
        }
        float gamma = 1.0f;
        if (hist > 0) {
            int li = head == 0 ? M - 1 : head - 1;
            float sy = 0.0f, yy = 0.0f;
            for (int j = 0; j < n; ++j) { sy += s[li][j] * yv[li][j]; yy += yv[li][j] * yv[li][j]; }
            if (yy > 1e-10f) gamma = sy / yy;
        }
        for (int j = 0; j < n; ++j) q[j] *= gamma;
        ci = hist < M ? 0 : head;
        for (int h = 0; h < hist; ++h) {
            float d = 0.0f;
            for (int j = 0; j < n; ++j) d += yv[ci][j] * q[j];
            float beta = rho[ci] * d;
            for (int j = 0; j < n; ++j) q[j] += s[ci][j] * (alpha[ci] - beta);
            ci = (ci + 1) % M;
        }
        float p[MAXP] = {0}, ddg = 0.0f;
        for (int j = 0; j < n; ++j) { p[j] = -q[j]; ddg += p[j] * cur_grad[j]; }
        if (ddg > -1e-8f) {
            for (int j = 0; j < n; ++j) p[j] = -cur_grad[j];
            ddg = -gn;
            hist = 0;
        }
        float step = 1.0f, c1 = 1e-4f, nx[MAXP] = {0}, nmse;
        int ls = 0;
        for (;;) {
            for (int j = 0; j < n; ++j) nx[j] = x[j] + step * p[j];
            unflatten(&o, nx, MAXP);
            float m = mse_of(&o);
            if (m <= cur_mse + c1 * step * ddg || ls > 10) { nmse = m; break; }
            step *= 0.5f;
            ls++;
        }
        float ngrad[MAXP];
        mse_grad_of(&o, ngrad);
        float sn[MAXP], yn[MAXP], sy = 0.0f;
        for (int j = 0; j < n; ++j) { sn[j] = nx[j] - x[j]; yn[j] = ngrad[j] - cur_grad[j]; sy += sn[j] * yn[j]; }
        if (sy > 1e-10f) {
            for (int j = 0; j < n; ++j) { s[head][j] = sn[j]; yv[head][j] = yn[j]; }
            rho[head] = 1.0f / sy;
            head = (head + 1) % M;
            if (hist < M) hist++;
        }
        for (int j = 0; j < n; ++j) x[j] = nx[j];
        cur_mse = nmse;
        for (int j = 0; j < MAXP; ++j) cur_grad[j] = ngrad[j];
    }
    unflatten(&o, x, MAXP);
    *fitness_out = cur_mse;
    return 0;
}

/* L-BFGS "config mode": history m, fixed iterations, the two-loop dot products,
 * curvature pairs (s, y, rho), gamma, directional derivative and the Armijo
 * sufficient-decrease test in FP64; parameters stay f32; the second loop starts
 * at the oldest live pair.  Otherwise identical control flow to run_lbfgs. */
ORC_API int orc_run_lbfgs_f64(const uint8_t *ops, const uint16_t *args, int n_code,
                              const uint8_t *ctypes, float *cvals_mut, int n_consts,
                              const float *feature_flat, const float *target_batches, int num_features,
                              int64_t num_batches, int64_t num_samples, int max_iterations, int M,
                              float *fitness_out) {
    OptCtx o = {mkprog(ops, args, n_code, ctypes, cvals_mut, n_consts), cvals_mut,
                feature_flat, target_batches, num_features, num_batches, num_samples};
    int n = param_count(&o.p);
    if (n > MAXP) n = MAXP;
    if (n == 0) { *fitness_out = NAN; return 1; }
    if (M > 8) M = 8;
    float x[MAXP] = {0};
    double s[8][MAXP] = {{0}}, yv[8][MAXP] = {{0}}, rho[8] = {0}, alpha[8] = {0}, q[MAXP] = {0};
    flatten(&o, x, MAXP);
    float cur_grad[MAXP];
    float cur_mse = mse_grad_of(&o, cur_grad);
    int hist = 0, head = 0;
    for (int it = 0; it < max_iterations; ++it) {
        double gn = 0.0;
        for (int j = 0; j < n; ++j) gn += (double)cur_grad[j] * (double)cur_grad[j];
        if (sqrt(gn) < 1e-5) break;
        for (int j = 0; j < n; ++j) q[j] = cur_grad[j];
        int ci = head;
        for (int h = 0; h < hist; ++h) {
            ci = ci == 0 ? M - 1 : ci - 1;
            double d = 0.0;
            for (int j = 0; j < n; ++j) d += s[ci][j] * q[j];
            alpha[ci] = rho[ci] * d;
            for (int j = 0; j < n; ++j) q[j] -= alpha[ci] * yv[ci][j];
        }
        double gamma = 1.0;
        if (hist > 0) {
            int li = head == 0 ? M - 1 : head - 1;
            double sy = 0.0, yy = 0.0;
            for (int j = 0; j < n; ++j) { sy += s[li][j] * yv[li][j]; yy += yv[li][j] * yv[li][j]; }
            if (yy > 1e-10) gamma = sy / yy;
        }
        for (int j = 0; j < n; ++j) q[j] *= gamma;
        ci = (head - hist + M) % M;  /* oldest live pair */
        for (int h = 0; h < hist; ++h) {
            double d = 0.0;
            for (int j = 0; j < n; ++j) d += yv[ci][j] * q[j];
            double beta = rho[ci] * d;
            for (int j = 0; j < n; ++j) q[j] += s[ci][j] * (alpha[ci] - beta);
            ci = (ci + 1) % M;
        }
        float p[MAXP] = {0};
        double ddg = 0.0;
        for (int j = 0; j < n; ++j) { p[j] = (float)(-q[j]); ddg += (double)p[j] * (double)cur_grad[j]; }
        if (ddg > -1e-8) {
            for (int j = 0; j < n; ++j) p[j] = -cur_grad[j];
            ddg = -gn;
            hist = 0;
        }
        float step = 1.0f, nx[MAXP] = {0}, nmse;
        int ls = 0;
        for (;;) {
            for (int j = 0; j < n; ++j) nx[j] = x[j] + step * p[j];
            unflatten(&o, nx, MAXP);
            float m = mse_of(&o);
            if ((double)m <= (double)cur_mse + 1e-4 * (double)step * ddg || ls > 10) { nmse = m; break; }
            step *= 0.5f;
            ls++;
        }
        float ngrad[MAXP];
        mse_grad_of(&o, ngrad);
        double sn[MAXP], yn[MAXP], sy = 0.0;
        for (int j = 0; j < n; ++j) { sn[j] = (double)nx[j] - (double)x[j]; yn[j] = (double)ngrad[j] - (double)cur_grad[j]; sy += sn[j] * yn[j]; }
        if (sy > 1e-10) {
            for (int j = 0; j < n; ++j) { s[head][j] = sn[j]; yv[head][j] = yn[j]; }
            rho[head] = 1.0 / sy;
            head = (head + 1) % M;
            if (hist < M) hist++;
        }
        for (int j = 0; j < n; ++j) x[j] = nx[j];
        cur_mse = nmse;
        for (int j = 0; j < MAXP; ++j) cur_grad[j] = ngrad[j];
    }
    unflatten(&o, x, MAXP);
    *fitness_out = cur_mse;
    return 0;
}

/* run_cma_es (cmaes.rs:19-192) with an injected stream of standard normals
 * (normals[it * lambda * 32 + i * 32 + j]) replacing the unseeded thread RNG.
 * lambda/mu generalise LAMBDA=16 / MU=8.  Includes the doubled-dimension
 * quirk (cmaes.rs:31-75).  The two ln/exp constants that the reference
 * computes with Rust std (weights, exp(-0.01)) are passed in precomputed. */
ORC_API int orc_run_cmaes(const uint8_t *ops, const uint16_t *args, int n_code,
                          const uint8_t *ctypes, float *cvals_mut, int n_consts,
                          const float *feature_flat, const float *target_batches, int num_features,
                          int64_t num_batches, int64_t num_samples, int max_iterations,
                          int lambda, const float *weights /* [mu], normalised */, int mu, float mu_eff,
                          float sigma_decay, const float *normals, float *fitness_out, int *iters_run) {
    OptCtx o = {mkprog(ops, args, n_code, ctypes, cvals_mut, n_consts), cvals_mut,
                feature_flat, target_batches, num_features, num_batches, num_samples};
    int n = param_count(&o.p);
    if (n > MAXP) n = MAXP;
    if (n == 0) { *fitness_out = NAN; *iters_run = 0; return 1; }
    if (lambda > 64 || mu > lambda) return ORC_E_ARG;
    float mean[MAXP] = {0};
    flatten(&o, mean, MAXP);
    /* quirk: constants appended a second time (cmaes.rs:33-75) */
    {
        int off = 0;
        for (int i = 0; i < o.p.n_consts; ++i) {
            int sz = vt_size(o.p.ctypes[i]);
            for (int k = 0; k < sz; ++k) if (n < MAXP) mean[n++] = cvals_mut[off + k];
            off += sz ? sz : 1;
        }
    }
    float sigma = 0.5f, cd[MAXP], pc[MAXP] = {0};
    for (int j = 0; j < MAXP; ++j) cd[j] = 1.0f;
    float nf = (float)n;
    float c_c = (4.0f + mu_eff / nf) / (nf + 4.0f + 2.0f * mu_eff / nf);
    float t1 = nf + 1.414f, t2 = nf + 2.0f;
    float c_cov = (1.0f / mu_eff) * (2.0f / (t1 * t1))
                + (1.0f - 1.0f / mu_eff) * ((2.0f * mu_eff - 1.0f) / (t2 * t2 + mu_eff));
    float init_mse = mse_of(&o);
    float l1 = 0.0f;
    for (int j = 0; j < n; ++j) l1 += fabsf(mean[j]);
    float best_pt[MAXP];
    memcpy(best_pt, mean, sizeof(best_pt));
    float best_mse = init_mse, best_fit = init_mse + 0.01f * l1;
    static __thread float pop[64][MAXP], stepv[64][MAXP];
    float fit[64], msev[64];
    int idx[64];
    int it;
    for (it = 0; it < max_iterations; ++it) {
        for (int i = 0; i < lambda; ++i) {
            float l1n = 0.0f;
            for (int j = 0; j < n; ++j) {
                float st = normals[((int64_t)it * lambda + i) * MAXP + j] * sqrtf(cd[j]);
                stepv[i][j] = st;
                float v = mean[j] + sigma * st;
                pop[i][j] = v;
                l1n += fabsf(v);
            }
            for (int j = n; j < MAXP; ++j) pop[i][j] = 0.0f;
            unflatten(&o, pop[i], MAXP);
            float m = mse_of(&o);
            fit[i] = m + 0.01f * l1n;
            msev[i] = m;
            idx[i] = i;
        }
        /* sort by fitness ascending; ties broken by index (deterministic; the
         * reference's sort_unstable_by leaves tie order unspecified) */
        for (int i = 1; i < lambda; ++i) {
            int k = idx[i], j = i - 1;
            while (j >= 0 && (fit[idx[j]] > fit[k] || (fit[idx[j]] == fit[k] && idx[j] > k))) { idx[j + 1] = idx[j]; --j; }
            idx[j + 1] = k;
        }
        int bi = idx[0];
        if (fit[bi] < best_fit) { best_fit = fit[bi]; best_mse = msev[bi]; memcpy(best_pt, pop[bi], sizeof(best_pt)); }
        if (best_mse < 1e-8f) { it++; break; }
        float sm[MAXP] = {0}, old[MAXP];
        memcpy(old, mean, sizeof(old));
        for (int i = 0; i < mu; ++i) {
            int k = idx[i];
            for (int j = 0; j < n; ++j) { sm[j] += weights[i] * stepv[k][j]; mean[j] += weights[i] * (pop[k][j] - old[j]); }
        }
        float sq = sqrtf(c_c * (2.0f - c_c) * mu_eff);
        for (int j = 0; j < n; ++j) {
            pc[j] = (1.0f - c_c) * pc[j] + sq * sm[j];
            float cu = 0.0f;
            for (int i = 0; i < mu; ++i) { int k = idx[i]; cu += weights[i] * (stepv[k][j] * stepv[k][j]); }
            cd[j] = (1.0f - c_cov) * cd[j] + (c_cov / mu_eff) * (pc[j] * pc[j]) + c_cov * (1.0f - 1.0f / mu_eff) * cu;
            cd[j] = cd[j] < 1e-6f ? 1e-6f : (cd[j] > 1e6f ? 1e6f : cd[j]);
        }
        sigma *= sigma_decay;
    }
    unflatten(&o, best_pt, MAXP);
    *fitness_out = best_mse;
    *iters_run = it;
    return 0;
}

/* CMA-ES strategy constants as cmaes.rs:85-98 with libm logf / expf (Rust f32::ln / exp). */
ORC_API void orc_cmaes_constants(int mu, float *weights, float *mu_eff, float *sigma_decay) {
    float sum = 0.0f, lm = logf((float)mu + 0.5f);
    for (int i = 0; i < mu; ++i) {
        float v = lm - logf((float)(i + 1));
        weights[i] = v > 0.0f ? v : 0.0f;
        sum += weights[i];
    }
    float me = 0.0f;
    for (int i = 0; i < mu; ++i) { weights[i] = weights[i] / sum; me += weights[i] * weights[i]; }
    *mu_eff = 1.0f / me;
    *sigma_decay = expf(-0.01f);
}

/* run_nelder_mead (nelder_mead.rs:9-165), deterministic. Only Float constants are optimised. */
ORC_API int orc_run_nelder_mead(const uint8_t *ops, const uint16_t *args, int n_code,
                                const uint8_t *ctypes, float *cvals_mut, int n_consts,
                                const float *feature_flat, const float *target_batches, int num_features,
                                int64_t num_batches, int64_t num_samples, int max_iterations,
                                float *fitness_out) {
    OptCtx o = {mkprog(ops, args, n_code, ctypes, cvals_mut, n_consts), cvals_mut,
                feature_flat, target_batches, num_features, num_batches, num_samples};
    int opt_off[32], n = 0;
    {
        int off = 0;
        for (int i = 0; i < n_consts; ++i) {
            int sz = vt_size(ctypes[i]);
            if (ctypes[i] == VT_FLOAT && n < 32) opt_off[n++] = off;
            off += sz ? sz : 1;
        }
    }
    if (n == 0) { *fitness_out = NAN; return 1; }
    typedef struct { float fit, mse, pt[32]; } V;
    V sx[33];
    memset(sx, 0, sizeof(sx));
#define NM_EVAL(vals, F, MS) do { float l1_ = 0.0f; for (int j_ = 0; j_ < n; ++j_) { cvals_mut[opt_off[j_]] = (vals)[j_]; } \
        float m_ = mse_of(&o); for (int j_ = 0; j_ < n; ++j_) l1_ += fabsf((vals)[j_]); (F) = m_ + 0.01f * l1_; (MS) = m_; } while (0)
    float start[32] = {0};
    for (int j = 0; j < n; ++j) start[j] = cvals_mut[opt_off[j]];
    NM_EVAL(start, sx[0].fit, sx[0].mse);
    memcpy(sx[0].pt, start, sizeof(start));
    for (int i = 0; i < n; ++i) {
        float np[32];
        memcpy(np, start, sizeof(np));
        float v = np[i];
        np[i] += fabsf(v) < 1e-4f ? 0.01f : v * 0.10f;
        NM_EVAL(np, sx[i + 1].fit, sx[i + 1].mse);
        memcpy(sx[i + 1].pt, np, sizeof(np));
    }
    float cen[32], refl[32], expd[32], con[32];
    for (int it = 0; it < max_iterations; ++it) {
        /* stable insertion sort by fitness (sort_unstable_by tie order unspecified) */
        for (int i = 1; i <= n; ++i) { V k = sx[i]; int j = i - 1; while (j >= 0 && sx[j].fit > k.fit) { sx[j + 1] = sx[j]; --j; } sx[j + 1] = k; }
        float bf = sx[0].fit, bm = sx[0].mse, wf = sx[n].fit;
        if (fabsf(wf - bf) < 1e-7f || bm < 1e-8f) break;
        for (int j = 0; j < n; ++j) cen[j] = 0.0f;
        for (int i = 0; i < n; ++i) for (int j = 0; j < n; ++j) cen[j] += sx[i].pt[j];
        for (int j = 0; j < n; ++j) cen[j] /= (float)n;
        float wp[32];
        memcpy(wp, sx[n].pt, sizeof(wp));
        float swf = sx[n - 1].fit;
        for (int j = 0; j < n; ++j) refl[j] = cen[j] + 1.0f * (cen[j] - wp[j]);
        float rf, rm;
        NM_EVAL(refl, rf, rm);
        if (rf >= bf && rf < swf) { sx[n].fit = rf; sx[n].mse = rm; memcpy(sx[n].pt, refl, sizeof(refl)); continue; }
        if (rf < bf) {
            for (int j = 0; j < n; ++j) expd[j] = cen[j] + 2.0f * (refl[j] - cen[j]);
            float ef, em;
            NM_EVAL(expd, ef, em);
            if (ef < rf) { sx[n].fit = ef; sx[n].mse = em; memcpy(sx[n].pt, expd, sizeof(expd)); }
            else { sx[n].fit = rf; sx[n].mse = rm; memcpy(sx[n].pt, refl, sizeof(refl)); }
            continue;
        }
        float lim;
        if (rf < wf) { for (int j = 0; j < n; ++j) con[j] = cen[j] + 0.5f * (refl[j] - cen[j]); lim = rf; }
        else { for (int j = 0; j < n; ++j) con[j] = cen[j] + 0.5f * (wp[j] - cen[j]); lim = wf; }
        float cf, cm;
        NM_EVAL(con, cf, cm);
        if (cf < lim) { sx[n].fit = cf; sx[n].mse = cm; memcpy(sx[n].pt, con, sizeof(con)); continue; }
        float bp[32];
        memcpy(bp, sx[0].pt, sizeof(bp));
        for (int i = 1; i <= n; ++i) {
            for (int j = 0; j < n; ++j) sx[i].pt[j] = bp[j] + 0.5f * (sx[i].pt[j] - bp[j]);
            NM_EVAL(sx[i].pt, sx[i].fit, sx[i].mse);
        }
    }
    for (int i = 1; i <= n; ++i) { V k = sx[i]; int j = i - 1; while (j >= 0 && sx[j].fit > k.fit) { sx[j + 1] = sx[j]; --j; } sx[j + 1] = k; }
    for (int j = 0; j < n; ++j) cvals_mut[opt_off[j]] = sx[0].pt[j];
    *fitness_out = sx[0].mse;
#undef NM_EVAL
    return 0;
}

/* =====================================================================
 * Selection (selection.rs:5-134).  Individuals given as (fitness, age).
 * ===================================================================== */
static int dominates(float fa, uint64_t aa, float fb, uint64_t ab) {
    return (fa <= fb && aa <= ab) && (fa < fb || aa < ab);
}

ORC_API int orc_assign_rank_and_crowding(int n, const float *fitness, const uint64_t *age,
                                         uint32_t *rank, float *crowding) {
    if (n <= 0) return 0;
    int *dc = (int *)calloc((size_t)n, sizeof(int));
    int *dom = (int *)malloc((size_t)n * (size_t)n * sizeof(int));
    int *ndom = (int *)calloc((size_t)n, sizeof(int));
    int *fronts = (int *)malloc((size_t)n * sizeof(int)); /* concatenated fronts */
    int *fstart = (int *)malloc(((size_t)n + 2) * sizeof(int));
    int nf = 0, cnt = 0;
    fstart[0] = 0;
    for (int i = 0; i < n; ++i) {
        for (int j = 0; j < n; ++j) {
            if (i == j) continue;
            if (dominates(fitness[i], age[i], fitness[j], age[j])) dom[i * n + ndom[i]++] = j;
            else if (dominates(fitness[j], age[j], fitness[i], age[i])) dc[i]++;
        }
        if (dc[i] == 0) { rank[i] = 0; fronts[cnt++] = i; }
    }
    fstart[1] = cnt; nf = 1;
    int fi = 0;
    while (fi < nf) {
        int before = cnt;
        for (int a = fstart[fi]; a < fstart[fi + 1]; ++a) {
            int p = fronts[a];
            for (int b = 0; b < ndom[p]; ++b) {
                int q = dom[p * n + b];
                if (--dc[q] == 0) { rank[q] = (uint32_t)(fi + 1); fronts[cnt++] = q; }
            }
        }
        if (cnt == before) break;
        fstart[++nf] = cnt;
        fi++;
    }
    for (int f = 0; f < nf; ++f) {
        int s0 = fstart[f], len = fstart[f + 1] - s0;
        if (len == 0) continue;
        int *fr = fronts + s0;
        for (int k = 0; k < len; ++k) crowding[fr[k]] = 0.0f;
        if (len < 3) { for (int k = 0; k < len; ++k) crowding[fr[k]] = FLT_MAX; continue; }
        int *srt = (int *)malloc((size_t)len * sizeof(int));
        memcpy(srt, fr, (size_t)len * sizeof(int));
        /* stable sort by fitness (Vec::sort_by is stable) */
        for (int i = 1; i < len; ++i) { int k = srt[i], j = i - 1; while (j >= 0 && fitness[srt[j]] > fitness[k]) { srt[j + 1] = srt[j]; --j; } srt[j + 1] = k; }
        crowding[srt[0]] = FLT_MAX;
        crowding[srt[len - 1]] = FLT_MAX;
        float rng = fitness[srt[len - 1]] - fitness[srt[0]];
        if (rng > 1e-9f)
            for (int k = 1; k < len - 1; ++k) {
                float d = (fitness[srt[k + 1]] - fitness[srt[k - 1]]) / rng;
                if (crowding[srt[k]] != FLT_MAX) crowding[srt[k]] += d;
            }
        /* stable sort by age (sort_by_key is stable) */
        for (int i = 1; i < len; ++i) { int k = srt[i], j = i - 1; while (j >= 0 && age[srt[j]] > age[k]) { srt[j + 1] = srt[j]; --j; } srt[j + 1] = k; }
        crowding[srt[0]] = FLT_MAX;
        crowding[srt[len - 1]] = FLT_MAX;
        float amin = (float)age[srt[0]], amax = (float)age[srt[len - 1]];
        float ar = amax - amin;
        if (ar > 1e-9f)
            for (int k = 1; k < len - 1; ++k) {
                float d = ((float)age[srt[k + 1]] - (float)age[srt[k - 1]]) / ar;
                if (crowding[srt[k]] != FLT_MAX) crowding[srt[k]] += d;
            }
        free(srt);
    }
    free(dc); free(dom); free(ndom); free(fronts); free(fstart);
    return 0;
}

/* tournament_selection_pareto (selection.rs:13-26) with the random indices supplied. */
ORC_API int orc_tournament(int k, const int *cand, const uint32_t *rank, const float *crowding) {
    int best = cand[0];
    for (int i = 1; i < k; ++i) {
        int c = cand[i];
        int less = rank[c] < rank[best] || (rank[c] == rank[best] && crowding[best] < crowding[c]);
        if (less) best = c;
    }
    return best;
}

/* fitness = mse + complexity * parsimony; non-finite -> f32::MAX (island.rs:241-255) */
ORC_API float orc_fitness(float mse, int complexity, float parsimony) {
    if (!isfinite(mse)) return FLT_MAX;
    return mse + (float)complexity * parsimony;
}
