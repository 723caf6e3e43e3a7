// devmath.h -- single-precision sin / cos / exp / ln for the device interpreter.
//
// The reference gets these from the third-party crate `wide 1.1.1` (call sites
// src/domains/basic/eval.rs:36,41,46,62 and src/engine/eval/autodiff.rs:34-35,42-43,49,76),
// whose source is not available, so their bits are ours to define; the accuracy bound
// (at most 2 ulp from the exact value) is pinned by the test-suite against FP64 libm.
//
// sin / cos / exp: Cody-Waite argument reduction written with fused multiply-adds (the
// first step x - k*C1 is exact, as k*C1 is within a factor of two of x), then the Cephes
// single-precision minimax polynomials (S. L. Moshier, Cephes Math Library, sinf.c,
// cosf.c, expf.c) evaluated in Horner form with fused multiply-adds.  sin and cos reduce
// by pi/2 with a three-part constant (float(pi/2) and the next two float remainders); exp
// reduces by ln 2 with a two-part constant and scales by 2^k with ldexp, which rounds once
// also where the result is subnormal.  ln is Cephes logf with plain IEEE mul/add.
//
// The fused multiply-adds are explicit calls: the kernels are built with -ffp-contract=off,
// so no other expression is contracted.  The same sequences, operation for operation,
// are the oracle's (C fmaf / rintf / ldexpf, all correctly rounded), so device and oracle
// values agree bit for bit.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace rsr {

// Reduction shared by sin and cos: |x| = k*pi/2 + r with k = rint(|x| * 2/pi), r in
// about [-pi/4, pi/4]; s ~ sin(r), c ~ cos(r); n = k as an integer (its low two bits pick
// the quadrant).  The results for arguments beyond 2^24 - 1 (and inf, NaN) are selected
// away by the callers.
__host__ __device__ __forceinline__ void dm_sincos_reduced(float ax, float& s, float& c, uint32_t& n) {
    const float k = __builtin_rintf(ax * 0.636619746685028076171875f);
    float r = __builtin_fmaf(-k, 1.57079637050628662109375f, ax);
    r = __builtin_fmaf(-k, -4.37113882867379300296e-08f, r);
    r = __builtin_fmaf(-k, -1.71512451000588190000e-15f, r);
    const float z = r * r;
    float p = __builtin_fmaf(z, -1.9515295891E-4f, 8.3321608736E-3f);
    p = __builtin_fmaf(z, p, -1.6666654611E-1f);
    s = __builtin_fmaf(r * z, p, r);
    float q = __builtin_fmaf(z, 2.443315711809948E-5f, -1.388731625493765E-3f);
    q = __builtin_fmaf(z, q, 4.166664568298827E-2f);
    c = __builtin_fmaf(z * z, q, __builtin_fmaf(-0.5f, z, 1.0f));
    n = (uint32_t)(int)__builtin_fminf(k, 16777216.0f);  // in range also for the arguments selected away
}

__host__ __device__ __forceinline__ float dm_flip(float y, uint32_t sign_bit31) {
    return __builtin_bit_cast(float, __builtin_bit_cast(uint32_t, y) ^ (sign_bit31 & 0x80000000u));
}

// The functions below are branch-free: every input runs the main path and the special
// results (|x| > 2^24 - 1, inf, NaN) are selected at the end, so a wavefront never splits
// on the data.  Beyond 2^24 - 1, sin gives +-0 and cos 1 (Cephes' total loss of
// significance); inf and NaN give NaN.
__host__ __device__ __forceinline__ float dm_sin(float xx) {
    const float ax = __builtin_fabsf(xx);
    float s, c;
    uint32_t n;
    dm_sincos_reduced(ax, s, c, n);
    const float y = dm_flip((n & 1u) ? c : s, (n << 30) ^ __builtin_bit_cast(uint32_t, xx));
    return ax <= 16777215.0f ? y : 0.0f * xx;
}

__host__ __device__ __forceinline__ float dm_cos(float xx) {
    const float ax = __builtin_fabsf(xx);
    float s, c;
    uint32_t n;
    dm_sincos_reduced(ax, s, c, n);
    const float y = dm_flip((n & 1u) ? s : c, (n + 1u) << 30);
    return ax <= 16777215.0f ? y : (ax == __builtin_inff() || ax != ax) ? __builtin_nanf("") : 1.0f;
}

// sin and cos of one argument from one reduction: the same bits as dm_sin and dm_cos.
__host__ __device__ __forceinline__ void dm_sincos(float xx, float& so, float& co) {
    const float ax = __builtin_fabsf(xx);
    float s, c;
    uint32_t n;
    dm_sincos_reduced(ax, s, c, n);
    const bool odd = (n & 1u) != 0u;
    const float ys = dm_flip(odd ? c : s, (n << 30) ^ __builtin_bit_cast(uint32_t, xx));
    const float yc = dm_flip(odd ? s : c, (n + 1u) << 30);
    const bool ok = ax <= 16777215.0f;
    so = ok ? ys : 0.0f * xx;
    co = ok ? yc : (ax == __builtin_inff() || ax != ax) ? __builtin_nanf("") : 1.0f;
}

// e^x: x is clamped to [-104, 89] (beyond it the result is 0 or inf, which ldexp then
// produces); NaN is returned as it came.
__host__ __device__ __forceinline__ float dm_exp(float x) {
    const float xc = __builtin_fminf(__builtin_fmaxf(x, -104.0f), 89.0f);
    const float k = __builtin_rintf(xc * 1.44269502162933349609375f);
    float r = __builtin_fmaf(-k, 0.693147182464599609375f, xc);
    r = __builtin_fmaf(-k, -1.904654212125933566e-09f, r);
    float p = __builtin_fmaf(1.9875691500E-4f, r, 1.3981999507E-3f);
    p = __builtin_fmaf(p, r, 8.3334519073E-3f);
    p = __builtin_fmaf(p, r, 4.1665795894E-2f);
    p = __builtin_fmaf(p, r, 1.6666665459E-1f);
    p = __builtin_fmaf(p, r, 5.0000001201E-1f);
    const float y = __builtin_fmaf(r * r, p, r) + 1.0f;
    const float res = __builtin_ldexpf(y, (int)k);
    return x != x ? x : res;
}

__host__ __device__ __forceinline__ float dm_ln(float xx) {
    const bool nan = xx != xx, negv = xx < 0.0f, zero = xx == 0.0f, inf = xx == __builtin_inff();
    const float xs = (nan || negv || zero || inf) ? 1.0f : xx;
    // frexp: xs = m * 2^e, m in [0.5, 1); subnormals are scaled by 2^23 first
    const uint32_t b0 = __builtin_bit_cast(uint32_t, xs);
    const bool sub = (b0 >> 23) == 0u;
    const uint32_t b1 = __builtin_bit_cast(uint32_t, xs * 8388608.0f);
    const uint32_t b = sub ? b1 : b0;
    int e = (int)(b >> 23) - 126 - (sub ? 23 : 0);
    float x = __builtin_bit_cast(float, (b & 0x807fffffu) | 0x3f000000u);
    const bool lt = x < 0.707106781186547524f;
    e -= lt ? 1 : 0;
    x = lt ? x + x - 1.0f : x - 1.0f;
    float z = x * x;
    float y = ((((((((7.0376836292E-2f * x - 1.1514610310E-1f) * x + 1.1676998740E-1f) * x
                    - 1.2420140846E-1f) * x + 1.4249322787E-1f) * x - 1.6668057665E-1f) * x
                 + 2.0000714765E-1f) * x - 2.4999993993E-1f) * x + 3.3333331174E-1f) * x * z;
    const float fe = (float)e;
    y = e ? y + -2.12194440e-4f * fe : y;
    y = y + -0.5f * z;
    z = x + y;
    z = e ? z + 0.693359375f * fe : z;
    return nan ? xx : negv ? __builtin_nanf("") : zero ? -__builtin_inff() : inf ? xx : z;
}

}  // namespace rsr
