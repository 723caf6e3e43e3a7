// devmath.h -- single-precision sin / cos / exp / ln for the device interpreter.
//
// The reference gets these from the third-party crate `wide 1.1.1` (call sites
// src/domains/basic/eval.rs:36,41,46,62 and src/engine/eval/autodiff.rs:34-35,42-43,49,76),
// whose source is not available.  They are implemented here with the Cephes
// single-precision methods (S. L. Moshier, Cephes Math Library, sinf.c / cosf.c /
// expf.c / logf.c): Cody-Waite reduction by pi/4 (three-part constant) or ln 2
// (two-part constant) followed by the published minimax polynomials, evaluated
// with plain IEEE mul/add (the kernels are built with -ffp-contract=off, so no
// FMA is introduced).  Accuracy is pinned by the test-suite against FP64 libm.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace rsr {

__host__ __device__ __forceinline__ float dm_sin_poly(float x, float z) {
    float y = ((-1.9515295891E-4f * z + 8.3321608736E-3f) * z - 1.6666654611E-1f) * z * x;
    return y + x;
}

__host__ __device__ __forceinline__ float dm_cos_poly(float z) {
    float y = 2.443315711809948E-005f;
    y = y * z + -1.388731625493765E-003f;
    y = y * z + 4.166664568298827E-002f;
    y = y * (z * z);
    y = y - 0.5f * z;
    return y + 1.0f;
}

// |x| = j*pi/4 + r with j rounded up to even, r in about [-pi/4, pi/4].
__host__ __device__ __forceinline__ float dm_octant(float x, unsigned& j) {
    unsigned q = (unsigned)(1.27323954473516f * x);
    float y = (float)q;
    if (q & 1u) { q += 1u; y += 1.0f; }
    j = q & 7u;
    return ((x - y * 0.78515625f) - y * 2.4187564849853515625e-4f) - y * 3.77489497744594108e-8f;
}

__host__ __device__ __forceinline__ float dm_sin(float xx) {
    float x = __builtin_fabsf(xx);
    if (!(x <= 16777215.0f)) {            // NaN, inf, or total loss of significance
        return (x != x || x == __builtin_inff()) ? __builtin_nanf("") : 0.0f * xx;
    }
    unsigned j;
    float r = dm_octant(x, j);
    bool neg = xx < 0.0f;
    if (j > 3u) { neg = !neg; j -= 4u; }
    float z = r * r;
    float y = (j == 1u || j == 2u) ? dm_cos_poly(z) : dm_sin_poly(r, z);
    return neg ? -y : y;
}

__host__ __device__ __forceinline__ float dm_cos(float xx) {
    float x = __builtin_fabsf(xx);
    if (!(x <= 16777215.0f)) {
        return (x != x || x == __builtin_inff()) ? __builtin_nanf("") : 1.0f;
    }
    unsigned j;
    float r = dm_octant(x, j);
    bool neg = false;
    if (j > 3u) { j -= 4u; neg = !neg; }
    if (j > 1u) neg = !neg;
    float z = r * r;
    float y = (j == 1u || j == 2u) ? dm_sin_poly(r, z) : dm_cos_poly(z);
    return neg ? -y : y;
}

__host__ __device__ __forceinline__ float dm_pow2i(int n) {  // n in [-149, 127]
    uint32_t b = n >= -126 ? (uint32_t)(n + 127) << 23 : 1u << (n + 149);
    return __builtin_bit_cast(float, b);
}

__host__ __device__ __forceinline__ float dm_exp(float xx) {
    if (xx != xx) return xx;
    if (xx > 88.72283905206835f) return __builtin_inff();
    if (xx < -103.278929903431851103f) return 0.0f;
    float x = xx;
    float z = __builtin_floorf(1.44269504088896341f * x + 0.5f);
    x = x - z * 0.693359375f;
    x = x - z * -2.12194440e-4f;
    int n = (int)z;
    z = x * x;
    z = (((((1.9875691500E-4f * x + 1.3981999507E-3f) * x + 8.3334519073E-3f) * x
           + 4.1665795894E-2f) * x + 1.6666665459E-1f) * x + 5.0000001201E-1f) * z + x + 1.0f;
    if (n > 127) { z = z * dm_pow2i(127); n -= 127; }
    if (n < -126) { z = z * dm_pow2i(-126); n += 126; }
    return z * dm_pow2i(n);
}

__host__ __device__ __forceinline__ float dm_ln(float xx) {
    if (xx != xx) return xx;
    if (xx < 0.0f) return __builtin_nanf("");
    if (xx == 0.0f) return -__builtin_inff();
    if (xx == __builtin_inff()) return xx;
    // frexp: xx = m * 2^e, m in [0.5, 1)
    uint32_t b = __builtin_bit_cast(uint32_t, xx);
    int e;
    if ((b >> 23) == 0u) {  // subnormal: scale by 2^23 first
        float s = xx * 8388608.0f;
        b = __builtin_bit_cast(uint32_t, s);
        e = (int)(b >> 23) - 126 - 23;
    } else {
        e = (int)(b >> 23) - 126;
    }
    float x = __builtin_bit_cast(float, (b & 0x807fffffu) | 0x3f000000u);
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

}  // namespace rsr
