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

// The functions below are branch-free: the main path runs on every input (special inputs
// are first replaced by a harmless value) and the special results are selected at the
// end, so a wavefront never splits on the data.  The arithmetic on every ordinary input
// is exactly the Cephes sequence.

// |x| = j*pi/4 + r with j rounded up to even, r in about [-pi/4, pi/4].
__host__ __device__ __forceinline__ float dm_octant(float x, unsigned& j) {
    unsigned q = (unsigned)(1.27323954473516f * x);
    float y = (float)q;
    const bool odd = (q & 1u) != 0u;
    q += odd ? 1u : 0u;
    y = odd ? y + 1.0f : y;
    j = q & 7u;
    return ((x - y * 0.78515625f) - y * 2.4187564849853515625e-4f) - y * 3.77489497744594108e-8f;
}

__host__ __device__ __forceinline__ float dm_sin(float xx) {
    const float ax = __builtin_fabsf(xx);
    const bool big = !(ax <= 16777215.0f);  // NaN, inf, or total loss of significance
    const float x = big ? 0.0f : ax;
    unsigned j;
    const float r = dm_octant(x, j);
    const bool neg = (xx < 0.0f) != (j > 3u);
    j &= 3u;
    const float z = r * r;
    const float yc = dm_cos_poly(z), ys = dm_sin_poly(r, z);
    const float y = (j == 1u || j == 2u) ? yc : ys;
    const float res = neg ? -y : y;
    const float special = (ax != ax || ax == __builtin_inff()) ? __builtin_nanf("") : 0.0f * xx;
    return big ? special : res;
}

__host__ __device__ __forceinline__ float dm_cos(float xx) {
    const float ax = __builtin_fabsf(xx);
    const bool big = !(ax <= 16777215.0f);
    const float x = big ? 0.0f : ax;
    unsigned j;
    const float r = dm_octant(x, j);
    bool neg = j > 3u;
    j &= 3u;
    neg = neg != (j > 1u);
    const float z = r * r;
    const float ys = dm_sin_poly(r, z), yc = dm_cos_poly(z);
    const float y = (j == 1u || j == 2u) ? ys : yc;
    const float res = neg ? -y : y;
    const float special = (ax != ax || ax == __builtin_inff()) ? __builtin_nanf("") : 1.0f;
    return big ? special : res;
}

__host__ __device__ __forceinline__ float dm_pow2i(int n) {  // n in [-149, 127]
    const uint32_t b = n >= -126 ? (uint32_t)(n + 127) << 23 : 1u << ((n + 149) & 31);
    return __builtin_bit_cast(float, b);
}

__host__ __device__ __forceinline__ float dm_exp(float xx) {
    const bool nan = xx != xx;
    const bool hi = xx > 88.72283905206835f;
    const bool lo = xx < -103.278929903431851103f;
    float x = (nan || hi || lo) ? 0.0f : xx;
    float z = __builtin_floorf(1.44269504088896341f * x + 0.5f);
    x = x - z * 0.693359375f;
    x = x - z * -2.12194440e-4f;
    int n = (int)z;
    z = x * x;
    z = (((((1.9875691500E-4f * x + 1.3981999507E-3f) * x + 8.3334519073E-3f) * x
           + 4.1665795894E-2f) * x + 1.6666665459E-1f) * x + 5.0000001201E-1f) * z + x + 1.0f;
    const bool up = n > 127;
    z = up ? z * dm_pow2i(127) : z;
    n = up ? n - 127 : n;
    const bool dn = n < -126;
    z = dn ? z * dm_pow2i(-126) : z;
    n = dn ? n + 126 : n;
    const float res = z * dm_pow2i(n);
    return nan ? xx : hi ? __builtin_inff() : lo ? 0.0f : res;
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
