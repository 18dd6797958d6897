// bm_math.hpp -- deterministic double-precision log / sin / cos for the
// Box-Muller transform of rand_gen (xmd/include/xmd/random/rand_gen.h:42-58).
//
// The reference computes its Langevin noise with glibc's float logf, sinf
// and cosf.  The engine reproduces those results bit for bit on the GPU in
// two parts:
//   1. the functions below evaluate log/sin/cos of a float argument in
//      double precision with only +, -, *, /, fma and exact bit operations,
//      so they give the same bits on the host and on the device (both
//      compiled with -ffp-contract=off); rounded to float they are the
//      correctly rounded result except in rare hard cases;
//   2. a correction table, built on the host by running these functions and
//      glibc over every float the uniform draw can produce (5*2^24 + 1
//      values), records where glibc's float result is one ulp above or below.
// Algorithms (textbook):
//   log: x = m 2^e, m in [sqrt(1/2), sqrt(2)), log m = 2 atanh(f/(2+f)),
//        f = m - 1, odd series to s^23; e ln2 with a two-part ln2.
//   sin/cos: Cody-Waite reduction by pi/2 with a three-part pi/2
//        (fdlibm's split), Taylor series to r^19 / r^18 on |r| <= pi/4.
#pragma once
#include <stdint.h>
#include <string.h>

#if defined(__HIPCC__)
#define XMD_HD __host__ __device__
#else
#define XMD_HD
#endif

namespace xmd_bm {

XMD_HD inline uint64_t dbits(double x) {
    uint64_t u;
    memcpy(&u, &x, 8);
    return u;
}
XMD_HD inline double bitsd(uint64_t u) {
    double x;
    memcpy(&x, &u, 8);
    return x;
}
XMD_HD inline uint32_t fbits(float x) {
    uint32_t u;
    memcpy(&u, &x, 4);
    return u;
}
XMD_HD inline float bitsf(uint32_t u) {
    float x;
    memcpy(&x, &u, 4);
    return x;
}

// log(x) for finite x > 0 given as a float widened to double.
XMD_HD inline double log_d(double x) {
    const double ln2_hi = 6.93147180369123816490e-01; /* 0x3fe62e42fee00000 */
    const double ln2_lo = 1.90821492927058770002e-10; /* 0x3dea39ef35793c76 */
    uint64_t u = dbits(x);
    int e = (int)((u >> 52) & 0x7ff);
    if (e == 0) {  // subnormal double cannot come from a float of (0, 1]
        x = x * 18014398509481984.0;  // 2^54
        u = dbits(x);
        e = (int)((u >> 52) & 0x7ff) - 54;
    }
    e -= 1023;
    double m = bitsd((u & 0x000fffffffffffffull) | 0x3ff0000000000000ull);  // [1,2)
    if (m > 1.4142135623730951) {
        m = m * 0.5;
        e += 1;
    }
    double f = m - 1.0;
    double s = f / (2.0 + f);
    double z = s * s;
    double p = 1.0 / 23.0;
    p = fma(p, z, 1.0 / 21.0);
    p = fma(p, z, 1.0 / 19.0);
    p = fma(p, z, 1.0 / 17.0);
    p = fma(p, z, 1.0 / 15.0);
    p = fma(p, z, 1.0 / 13.0);
    p = fma(p, z, 1.0 / 11.0);
    p = fma(p, z, 1.0 / 9.0);
    p = fma(p, z, 1.0 / 7.0);
    p = fma(p, z, 1.0 / 5.0);
    p = fma(p, z, 1.0 / 3.0);
    double logm = fma(2.0 * s, z * p, 2.0 * s);
    double de = (double)e;
    return fma(de, ln2_hi, fma(de, ln2_lo, logm));
}

XMD_HD inline void sincos_red(double t, double &s, double &c) {
    // t in [0, 8): reduce by k*pi/2
    const double two_over_pi = 6.36619772367581382433e-01;
    const double pio2_1 = 1.57079632673412561417e+00;  /* 0x3FF921FB54400000 */
    const double pio2_2 = 6.07710050650619224932e-11;  /* 0x3DD0B4611A600000 */
    const double pio2_3 = 2.02226624879595063154e-21;  /* 0x3BA3198A2E037073 */
    double kd = t * two_over_pi;
    // round half to even by adding/subtracting 2^52 (|kd| < 2^51)
    double big = 4503599627370496.0;
    kd = (kd + big) - big;
    int k = (int)kd;
    double r = t - kd * pio2_1;
    r = r - kd * pio2_2;
    r = r - kd * pio2_3;
    double z = r * r;
    // sin r = r + r^3 * S(z)
    double sp = -1.0 / 121645100408832000.0;  // -1/19!
    sp = fma(sp, z, 1.0 / 355687428096000.0);  // 1/17!
    sp = fma(sp, z, -1.0 / 1307674368000.0);   // -1/15!
    sp = fma(sp, z, 1.0 / 6227020800.0);       // 1/13!
    sp = fma(sp, z, -1.0 / 39916800.0);        // -1/11!
    sp = fma(sp, z, 1.0 / 362880.0);           // 1/9!
    sp = fma(sp, z, -1.0 / 5040.0);            // -1/7!
    sp = fma(sp, z, 1.0 / 120.0);              // 1/5!
    sp = fma(sp, z, -1.0 / 6.0);               // -1/3!
    double sr = fma(r * z, sp, r);
    // cos r = 1 - z/2 + z^2 C(z), C(z) = 1/4! - z/6! + ... - z^7/18!
    double cp = -1.0 / 6402373705728000.0;      // -1/18!
    cp = fma(cp, z, 1.0 / 20922789888000.0);    // 1/16!
    cp = fma(cp, z, -1.0 / 87178291200.0);      // -1/14!
    cp = fma(cp, z, 1.0 / 479001600.0);         // 1/12!
    cp = fma(cp, z, -1.0 / 3628800.0);          // -1/10!
    cp = fma(cp, z, 1.0 / 40320.0);             // 1/8!
    cp = fma(cp, z, -1.0 / 720.0);              // -1/6!
    cp = fma(cp, z, 1.0 / 24.0);                // 1/4!
    double hz = 0.5 * z;
    double w = 1.0 - hz;
    double cr = w + (((1.0 - w) - hz) + z * z * cp);
    switch (k & 3) {
    case 0: s = sr; c = cr; break;
    case 1: s = cr; c = -sr; break;
    case 2: s = -sr; c = -cr; break;
    default: s = -cr; c = sr; break;
    }
}

// Index of a uniform draw's float value v = (float)(u >> 32) (before the
// 2^-32 scale) in [0, 5*2^24].
XMD_HD inline uint32_t draw_index(float v) {
    if (v < 16777216.0f) return (uint32_t)v;
    uint32_t b = fbits(v);
    uint32_t e = (b >> 23) & 0xff;  // 24+127 .. 32+127
    return 16777216u + (e - 151u) * 8388608u + (b & 0x7fffffu);
}
static const uint32_t NUM_DRAW_VALUES = 5u * 16777216u + 1u;

XMD_HD inline float next_up(float v) {
    uint32_t b = fbits(v);
    if (v == 0.0f) return bitsf(1u);
    return bitsf(v > 0.0f ? b + 1u : b - 1u);
}
XMD_HD inline float next_down(float v) {
    uint32_t b = fbits(v);
    if (v == 0.0f) return bitsf(0x80000001u);
    return bitsf(v > 0.0f ? b - 1u : b + 1u);
}
// code: 0 keep, 1 one ulp up, 2 one ulp down
XMD_HD inline float apply_fix(float v, uint32_t code) {
    return code == 1u ? next_up(v) : (code == 2u ? next_down(v) : v);
}

// float logf / sinf / cosf of the engine before correction
XMD_HD inline float logf_cr(float x) {
    if (x == 0.0f) return bitsf(0xff800000u);  // -inf
    return (float)log_d((double)x);
}

}  // namespace xmd_bm
