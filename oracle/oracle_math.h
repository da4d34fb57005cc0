/*
 * oracle_math.h -- binary32 elementary functions of the oracle.
 *
 * Test infrastructure: only the oracle (oracle/egs_oracle.c) includes this.
 * The transport is evaluated with these functions instead of libm so that
 * its arithmetic is fully specified by IEEE-754 +,-,*,/,sqrt: the same
 * algorithms, written independently in the HIP path
 * (gegs_amd/csrc/gmath.hpp), give bit-identical results on the GPU, which
 * lets parity be checked per history.  Methods:
 *   log  : x = 2^k m, m in [sqrt(1/2), sqrt(2)), f = (m-1)/(m+1),
 *          ln m = 2 atanh f = 2 (f + f^3/3 + ... + f^11/11)   (Taylor series)
 *          ln x = k ln2 + ln m, ln2 split in a Cody-Waite hi/lo pair.
 *   exp  : k = nearest integer of x / ln2, r = x - k ln2 (Cody-Waite),
 *          e^r by its Taylor series to r^8, scaled by 2^k built from bits.
 *   sin/cos(2 pi t): t reduced modulo 1 exactly, then to an octant
 *          |y| <= 1/8 so that 2 pi y <= pi/4; Taylor series of sin and cos.
 * All functions are accurate to a few ulp over the ranges the transport uses
 * (tests/test_oracle_core.py checks them against double-precision libm).
 */
#ifndef ORACLE_MATH_H
#define ORACLE_MATH_H
#include <math.h>
#include <stdint.h>
#include <string.h>

static inline float om_bits2f(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }
static inline uint32_t om_f2bits(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }

/* 2^k for integer k in [-149, 127], exact */
static inline float om_pow2i(int k) {
    if (k >= -126) return om_bits2f((uint32_t)(k + 127) << 23);
    return om_bits2f(1u << (k + 149));
}

static inline float om_logf(float x) {
    uint32_t u = om_f2bits(x);
    int k = (int)((u >> 23) & 0xffu) - 127;
    if (k == -127) {                      /* subnormal input */
        x = x * 8388608.0f;               /* 2^23 */
        u = om_f2bits(x);
        k = (int)((u >> 23) & 0xffu) - 127 - 23;
    }
    float m = om_bits2f((u & 0x007fffffu) | 0x3f800000u);   /* [1,2) */
    if (m > 1.41421356f) { m = m * 0.5f; k += 1; }
    float f = (m - 1.0f) / (m + 1.0f);
    float f2 = f * f;
    float p = 0.0909090909f;              /* 1/11 */
    p = p * f2 + 0.111111111f;            /* 1/9  */
    p = p * f2 + 0.142857143f;            /* 1/7  */
    p = p * f2 + 0.2f;
    p = p * f2 + 0.333333333f;
    float lm = 2.0f * f + 2.0f * f * f2 * p;
    float kf = (float)k;
    return kf * 0.693145752f + (lm + kf * 1.42860677e-06f);
}

static inline float om_expf(float x) {
    if (x > 88.7f) return om_bits2f(0x7f800000u);
    if (x < -103.9f) return 0.0f;
    float kf = floorf(x * 1.44269504f + 0.5f);
    int k = (int)kf;
    float r = x - kf * 0.693145752f;
    r = r - kf * 1.42860677e-06f;
    float p = 2.48015873e-05f;            /* 1/8! */
    p = p * r + 1.98412698e-04f;          /* 1/7! */
    p = p * r + 1.38888889e-03f;          /* 1/6! */
    p = p * r + 8.33333333e-03f;          /* 1/5! */
    p = p * r + 4.16666667e-02f;          /* 1/4! */
    p = p * r + 0.166666667f;
    p = p * r + 0.5f;
    p = p * r + 1.0f;
    p = p * r + 1.0f;
    if (k > 127) { p = p * 2.0f; k -= 1; }
    if (k < -126) { p = p * om_pow2i(-126); k += 126; }
    return p * om_pow2i(k);
}

/* sin and cos of 2 pi t for finite t */
static inline void om_sincos2pi(float t, float *s, float *c) {
    float y = t - floorf(t);              /* [0,1) exact */
    float q = floorf(y * 8.0f + 0.5f);    /* octant 0..8 */
    y = y - q * 0.125f;                   /* |y| <= 1/16 * 2, exact */
    int oct = ((int)q) & 7;
    float a = y * 6.28318531f;            /* |a| <= pi/8 */
    float a2 = a * a;
    float sp = -2.50521084e-08f;          /* -1/11! */
    sp = sp * a2 + 2.75573192e-06f;       /*  1/9!  */
    sp = sp * a2 - 1.98412698e-04f;       /* -1/7!  */
    sp = sp * a2 + 8.33333333e-03f;       /*  1/5!  */
    sp = sp * a2 - 0.166666667f;
    float sa = a + a * a2 * sp;
    float cp = 2.08767570e-09f;           /*  1/12! */
    cp = cp * a2 - 2.75573192e-07f;       /* -1/10! */
    cp = cp * a2 + 2.48015873e-05f;       /*  1/8!  */
    cp = cp * a2 - 1.38888889e-03f;       /* -1/6!  */
    cp = cp * a2 + 4.16666667e-02f;       /*  1/4!  */
    cp = cp * a2 - 0.5f;
    float ca = 1.0f + a2 * cp;
    /* rotate by oct * 45 degrees */
    const float h = 0.707106781f;
    float ss, cc;
    switch (oct) {
    case 0: ss = sa; cc = ca; break;
    case 1: ss = h * (sa + ca); cc = h * (ca - sa); break;
    case 2: ss = ca; cc = -sa; break;
    case 3: ss = h * (ca - sa); cc = -h * (sa + ca); break;
    case 4: ss = -sa; cc = -ca; break;
    case 5: ss = -h * (sa + ca); cc = h * (sa - ca); break;
    case 6: ss = -ca; cc = sa; break;
    default: ss = h * (sa - ca); cc = h * (sa + ca); break;
    }
    *s = ss;
    *c = cc;
}

#endif
