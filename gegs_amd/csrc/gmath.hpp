// gmath.hpp -- binary32 elementary functions of the transport kernels.
//
// The transport evaluates log, exp and sin/cos(2 pi t) with these routines
// instead of the device math library so that every result is a fixed
// sequence of IEEE-754 +, -, *, / (compiled with -ffp-contract=off and
// correctly rounded division/sqrt): the result is then the same on every
// device and on the host.  Published methods:
//   log : Cody-Waite: x = 2^k m with m in [sqrt(1/2), sqrt(2)),
//         ln m = 2 atanh((m-1)/(m+1)) by its Taylor series to the 11th power.
//   exp : Cody-Waite reduction r = x - k ln2 (ln2 split hi/lo), Taylor series
//         of e^r to r^8, scaling by 2^k assembled from exponent bits.
//   sincos2pi: exact reduction of t modulo 1 and to an octant, Taylor series
//         of sin and cos on |2 pi y| <= pi/8, rotation by multiples of 45 deg.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace gegs {

__host__ __device__ __forceinline__ float pow2i(int k) {
    return k >= -126 ? __uint_as_float((uint32_t)(k + 127) << 23) : __uint_as_float(1u << (k + 149));
}

__host__ __device__ __forceinline__ float glog(float x) {
    uint32_t bits = __float_as_uint(x);
    int k = (int)((bits >> 23) & 0xffu) - 127;
    if (k == -127) {
        x = x * 8388608.0f;
        bits = __float_as_uint(x);
        k = (int)((bits >> 23) & 0xffu) - 150;
    }
    float m = __uint_as_float((bits & 0x007fffffu) | 0x3f800000u);
    if (m > 1.41421356f) {
        m = m * 0.5f;
        k += 1;
    }
    const float f = (m - 1.0f) / (m + 1.0f);
    const float f2 = f * f;
    float s = 0.0909090909f;
    s = s * f2 + 0.111111111f;
    s = s * f2 + 0.142857143f;
    s = s * f2 + 0.2f;
    s = s * f2 + 0.333333333f;
    const float lnm = 2.0f * f + 2.0f * f * f2 * s;
    const float kf = (float)k;
    return kf * 0.693145752f + (lnm + kf * 1.42860677e-06f);
}

__host__ __device__ __forceinline__ float gexp(float x) {
    if (x > 88.7f) return __uint_as_float(0x7f800000u);
    if (x < -103.9f) return 0.0f;
    const float kf = floorf(x * 1.44269504f + 0.5f);
    int k = (int)kf;
    float r = x - kf * 0.693145752f;
    r = r - kf * 1.42860677e-06f;
    float s = 2.48015873e-05f;
    s = s * r + 1.98412698e-04f;
    s = s * r + 1.38888889e-03f;
    s = s * r + 8.33333333e-03f;
    s = s * r + 4.16666667e-02f;
    s = s * r + 0.166666667f;
    s = s * r + 0.5f;
    s = s * r + 1.0f;
    s = s * r + 1.0f;
    if (k > 127) {
        s = s * 2.0f;
        k -= 1;
    }
    if (k < -126) {
        s = s * pow2i(-126);
        k += 126;
    }
    return s * pow2i(k);
}

__host__ __device__ __forceinline__ void gsincos2pi(float t, float &sn, float &cs) {
    float y = t - floorf(t);
    const float q = floorf(y * 8.0f + 0.5f);
    y = y - q * 0.125f;
    const int oct = ((int)q) & 7;
    const float a = y * 6.28318531f;
    const float a2 = a * a;
    float ps = -2.50521084e-08f;
    ps = ps * a2 + 2.75573192e-06f;
    ps = ps * a2 - 1.98412698e-04f;
    ps = ps * a2 + 8.33333333e-03f;
    ps = ps * a2 - 0.166666667f;
    const float sa = a + a * a2 * ps;
    float pc = 2.08767570e-09f;
    pc = pc * a2 - 2.75573192e-07f;
    pc = pc * a2 + 2.48015873e-05f;
    pc = pc * a2 - 1.38888889e-03f;
    pc = pc * a2 + 4.16666667e-02f;
    pc = pc * a2 - 0.5f;
    const float ca = 1.0f + a2 * pc;
    const float h = 0.707106781f;
    switch (oct) {
    case 0: sn = sa; cs = ca; break;
    case 1: sn = h * (sa + ca); cs = h * (ca - sa); break;
    case 2: sn = ca; cs = -sa; break;
    case 3: sn = h * (ca - sa); cs = -h * (sa + ca); break;
    case 4: sn = -sa; cs = -ca; break;
    case 5: sn = -h * (sa + ca); cs = h * (sa - ca); break;
    case 6: sn = -ca; cs = sa; break;
    default: sn = h * (sa - ca); cs = h * (sa + ca); break;
    }
}

}  // namespace gegs
