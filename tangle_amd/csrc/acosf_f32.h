// Single-precision arc cosine, evaluated the way the host's libm evaluates
// acosf: the fdlibm algorithm (e_acosf.c, Sun Microsystems 1993, float
// version by Ian Lance Taylor), which glibc ships for binary32. One
// rational approximation of asin on [0, 1/2], R(z) = P(z) / Q(z), is applied
// to x for |x| < 1/2 and to sqrt((1 -+ x) / 2) otherwise:
//   |x| < 1/2:  acos x = pi/2 - (x + x R(x^2))
//   x < -1/2:   acos x = pi - 2 (s + s R(z)),       z = (1 + x)/2, s = sqrt z
//   x > 1/2:    acos x = 2 (s_hi + (s R(z) + c)),   z = (1 - x)/2, s_hi = s
//               with its low 12 bits cleared, c = (z - s_hi^2) / (s + s_hi)
// pi/2 is split into a high and a low part. Every operation is a correctly
// rounded binary32 operation (no contraction), in the order written, so the
// result is the host libm's, bit for bit. The three cases are evaluated
// without branches and selected, which keeps lanes of a wave together.
#pragma once

#include <math.h>
#include <stdint.h>
#include <string.h>

#if defined(__HIPCC__)
#define TGH_HOST_DEVICE __host__ __device__
#else
#define TGH_HOST_DEVICE
#endif

namespace tgh {

TGH_HOST_DEVICE inline float f32_from_bits(uint32_t u) {
    float f;
    memcpy(&f, &u, sizeof f);
    return f;
}

TGH_HOST_DEVICE inline uint32_t f32_bits(float f) {
    uint32_t u;
    memcpy(&u, &f, sizeof u);
    return u;
}

// `Sqrt` is the correctly rounded square root to use for s = sqrt(z); z is
// 0 or at least 2^-25 whenever s is used (|x| >= 1/2).
template <typename Sqrt>
TGH_HOST_DEVICE inline float acosf_f32_with(float x, Sqrt sqrt_rn) {
    const float pi = f32_from_bits(0x40490fdau);
    const float pio2_hi = f32_from_bits(0x3fc90fdau);
    const float pio2_lo = f32_from_bits(0x33a22168u);
    const float p0 = f32_from_bits(0x3e2aaaabu), p1 = f32_from_bits(0xbea6b090u), p2 = f32_from_bits(0x3e4e0aa8u),
                p3 = f32_from_bits(0xbd241146u), p4 = f32_from_bits(0x3a4f7f04u), p5 = f32_from_bits(0x3811ef08u);
    const float q1 = f32_from_bits(0xc019d139u), q2 = f32_from_bits(0x4001572du), q3 = f32_from_bits(0xbf303361u),
                q4 = f32_from_bits(0x3d9dc62eu);
    const uint32_t hx = f32_bits(x), ix = hx & 0x7fffffffu;
    const bool negative = (hx >> 31) != 0u;
    const bool small = ix < 0x3f000000u;
    const float z = small ? x * x : (negative ? (1.0f + x) : (1.0f - x)) * 0.5f;
    const float p = z * (p0 + z * (p1 + z * (p2 + z * (p3 + z * (p4 + z * p5)))));
    const float q = 1.0f + z * (q1 + z * (q2 + z * (q3 + z * q4)));
    const float r = p / q;
    const float s = sqrt_rn(z);
    const float near_zero = pio2_hi - (x - (pio2_lo - x * r));
    const float below = pi - 2.0f * (s + (r * s - pio2_lo));
    const float s_hi = f32_from_bits(f32_bits(s) & 0xfffff000u);
    const float c = (z - s_hi * s_hi) / (s + s_hi);
    const float above = 2.0f * (s_hi + (r * s + c));
    float result = small ? near_zero : (negative ? below : above);
    if (ix <= 0x32800000u) result = pio2_hi + pio2_lo;  // |x| <= 2^-26
    if (ix == 0x3f800000u) result = negative ? pi + 2.0f * pio2_lo : 0.0f;
    if (ix > 0x3f800000u) result = (x - x) / (x - x);  // |x| > 1, NaN: NaN
    return result;
}

struct SqrtRn {
    TGH_HOST_DEVICE float operator()(float z) const { return sqrtf(z); }
};

TGH_HOST_DEVICE inline float acosf_f32(float x) { return acosf_f32_with(x, SqrtRn{}); }

}  // namespace tgh
