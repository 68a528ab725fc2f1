// The nearest periodic image index roundf(x / length) of the contact, list
// and capture tests, without a division in the common case.
#pragma once

#include <hip/hip_runtime.h>

namespace tgh {

// roundf(x / length), bit for bit. q = x * inv_length (inv_length the
// correctly rounded 1 / length) lies within 1.6 * 2^-23 |q| of the correctly
// rounded quotient, so both round to the same integer unless a half-integer
// lies between them; that can only happen when q is within 2^-20 |q| of a
// half-integer (or q is not finite), and then the division is done.
__device__ __forceinline__ float image_round(float x, float length, float inv_length) {
    const float q = x * inv_length;
    const float h = fabsf(q);
    const float off = fabsf((h - truncf(h)) - 0.5f);
    float r = roundf(q);
    if (!(off > 0x1p-20f * h)) {
        // the rare case: divide. The empty asm keeps the division inside the
        // branch (the compiler would otherwise compute it always and select).
        float exact = x / length;
        __asm__ volatile("" : "+v"(exact));
        r = roundf(exact);
    }
    return r;
}

}  // namespace tgh
