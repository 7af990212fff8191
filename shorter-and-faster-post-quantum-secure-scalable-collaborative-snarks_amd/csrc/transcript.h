// transcript.h -- Fiat-Shamir chain of the reference (src/Fiat_Shamir.cpp).
//
// hash_F(x, y) (:8-27): BLAKE3 of the 64-byte block
// {x.real, x.img, 0, 0, y.real, y.img, 0, 0} (little-endian u64 words);
// digest bytes 0..7 give real = sum_i H[i] * 2^(7i), bytes 8..15 give img
// the same way (a 7-bit stride, so both stay below 2^57.1 < p).
// hash_to_field (:29-34) folds each absorbed element into the running value.
#pragma once
#include "gf.h"
#include "blake3.h"

namespace ts {

GF_HD gf::fe hash_F(const gf::fe &x, const gf::fe &y) {
    uint32_t m[16], h[8];
    m[0] = (uint32_t)x.re; m[1] = (uint32_t)(x.re >> 32);
    m[2] = (uint32_t)x.im; m[3] = (uint32_t)(x.im >> 32);
    m[4] = 0; m[5] = 0; m[6] = 0; m[7] = 0;
    m[8] = (uint32_t)y.re; m[9] = (uint32_t)(y.re >> 32);
    m[10] = (uint32_t)y.im; m[11] = (uint32_t)(y.im >> 32);
    m[12] = 0; m[13] = 0; m[14] = 0; m[15] = 0;
    b3::compress64(m, h);
    uint64_t re = 0, im = 0;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        const uint64_t br = (h[i >> 2] >> (8 * (i & 3))) & 0xFF;          // byte i
        const uint64_t bi = (h[2 + (i >> 2)] >> (8 * (i & 3))) & 0xFF;    // byte 8+i
        re += br << (7 * i);
        im += bi << (7 * i);
    }
    return gf::mk(re, im);
}

}  // namespace ts
