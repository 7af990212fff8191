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

__host__ __device__ __forceinline__ gf::fe hash_F(const gf::fe &x, const gf::fe &y) {
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

// ---- the same hash_F on four lanes (device only) ----
// Lanes 4j..4j+3 of a wave call this together with identical arguments.
// Lane q holds column q of the BLAKE3 state (v[q], v[4+q], v[8+q],
// v[12+q]); the column half-round is one G per lane, the diagonal half-round
// rotates rows 1..3 by 1..3 lanes with DPP quad permutes, runs one G per
// lane and rotates back.  About 300 instructions on the chain instead of
// ~700 for the one-lane version.  Returns the same value on all four lanes.

// message-word index used by round r at slot i: the BLAKE3 permutation
// (2,6,3,10,7,0,4,13,1,11,12,5,9,14,15,8) applied r times
struct B3Schedule {
    int s[7][16];
    constexpr B3Schedule() : s() {
        constexpr int perm[16] = {2, 6, 3, 10, 7, 0, 4, 13, 1, 11, 12, 5, 9, 14, 15, 8};
        for (int i = 0; i < 16; ++i) s[0][i] = i;
        for (int r = 1; r < 7; ++r)
            for (int i = 0; i < 16; ++i) s[r][i] = s[r - 1][perm[i]];
    }
};
constexpr B3Schedule B3_SCHED{};

// DPP quad_perm control: lane q reads lane p[q] of its quad
constexpr int qp(int p0, int p1, int p2, int p3) { return p0 | (p1 << 2) | (p2 << 4) | (p3 << 6); }

template <int CTRL>
__device__ __forceinline__ uint32_t quad_perm(uint32_t x) {
    return (uint32_t)__builtin_amdgcn_mov_dpp((int)x, CTRL, 0xF, 0xF, true);
}

// word w_q for lane q, branch-free: k0 / k1 are all-ones where bit 0 / bit 1
// of q is set (nested ?: on q compiles to divergent branches)
__device__ __forceinline__ uint32_t sel4(uint32_t k0, uint32_t k1, uint32_t w0, uint32_t w1, uint32_t w2,
                                         uint32_t w3) {
    const uint32_t a = (w0 & ~k0) | (w1 & k0), b = (w2 & ~k0) | (w3 & k0);
    return (a & ~k1) | (b & k1);
}

__device__ __forceinline__ void g1(uint32_t &a, uint32_t &b, uint32_t &c, uint32_t &d, uint32_t x, uint32_t y) {
    a = a + b + x; d = b3::rotr(d ^ a, 16);
    c = c + d;     b = b3::rotr(b ^ c, 12);
    a = a + b + y; d = b3::rotr(d ^ a, 8);
    c = c + d;     b = b3::rotr(b ^ c, 7);
}

__device__ __forceinline__ gf::fe hash_F_quad(const gf::fe &x, const gf::fe &y) {
    const uint32_t q = threadIdx.x & 3;
    const uint32_t k0 = 0u - (q & 1), k1 = 0u - (q >> 1);
    const uint32_t m[16] = {(uint32_t)x.re, (uint32_t)(x.re >> 32), (uint32_t)x.im, (uint32_t)(x.im >> 32),
                            0u, 0u, 0u, 0u,
                            (uint32_t)y.re, (uint32_t)(y.re >> 32), (uint32_t)y.im, (uint32_t)(y.im >> 32),
                            0u, 0u, 0u, 0u};
    uint32_t a = sel4(k0, k1, b3::IV0, b3::IV1, b3::IV2, b3::IV3);
    uint32_t b = sel4(k0, k1, b3::IV4, b3::IV5, b3::IV6, b3::IV7);
    uint32_t c = a;
    uint32_t d = sel4(k0, k1, 0u, 0u, 64u, b3::FLAGS_ROOT_BLOCK);
#pragma unroll
    for (int r = 0; r < 7; ++r) {
        const int *S = B3_SCHED.s[r];
        g1(a, b, c, d, sel4(k0, k1, m[S[0]], m[S[2]], m[S[4]], m[S[6]]), sel4(k0, k1, m[S[1]], m[S[3]], m[S[5]], m[S[7]]));
        b = quad_perm<qp(1, 2, 3, 0)>(b);
        c = quad_perm<qp(2, 3, 0, 1)>(c);
        d = quad_perm<qp(3, 0, 1, 2)>(d);
        g1(a, b, c, d, sel4(k0, k1, m[S[8]], m[S[10]], m[S[12]], m[S[14]]), sel4(k0, k1, m[S[9]], m[S[11]], m[S[13]], m[S[15]]));
        b = quad_perm<qp(3, 0, 1, 2)>(b);
        c = quad_perm<qp(2, 3, 0, 1)>(c);
        d = quad_perm<qp(1, 2, 3, 0)>(d);
    }
    const uint32_t h = a ^ c;   // digest word q (words 0..3 live on lanes 0..3)
    const uint32_t w0 = quad_perm<qp(0, 0, 0, 0)>(h), w1 = quad_perm<qp(1, 1, 1, 1)>(h);
    const uint32_t w2 = quad_perm<qp(2, 2, 2, 2)>(h), w3 = quad_perm<qp(3, 3, 3, 3)>(h);
    uint64_t re = 0, im = 0;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        const uint32_t wr = i < 4 ? w0 : w1, wi = i < 4 ? w2 : w3;
        re += (uint64_t)((wr >> (8 * (i & 3))) & 0xFF) << (7 * i);
        im += (uint64_t)((wi >> (8 * (i & 3))) & 0xFF) << (7 * i);
    }
    return gf::mk(re, im);
}

}  // namespace ts
