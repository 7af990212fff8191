// blake3.h -- single-block BLAKE3 compression (host and device).
//
// The reference hashes exactly 64 bytes per call (src/Blake3_hash.cpp:5-9):
// one chunk of one block, so the digest is the first 8 words of one
// compression with chaining value IV, counter 0, block_len 64 and flags
// CHUNK_START|CHUNK_END|ROOT.  Written from the BLAKE3 specification
// (round function G, 7 rounds, fixed message permutation between rounds).
#pragma once
#include <stdint.h>
#include <hip/hip_runtime.h>

#define B3_HD __host__ __device__ __forceinline__

namespace b3 {

constexpr uint32_t IV0 = 0x6A09E667u, IV1 = 0xBB67AE85u, IV2 = 0x3C6EF372u, IV3 = 0xA54FF53Au;
constexpr uint32_t IV4 = 0x510E527Fu, IV5 = 0x9B05688Cu, IV6 = 0x1F83D9ABu, IV7 = 0x5BE0CD19u;
constexpr uint32_t FLAGS_ROOT_BLOCK = 0x0Bu;   // CHUNK_START | CHUNK_END | ROOT

B3_HD uint32_t rotr(uint32_t x, int c) { return (x >> c) | (x << (32 - c)); }

#define B3_G(a, b, c, d, x, y)                   \
    do {                                         \
        a = a + b + (x); d = rotr(d ^ a, 16);    \
        c = c + d;       b = rotr(b ^ c, 12);    \
        a = a + b + (y); d = rotr(d ^ a, 8);     \
        c = c + d;       b = rotr(b ^ c, 7);     \
    } while (0)

// one round on state v with message words m0..m15 in the order given
#define B3_ROUND(m0, m1, m2, m3, m4, m5, m6, m7, m8, m9, m10, m11, m12, m13, m14, m15) \
    do {                                                  \
        B3_G(v0, v4, v8, v12, m0, m1);                     \
        B3_G(v1, v5, v9, v13, m2, m3);                     \
        B3_G(v2, v6, v10, v14, m4, m5);                    \
        B3_G(v3, v7, v11, v15, m6, m7);                    \
        B3_G(v0, v5, v10, v15, m8, m9);                    \
        B3_G(v1, v6, v11, v12, m10, m11);                  \
        B3_G(v2, v7, v8, v13, m12, m13);                   \
        B3_G(v3, v4, v9, v14, m14, m15);                   \
    } while (0)

// Compress one 64-byte block given as 16 little-endian words; writes the
// 8-word digest.  The message permutation (2,6,3,10,7,0,4,13,1,11,12,5,9,
// 14,15,8) is applied between rounds by renaming, so round k reads the words
// at the k-fold composed permutation.
B3_HD void compress64(const uint32_t m[16], uint32_t out[8]) {
    uint32_t v0 = IV0, v1 = IV1, v2 = IV2, v3 = IV3, v4 = IV4, v5 = IV5, v6 = IV6, v7 = IV7;
    uint32_t v8 = IV0, v9 = IV1, v10 = IV2, v11 = IV3;
    uint32_t v12 = 0, v13 = 0, v14 = 64, v15 = FLAGS_ROOT_BLOCK;
    uint32_t a0 = m[0], a1 = m[1], a2 = m[2], a3 = m[3], a4 = m[4], a5 = m[5], a6 = m[6], a7 = m[7];
    uint32_t a8 = m[8], a9 = m[9], a10 = m[10], a11 = m[11], a12 = m[12], a13 = m[13], a14 = m[14], a15 = m[15];
    B3_ROUND(a0, a1, a2, a3, a4, a5, a6, a7, a8, a9, a10, a11, a12, a13, a14, a15);
    // each subsequent round uses the previous round's words permuted by
    // perm = (2,6,3,10,7,0,4,13,1,11,12,5,9,14,15,8)
#define B3_PERMUTED_ROUND()                                                            \
    do {                                                                               \
        uint32_t t0 = a2, t1 = a6, t2 = a3, t3 = a10, t4 = a7, t5 = a0, t6 = a4, t7 = a13; \
        uint32_t t8 = a1, t9 = a11, t10 = a12, t11 = a5, t12 = a9, t13 = a14, t14 = a15, t15 = a8; \
        a0 = t0; a1 = t1; a2 = t2; a3 = t3; a4 = t4; a5 = t5; a6 = t6; a7 = t7;        \
        a8 = t8; a9 = t9; a10 = t10; a11 = t11; a12 = t12; a13 = t13; a14 = t14; a15 = t15; \
        B3_ROUND(a0, a1, a2, a3, a4, a5, a6, a7, a8, a9, a10, a11, a12, a13, a14, a15); \
    } while (0)
    B3_PERMUTED_ROUND();
    B3_PERMUTED_ROUND();
    B3_PERMUTED_ROUND();
    B3_PERMUTED_ROUND();
    B3_PERMUTED_ROUND();
    B3_PERMUTED_ROUND();
#undef B3_PERMUTED_ROUND
    out[0] = v0 ^ v8;  out[1] = v1 ^ v9;  out[2] = v2 ^ v10; out[3] = v3 ^ v11;
    out[4] = v4 ^ v12; out[5] = v5 ^ v13; out[6] = v6 ^ v14; out[7] = v7 ^ v15;
}

#undef B3_ROUND
#undef B3_G

// ---- one compression on four lanes (device only) ----
// Lanes 4j..4j+3 of a wave call this together, each with the whole 16-word
// message.  Lane q holds column q of the state (v[q], v[4+q], v[8+q],
// v[12+q]): the column half-round is one G per lane, the diagonal half-round
// rotates rows 1..3 by 1..3 lanes with DPP quad permutes, runs one G per lane
// and rotates back.  The chain is about 300 instructions instead of about
// 700 on one lane.  Lane q returns digest words q and q + 4.

// message-word index used by round r at slot i: the permutation
// (2,6,3,10,7,0,4,13,1,11,12,5,9,14,15,8) applied r times
struct QuadSchedule {
    int s[7][16];
    constexpr QuadSchedule() : s() {
        constexpr int perm[16] = {2, 6, 3, 10, 7, 0, 4, 13, 1, 11, 12, 5, 9, 14, 15, 8};
        for (int i = 0; i < 16; ++i) s[0][i] = i;
        for (int r = 1; r < 7; ++r)
            for (int i = 0; i < 16; ++i) s[r][i] = s[r - 1][perm[i]];
    }
};
constexpr QuadSchedule QUAD_SCHED{};

template <int CTRL>
__device__ __forceinline__ uint32_t qperm(uint32_t x) {
    return (uint32_t)__builtin_amdgcn_mov_dpp((int)x, CTRL, 0xF, 0xF, true);
}
// DPP quad_perm control: lane q reads lane p[q] of its quad
constexpr int qctl(int p0, int p1, int p2, int p3) { return p0 | (p1 << 2) | (p2 << 4) | (p3 << 6); }

// word w_q for lane q, branch-free: k0 / k1 all-ones where bit 0 / bit 1 of q is set
__device__ __forceinline__ uint32_t qsel(uint32_t k0, uint32_t k1, uint32_t w0, uint32_t w1, uint32_t w2,
                                         uint32_t w3) {
    const uint32_t a = (w0 & ~k0) | (w1 & k0), b = (w2 & ~k0) | (w3 & k0);
    return (a & ~k1) | (b & k1);
}

__device__ __forceinline__ void qg(uint32_t &a, uint32_t &b, uint32_t &c, uint32_t &d, uint32_t x, uint32_t y) {
    a = a + b + x; d = rotr(d ^ a, 16);
    c = c + d;     b = rotr(b ^ c, 12);
    a = a + b + y; d = rotr(d ^ a, 8);
    c = c + d;     b = rotr(b ^ c, 7);
}

__device__ __forceinline__ void compress64_quad(const uint32_t m[16], uint32_t &h_lo, uint32_t &h_hi) {
    const uint32_t q = threadIdx.x & 3;
    const uint32_t k0 = 0u - (q & 1), k1 = 0u - (q >> 1);
    uint32_t a = qsel(k0, k1, IV0, IV1, IV2, IV3);
    uint32_t b = qsel(k0, k1, IV4, IV5, IV6, IV7);
    uint32_t c = a;
    uint32_t d = qsel(k0, k1, 0u, 0u, 64u, FLAGS_ROOT_BLOCK);
#pragma unroll
    for (int r = 0; r < 7; ++r) {
        const int *S = QUAD_SCHED.s[r];
        qg(a, b, c, d, qsel(k0, k1, m[S[0]], m[S[2]], m[S[4]], m[S[6]]),
           qsel(k0, k1, m[S[1]], m[S[3]], m[S[5]], m[S[7]]));
        b = qperm<qctl(1, 2, 3, 0)>(b);
        c = qperm<qctl(2, 3, 0, 1)>(c);
        d = qperm<qctl(3, 0, 1, 2)>(d);
        qg(a, b, c, d, qsel(k0, k1, m[S[8]], m[S[10]], m[S[12]], m[S[14]]),
           qsel(k0, k1, m[S[9]], m[S[11]], m[S[13]], m[S[15]]));
        b = qperm<qctl(3, 0, 1, 2)>(b);
        c = qperm<qctl(2, 3, 0, 1)>(c);
        d = qperm<qctl(1, 2, 3, 0)>(d);
    }
    h_lo = a ^ c;   // word q: v[q] ^ v[8 + q]
    h_hi = b ^ d;   // word 4 + q: v[4 + q] ^ v[12 + q]
}

}  // namespace b3
