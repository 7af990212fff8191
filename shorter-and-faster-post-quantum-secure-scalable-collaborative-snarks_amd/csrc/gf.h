// gf.h -- GF((2^61-1)^2) arithmetic for gfx950 kernels and host code.
//
// Element layout matches the reference's virgo::fieldElement
// (src/fieldElement.hpp:96-97): {u64 real; u64 img;}, 16 bytes, values
// canonical in [0, p).  The extension is F_p[I]/(I^2 + 1), p = 2^61 - 1
// (src/fieldElement.cpp:49-78).  Exact arithmetic: any correct canonical
// implementation yields the reference's bits.
//
// The 61x61-bit product is built from four 32x32->64 partial products
// (v_mad_u64_u32 on CDNA) and folded with 2^61 == 1, 2^64 == 8 (mod p).
#pragma once
#include <stdint.h>
#include <hip/hip_runtime.h>

#define GF_HD __host__ __device__ __forceinline__ constexpr

namespace gf {

constexpr uint64_t P = 0x1FFFFFFFFFFFFFFFull;   // 2^61 - 1
constexpr uint64_t M29 = 0x1FFFFFFFull;

struct alignas(16) fe {
    uint64_t re, im;
};

GF_HD fe mk(uint64_t re, uint64_t im) { return fe{re, im}; }
GF_HD fe zero() { return mk(0, 0); }
GF_HD fe one() { return mk(1, 0); }
GF_HD bool eq(const fe &a, const fe &b) { return a.re == b.re && a.im == b.im; }

// x < 2p  ->  canonical
GF_HD uint64_t canon(uint64_t x) { return x >= P ? x - P : x; }

// a, b < 2^62  ->  a*b mod p, canonical
GF_HD uint64_t mul61(uint64_t a, uint64_t b) {
    uint32_t a0 = (uint32_t)a, a1 = (uint32_t)(a >> 32);
    uint32_t b0 = (uint32_t)b, b1 = (uint32_t)(b >> 32);
    uint64_t p00 = (uint64_t)a0 * b0;                       // < 2^64
    uint64_t p01 = (uint64_t)a0 * b1 + (uint64_t)a1 * b0;   // < 2^63
    uint64_t p11 = (uint64_t)a1 * b1;                       // < 2^60
    // a*b = p11*2^64 + p01*2^32 + p00
    //     == 8*p11 + (p01 >> 29) + (p01 & M29)*2^32 + (p00 >> 61) + (p00 & P)
    uint64_t s = (p00 & P) + (p00 >> 61) + (p11 << 3) + (p01 >> 29) + ((p01 & M29) << 32);
    uint64_t r = (s & P) + (s >> 61);                       // <= p + 7
    return canon(r);
}

GF_HD uint64_t add61(uint64_t a, uint64_t b) { return canon(a + b); }
GF_HD uint64_t sub61(uint64_t a, uint64_t b) { return a >= b ? a - b : a + P - b; }

// x * 2^k mod p for canonical x, 0 < k < 61: a 61-bit rotation (the
// all-ones pattern p is the only fixed point, so canonical stays canonical)
GF_HD uint64_t shl61(uint64_t x, int k) {
    return ((x << k) & P) | (x >> (61 - k));
}

GF_HD fe add(const fe &a, const fe &b) { return mk(add61(a.re, b.re), add61(a.im, b.im)); }
GF_HD fe sub(const fe &a, const fe &b) { return mk(sub61(a.re, b.re), sub61(a.im, b.im)); }
GF_HD fe neg(const fe &a) { return mk(sub61(0, a.re), sub61(0, a.im)); }

// (a + bI)(c + dI) = (ac - bd) + (ad + bc) I.  Both output coordinates are
// sums of two 61x61-bit products (the real one with -b = p XOR b for
// canonical b), accumulated from 31/30-bit limbs with one Mersenne fold per
// coordinate: 16 v_mad_u64_u32 and two reductions (60 VALU instructions on
// gfx950, against 92 for a three-multiply Karatsuba with three reductions).
//   x = x1*2^31 + x0 (x0 < 2^31, x1 < 2^30); 2^62 == 2, 2^61 == 1 (mod p)
// a*b + c*d for canonical a, c and b, d < 2^61 + 2^33 (weak, see below),
// folded once: the result is congruent mod p and <= p + 7
GF_HD uint64_t dot2_61_fold(uint64_t a, uint64_t b, uint64_t c, uint64_t d) {
    const uint64_t M31 = 0x7FFFFFFFull;
    const uint32_t a0 = (uint32_t)(a & M31), a1 = (uint32_t)(a >> 31);
    const uint32_t b0 = (uint32_t)(b & M31), b1 = (uint32_t)(b >> 31);
    const uint32_t c0 = (uint32_t)(c & M31), c1 = (uint32_t)(c >> 31);
    const uint32_t d0 = (uint32_t)(d & M31), d1 = (uint32_t)(d >> 31);
    const uint64_t p00 = (uint64_t)a0 * b0 + (uint64_t)c0 * d0;                        // < 2^63
    const uint64_t p01 = (uint64_t)a0 * b1 + (uint64_t)a1 * b0 + (uint64_t)c0 * d1 + (uint64_t)c1 * d0;  // < 2^63
    const uint64_t p11 = (uint64_t)a1 * b1 + (uint64_t)c1 * d1;                        // < 2^61
    // value = p11*2^62 + p01*2^31 + p00; p01*2^31 = (p01>>30)*2^61 + (p01 & M30)*2^31
    const uint64_t s = (p11 << 1) + p00 + (p01 >> 30) + ((p01 & 0x3FFFFFFFull) << 31);  // < 2^64
    return (s & P) + (s >> 61);                                                       // <= p + 7
}

// a*b + c*d mod p, canonical
GF_HD uint64_t dot2_61(uint64_t a, uint64_t b, uint64_t c, uint64_t d) { return canon(dot2_61_fold(a, b, c, d)); }

GF_HD fe mul(const fe &x, const fe &y) {
    return mk(dot2_61(x.re, y.re, x.im ^ P, y.im), dot2_61(x.re, y.im, x.im, y.re));
}

// The same product left folded once (components <= p + 7) for consumers that
// take weak values; x's components must be <= p (x.im goes through P ^ x.im).
GF_HD fe mul_w(const fe &x, const fe &y) {
    return mk(dot2_61_fold(x.re, y.re, x.im ^ P, y.im), dot2_61_fold(x.re, y.im, x.im, y.re));
}

// multiply by I: (a + bI) I = -b + aI
GF_HD fe mul_i(const fe &x) { return mk(sub61(0, x.im), x.re); }
// multiply by -I: (a + bI)(-I) = b - aI
GF_HD fe mul_neg_i(const fe &x) { return mk(x.im, sub61(0, x.re)); }

GF_HD fe from_u64(uint64_t x) { return mk(x, 0); }

GF_HD fe pow(fe x, uint64_t e) {
    fe r = mk(1, 0);
    while (e) {
        if (e & 1) r = mul(r, x);
        x = mul(x, x);
        e >>= 1;
    }
    return r;
}

// a^(p-2) in F_p
GF_HD uint64_t inv61(uint64_t a) {
    uint64_t r = 1, x = a, e = P - 2;
    while (e) {
        if (e & 1) r = mul61(r, x);
        x = mul61(x, x);
        e >>= 1;
    }
    return r;
}

// 1/(a + bI) = (a - bI)/(a^2 + b^2); 0 maps to 0 as x^(p^2-2) does
GF_HD fe inv(const fe &x) {
    uint64_t n = add61(mul61(x.re, x.re), mul61(x.im, x.im));
    uint64_t ni = inv61(n);
    return mk(mul61(x.re, ni), mul61(sub61(0, x.im), ni));
}

// generator of the 2^62-order subgroup used by the reference
// (src/fieldElement.cpp:238-250, src/utils.cpp:706-717)
constexpr uint64_t G_RE = 2147483648ull, G_IM = 1033321771269002680ull;

GF_HD fe root_of_unity(int logn) {
    fe g{G_RE, G_IM};
    for (int i = 0; i < 62 - logn; ++i) g = mul(g, g);
    return g;
}

// ---- multiplication by small roots of unity, for transform kernels ----
// The reference's roots (src/utils.cpp:706-717) give w_4 = -I and
// w_8 = 2^30 (1 - I); both are checked at compile time below.
static_assert(root_of_unity(2).re == 0 && root_of_unity(2).im == P - 1, "w4 = -I");
static_assert(root_of_unity(3).re == (1ull << 30) && root_of_unity(3).im == P - (1ull << 30),
              "w8 = 2^30 (1 - I)");

// x * w_8^k for odd k: w_8^k = 2^30 (a + b I) with (a, b) = (1,-1), (-1,-1),
// (-1,1), (1,1) for k = 1, 3, 5, 7
template <int K>
GF_HD fe mul_w8(const fe &x) {
    static_assert(K == 1 || K == 3 || K == 5 || K == 7, "odd power of w8");
    const uint64_t s = add61(x.re, x.im);   // x + y
    const uint64_t d = sub61(x.im, x.re);   // y - x
    const uint64_t rs = shl61(s, 30), rd = shl61(d, 30);
    if (K == 1) return mk(rs, rd);
    if (K == 3) return mk(rd, sub61(0, rs));
    if (K == 5) return mk(sub61(0, rs), sub61(0, rd));
    return mk(sub61(0, rd), rs);
}

// x * w_M^E (w_M = root_of_unity(log2 M)), E reduced mod M at compile time;
// powers that are +-1, +-I or odd powers of w_8 avoid the general multiply
template <int M, int E>
GF_HD fe mul_root(const fe &x) {
    constexpr int e = ((E % M) + M) % M;
    if constexpr (e == 0) {
        return x;
    } else if constexpr (2 * e == M) {
        return neg(x);
    } else if constexpr (4 * e == M) {
        return mk(x.im, sub61(0, x.re));              // x * (-I)
    } else if constexpr (4 * e == 3 * M) {
        return mk(sub61(0, x.im), x.re);              // x * I
    } else if constexpr ((8 * e) % M == 0) {
        return mul_w8<(8 * e) / M>(x);
    } else {
        constexpr int lg = (M <= 1) ? 0 : (M == 2 ? 1 : M == 4 ? 2 : M == 8 ? 3 : M == 16 ? 4 : M == 32 ? 5 :
                                         M == 64 ? 6 : M == 128 ? 7 : M == 256 ? 8 : M == 512 ? 9 : 10);
        constexpr fe w = pow(root_of_unity(lg), (uint64_t)e);
        return mul(x, w);
    }
}

// ---- weak forms for the transform kernels ----
// A weak component lies in [0, W], W = 2^61 + 2^32: congruent to its value
// mod p, not reduced.  mul(w, x) with canonical w and weak x is exact and
// canonical (x's high limb x >> 31 <= 2^30 + 2 keeps dot2_61's partial sums
// below 2^64; only the first operand goes through the P ^ b negation).  The
// butterflies add and subtract without reducing and fold once per radix-4
// layer; canon_w gives the canonical form for storage.
constexpr uint64_t P2 = 2 * P, P4 = 4 * P;

// x < 2^64  ->  (x & p) + (x >> 61) <= p + 7
GF_HD uint64_t fold(uint64_t x) { return (x & P) + (x >> 61); }
// weak -> canonical: one fold leaves [0, p]
GF_HD uint64_t canon_w(uint64_t x) {
    x = fold(x);
    return x == P ? 0 : x;
}
GF_HD fe canon_w(const fe &a) { return mk(canon_w(a.re), canon_w(a.im)); }
// y < 2^63  ->  y * 2^30 as a weak value: y = h 2^31 + l, 2^61 == 1
GF_HD uint64_t mul2_30(uint64_t y) { return ((y & 0x7FFFFFFFull) << 30) + (y >> 31); }

// y * 2^K mod p (0 < K < 61) on a weak y: with y = a + b 2^(61-K), a < 2^(61-K),
// y 2^K = a 2^K + b 2^61 = a 2^K + b.  For y <= 2^61 + 2^32, b <= 2^K + 2^(K-29)
// and the sum is at most (2^(61-K) - 1) 2^K + 2^K + 2^(K-29) <= 2^61 + 2^31: weak
template <int K>
GF_HD uint64_t mul2k_w(uint64_t y) {
    static_assert(K > 0 && K < 61, "0 < K < 61");
    return ((y & ((1ull << (61 - K)) - 1)) << K) + (y >> (61 - K));
}
// x * 2^K on a weak x (2^K is real; 1 / 2^k = 2^(61 - k)), weak result
template <int K>
GF_HD fe mul_pow2_w(const fe &x) { return mk(mul2k_w<K>(x.re), mul2k_w<K>(x.im)); }

// x * w_M^E on a weak x, weak result; E reduced mod M at compile time
template <int M, int E>
GF_HD fe mul_root_w(const fe &x) {
    constexpr int e = ((E % M) + M) % M;
    if constexpr (e == 0) {
        return x;
    } else if constexpr (2 * e == M) {
        return mk(fold(P2 - x.re), fold(P2 - x.im));
    } else if constexpr (4 * e == M) {
        return mk(x.im, fold(P2 - x.re));             // x * (-I)
    } else if constexpr (4 * e == 3 * M) {
        return mk(fold(P2 - x.im), x.re);             // x * I
    } else if constexpr ((8 * e) % M == 0) {
        // w_8^K = 2^30 (a + b I), (a, b) = (1,-1), (-1,-1), (-1,1), (1,1) for K = 1, 3, 5, 7;
        // negations enter as swapped differences or 4p - sum (all < 2^63)
        constexpr int K = (8 * e) / M;
        if constexpr (K == 1) return mk(mul2_30(x.re + x.im), mul2_30(x.im + (P2 - x.re)));
        if constexpr (K == 3) return mk(mul2_30(x.im + (P2 - x.re)), mul2_30(P4 - (x.re + x.im)));
        if constexpr (K == 5) return mk(mul2_30(P4 - (x.re + x.im)), mul2_30(x.re + (P2 - x.im)));
        return mk(mul2_30(x.re + (P2 - x.im)), mul2_30(x.re + x.im));
    } else {
        constexpr int lg = (M <= 1) ? 0 : (M == 2 ? 1 : M == 4 ? 2 : M == 8 ? 3 : M == 16 ? 4 : M == 32 ? 5 :
                                         M == 64 ? 6 : M == 128 ? 7 : M == 256 ? 8 : M == 512 ? 9 : 10);
        constexpr fe w = pow(root_of_unity(lg), (uint64_t)e);
        return mul_w(w, x);
    }
}

}  // namespace gf
