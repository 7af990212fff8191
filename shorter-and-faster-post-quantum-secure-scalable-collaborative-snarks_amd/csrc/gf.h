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

#define GF_HD __host__ __device__ __forceinline__

namespace gf {

constexpr uint64_t P = 0x1FFFFFFFFFFFFFFFull;   // 2^61 - 1
constexpr uint64_t M29 = 0x1FFFFFFFull;

struct alignas(16) fe {
    uint64_t re, im;
};

GF_HD fe mk(uint64_t re, uint64_t im) { fe r; r.re = re; r.im = im; return r; }
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

// x * 2^k mod p for canonical x, 0 <= k < 61: a 61-bit rotation
GF_HD uint64_t shl61(uint64_t x, int k) {
    return canon(((x << k) & P) | (x >> (61 - k)));
}

GF_HD fe add(const fe &a, const fe &b) { return mk(add61(a.re, b.re), add61(a.im, b.im)); }
GF_HD fe sub(const fe &a, const fe &b) { return mk(sub61(a.re, b.re), sub61(a.im, b.im)); }
GF_HD fe neg(const fe &a) { return mk(sub61(0, a.re), sub61(0, a.im)); }

// (a + bI)(c + dI) = (ac - bd) + ((a+b)(c+d) - ac - bd) I
GF_HD fe mul(const fe &x, const fe &y) {
    uint64_t ac = mul61(x.re, y.re);
    uint64_t bd = mul61(x.im, y.im);
    uint64_t t = mul61(x.re + x.im, y.re + y.im);
    uint64_t im = t + (P - ac) + (P - bd);                  // < 3p
    im = canon((im & P) + (im >> 61));
    return mk(sub61(ac, bd), im);
}

// multiply by I: (a + bI) I = -b + aI
GF_HD fe mul_i(const fe &x) { return mk(sub61(0, x.im), x.re); }
// multiply by -I: (a + bI)(-I) = b - aI
GF_HD fe mul_neg_i(const fe &x) { return mk(x.im, sub61(0, x.re)); }

GF_HD fe from_u64(uint64_t x) { return mk(x, 0); }

GF_HD fe pow(fe x, uint64_t e) {
    fe r = one();
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
    fe g = mk(G_RE, G_IM);
    for (int i = 0; i < 62 - logn; ++i) g = mul(g, g);
    return g;
}

}  // namespace gf
