// ref_shim.cpp -- C entry points over the reference's own compiled sources.
//
// TEST INFRASTRUCTURE ONLY.  Built by oracle/Makefile together with the
// reference files that compile from their own sources in this image:
//   src/fieldElement.cpp, src/polynomial.cpp, src/Blake3_hash.cpp and
//   Blake/blake3.c, Blake/blake3_dispatch.c, Blake/blake3_portable.c.
// The output (oracle/_ref/libref.so) pins the oracle: tests compare the
// oracle's field, BLAKE3, polynomial and glibc-challenge functions against
// these on identical inputs.  No reference source is copied here; this file
// only calls into the compiled reference objects.
#include <cstddef>
#include <cstdint>
#include <cstring>
#include "config_pc.hpp"
#include "polynomial.h"
#include "Blake3_hash.h"

using FE = virgo::fieldElement;

static inline FE ld(const uint64_t *p) { FE x; x.real = p[0]; x.img = p[1]; return x; }
static inline void st(uint64_t *p, const FE &x) { p[0] = x.real; p[1] = x.img; }

extern "C" {

// layout of the reference element: sizeof and member offsets
void ref_fe_layout(uint64_t out[3]) {
    out[0] = sizeof(FE);
    out[1] = offsetof(FE, real);
    out[2] = offsetof(FE, img);
}

void ref_mul(const uint64_t *a, const uint64_t *b, uint64_t *o, uint64_t n) {
    for (uint64_t i = 0; i < n; ++i) st(o + 2 * i, ld(a + 2 * i) * ld(b + 2 * i));
}
void ref_add(const uint64_t *a, const uint64_t *b, uint64_t *o, uint64_t n) {
    for (uint64_t i = 0; i < n; ++i) st(o + 2 * i, ld(a + 2 * i) + ld(b + 2 * i));
}
void ref_sub(const uint64_t *a, const uint64_t *b, uint64_t *o, uint64_t n) {
    for (uint64_t i = 0; i < n; ++i) st(o + 2 * i, ld(a + 2 * i) - ld(b + 2 * i));
}
void ref_inv(const uint64_t *a, uint64_t *o, uint64_t n) {
    for (uint64_t i = 0; i < n; ++i) st(o + 2 * i, ld(a + 2 * i).inv());
}
void ref_pow(const uint64_t *a, uint64_t e_lo, uint64_t e_hi, uint64_t *o) {
    __uint128_t e = ((__uint128_t)e_hi << 64) | e_lo;
    st(o, FE::fastPow(ld(a), e));
}
void ref_root_of_unity(int logn, uint64_t *o) { st(o, FE::getRootOfUnity(logn)); }
void ref_from_ll(long long x, uint64_t *o) { st(o, FE(x)); }
void ref_random_elements(uint64_t *o, uint64_t n) {
    for (uint64_t i = 0; i < n; ++i) st(o + 2 * i, FE::_random());
}
void ref_blake3_hash64(const uint8_t *src, uint8_t *dst) {
    uint8_t buf[64];
    std::memcpy(buf, src, 64);
    blake3_hash(buf, dst);
}
void ref_blake3_any(const uint8_t *src, size_t len, uint8_t *dst) {
    blake3_hasher h;
    blake3_hasher_init(&h);
    blake3_hasher_update(&h, src, len);
    blake3_hasher_finalize(&h, dst, 32);
}
// (a1 x + b1)(a2 x + b2) through the reference's linear_poly product:
// out = quadratic_poly {a, b, c}
void ref_linear_mul(const uint64_t *l1, const uint64_t *l2, uint64_t *o) {
    linear_poly p1(ld(l1), ld(l1 + 2)), p2(ld(l2), ld(l2 + 2));
    quadratic_poly q = p1 * p2;
    st(o, q.a); st(o + 2, q.b); st(o + 4, q.c);
}
void ref_quadratic_eval(const uint64_t *poly, const uint64_t *x, uint64_t *o) {
    quadratic_poly q(ld(poly), ld(poly + 2), ld(poly + 4));
    st(o, q.eval(ld(x)));
}

}  // extern "C"
