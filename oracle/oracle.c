/*
 * oracle.c -- CPU restatement of the reference prover's rank-local path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker).  Loaded by tests/, by
 * __graft_entry__.smoke() and by bench.py's cpu_baseline leg; never by the
 * product library.  Each function cites the reference file:line it follows
 * (paths relative to the reference repository root).
 *
 * Pinning (what checks each function independently of this text):
 *   field ops     -- oracle/_ref (reference src/fieldElement.cpp compiled),
 *                    Python big-int brute force, algebraic identities
 *   root of unity -- oracle/_ref getRootOfUnity; order 2^n exactly
 *   blake3        -- oracle/_ref (reference Blake/ C compiled), published
 *                    BLAKE3 known answers
 *   fft / pfft    -- O(n^2) DFT by definition, inverse round trip
 *   merkle        -- dup-left closed form (level-l node i is the l-fold
 *                    self-hash chain of leaf 2^l*i), leaf digests vs _ref
 *   hash_F        -- digest bytes vs _ref blake3 on the documented layout;
 *                    7-bit packing bound; chain order (partially pinned)
 *   sumcheck      -- H(0)+H(1)=y each round, final values = multilinear
 *                    evaluation at the challenges (brute force)
 *   beta/eval     -- brute-force eq polynomial, Lagrange identity
 *   _random       -- oracle/_ref F::_random() on the same libc seed
 *   whir          -- internal consistency (the reference's own verifier
 *                    checks 1-2 run inside or_whir_open), query-fold closed
 *                    form; exact byte order of the record: parity unpinned
 *                    beyond those checks (reference Whir.cpp unbuildable).
 */
#include "oracle.h"

#include <stdlib.h>
#include <string.h>
#include <math.h>

#define P61 0x1FFFFFFFFFFFFFFFULL

/* ------------------------------------------------------------------ */
/* field: src/fieldElement.cpp                                          */
/* ------------------------------------------------------------------ */

static uint64_t modp(unsigned __int128 t) {
    /* 2^61 == 1 (mod p): fold the high part onto the low 61 bits */
    uint64_t lo = (uint64_t)t & P61;
    uint64_t hi = (uint64_t)(t >> 61);
    uint64_t s = lo + hi;               /* < 2^64 for t < 2^125 */
    s = (s & P61) + (s >> 61);
    if (s >= P61) s -= P61;
    return s;
}

static uint64_t addp(uint64_t a, uint64_t b) {
    uint64_t s = a + b;
    return s >= P61 ? s - P61 : s;
}

static uint64_t subp(uint64_t a, uint64_t b) {
    return a >= b ? a - b : a + P61 - b;
}

or_fe or_add(or_fe a, or_fe b) {            /* fieldElement.cpp:34-47 */
    or_fe r = { addp(a.real, b.real), addp(a.img, b.img) };
    return r;
}

or_fe or_sub(or_fe a, or_fe b) {            /* fieldElement.cpp:80-96 */
    or_fe r = { subp(a.real, b.real), subp(a.img, b.img) };
    return r;
}

or_fe or_mul(or_fe a, or_fe b) {            /* fieldElement.cpp:49-78 */
    /* (a.r + a.i I)(b.r + b.i I), I^2 = -1:
     * real = ar*br - ai*bi, img = ar*bi + ai*br  (all mod p, canonical) */
    uint64_t ac = modp((unsigned __int128)a.real * b.real);
    uint64_t bd = modp((unsigned __int128)a.img * b.img);
    uint64_t ad = modp((unsigned __int128)a.real * b.img);
    uint64_t bc = modp((unsigned __int128)a.img * b.real);
    or_fe r = { subp(ac, bd), addp(ad, bc) };
    return r;
}

or_fe or_from_u64(uint64_t x) {             /* fieldElement.cpp:24-27 */
    or_fe r = { x, 0 };
    return r;
}

or_fe or_pow(or_fe x, unsigned __int128 e) {   /* fieldElement.cpp:323-335 */
    or_fe ret = { 1, 0 }, t = x;
    while (e) {
        if (e & 1) ret = or_mul(ret, t);
        t = or_mul(t, t);
        e >>= 1;
    }
    return ret;
}

or_fe or_inv(or_fe a) {                     /* fieldElement.cpp:207-210 */
    unsigned __int128 p = P61;
    return or_pow(a, p * p - 2);
}

or_fe or_root_of_unity(int logn) {          /* utils.cpp:706-717 */
    /* generator of order 2^62: 2^31 + 1033321771269002680 I */
    or_fe g = { 2147483648ULL, 1033321771269002680ULL };
    for (int i = 0; i < 62 - logn; ++i) g = or_mul(g, g);
    return g;
}

void or_vec_mul(const or_fe *a, const or_fe *b, or_fe *o, uint64_t n) { for (uint64_t i = 0; i < n; ++i) o[i] = or_mul(a[i], b[i]); }
void or_vec_add(const or_fe *a, const or_fe *b, or_fe *o, uint64_t n) { for (uint64_t i = 0; i < n; ++i) o[i] = or_add(a[i], b[i]); }
void or_vec_sub(const or_fe *a, const or_fe *b, or_fe *o, uint64_t n) { for (uint64_t i = 0; i < n; ++i) o[i] = or_sub(a[i], b[i]); }
void or_vec_inv(const or_fe *a, or_fe *o, uint64_t n) { for (uint64_t i = 0; i < n; ++i) o[i] = or_inv(a[i]); }

/* ------------------------------------------------------------------ */
/* BLAKE3 single-block compression (BLAKE3 specification; the reference */
/* vendors BLAKE3 1.8.2 in Blake/, blake3_portable.c:8-82)              */
/* ------------------------------------------------------------------ */

static const uint32_t B3_IV[8] = {
    0x6A09E667u, 0xBB67AE85u, 0x3C6EF372u, 0xA54FF53Au,
    0x510E527Fu, 0x9B05688Cu, 0x1F83D9ABu, 0x5BE0CD19u };
/* per-round message word permutation from the BLAKE3 paper, sec. 2.2 */
static const uint8_t B3_PERM[16] = { 2, 6, 3, 10, 7, 0, 4, 13, 1, 11, 12, 5, 9, 14, 15, 8 };

static inline uint32_t rotr(uint32_t x, int c) { return (x >> c) | (x << (32 - c)); }

static inline void b3_g(uint32_t *v, int a, int b, int c, int d, uint32_t mx, uint32_t my) {
    v[a] = v[a] + v[b] + mx; v[d] = rotr(v[d] ^ v[a], 16);
    v[c] = v[c] + v[d];      v[b] = rotr(v[b] ^ v[c], 12);
    v[a] = v[a] + v[b] + my; v[d] = rotr(v[d] ^ v[a], 8);
    v[c] = v[c] + v[d];      v[b] = rotr(v[b] ^ v[c], 7);
}

/* one compression with chaining value IV, counter 0; out = first 8 words */
static void b3_compress(const uint32_t m_in[16], uint32_t block_len, uint32_t flags, uint32_t out[8]) {
    uint32_t v[16], m[16], t[16];
    memcpy(m, m_in, sizeof(m));
    for (int i = 0; i < 8; ++i) v[i] = B3_IV[i];
    v[8] = B3_IV[0]; v[9] = B3_IV[1]; v[10] = B3_IV[2]; v[11] = B3_IV[3];
    v[12] = 0; v[13] = 0; v[14] = block_len; v[15] = flags;
    for (int round = 0; round < 7; ++round) {
        b3_g(v, 0, 4, 8, 12, m[0], m[1]);
        b3_g(v, 1, 5, 9, 13, m[2], m[3]);
        b3_g(v, 2, 6, 10, 14, m[4], m[5]);
        b3_g(v, 3, 7, 11, 15, m[6], m[7]);
        b3_g(v, 0, 5, 10, 15, m[8], m[9]);
        b3_g(v, 1, 6, 11, 12, m[10], m[11]);
        b3_g(v, 2, 7, 8, 13, m[12], m[13]);
        b3_g(v, 3, 4, 9, 14, m[14], m[15]);
        for (int i = 0; i < 16; ++i) t[i] = m[B3_PERM[i]];
        memcpy(m, t, sizeof(m));
    }
    for (int i = 0; i < 8; ++i) out[i] = v[i] ^ v[i + 8];
}

/* flags CHUNK_START|CHUNK_END|ROOT: a whole message that fits one block */
#define B3_FLAGS_SINGLE 0x0Bu

void or_blake3_small(const uint8_t *src, size_t len, uint8_t dst[32]) {
    uint8_t blk[64];
    uint32_t m[16], out[8];
    memset(blk, 0, sizeof(blk));
    if (len > 64) len = 64;
    memcpy(blk, src, len);
    for (int i = 0; i < 16; ++i)
        m[i] = (uint32_t)blk[4 * i] | ((uint32_t)blk[4 * i + 1] << 8) |
               ((uint32_t)blk[4 * i + 2] << 16) | ((uint32_t)blk[4 * i + 3] << 24);
    b3_compress(m, (uint32_t)len, B3_FLAGS_SINGLE, out);
    for (int i = 0; i < 8; ++i) {
        dst[4 * i] = (uint8_t)out[i];
        dst[4 * i + 1] = (uint8_t)(out[i] >> 8);
        dst[4 * i + 2] = (uint8_t)(out[i] >> 16);
        dst[4 * i + 3] = (uint8_t)(out[i] >> 24);
    }
}

void or_blake3_hash64(const uint8_t src[64], uint8_t dst[32]) {   /* Blake3_hash.cpp:5-9 */
    or_blake3_small(src, 64, dst);
}

void or_blake3_hash64_many(const uint8_t *src, uint8_t *dst, uint64_t n) {
    for (uint64_t i = 0; i < n; ++i) or_blake3_hash64(src + 64 * i, dst + 32 * i);
}

/* ------------------------------------------------------------------ */
/* Fiat-Shamir: src/Fiat_Shamir.cpp                                     */
/* ------------------------------------------------------------------ */

or_fe or_hash_F(or_fe x, or_fe y) {         /* Fiat_Shamir.cpp:8-27 */
    /* 64-byte block: {x.real, x.img, 0, 0, y.real, y.img, 0, 0} as LE u64 */
    uint64_t w[8] = { x.real, x.img, 0, 0, y.real, y.img, 0, 0 };
    uint8_t data[64], h[32];
    memcpy(data, w, 64);                    /* x86 is little-endian */
    or_blake3_hash64(data, h);
    /* digest bytes 0..7 -> real, 8..15 -> img, with a 7-bit stride */
    uint64_t real = 0, img = 0;
    for (int i = 0; i < 8; ++i) {
        real += (1ULL << (7 * i)) * h[i];
        img += (1ULL << (7 * i)) * h[i + 8];
    }
    or_fe r = { real, img };
    return r;
}

or_fe or_hash_to_field(or_fe *prev, const or_fe *elems, uint64_t n) {  /* :29-34 */
    for (uint64_t i = 0; i < n; ++i) *prev = or_hash_F(*prev, elems[i]);
    return *prev;
}

void or_get_indexes(or_fe *prev, int l, uint32_t N, uint32_t M, uint32_t *idx) {  /* :37-56 */
    for (int i = 0; i < l; ++i) {
        *prev = or_hash_F(*prev, *prev);
        idx[2 * i] = (uint32_t)(prev->real % N);
        idx[2 * i + 1] = (uint32_t)(prev->img % M);
    }
}

/* ------------------------------------------------------------------ */
/* NTT: src/utils.cpp:904-964 (fft), :1583-1673 (pfft)                  */
/* ------------------------------------------------------------------ */

void or_fft(or_fe *arr, int logn, int inverse) {
    uint64_t len = 1ULL << logn;
    if (logn == 0) return;
    /* bit-reversal permutation (utils.cpp:918-919,938-939) */
    for (uint64_t i = 1, j = 0; i < len; ++i) {
        uint64_t bit = len >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) { or_fe t = arr[i]; arr[i] = arr[j]; arr[j] = t; }
    }
    or_fe w1 = or_root_of_unity(logn);
    if (inverse) w1 = or_inv(w1);
    /* twiddle table w[k] = w1^k (utils.cpp:931-937) */
    or_fe *w = (or_fe *)malloc(sizeof(or_fe) * (len > 1 ? len / 2 : 1));
    w[0].real = 1; w[0].img = 0;
    for (uint64_t k = 1; k < len / 2; ++k) w[k] = or_mul(w[k - 1], w1);
    /* butterflies (utils.cpp:943-950) */
    for (uint64_t i = 2; i <= len; i <<= 1) {
        uint64_t h = i >> 1, step = len / i;
        for (uint64_t j = 0; j < len; j += i)
            for (uint64_t k = 0; k < h; ++k) {
                or_fe u = arr[j + k];
                or_fe v = or_mul(arr[j + k + h], w[step * k]);
                arr[j + k] = or_add(u, v);
                arr[j + k + h] = or_sub(u, v);
            }
    }
    free(w);
    if (inverse) {                          /* utils.cpp:952-963 */
        or_fe ilen = or_inv(or_from_u64(len));
        for (uint64_t i = 0; i < len; ++i) arr[i] = or_mul(arr[i], ilen);
    }
}

/* pfft is a four-step evaluation of the same forward DFT (16 row transforms,
 * twiddle, 16-point column transforms; utils.cpp:1583-1673) and falls back to
 * fft for len <= 1024.  Its result equals fft(arr, logn, false). */
void or_pfft(or_fe *arr, int logn) { or_fft(arr, logn, 0); }

void or_rs_encode(const or_fe *arr, uint64_t n, int rate, or_fe *cw) {  /* Whir.cpp:158-163 */
    uint64_t m = n * (uint64_t)rate;
    memcpy(cw, arr, n * sizeof(or_fe));
    memset(cw + n, 0, (m - n) * sizeof(or_fe));
    int logm = 0;
    while ((1ULL << logm) < m) ++logm;
    or_pfft(cw, logm);
}

/* ------------------------------------------------------------------ */
/* Merkle tree: src/merkle_tree.cpp                                     */
/* ------------------------------------------------------------------ */

uint64_t or_merkle_node_count(uint64_t n_elems) {
    uint64_t leaves = n_elems / 4, total = 0;
    for (uint64_t s = leaves; s >= 1; s /= 2) total += s;
    return total;
}

void or_merkle_commit(const or_fe *elems, uint64_t n_elems, uint8_t *nodes, int dup_left) {
    uint64_t leaves = n_elems / 4;
    /* leaf i = BLAKE3(F[4i] || F[4i+1] || F[4i+2] || F[4i+3]), raw 64 bytes
     * of {real,img} pairs (merkle_tree.cpp:211-220) */
    for (uint64_t i = 0; i < leaves; ++i)
        or_blake3_hash64((const uint8_t *)(elems + 4 * i), nodes + 32 * i);
    /* levels (create_tree_blake, merkle_tree.cpp:267-299) */
    uint8_t *prev = nodes, *cur = nodes + 32 * leaves;
    for (uint64_t s = leaves / 2; s >= 1; s /= 2) {
        for (uint64_t i = 0; i < s; ++i) {
            uint8_t data[64];
            memcpy(data, prev + 32 * (2 * i), 32);
            /* the reference copies the left child into both halves (:287-292) */
            memcpy(data + 32, prev + 32 * (dup_left ? 2 * i : 2 * i + 1), 32);
            or_blake3_hash64(data, cur + 32 * i);
        }
        prev = cur;
        cur += 32 * s;
    }
}

int or_merkle_open(const uint8_t *nodes, uint64_t n_elems, uint64_t pos, uint8_t *path) {
    /* open_tree_blake (merkle_tree.cpp:320-335): per level below the root,
     * the sibling 2*(pos/2) + (1 - pos%2), then pos /= 2 */
    uint64_t s = n_elems / 4, off = 0;
    int lvl = 0;
    if (pos >= s) return -1;
    while (s > 1) {
        uint64_t sib = 2 * (pos / 2) + (1 - (pos % 2));
        memcpy(path + 32 * lvl, nodes + 32 * (off + sib), 32);
        off += s;
        s /= 2;
        pos /= 2;
        ++lvl;
    }
    return lvl;
}

/* ------------------------------------------------------------------ */
/* multilinear helpers: src/utils.cpp                                   */
/* ------------------------------------------------------------------ */

void or_precompute_beta(const or_fe *r, int m, or_fe *B) {   /* utils.cpp:198-243 */
    /* B[x] = prod_k (bit_k(x) ? r[k] : 1 - r[k]); bit 0 pairs with r[0]:
     * the table doubles once per coordinate, consuming r[m-1] first */
    B[0].real = 1; B[0].img = 0;
    for (int i = 0; i < m; ++i) {
        uint64_t sz = 1ULL << i;
        or_fe ri = r[m - 1 - i];
        for (uint64_t j = sz; j-- > 0;) {
            or_fe t = or_mul(ri, B[j]);
            B[2 * j] = or_sub(B[j], t);
            B[2 * j + 1] = t;
        }
    }
}

or_fe or_evaluate_vector(const or_fe *v, uint64_t n, const or_fe *r) {  /* utils.cpp:1201-1214 */
    or_fe *t = (or_fe *)malloc(sizeof(or_fe) * n);
    memcpy(t, v, sizeof(or_fe) * n);
    int m = 0;
    while ((1ULL << m) < n) ++m;
    or_fe one = { 1, 0 };
    for (int i = 0; i < m; ++i) {
        uint64_t L = 1ULL << (m - 1 - i);
        for (uint64_t j = 0; j < L; ++j)
            t[j] = or_add(or_mul(or_sub(one, r[i]), t[2 * j]), or_mul(r[i], t[2 * j + 1]));
    }
    or_fe res = t[0];
    free(t);
    return res;
}

void or_compute_lagrange_coeff(or_fe r, int degree, or_fe *L) {   /* utils.cpp:169-197 */
    /* inverse DFT of (1, r, r^2, ..., r^{d-1}) */
    L[0].real = 1; L[0].img = 0;
    for (int i = 1; i < degree; ++i) L[i] = or_mul(L[i - 1], r);
    int logd = 0;
    while ((1 << logd) < degree) ++logd;
    or_fft(L, logd, 1);
}

/* ------------------------------------------------------------------ */
/* glibc challenge stream: src/fieldElement.cpp:119-125, 363-368        */
/* ------------------------------------------------------------------ */

static uint64_t random_number(void) {
    uint64_t ret = (uint64_t)(random() % 10);
    for (int i = 1; i < 20; ++i) ret = (ret * 10ULL + (uint64_t)(random() % 10)) % P61;
    return ret;
}

or_fe or_random_element(void) {
    or_fe r;
    r.real = random_number() % P61;
    r.img = random_number() % P61;
    return r;
}

/* ------------------------------------------------------------------ */
/* product sumcheck: src/Distributed_Sumcheck.cpp:204-348 at N = 1     */
/* ------------------------------------------------------------------ */

/* one round's polynomial: sum over pairs of (v[2j+1]-v[2j]) x + v[2j]
 * multiplied for the two vectors (Distributed_Sumcheck.cpp:224-233,
 * polynomial.cpp:133-135: (a x + b)(c x + d) = ac x^2 + (ad + bc) x + bd) */
static void round_poly(const or_fe *v1, const or_fe *v2, uint64_t L, or_fe out[3]) {
    or_fe a = { 0, 0 }, b = { 0, 0 }, c = { 0, 0 };
    for (uint64_t j = 0; j < L; ++j) {
        or_fe d1 = or_sub(v1[2 * j + 1], v1[2 * j]);
        or_fe d2 = or_sub(v2[2 * j + 1], v2[2 * j]);
        a = or_add(a, or_mul(d1, d2));
        b = or_add(b, or_add(or_mul(d1, v2[2 * j]), or_mul(v1[2 * j], d2)));
        c = or_add(c, or_mul(v1[2 * j], v2[2 * j]));
    }
    out[0] = a; out[1] = b; out[2] = c;
}

static or_fe quad_eval(const or_fe p[3], or_fe x) {   /* polynomial.cpp:111-113 */
    return or_add(or_mul(or_add(or_mul(p[0], x), p[1]), x), p[2]);
}

static void fold(or_fe *v, uint64_t L, or_fe r) {     /* :257-261 */
    for (uint64_t j = 0; j < L; ++j)
        v[j] = or_add(or_mul(r, or_sub(v[2 * j + 1], v[2 * j])), v[2 * j]);
}

int or_product_sumcheck(or_fe y, or_fe *v1, or_fe *v2, uint64_t n, or_fe *prev,
                        int offset, or_fe *polys, or_fe *r, or_fe *final_out) {
    int m = 0;
    while ((1ULL << m) < n) ++m;
    int fs_rounds = m - offset;
    if (fs_rounds < 0) fs_rounds = 0;
    or_fe zero = { 0, 0 }, one = { 1, 0 };
    for (int i = 0; i < m; ++i) {
        uint64_t L = 1ULL << (m - 1 - i);
        or_fe p[3];
        round_poly(v1, v2, L, p);
        /* consistency check (:243-247; sparse_eval.cpp:160-163) */
        if (!(or_add(quad_eval(p, zero), quad_eval(p, one)).real == y.real &&
              or_add(quad_eval(p, zero), quad_eval(p, one)).img == y.img))
            return -(i + 1);
        or_fe ch;
        if (i < fs_rounds) {
            ch = or_hash_to_field(prev, p, 3);       /* :248 */
        } else {
            /* tail rounds: four F::_random() draws, the last is kept
             * (sparse_eval.cpp:164-168) */
            ch = or_random_element();
            ch = or_random_element();
            ch = or_random_element();
            ch = or_random_element();
        }
        y = quad_eval(p, ch);
        polys[3 * i] = p[0]; polys[3 * i + 1] = p[1]; polys[3 * i + 2] = p[2];
        r[i] = ch;
        fold(v1, L, ch);
        fold(v2, L, ch);
    }
    final_out[0] = v1[0];
    final_out[1] = v2[0];
    return 0;
}

/* ------------------------------------------------------------------ */
/* synthetic inputs                                                     */
/* ------------------------------------------------------------------ */

static uint64_t splitmix64(uint64_t *s) {
    uint64_t z = (*s += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

static inline uint64_t rotl64(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

static uint64_t xoshiro_next(uint64_t s[4]) {
    uint64_t result = rotl64(s[1] * 5, 7) * 9;
    uint64_t t = s[1] << 17;
    s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl64(s[3], 45);
    return result;
}

void or_synth_fill(uint64_t seed, uint64_t stream, or_fe *out, uint64_t n) {
    uint64_t sm = seed ^ (stream * 0xD1B54A32D192ED03ULL), s[4];
    for (int i = 0; i < 4; ++i) s[i] = splitmix64(&sm);
    for (uint64_t i = 0; i < n; ++i) {
        uint64_t x;
        do { x = xoshiro_next(s) & P61; } while (x == P61);
        out[i].real = x;
        do { x = xoshiro_next(s) & P61; } while (x == P61);
        out[i].img = x;
    }
}

/* ------------------------------------------------------------------ */
/* WHIR: src/Whir.cpp, whir_k = 4                                       */
/* ------------------------------------------------------------------ */

#define WHIR_K 4

int or_whir_query_count(int rate) {         /* Whir.cpp:253-258 (conjectured_security false) */
    return (int)(-100.0 / log2(sqrt(1.0 / (double)rate)));
}

static int ilog2u(uint64_t x) { int l = 0; while ((1ULL << l) < x) ++l; return l; }

or_fe or_rs_multifold_segment_tensor(const or_fe *seg, int seg_n, const or_fe *r, int nr,
                                     uint64_t n, uint64_t idx) {   /* coPCS_utils.cpp:333-383 */
    or_fe one = { 1, 0 };
    or_fe len_inv = or_inv(or_from_u64((uint64_t)seg_n));
    int fold_degree = seg_n, log_fold = ilog2u((uint64_t)fold_degree);
    int logn = ilog2u(n);
    or_fe omega = or_inv(or_root_of_unity(logn));
    omega = or_pow(omega, n / (uint64_t)seg_n);
    or_fe inv_omegas = or_pow(or_inv(or_root_of_unity(logn)), idx);
    or_fe part[64];
    for (int j = 0; j < seg_n / 2; ++j) {
        or_fe s = or_add(seg[j], seg[j + seg_n / 2]);
        or_fe d = or_sub(seg[j], seg[j + seg_n / 2]);
        part[j] = or_add(or_mul(or_sub(one, r[0]), s), or_mul(or_mul(r[0], inv_omegas), d));
        inv_omegas = or_mul(inv_omegas, omega);
    }
    if (nr > 1) {
        or_fe pw[16];
        for (int i = 0; i < log_fold; ++i) pw[i] = one;
        or_fe om = or_inv(or_root_of_unity(logn));
        if (log_fold > 1) pw[1] = or_mul(om, om);
        for (int i = 2; i < log_fold; ++i) pw[i] = or_mul(pw[i - 1], pw[i - 1]);
        for (int i = 1; i < log_fold; ++i) pw[i] = or_pow(pw[i], n / (uint64_t)seg_n);
        int sn = seg_n;
        for (int i = 1; i < log_fold; ++i) {
            sn /= 2;
            or_fe om_i = or_inv(or_root_of_unity(ilog2u(n >> i)));
            or_fe pow_omega = or_pow(om_i, idx);
            for (int j = 0; j < sn / 2; ++j) {
                or_fe s = or_add(part[j], part[j + sn / 2]);
                or_fe d = or_sub(part[j], part[j + sn / 2]);
                part[j] = or_add(or_mul(or_sub(one, r[i]), s), or_mul(r[i], or_mul(pow_omega, d)));
                pow_omega = or_mul(pow_omega, pw[i]);
            }
        }
    }
    return or_mul(len_inv, part[0]);
}

/* record writer */
typedef struct { uint64_t *p; uint64_t n, cap; int overflow; } rec_t;
static void rec_u64(rec_t *R, uint64_t x) { if (R->n < R->cap) R->p[R->n] = x; else R->overflow = 1; R->n++; }
static void rec_fe(rec_t *R, or_fe x) { rec_u64(R, x.real); rec_u64(R, x.img); }
static void rec_bytes32(rec_t *R, const uint8_t *h) { uint64_t w[4]; memcpy(w, h, 32); for (int i = 0; i < 4; ++i) rec_u64(R, w[i]); }

/*
 * WHIR opening record (u64 words):
 *   [n_iter]
 *   per iteration:
 *     polys  2 x (a,b,c)           12 words
 *     r      2 challenges           4 words
 *     root   new commitment root    4 words
 *     l      query count            1 word
 *     q[l]   query positions as drawn (before scaling)   l words
 *     replies l x 4 elements       8l words
 *     y      l folded values       2l words
 *     paths  l x depth digests     4*l*depth words (depth = log2(|cw_old|/4))
 *     rb     update_betas challenge 2 words
 *   final: arr (4 elements, 8 words), K (2 words), final root (4 words),
 *          final r (log2|arr| elements)
 * With a general fold factor kappa an iteration holds log2(kappa) polys and
 * challenges and kappa replies per query (or_whir_batch_open's record).
 */

/* A WHIR prover's running state: arr and its weights b (n entries), the
 * codeword (cw_n entries) and the tree it is opened against: the dup-left
 * nodes mt, or, after aggregate_commitments, a tree over all-zero leaf
 * digests (mt NULL) with mt_leaves leaves. */
typedef struct {
    or_fe *arr, *b;
    uint64_t n;
    or_fe *cw;
    uint64_t cw_n;
    uint8_t *mt;
    uint64_t mt_leaves;
    or_fe K;
    int rate;
    or_fe *chal;        /* when set: every challenge of the opening, in order (_Whir_Open's total_r) */
    int nchal;
} whir_state;

/* node of level l of a dup-left tree over zero digests: the l-fold
 * self-hash of the zero digest (levels 0..levels) */
static void zero_chain(int levels, uint8_t *out) {
    uint8_t d[64];
    memset(out, 0, 32);
    for (int l = 1; l <= levels; ++l) {
        memcpy(d, out + 32 * (l - 1), 32);
        memcpy(d + 32, out + 32 * (l - 1), 32);
        or_blake3_hash64(d, out + 32 * l);
    }
}

/* open_tree_blake (merkle_tree.cpp:320-335) on the state's tree; -1 where
 * the reference exits (position beyond the leaves, :321-330) */
static int state_path(const whir_state *S, uint64_t pos, uint8_t *path) {
    if (pos >= S->mt_leaves) return -1;
    if (S->mt) return or_merkle_open(S->mt, S->cw_n, pos, path);
    int d = ilog2u(S->mt_leaves);
    uint8_t *z = (uint8_t *)malloc(32 * (d + 1));
    zero_chain(d, z);
    memcpy(path, z, 32 * (size_t)d);
    free(z);
    return d;
}

static int fe_eq(or_fe a, or_fe b) { return a.real == b.real && a.img == b.img; }

/* Thread count of the WHIR restatement's heavy steps: 1 runs the
 * single-thread functions of this file; more runs their oracle_mt.c twins,
 * which give the same bytes (tests/test_oracle_mt.py), so a full-size
 * opening can be checked in seconds. */
static int whir_threads = 1;
void or_whir_set_threads(int t) { whir_threads = t < 1 ? 1 : t; }

static void w_encode(const or_fe *arr, uint64_t n, int rate, or_fe *cw) {
    if (whir_threads > 1) or_rs_encode_mt(arr, n, rate, cw, whir_threads);
    else or_rs_encode(arr, n, rate, cw);
}

static void w_merkle(const or_fe *elems, uint64_t n_elems, uint8_t *nodes) {
    if (whir_threads > 1) or_merkle_commit_mt(elems, n_elems, nodes, 1, whir_threads);
    else or_merkle_commit(elems, n_elems, nodes, 1);
}

/* Whir_Iteration (Whir.cpp:354-386) with whir_k = kappa; 0, or a negative
 * status where the reference reports an error */
static int whir_iteration(whir_state *S, int kappa, rec_t *R, int iter) {
    or_fe zero = { 0, 0 }, one = { 1, 0 };
    int k = ilog2u((uint64_t)kappa);
    or_fe rr[32];
    for (int rd = 0; rd < k; ++rd) {                       /* whir_sumcheck :233-249 */
        uint64_t L = S->n / 2;
        or_fe p[3];
        round_poly(S->arr, S->b, L, p);
        if (!fe_eq(or_add(quad_eval(p, zero), quad_eval(p, one)), S->K)) return -10 - iter;
        rr[rd] = or_random_element();
        if (S->chal) S->chal[S->nchal++] = rr[rd];
        S->K = quad_eval(p, rr[rd]);
        fold(S->arr, L, rr[rd]);
        fold(S->b, L, rr[rd]);
        S->n = L;
        rec_fe(R, p[0]); rec_fe(R, p[1]); rec_fe(R, p[2]);
    }
    for (int rd = 0; rd < k; ++rd) rec_fe(R, rr[rd]);

    /* Whir_Commit_initial at 2*rate (:158-190) */
    uint64_t ncw_n = S->n * (uint64_t)(2 * S->rate);
    or_fe *ncw = (or_fe *)malloc(sizeof(or_fe) * ncw_n);
    w_encode(S->arr, S->n, 2 * S->rate, ncw);
    uint64_t nnodes = or_merkle_node_count(ncw_n);
    uint8_t *nmt = (uint8_t *)malloc(32 * nnodes);
    w_merkle(ncw, ncw_n, nmt);
    rec_bytes32(R, nmt + 32 * (nnodes - 1));

    /* whir_query (:251-314) on the current codeword */
    int l = or_whir_query_count(S->rate);
    rec_u64(R, (uint64_t)l);
    uint64_t *q = (uint64_t *)malloc(sizeof(uint64_t) * l);
    uint64_t qmod = kappa > 4 ? S->cw_n / kappa : S->cw_n / 4;   /* :262-267 */
    for (int i = 0; i < l; ++i) {
        q[i] = (uint64_t)(rand() % (int)qmod);
        rec_u64(R, q[i]);
    }
    or_fe *y = (or_fe *)malloc(sizeof(or_fe) * l);
    or_fe *rep = (or_fe *)malloc(sizeof(or_fe) * kappa);
    for (int i = 0; i < l; ++i) {
        for (int j = 0; j < kappa; ++j) {
            rep[j] = S->cw[q[i] + (uint64_t)j * S->cw_n / kappa];
            rec_fe(R, rep[j]);
        }
        y[i] = or_rs_multifold_segment_tensor(rep, kappa, rr, k, S->cw_n, q[i]);
    }
    free(rep);
    for (int i = 0; i < l; ++i) rec_fe(R, y[i]);
    int status = 0;
    {
        uint8_t path[32 * 64];
        for (int i = 0; i < l && !status; ++i) {
            int d = state_path(S, q[i], path);
            if (d < 0) { status = -20 - iter; break; }
            for (int t = 0; t < d; ++t) rec_bytes32(R, path + 32 * t);
        }
    }
    if (!status) {
        uint64_t scale = (uint64_t)kappa * ncw_n / S->cw_n;      /* :373-374 */
        for (int i = 0; i < l; ++i) q[i] *= scale;
        or_fe rb = or_random_element();                          /* update_betas :325-342 */
        rec_fe(R, rb);
        or_fe ch = rb;
        or_fe omega = or_root_of_unity(ilog2u(ncw_n));
        if (whir_threads > 1) {
            /* b[j] += sum_i rb^(i+1) (w^q_i)^j on all cores */
            or_fe *chs = (or_fe *)malloc(sizeof(or_fe) * l), *ws = (or_fe *)malloc(sizeof(or_fe) * l);
            for (int i = 0; i < l; ++i) {
                if (i > 0) ch = or_mul(ch, rb);
                chs[i] = ch;
                ws[i] = or_pow(omega, q[i]);
                S->K = or_add(S->K, or_mul(ch, y[i]));           /* :383-384 */
            }
            or_update_betas_mt(S->b, S->n, chs, ws, l, whir_threads);
            free(chs); free(ws);
        } else {
            for (int i = 0; i < l; ++i) {
                if (i > 0) ch = or_mul(ch, rb);
                or_fe w = or_pow(omega, q[i]), pw = one;
                for (uint64_t j = 0; j < S->n; ++j) {
                    S->b[j] = or_add(S->b[j], or_mul(ch, pw));
                    pw = or_mul(pw, w);
                }
                S->K = or_add(S->K, or_mul(ch, y[i]));           /* :383-384 */
            }
        }
    }
    free(q); free(y);
    free(S->cw); free(S->mt);
    S->cw = ncw; S->mt = nmt; S->cw_n = ncw_n; S->mt_leaves = ncw_n / 4;
    return status;
}

/* The closing steps of Whir_Open / Whir_Batch_Open (Whir.cpp:425-449,
 * :844-863) with whir_k = kappa: arr, K, the root of Whir_Commit (:192-213,
 * a tree over |arr|*rate/kappa all-zero leaf digests), and the final
 * challenges; verification steps 1 (the codeword is arr's encoding) and 2
 * (K = <arr, b>) */
static int whir_final(whir_state *S, int kappa, rec_t *R) {
    for (uint64_t i = 0; i < S->n; ++i) rec_fe(R, S->arr[i]);
    rec_fe(R, S->K);
    {
        uint64_t leaves = S->n * (uint64_t)S->rate / kappa;
        int d = ilog2u(leaves);
        uint8_t *z = (uint8_t *)malloc(32 * (d + 1));
        zero_chain(d, z);
        rec_bytes32(R, z + 32 * d);
        free(z);
    }
    {
        uint64_t ncw_n = S->n * (uint64_t)S->rate;
        or_fe *ncw = (or_fe *)malloc(sizeof(or_fe) * ncw_n);
        w_encode(S->arr, S->n, S->rate, ncw);
        int same = ncw_n == S->cw_n && !memcmp(ncw, S->cw, sizeof(or_fe) * ncw_n);
        free(ncw);
        if (!same) return -3;
    }
    or_fe Kc = S->K;
    for (uint64_t i = 0; i < S->n; ++i) Kc = or_sub(Kc, or_mul(S->arr[i], S->b[i]));
    if (Kc.real != 0 || Kc.img != 0) return -2;
    for (int i = 0; i < ilog2u(S->n); ++i) {
        const or_fe r = or_random_element();
        if (S->chal) S->chal[S->nchal++] = r;
        rec_fe(R, r);
    }
    return 0;
}

static void state_free(whir_state *S) { free(S->arr); free(S->b); free(S->cw); free(S->mt); }

int64_t or_whir_open(const or_fe *arr_in, uint64_t n, const or_fe *z, or_fe K,
                     int rate, int stop_log, uint64_t *record, uint64_t record_cap) {
    rec_t R = { record, 0, record_cap, 0 };
    uint64_t stop = WHIR_K;
    if (stop_log > 0 && (1ULL << stop_log) > stop) stop = 1ULL << stop_log;
    whir_state S;
    memset(&S, 0, sizeof(S));
    S.n = n;
    S.arr = (or_fe *)malloc(sizeof(or_fe) * n);
    S.b = (or_fe *)malloc(sizeof(or_fe) * n);
    memcpy(S.arr, arr_in, sizeof(or_fe) * n);
    or_precompute_beta(z, ilog2u(n), S.b);                 /* Whir.cpp:395 */
    S.K = K;
    S.rate = rate;
    /* the initial commitment (Whir_Commit_initial with the caller's rate) */
    S.cw_n = n * (uint64_t)rate;
    S.cw = (or_fe *)malloc(sizeof(or_fe) * S.cw_n);
    w_encode(S.arr, n, rate, S.cw);
    S.mt = (uint8_t *)malloc(32 * or_merkle_node_count(S.cw_n));
    w_merkle(S.cw, S.cw_n, S.mt);
    S.mt_leaves = S.cw_n / 4;

    uint64_t iters_pos = R.n;
    rec_u64(&R, 0);
    uint64_t iters = 0;
    int status = 0;
    while (1) {                                            /* :413-422 */
        status = whir_iteration(&S, WHIR_K, &R, (int)iters);
        if (status) break;
        ++iters;
        S.rate *= 2;
        if (S.n <= stop) break;
    }
    if (iters_pos < record_cap) record[iters_pos] = iters;
    if (!status) status = whir_final(&S, WHIR_K, &R);
    state_free(&S);
    if (status) return status;
    if (R.overflow) return -1;
    return (int64_t)R.n;
}

void or_whir_rates(const uint64_t *sizes, int count, int rate, int whir_k, int *rates) {
    /* get_whir_rates (Whir.cpp:18-137) for sizes already sorted largest
     * first: the rate each array is committed at */
    int w = 0, i = 1;
    uint64_t size = sizes[0];
    rates[w++] = rate;
    while (i < count && sizes[i] == size) { rates[w++] = rate; ++i; }      /* :43-51 */
    while (i < count) {                                                    /* :53-73 */
        if (size / (uint64_t)whir_k > sizes[i]) {
            size /= (uint64_t)whir_k;
            rate *= 2;
        } else {
            size = sizes[i];
            rate *= 2;
            rates[w++] = rate;
            uint64_t s2 = sizes[i++];
            while (i < count && sizes[i] == s2) { rates[w++] = rate; ++i; }
        }
    }
}

int or_whir_kappas(const uint64_t *sizes, int count, int whir_k, int *kappas, int *aggr, int cap) {
    /* the fold schedule of Whir_Batch_Open (Whir.cpp:736-755) */
    int nk = 0, i = 1;
    uint64_t size = sizes[0];
    for (int t = 1; t < count; ++t)
        if (sizes[t] > sizes[t - 1] || sizes[t] == 0) return -2;      /* not largest first */
    while (1) {
        if (nk >= cap) return -1;
        if (i >= count || size / (uint64_t)whir_k > sizes[i]) {
            kappas[nk] = whir_k;
            size /= (uint64_t)whir_k;
            aggr[nk++] = 0;
        } else {
            kappas[nk] = (int)(size / sizes[i]);
            uint64_t s2 = sizes[i];
            while (i < count && sizes[i] == s2) ++i;
            size /= (uint64_t)kappas[nk];
            aggr[nk++] = 1;
        }
        if (size <= (uint64_t)whir_k) break;
    }
    return nk;
}

/*
 * Whir_Batch_Open record (u64 words):
 *   [n_steps]
 *   per step [kind]:
 *     0 iteration:   kappa, then the iteration as above with log2(kappa)
 *                    polys and challenges and kappa replies per query
 *     1 aggregation: cnt (running array + cnt-1 arrays of its size),
 *                    challenge1[cnt], challenge2[cnt] (entry 0 set to 1),
 *                    evals[cnt x cnt] (<arr_i, b_j>), K, l, q[l], l paths
 *                    on the running tree, root of the new (zero-digest) tree
 *   final as in Whir_Open, with whir_k as it stands at the end.
 */
/* Whir_Batch_Open's loop (Whir.cpp:762-899) from the running state S:
 * iterations, aggregations, the closing steps */
static int64_t batch_open_from(whir_state *S_, int count, const or_fe *const *arrays, const uint64_t *sizes,
                               const or_fe *const *codewords, const uint64_t *cw_sizes, const or_fe *const *betas,
                               int whir_k, uint64_t *record, uint64_t record_cap);

int64_t or_whir_batch_open(int count, const or_fe *const *arrays, const uint64_t *sizes,
                           const or_fe *const *codewords, const uint64_t *cw_sizes, const uint8_t *const *trees,
                           const or_fe *evals, const or_fe *const *betas, int rate, int whir_k,
                           uint64_t *record, uint64_t record_cap) {
    whir_state S;
    memset(&S, 0, sizeof(S));
    S.n = sizes[0];
    S.arr = (or_fe *)malloc(sizeof(or_fe) * S.n);
    S.b = (or_fe *)malloc(sizeof(or_fe) * S.n);
    memcpy(S.arr, arrays[0], sizeof(or_fe) * S.n);
    memcpy(S.b, betas[0], sizeof(or_fe) * S.n);
    S.cw_n = cw_sizes[0];
    S.cw = (or_fe *)malloc(sizeof(or_fe) * S.cw_n);
    memcpy(S.cw, codewords[0], sizeof(or_fe) * S.cw_n);
    S.mt = (uint8_t *)malloc(32 * or_merkle_node_count(S.cw_n));
    memcpy(S.mt, trees[0], 32 * or_merkle_node_count(S.cw_n));
    S.mt_leaves = S.cw_n / 4;
    S.K = evals[0];                                        /* :714 */
    S.rate = rate;
    return batch_open_from(&S, count, arrays, sizes, codewords, cw_sizes, betas, whir_k, record, record_cap);
}

/* _Whir_Open (Whir.cpp:469-545): the iterations of one array arr with its
 * weights b and claim K down to |arr| <= whir_k, then the closing steps.
 * The codeword is opened against its dup-left tree (nodes), or, when nodes
 * is NULL, against the tree over zero_leaves all-zero leaf digests that
 * aggregate_commitments leaves (:590-605).  chal receives total_r: every
 * challenge in the order drawn (log2 n of them); final_claim the MLE of the
 * initial b at total_r (:517-518).  Returns the record's word count (the
 * layout of or_whir_batch_open's record) or a negative status. */
int64_t or_whir_open_weights(const or_fe *arr, uint64_t n, const or_fe *b, or_fe K, const or_fe *cw, uint64_t cw_n,
                             const uint8_t *nodes, uint64_t zero_leaves, int rate, int whir_k, uint64_t *record,
                             uint64_t record_cap, or_fe *chal, or_fe *final_claim) {
    whir_state S;
    memset(&S, 0, sizeof(S));
    S.n = n;
    S.arr = (or_fe *)malloc(sizeof(or_fe) * n);
    S.b = (or_fe *)malloc(sizeof(or_fe) * n);
    memcpy(S.arr, arr, sizeof(or_fe) * n);
    memcpy(S.b, b, sizeof(or_fe) * n);
    S.cw_n = cw_n;
    S.cw = (or_fe *)malloc(sizeof(or_fe) * cw_n);
    memcpy(S.cw, cw, sizeof(or_fe) * cw_n);
    if (nodes) {
        S.mt = (uint8_t *)malloc(32 * or_merkle_node_count(cw_n));
        memcpy(S.mt, nodes, 32 * or_merkle_node_count(cw_n));
        S.mt_leaves = cw_n / 4;
    } else {
        S.mt = NULL;
        S.mt_leaves = zero_leaves;
    }
    S.K = K;
    S.rate = rate;
    or_fe *own = chal ? NULL : (or_fe *)malloc(sizeof(or_fe) * 64);
    S.chal = chal ? chal : own;
    S.nchal = 0;
    const uint64_t sizes[1] = { n };
    const int64_t w = batch_open_from(&S, 1, &arr, sizes, &cw, &cw_n, &b, whir_k, record, record_cap);
    if (w >= 0 && final_claim) *final_claim = or_evaluate_vector(b, n, S.chal);
    free(own);
    return w;
}

static int64_t batch_open_from(whir_state *S_, int count, const or_fe *const *arrays, const uint64_t *sizes,
                               const or_fe *const *codewords, const uint64_t *cw_sizes, const or_fe *const *betas,
                               int whir_k, uint64_t *record, uint64_t record_cap) {
    whir_state S = *S_;
    rec_t R = { record, 0, record_cap, 0 };
    int kap[64], agg[64];
    int nk = or_whir_kappas(sizes, count, whir_k, kap, agg, 64);
    if (nk < 0) { state_free(&S); return -4; }

    uint64_t steps_pos = R.n;
    rec_u64(&R, 0);
    uint64_t steps = 0;
    int status = 0, iter = 0, ri = 1, wk = whir_k;
    while (1) {                                            /* :762-843 */
        wk = kap[iter];
        if (!agg[iter] || wk != 1) {
            rec_u64(&R, 0); rec_u64(&R, (uint64_t)wk); ++steps;
            if ((status = whir_iteration(&S, wk, &R, iter))) break;
            S.rate *= 2;
        }
        ++iter;
        if (agg[iter - 1]) {
            /* arrays of the running array's size join it (:786-804) */
            int cnt = 1;
            while (ri + cnt - 1 < count && sizes[ri + cnt - 1] == S.n) ++cnt;
            for (int j = ri + cnt - 1; j < count; ++j)
                if (sizes[j] == S.n) { status = -5; break; }      /* a later equal size: input not sorted */
            if (status) break;
            rec_u64(&R, 1); rec_u64(&R, (uint64_t)cnt); ++steps;
            or_fe *c1 = (or_fe *)malloc(sizeof(or_fe) * cnt), *c2 = (or_fe *)malloc(sizeof(or_fe) * cnt);
            for (int i = 0; i < cnt; ++i) c1[i] = or_random_element();    /* :809-812 */
            for (int i = 0; i < cnt; ++i) c2[i] = or_random_element();
            c1[0].real = 1; c1[0].img = 0; c2[0].real = 1; c2[0].img = 0;
            for (int i = 0; i < cnt; ++i) rec_fe(&R, c1[i]);
            for (int i = 0; i < cnt; ++i) rec_fe(&R, c2[i]);
            /* aggregate_arrays(arr, b, arr_aggr, B_aggr, c2, c1) (:670-703) */
            const or_fe **A = (const or_fe **)malloc(sizeof(or_fe *) * cnt);
            const or_fe **Bv = (const or_fe **)malloc(sizeof(or_fe *) * cnt);
            or_fe *a0 = (or_fe *)malloc(sizeof(or_fe) * S.n), *b0 = (or_fe *)malloc(sizeof(or_fe) * S.n);
            memcpy(a0, S.arr, sizeof(or_fe) * S.n);
            memcpy(b0, S.b, sizeof(or_fe) * S.n);
            A[0] = a0; Bv[0] = b0;
            for (int i = 1; i < cnt; ++i) { A[i] = arrays[ri + i - 1]; Bv[i] = betas[ri + i - 1]; }
            or_fe K = { 0, 0 };
            for (int i = 0; i < cnt; ++i)
                for (int j = 0; j < cnt; ++j) {
                    or_fe e = { 0, 0 };
                    for (uint64_t x = 0; x < S.n; ++x) e = or_add(e, or_mul(A[i][x], Bv[j][x]));
                    rec_fe(&R, e);
                    K = or_add(K, or_mul(or_mul(c2[i], c1[j]), e));
                }
            for (int i = 1; i < cnt; ++i)
                for (uint64_t x = 0; x < S.n; ++x) {
                    S.arr[x] = or_add(S.arr[x], or_mul(c2[i], A[i][x]));
                    S.b[x] = or_add(S.b[x], or_mul(c1[i], Bv[i][x]));
                }
            S.K = K;
            rec_fe(&R, K);
            /* the next fold factor (:826-829); past the end of the schedule
             * the reference reads beyond its vector, here the last entry */
            wk = iter + 1 < nk ? kap[iter + 1] : kap[nk - 1];
            if (ri + cnt - 1 < count && S.n / (uint64_t)wk <= sizes[ri + cnt - 1]) wk = (int)(S.n / sizes[ri + cnt - 1]);
            /* aggregate_commitments(codeword, MT, cw_aggr, MT_aggr, c2, rate') (:547-606) */
            int rate2 = (int)(S.cw_n / S.n);
            for (int i = 1; i < cnt && !status; ++i)
                if (cw_sizes[ri + i - 1] != S.cw_n) status = -6;
            if (!status) {
                for (int i = 1; i < cnt; ++i) {
                    const or_fe *ci = codewords[ri + i - 1];
                    for (uint64_t x = 0; x < S.cw_n; ++x) S.cw[x] = or_add(S.cw[x], or_mul(c2[i], ci[x]));
                }
                int l = or_whir_query_count(rate2);
                rec_u64(&R, (uint64_t)l);
                uint64_t *q = (uint64_t *)malloc(sizeof(uint64_t) * l);
                for (int i = 0; i < l; ++i) { q[i] = (uint64_t)(rand() % (int)(S.cw_n / 4)); rec_u64(&R, q[i]); }
                uint8_t path[32 * 64];
                for (int i = 0; i < l && !status; ++i) {
                    int d = state_path(&S, q[i], path);
                    if (d < 0) { status = -7; break; }
                    for (int t = 0; t < d; ++t) rec_bytes32(&R, path + 32 * t);
                }
                free(q);
                free(S.mt);
                S.mt = NULL;
                S.mt_leaves = S.cw_n / (uint64_t)wk;
                int d = ilog2u(S.mt_leaves);
                uint8_t *z = (uint8_t *)malloc(32 * (d + 1));
                zero_chain(d, z);
                rec_bytes32(&R, z + 32 * d);
                free(z);
            }
            ri += cnt - 1;
            free(c1); free(c2); free(A); free(Bv); free(a0); free(b0);
            if (status) break;
        }
        if (S.n <= (uint64_t)wk) break;
        if (iter >= nk) { status = -8; break; }
    }
    if (steps_pos < record_cap) record[steps_pos] = steps;
    if (!status) status = whir_final(&S, wk, &R);
    state_free(&S);
    if (status) return status;
    if (R.overflow) return -1;
    return (int64_t)R.n;
}

void or_whir_aggregate_claims(int narrays, const or_fe *const *arrays, uint64_t n, int npoints, const or_fe *points,
                              or_fe *arr, or_fe *b_aggr, or_fe *K, or_fe *beta_ch, or_fe *arr_ch) {
    /* aggregate_claims (Whir.cpp:608-668) with arr and b_aggr empty on entry */
    int m = ilog2u(n);
    or_fe **betas = (or_fe **)malloc(sizeof(or_fe *) * npoints);
    for (int j = 0; j < npoints; ++j) {                     /* :615-622 */
        betas[j] = (or_fe *)malloc(sizeof(or_fe) * n);
        or_precompute_beta(points + (size_t)j * m, m, betas[j]);
    }
    or_fe *sums = (or_fe *)malloc(sizeof(or_fe) * narrays * npoints);
    for (int i = 0; i < narrays; ++i)                       /* :625-633 */
        for (int j = 0; j < npoints; ++j) {
            or_fe s = { 0, 0 };
            for (uint64_t k = 0; k < n; ++k) s = or_add(s, or_mul(betas[j][k], arrays[i][k]));
            sums[i * npoints + j] = s;
        }
    for (int j = 0; j < npoints; ++j) beta_ch[j] = or_random_element();   /* :638-644 */
    for (int i = 0; i < narrays; ++i) arr_ch[i] = or_random_element();
    or_fe Ka = { 0, 0 };
    for (int i = 0; i < narrays; ++i)                       /* :645-651 */
        for (int j = 0; j < npoints; ++j)
            Ka = or_add(Ka, or_mul(or_mul(beta_ch[j], arr_ch[i]), sums[i * npoints + j]));
    for (uint64_t k = 0; k < n; ++k) { b_aggr[k].real = b_aggr[k].img = 0; arr[k].real = arr[k].img = 0; }
    for (int j = 0; j < npoints; ++j)                       /* :652-656 */
        for (uint64_t k = 0; k < n; ++k) b_aggr[k] = or_add(b_aggr[k], or_mul(beta_ch[j], betas[j][k]));
    for (int i = 0; i < narrays; ++i)                       /* :657-661 */
        for (uint64_t k = 0; k < n; ++k) arr[k] = or_add(arr[k], or_mul(arr_ch[i], arrays[i][k]));
    *K = Ka;
    for (int j = 0; j < npoints; ++j) free(betas[j]);
    free(betas); free(sums);
}

/* ------------------------------------------------------------------ */
/* coPCS rank-local kernels                                             */
/* ------------------------------------------------------------------ */

void or_copcs_encode(const or_fe *rows, int nrows, uint64_t L, int rate, int log_rate_offset, or_fe *out) {
    int logL = ilog2u(L);
    uint64_t M = L * (uint64_t)rate;
    or_fe *tmp = (or_fe *)malloc(sizeof(or_fe) * M);
    for (int i = 0; i < nrows; ++i) {                     /* coPCS.cpp:132-146 */
        memcpy(tmp, rows + (uint64_t)i * L, sizeof(or_fe) * L);
        or_fft(tmp, logL, 1);
        or_fe omega = or_inv(or_root_of_unity(1 + log_rate_offset + logL));
        or_fe mul = { 1, 0 };
        for (uint64_t j = 0; j < L; ++j) {
            tmp[j] = or_mul(mul, tmp[j]);
            mul = or_mul(mul, omega);
        }
        memset(tmp + L, 0, sizeof(or_fe) * (M - L));
        or_fft(tmp, ilog2u(M), 0);
        memcpy(out + (uint64_t)i * M, tmp, sizeof(or_fe) * M);
    }
    free(tmp);
}

void or_distributed_mt(const or_fe *cw, int nrows, uint64_t Lc, uint8_t *base_nodes) {
    /* column roots (coPCS.cpp:167-177) */
    or_fe col[32];
    uint8_t nodes[32 * 15];
    for (uint64_t j = 0; j < Lc; ++j) {
        for (int k = 0; k < 32; ++k) {
            if (k < nrows) col[k] = cw[(uint64_t)k * Lc + j];
            else { col[k].real = 0; col[k].img = 0; }
        }
        or_merkle_commit(col, 32, nodes, 1);
        memcpy(base_nodes + 32 * j, nodes + 32 * 14, 32);
    }
    /* create_tree_blake over the column roots (dup-left) */
    uint8_t *prev = base_nodes, *cur = base_nodes + 32 * Lc;
    for (uint64_t s = Lc / 2; s >= 1; s /= 2) {
        for (uint64_t i = 0; i < s; ++i) {
            uint8_t data[64];
            memcpy(data, prev + 32 * (2 * i), 32);
            memcpy(data + 32, prev + 32 * (2 * i), 32);
            or_blake3_hash64(data, cur + 32 * i);
        }
        prev = cur;
        cur += 32 * s;
    }
}

void or_pcs_sumcheck_phase1(const or_fe *data, const or_fe *B, uint64_t M, int d, or_fe *poly) {
    int logd = ilog2u((uint64_t)d);
    or_fe *h1 = (or_fe *)malloc(sizeof(or_fe) * 2 * d), *h2 = (or_fe *)malloc(sizeof(or_fe) * 2 * d);
    for (int k = 0; k < 2 * d; ++k) { poly[k].real = 0; poly[k].img = 0; }
    for (uint64_t j = 0; j < M / d; ++j) {                /* coPCS.cpp:553-567 */
        memcpy(h1, data + j * d, sizeof(or_fe) * d);
        memcpy(h2, B + j * d, sizeof(or_fe) * d);
        or_fft(h1, logd, 1);
        or_fft(h2, logd, 1);
        memset(h1 + d, 0, sizeof(or_fe) * d);
        memset(h2 + d, 0, sizeof(or_fe) * d);
        /* coPCS.cpp:561 resizes h2 to 4d (it reads h1's new size), and
         * :562 transforms its first 2d entries: fft's length assert
         * (utils.cpp:912) is compiled out in the Release build (setup.sh:3
         * -> -DNDEBUG), so the products below are of the 2d-point images. */
        or_fft(h1, logd + 1, 0);
        or_fft(h2, logd + 1, 0);
        for (int k = 0; k < 2 * d; ++k) poly[k] = or_add(poly[k], or_mul(h1[k], h2[k]));
    }
    free(h1); free(h2);
}

void or_pcs_sumcheck_phase2(or_fe *data, or_fe *B, uint64_t M, int d, or_fe r) {
    or_fe *L = (or_fe *)malloc(sizeof(or_fe) * d);
    or_compute_lagrange_coeff(r, d, L);                   /* coPCS.cpp:574 */
    for (uint64_t j = 0; j < M / d; ++j) {
        or_fe s1 = { 0, 0 }, s2 = { 0, 0 };
        for (int k = 0; k < d; ++k) {
            s1 = or_add(s1, or_mul(L[k], data[j * d + k]));
            s2 = or_add(s2, or_mul(L[k], B[j * d + k]));
        }
        data[j] = s1;
        B[j] = s2;
    }
    free(L);
}

void or_rs_multifold5(const or_fe *cw, uint64_t n0, const or_fe *a, int fd, or_fe *out) {
    /* partial codewords of growing length, exactly as coPCS_utils.cpp:389-433 */
    uint64_t n = n0;
    int log_fd = ilog2u((uint64_t)fd);
    or_fe len_inv = or_inv(or_from_u64((uint64_t)fd));
    uint64_t np = n0 / 2;
    or_fe **part = (or_fe **)malloc(sizeof(or_fe *) * np);
    for (uint64_t i = 0; i < np; ++i) part[i] = (or_fe *)calloc((size_t)fd, sizeof(or_fe));
    or_fe omega = or_inv(or_root_of_unity(ilog2u(n)));
    or_fe inv_omegas = { 1, 0 };
    for (uint64_t j = 0; j < n0 / 2; ++j) {
        part[j][0] = or_add(cw[j], cw[j + n / 2]);
        part[j][1] = or_mul(inv_omegas, or_sub(cw[j], cw[j + n / 2]));
        inv_omegas = or_mul(inv_omegas, omega);
    }
    if (fd > 2) {
        or_fe pw[32];
        for (int i = 0; i < log_fd; ++i) { pw[i].real = 1; pw[i].img = 0; }
        pw[1] = or_mul(omega, omega);
        for (int i = 2; i < log_fd; ++i) pw[i] = or_mul(pw[i - 1], pw[i - 1]);
        for (int i = 1; i < log_fd; ++i) {
            n = n / 2;
            or_fe pow_omega = { 1, 0 };
            uint64_t half = 1ULL << i;                    /* old length of part[j] */
            for (uint64_t j = 0; j < n0 / (1ULL << (i + 1)); ++j) {
                for (uint64_t k = 0; k < half; ++k) {
                    or_fe v1 = or_add(part[j][k], part[j + n / 2][k]);
                    or_fe v2 = or_mul(pow_omega, or_sub(part[j][k], part[j + n / 2][k]));
                    part[j][k] = v1;
                    part[j][k + half] = v2;
                }
                pow_omega = or_mul(pow_omega, pw[i]);
            }
        }
    }
    for (uint64_t i = 0; i < n0 / (uint64_t)fd; ++i) {
        or_fe acc = { 0, 0 };
        for (int j = 0; j < fd; ++j) acc = or_add(acc, or_mul(a[j], part[i][j]));
        out[i] = or_mul(acc, len_inv);
    }
    for (uint64_t i = 0; i < np; ++i) free(part[i]);
    free(part);
}

void or_encode_concatenated(const or_fe *data, uint64_t n, int rate0, int rate1, int degree, or_fe *cprime, or_fe *C) {
    const uint64_t m = n * (uint64_t)rate0;              /* coPCS.cpp:658-668 */
    or_rs_encode(data, n, rate0, cprime);
    const uint64_t S = m / (uint64_t)degree, D = (uint64_t)degree * (uint64_t)rate1;
    or_fe *seg = (or_fe *)malloc(sizeof(or_fe) * D);
    for (uint64_t j = 0; j < S; ++j) {                   /* :669-676 */
        memset(seg, 0, sizeof(or_fe) * D);
        for (int k = 0; k < degree; ++k) seg[k] = cprime[j + (uint64_t)k * S];
        or_fft(seg, ilog2u(D), 0);
        memcpy(C + j * D, seg, sizeof(or_fe) * D);
    }
    free(seg);
}

or_fe or_rs_multifold_segment(const or_fe *segment, const or_fe *L, int fd, uint64_t n, uint64_t idx) {
    /* coPCS_utils.cpp:172-230, with partial codewords of growing length */
    const int log_fd = ilog2u((uint64_t)fd);
    uint64_t seg_n = (uint64_t)fd;
    or_fe len_inv = or_inv(or_from_u64((uint64_t)fd));
    or_fe w = or_inv(or_root_of_unity(ilog2u(n)));                   /* :177-179 */
    or_fe omega = or_pow(w, (unsigned __int128)(n / seg_n));
    or_fe inv_omegas = or_pow(w, (unsigned __int128)idx);           /* :182-184 */
    or_fe part[64][64];
    for (uint64_t j = 0; j < seg_n / 2; ++j) {                      /* :187-192 */
        part[j][0] = or_add(segment[j], segment[j + seg_n / 2]);
        part[j][1] = or_mul(inv_omegas, or_sub(segment[j], segment[j + seg_n / 2]));
        inv_omegas = or_mul(inv_omegas, omega);
    }
    if (fd > 2) {
        or_fe pw[32], pwo[32];
        for (int i = 0; i < log_fd; ++i) { pw[i].real = 1; pw[i].img = 0; pwo[i] = pw[i]; }
        pw[1] = or_mul(w, w);                                        /* :199-200 */
        for (int i = 2; i < log_fd; ++i) pw[i] = or_mul(pw[i - 1], pw[i - 1]);
        for (int i = 1; i < log_fd; ++i) pwo[i] = or_pow(pw[i], (unsigned __int128)idx);
        for (int i = 1; i < log_fd; ++i) pw[i] = or_pow(pw[i], (unsigned __int128)(n / (uint64_t)fd));
        for (int i = 1; i < log_fd; ++i) {                           /* :211-225 */
            seg_n /= 2;
            const uint64_t half = 1ULL << i;
            for (uint64_t j = 0; j < seg_n / 2; ++j) {
                for (uint64_t k = 0; k < half; ++k) {
                    or_fe v1 = or_add(part[j][k], part[j + seg_n / 2][k]);
                    or_fe v2 = or_mul(pwo[i], or_sub(part[j][k], part[j + seg_n / 2][k]));
                    part[j][k] = v1;
                    part[j][k + half] = v2;
                }
                pwo[i] = or_mul(pwo[i], pw[i]);
            }
        }
    }
    or_fe y = { 0, 0 };
    for (int i = 0; i < fd; ++i) y = or_add(y, or_mul(L[i], part[0][i]));
    return or_mul(len_inv, y);
}

void or_row_combine(const or_fe *rows, int nrows, uint64_t len, const or_fe *c, or_fe *out) {
    for (uint64_t k = 0; k < len; ++k) {                             /* coPCS.cpp:1349-1353 */
        or_fe acc = { 0, 0 };
        for (int i = 0; i < nrows; ++i) acc = or_add(acc, or_mul(c[i], rows[(uint64_t)i * len + k]));
        out[k] = acc;
    }
}

void or_cosumcheck3_phase1(const or_fe *p1, const or_fe *p2, const or_fe *p3, const or_fe *b1, uint64_t M, int d,
                           or_fe *H) {
    int logd = ilog2u((uint64_t)d);
    const or_fe *src[4] = { p1, p2, b1, p3 };                         /* h1, h2, h3, h4: coSumcheck.cpp:17-20 */
    or_fe *h[4];
    for (int v = 0; v < 4; ++v) h[v] = (or_fe *)malloc(sizeof(or_fe) * 4 * d);
    for (int k = 0; k < 4 * d; ++k) { H[k].real = 0; H[k].img = 0; }  /* :11 */
    for (uint64_t j = 0; j < M / d; ++j) {
        int zero = 1;                                                  /* :22-28 */
        for (int k = 0; k < d; ++k)
            if (p1[j * d + k].real || p1[j * d + k].img) { zero = 0; break; }
        if (zero) continue;
        for (int v = 0; v < 4; ++v) {
            memcpy(h[v], src[v] + j * d, sizeof(or_fe) * d);
            or_fft(h[v], logd, 1);                                     /* :29-32 */
            memset(h[v] + d, 0, sizeof(or_fe) * 3 * d);                /* :33-36 */
            or_fft(h[v], logd + 2, 0);                                 /* :37-40 */
        }
        for (int k = 0; k < 4 * d; ++k)                                /* :41-43 */
            H[k] = or_add(H[k], or_mul(or_sub(or_mul(h[0][k], h[1][k]), h[3][k]), h[2][k]));
    }
    for (int v = 0; v < 4; ++v) free(h[v]);
}

void or_cosumcheck3_phase2(or_fe *p1, or_fe *p2, or_fe *p3, or_fe *b1, uint64_t M, int d, const or_fe *L) {
    or_fe *vec[4] = { p1, p2, p3, b1 };                                /* coSumcheck.cpp:66-141 */
    for (uint64_t j = 0; j < M / d; ++j)
        for (int v = 0; v < 4; ++v) {
            or_fe *p = vec[v];
            int zero = 1;
            for (int k = 0; k < d; ++k)
                if (p[j * d + k].real || p[j * d + k].img) { zero = 0; break; }
            or_fe s = { 0, 0 };
            if (!zero) {
                s = or_mul(L[0], p[j * d]);
                for (int k = 1; k < d; ++k) s = or_add(s, or_mul(L[k], p[j * d + k]));
            }
            p[j] = s;
        }
}

void or_unpack_secret_share(or_fe *secret, int n, int kk, or_fe *v) {
    int logn = ilog2u((uint64_t)n);
    or_fft(secret, logn, 1);                                           /* MPI_utils.cpp:2771 */
    or_fe omega = or_inv(or_root_of_unity(1 + logn));              /* :2773-2774 */
    or_fe mul = { 1, 0 };
    for (int i = 0; i < n; ++i) {                                      /* :2775-2779 */
        secret[i] = or_mul(mul, secret[i]);
        mul = or_mul(omega, mul);
    }
    or_fft(secret, logn, 0);                                           /* :2781 */
    for (int j = 0; j < kk; ++j) v[j] = secret[(uint64_t)n * j / kk];  /* :2782-2783 */
}

void or_compute_secret_shares(const or_fe *v, uint64_t M, int N, int k, int kk, const or_fe *pads, or_fe *shares) {
    int B = 2 * N > kk ? 2 * N : kk;
    or_fe *buff = (or_fe *)malloc(sizeof(or_fe) * B);
    uint64_t counter = 0, groups = M / k;
    for (uint64_t i = 0; i < groups; ++i) {                            /* MPI_utils.cpp:2748-2766 */
        for (int j = 0; j < B; ++j) { buff[j].real = 0; buff[j].img = 0; }
        for (int j = 0; j < k; ++j) buff[j] = v[counter++];
        for (int j = k; j < kk; ++j)
            if (pads) buff[j] = pads[i * (uint64_t)(kk - k) + (j - k)];
        or_fft(buff, ilog2u((uint64_t)kk), 1);                         /* :2760 */
        for (int j = kk; j < 2 * N; ++j) { buff[j].real = 0; buff[j].img = 0; }   /* :2761 */
        or_fft(buff, ilog2u((uint64_t)(2 * N)), 0);                    /* :2762 */
        for (int j = 0; j < N; ++j) shares[(uint64_t)j * groups + i] = buff[2 * j + 1];
    }
    free(buff);
}

void or_sumcheck_aggr(const or_fe *polys, int N, int len, const or_fe *b, or_fe *out) {
    or_fe *P = (or_fe *)malloc(sizeof(or_fe) * N), *v = (or_fe *)malloc(sizeof(or_fe) * N);
    for (int i = 0; i < len; ++i) {                                    /* MPI_utils.cpp:2630-2644 */
        for (int m = 0; m < N; ++m) P[m] = polys[(uint64_t)m * len + i];   /* transpose */
        or_unpack_secret_share(P, N, N, v);
        or_fe acc = { 0, 0 };
        for (int j = 0; j < N / 4; ++j) acc = or_add(acc, or_mul(b[j], v[2 * j]));
        out[i] = acc;
    }
    free(P); free(v);
}

/* ------------------------------------------------------------------ */
/* coSumcheck drivers, all N ranks in lockstep (src/coSumcheck.cpp)     */
/* ------------------------------------------------------------------ */

/* quadratic_coSumcheck_phase2_sparse (coSumcheck.cpp:215-228) on one
 * rank's masks: s = sum_j L[j] R[iter][j]; the later masks scale by L[0];
 * R[iter+1][0] += s */
static or_fe phase2_sparse(or_fe **R, const uint64_t *R_sizes, int nR, const or_fe *L, int iter) {
    or_fe s = { 0, 0 };
    for (uint64_t j = 0; j < R_sizes[iter]; ++j) s = or_add(s, or_mul(L[j], R[iter][j]));
    for (int j = iter + 1; j < nR; ++j)
        for (uint64_t k = 0; k < R_sizes[j]; ++k) R[j][k] = or_mul(R[j][k], L[0]);
    if (iter + 1 < nR) R[iter + 1][0] = or_add(R[iter + 1][0], s);
    return s;
}

/* the mask of round iter (coSumcheck.cpp:182-193 with stride 2, :48-58 with
 * stride 4): H[t] += c H_r[t], H_r = R[iter] with H_r[0] += R[j][stride k]
 * for j > iter */
static void add_mask(or_fe *H, int len, or_fe **R, const uint64_t *R_sizes, int nR, int iter, or_fe c, int stride) {
    or_fe h0 = R[iter][0];
    for (int j = iter + 1; j < nR; ++j)
        for (uint64_t k = 0; k < R_sizes[j] / (uint64_t)stride; ++k) h0 = or_add(h0, R[j][stride * k]);
    for (int t = 0; t < len; ++t) H[t] = or_add(H[t], or_mul(c, t == 0 ? h0 : R[iter][t]));
}

/* ip_vec's update after a round (coSumcheck.cpp:268-276, :390-399) */
static void ip_update(or_fe *ip, const int *lens, int ndeg, int i, const or_fe *L) {
    int off = 0;
    for (int j = 0; j < ndeg; ++j) {
        for (int k = 0; k < lens[j]; ++k)
            ip[off + k] = or_mul(ip[off + k], j == i ? L[k] : L[0]);
        off += lens[j];
    }
}

/* send_final_sumcheck_poly's unpack (MPI_utils.cpp:2672-2736): coordinate
 * i's N shares (rank order) to its N/2 values */
static void final_unpack(const or_fe *mine, int N, int ncoord, or_fe *H) {
    or_fe *P = (or_fe *)malloc(sizeof(or_fe) * N);
    for (int i = 0; i < ncoord; ++i) {
        for (int m = 0; m < N; ++m) P[m] = mine[(uint64_t)m * ncoord + i];
        or_unpack_secret_share(P, N, N / 2, H + (uint64_t)i * (N / 2));
    }
    free(P);
}

/* the closing weights: 1 at N = 4, else compute_lagrange_coeff(r, N/4) */
static void closing_weights(or_fe r, int N, or_fe *L) {
    if (N / 4 == 1) { L[0].real = 1; L[0].img = 0; }
    else or_compute_lagrange_coeff(r, N / 4, L);
}

int or_quadratic_cosumcheck(int N, or_fe **p1, or_fe **p2, uint64_t M, or_fe ***R, const uint64_t *R_sizes, int nR,
                            or_fe K, or_fe K_r, const int *degrees, int ndeg, or_fe *prev,
                            or_fe *claims, or_fe *r_out, or_fe *ip_vec, int *failures) {
    if (N < 4 || (N & (N - 1)) || nR != ndeg) return -1;
    uint64_t Mc = M;
    for (int i = 0; i < ndeg; ++i) {
        if (R_sizes[i] != 2 * (uint64_t)degrees[i] || Mc % (uint64_t)degrees[i]) return -1;
        Mc /= (uint64_t)degrees[i];
    }
    if (Mc < 2) return -1;
    Mc = M;
    *failures = 0;
    or_fe kk[2] = { K, K_r };
    const or_fe c = or_hash_to_field(prev, kk, 2);                /* :234 */
    K = or_add(K, or_mul(c, K_r));                                 /* :235 */
    int *lens = (int *)malloc(sizeof(int) * ndeg), total = 0;
    for (int i = 0; i < ndeg; ++i) { lens[i] = 2 * degrees[i]; total += lens[i]; }
    for (int t = 0; t < total; ++t) { ip_vec[t].real = 1; ip_vec[t].img = 0; }  /* :244-247 */
    or_fe *fin_r = (or_fe *)malloc(sizeof(or_fe) * N);
    or_fe *ones = (or_fe *)malloc(sizeof(or_fe) * (N / 4));
    for (int j = 0; j < N / 4; ++j) { ones[j].real = 1; ones[j].img = 0; }
    for (int i = 0; i < ndeg; ++i) {
        const int d = degrees[i], len = 2 * d;
        or_fe *all = (or_fe *)malloc(sizeof(or_fe) * N * len), *H = (or_fe *)malloc(sizeof(or_fe) * len);
        for (int m = 0; m < N; ++m) {                              /* phase 1 on every rank (:150-198) */
            or_pcs_sumcheck_phase1(p1[m], p2[m], Mc, d, all + (uint64_t)m * len);
            add_mask(all + (uint64_t)m * len, len, R[m], R_sizes, nR, i, c, 2);
        }
        or_sumcheck_aggr(all, N, len, ones, H);                    /* sumcheck_aggr(H, ones, N) :197 */
        or_fe eval = { 0, 0 };
        for (int j = 0; j < len / 2; ++j) eval = or_add(eval, H[2 * j]);   /* :254-257 */
        if (eval.real != K.real || eval.img != K.img) *failures |= 1 << i;
        const or_fe r = or_hash_to_field(prev, H, len);            /* :264 */
        r_out[1 + i] = r;
        or_fe *L2 = (or_fe *)malloc(sizeof(or_fe) * len);
        or_compute_lagrange_coeff(r, len, L2);                     /* :267 */
        ip_update(ip_vec, lens, ndeg, i, L2);
        K.real = 0; K.img = 0;
        for (int j = 0; j < len; ++j) K = or_add(K, or_mul(L2[j], H[j]));   /* :280 */
        for (int m = 0; m < N; ++m) fin_r[m] = phase2_sparse(R[m], R_sizes, nR, L2, i);   /* :283 */
        for (int m = 0; m < N; ++m) or_pcs_sumcheck_phase2(p1[m], p2[m], Mc, d, r);      /* :287-289 */
        Mc /= (uint64_t)d;
        free(L2); free(all); free(H);
    }
    /* send_final_sumcheck_poly({p1[0], p2[0], p1[1], p2[1], final_eval_r}, N) (:293) */
    or_fe *mine = (or_fe *)malloc(sizeof(or_fe) * N * 5), *Hf = (or_fe *)malloc(sizeof(or_fe) * 5 * (N / 2));
    for (int m = 0; m < N; ++m) {
        mine[m * 5 + 0] = p1[m][0]; mine[m * 5 + 1] = p2[m][0];
        mine[m * 5 + 2] = p1[m][1]; mine[m * 5 + 3] = p2[m][1];
        mine[m * 5 + 4] = fin_r[m];
    }
    final_unpack(mine, N, 5, Hf);
    const or_fe *H0 = Hf, *H1 = Hf + N / 2, *H2 = Hf + N, *H3 = Hf + 3 * (N / 2), *H4 = Hf + 2 * N;
    for (int i = 0; i < N / 4; ++i)                                /* :298-300 */
        K = or_sub(K, or_add(or_add(or_mul(H0[i], H1[i]), or_mul(H2[i], H3[i])), or_mul(c, H4[i])));
    if (K.real || K.img) *failures |= 1 << ndeg;
    or_fe r = or_hash_to_field(prev, H0, N / 2);                   /* :306-307 */
    for (int i = 1; i < 5; ++i) r = or_hash_to_field(prev, Hf + (uint64_t)i * (N / 2), N / 2);
    r_out[0] = r;                                                  /* :309 */
    or_fe *L = (or_fe *)malloc(sizeof(or_fe) * (N / 4));
    closing_weights(r, N, L);                                      /* :311-316 */
    for (int i = 0; i < 5; ++i) {                                  /* :318-325 */
        or_fe e = { 0, 0 };
        for (int j = 0; j < N / 4; ++j) e = or_add(e, or_mul(L[j], Hf[(uint64_t)i * (N / 2) + j]));
        claims[i] = e;
    }
    free(L); free(mine); free(Hf); free(fin_r); free(ones); free(lens);
    return 0;
}

int or_cubic_cosumcheck(int N, or_fe **p1, or_fe **p2, or_fe **p3, uint64_t M, const or_fe *r1, int nr1,
                        const or_fe *r2, int nr2, or_fe ***R, const uint64_t *R_sizes, int nR, or_fe K, or_fe K_r,
                        const int *degrees, int ndeg, or_fe *prev, or_fe *claims, or_fe *r_out, or_fe *ip_vec,
                        int *failures) {
    if (N < 4 || (N & (N - 1)) || nR != ndeg || (1ULL << nr1) != M || (1 << nr2) < N / 4) return -1;
    uint64_t Mc = M;
    int logsum = 0;
    for (int i = 0; i < ndeg; ++i) {
        if (R_sizes[i] != 4 * (uint64_t)degrees[i] || Mc % (uint64_t)degrees[i]) return -1;
        Mc /= (uint64_t)degrees[i];
        logsum += ilog2u((uint64_t)degrees[i]);
    }
    if (Mc < 1 || logsum > nr1) return -1;
    Mc = M;
    *failures = 0;
    or_fe kk[2] = { K, K_r };
    const or_fe c = or_hash_to_field(prev, kk, 2);                /* :336 */
    K = or_add(K, or_mul(c, K_r));                                 /* :337 */
    or_fe **b1 = (or_fe **)malloc(sizeof(or_fe *) * N);
    for (int m = 0; m < N; ++m) {                                  /* precompute_beta(r1, b1) :342 */
        b1[m] = (or_fe *)malloc(sizeof(or_fe) * M);
        or_precompute_beta(r1, nr1, b1[m]);
    }
    or_fe *b2 = (or_fe *)malloc(sizeof(or_fe) << nr2);
    or_precompute_beta(r2, nr2, b2);                               /* :343 */
    int *lens = (int *)malloc(sizeof(int) * ndeg), total = 0;
    for (int i = 0; i < ndeg; ++i) { lens[i] = 4 * degrees[i]; total += lens[i]; }
    for (int t = 0; t < total; ++t) { ip_vec[t].real = 1; ip_vec[t].img = 0; }
    or_fe *fin_r = (or_fe *)malloc(sizeof(or_fe) * N);
    or_fe **LV = (or_fe **)malloc(sizeof(or_fe *) * ndeg);
    for (int i = 0; i < ndeg; ++i) {
        const int d = degrees[i], len = 4 * d;
        or_fe *all = (or_fe *)malloc(sizeof(or_fe) * N * len), *H = (or_fe *)malloc(sizeof(or_fe) * len);
        for (int m = 0; m < N; ++m) {                              /* :9-62 on every rank */
            or_cosumcheck3_phase1(p1[m], p2[m], p3[m], b1[m], Mc, d, all + (uint64_t)m * len);
            add_mask(all + (uint64_t)m * len, len, R[m], R_sizes, nR, i, c, 4);
        }
        or_sumcheck_aggr(all, N, len, b2, H);                      /* sumcheck_aggr(H, b2, N) :61 */
        or_fe eval = { 0, 0 };
        for (int j = 0; j < len / 4; ++j) eval = or_add(eval, H[4 * j]);   /* :358-361 */
        if (eval.real != K.real || eval.img != K.img) *failures |= 1 << i;
        const or_fe r = or_hash_to_field(prev, H, len);            /* :367 */
        r_out[1 + i] = r;
        or_fe *L4 = (or_fe *)malloc(sizeof(or_fe) * len);
        or_compute_lagrange_coeff(r, len, L4);                     /* :371 */
        ip_update(ip_vec, lens, ndeg, i, L4);
        K.real = 0; K.img = 0;
        for (int j = 0; j < len; ++j) K = or_add(K, or_mul(L4[j], H[j]));   /* :387 */
        for (int m = 0; m < N; ++m) fin_r[m] = phase2_sparse(R[m], R_sizes, nR, L4, i);   /* :388 */
        LV[i] = (or_fe *)malloc(sizeof(or_fe) * d);
        or_compute_lagrange_coeff(r, d, LV[i]);                    /* :390 */
        for (int m = 0; m < N; ++m) or_cosumcheck3_phase2(p1[m], p2[m], p3[m], b1[m], Mc, d, LV[i]);   /* :393 */
        Mc /= (uint64_t)d;
        free(L4); free(all); free(H);
    }
    /* send_final_sumcheck_poly({p1[0], p2[0], p3[0], final_eval_r}, N) (:398) */
    or_fe *mine = (or_fe *)malloc(sizeof(or_fe) * N * 4), *Hf = (or_fe *)malloc(sizeof(or_fe) * 4 * (N / 2));
    for (int m = 0; m < N; ++m) {
        mine[m * 4 + 0] = p1[m][0]; mine[m * 4 + 1] = p2[m][0];
        mine[m * 4 + 2] = p3[m][0]; mine[m * 4 + 3] = fin_r[m];
    }
    final_unpack(mine, N, 4, Hf);
    const or_fe *H0 = Hf, *H1 = Hf + N / 2, *H2 = Hf + N, *H3 = Hf + 3 * (N / 2);
    for (int i = 0; i < N / 4; ++i)                                /* :403-405 */
        K = or_sub(K, or_mul(b2[i], or_add(or_mul(b1[0][0], or_sub(or_mul(H0[i], H1[i]), H2[i])), or_mul(c, H3[i]))));
    if (K.real || K.img) *failures |= 1 << ndeg;
    or_fe r = or_hash_to_field(prev, H0, N / 2);                   /* :410: H[3] is not absorbed */
    r = or_hash_to_field(prev, H1, N / 2);
    r = or_hash_to_field(prev, H2, N / 2);
    r_out[0] = r;
    or_fe *L = (or_fe *)malloc(sizeof(or_fe) * (N / 4));
    closing_weights(r, N, L);                                      /* :413-418 */
    for (int i = 0; i < 4; ++i) {                                  /* :420-427 */
        or_fe e = { 0, 0 };
        for (int j = 0; j < N / 4; ++j) e = or_add(e, or_mul(L[j], Hf[(uint64_t)i * (N / 2) + j]));
        claims[i] = e;
    }
    /* verification 2: b1[0] = prod_i <beta(r1 slice i), L_i> (:429-443) */
    {
        or_fe y = { 1, 0 };
        int ctr = 0;
        for (int i = 0; i < ndeg; ++i) {
            const int lg = ilog2u((uint64_t)degrees[i]);
            or_fe *tb = (or_fe *)malloc(sizeof(or_fe) << lg);
            or_precompute_beta(r1 + ctr, lg, tb);
            ctr += lg;
            or_fe s = { 0, 0 };
            for (int j = 0; j < degrees[i]; ++j) s = or_add(s, or_mul(tb[j], LV[i][j]));
            y = or_mul(y, s);
            free(tb);
        }
        if (y.real != b1[0][0].real || y.img != b1[0][0].img) *failures |= 1 << (ndeg + 1);
    }
    for (int i = 0; i < ndeg; ++i) free(LV[i]);
    for (int m = 0; m < N; ++m) free(b1[m]);
    free(LV); free(b1); free(b2); free(L); free(mine); free(Hf); free(fin_r); free(lens);
    return 0;
}

/* ------------------------------------------------------------------ */
/* PCS_commit at one rank: src/coPCS.cpp:76-128, 236-250               */
/* ------------------------------------------------------------------ */

/* prepare_data(Commit_data, Shares, Random) with BATCH_FACTOR = 32
 * (src/New_PCS.cpp:9) and _queries[0] = q0: rows (32 x B, B = n/16) are
 * filled as the reference fills them; returns the non-empty row count, or
 * -1 where the reference exits or reads past its inputs */
int or_prepare_data(const or_fe *shares, uint64_t n, const or_fe *rnd, uint64_t nr, int q0, or_fe *rows) {
    const uint64_t B = n / 16, q = (uint64_t)q0;
    if (B <= q) return -1;
    memset(rows, 0, sizeof(or_fe) * 32 * B);
    uint64_t data_ctr = 0, rand_ctr = 0;
    int last_batch = 0, done = 0;
    for (int i = 0; i < 32; ++i) {                         /* :84-104 */
        for (uint64_t k = 0; k < B - q; ++k) {
            rows[i * B + k] = shares[data_ctr++];
            if (data_ctr == n) break;
        }
        for (uint64_t k = B - q; k < B; ++k) {
            if (rand_ctr >= nr) return -1;
            rows[i * B + k] = rnd[rand_ctr++];
        }
        if (data_ctr == n) { last_batch = i; done = 1; break; }
    }
    if (!done) return -1;                                  /* :105-108 */
    last_batch++;                                          /* :109 */
    if (last_batch >= 32) return -1;
    if (n - rand_ctr > B || rand_ctr > n) {                /* :112 (unsigned) */
        for (uint64_t i = 0; i < B; ++i) {
            if (rand_ctr >= nr) return -1;
            rows[last_batch * B + i] = rnd[rand_ctr++];
            if (rand_ctr == nr) break;
        }
    } else {
        for (uint64_t i = 0; i < B - (n - rand_ctr); ++i) {
            if (rand_ctr >= nr) return -1;
            rows[last_batch * B + i] = rnd[rand_ctr++];
        }
    }
    return last_batch + 1;
}
