/*
 * oracle_mt.c -- all-cores (OpenMP) variants of the oracle's plain-prover
 * stages: RS encode, dup-left Merkle commit, product sumcheck.
 *
 * TEST INFRASTRUCTURE ONLY, like oracle.c: loaded by tests/ (to check
 * full-size GPU outputs in seconds) and by bench.py's cpu_baseline leg (the
 * all-cores CPU figure).  Each function computes exactly what its
 * single-thread counterpart in oracle.c computes -- the arithmetic is exact,
 * so a different evaluation order gives identical bytes -- and
 * tests/test_oracle_mt.py checks them against those counterparts.
 *
 * The transform follows the reference's pfft (src/utils.cpp:1583-1673): a
 * four-step evaluation of the forward DFT with threads over rows and
 * columns; the reference fixes 16 threads and a 16-way split, here the split
 * is sqrt(n) by sqrt(n) and the thread count is the caller's.
 */
#include "oracle.h"

#include <omp.h>
#include <stdlib.h>
#include <string.h>

/* The field operations of oracle.c (src/fieldElement.cpp:34-96), inlined
 * here so the hot loops do not pay a call per operation. */
#define P61 0x1FFFFFFFFFFFFFFFULL

static inline uint64_t mt_modp(unsigned __int128 t) {
    uint64_t s = ((uint64_t)t & P61) + (uint64_t)(t >> 61);
    s = (s & P61) + (s >> 61);
    const uint64_t d = s - P61;             /* branch-free conditional subtract */
    return d + (P61 & (uint64_t)((int64_t)d >> 63));
}
/* branch-free: x - p when that does not go negative, else x (x < 2p) */
static inline uint64_t mt_red(uint64_t x) {
    const uint64_t d = x - P61;
    return d + (P61 & (uint64_t)((int64_t)d >> 63));
}
static inline or_fe mt_add(or_fe a, or_fe b) {
    or_fe o = { mt_red(a.real + b.real), mt_red(a.img + b.img) };
    return o;
}
static inline or_fe mt_sub(or_fe a, or_fe b) {
    or_fe o = { mt_red(a.real + P61 - b.real), mt_red(a.img + P61 - b.img) };
    return o;
}
static inline or_fe mt_mul(or_fe a, or_fe b) {
    /* real = ar br - ai bi, img = ar bi + ai br: both sums of two 122-bit
     * products folded once (each product < 2^122, their sum < 2^123) */
    const unsigned __int128 ac = (unsigned __int128)a.real * b.real;
    const unsigned __int128 bd = (unsigned __int128)a.img * b.img;
    const unsigned __int128 ad = (unsigned __int128)a.real * b.img;
    const unsigned __int128 bc = (unsigned __int128)a.img * b.real;
    const uint64_t re = mt_modp(ac), rb = mt_modp(bd);
    or_fe o = { mt_red(re + P61 - rb), mt_modp(ad + bc) };
    return o;
}
#define or_add mt_add
#define or_sub mt_sub
#define or_mul mt_mul

/* in-place radix-2 transform of 2^logn points whose twiddles w[k] = root^k
 * (k < 2^(logn-1), root of order 2^logn) are given; natural order in/out */
static void fft_tab(or_fe *a, int logn, const or_fe *w) {
    const uint64_t len = 1ULL << logn;
    for (uint64_t i = 1, j = 0; i < len; ++i) {
        uint64_t bit = len >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) { or_fe t = a[i]; a[i] = a[j]; a[j] = t; }
    }
    for (uint64_t i = 2; i <= len; i <<= 1) {
        const uint64_t h = i >> 1, step = len / i;
        for (uint64_t j = 0; j < len; j += i)
            for (uint64_t k = 0; k < h; ++k) {
                const or_fe u = a[j + k];
                const or_fe v = or_mul(a[j + k + h], w[step * k]);
                a[j + k] = or_add(u, v);
                a[j + k + h] = or_sub(u, v);
            }
    }
}

static or_fe *pow_table(or_fe root, uint64_t cnt) {
    or_fe *t = (or_fe *)malloc(sizeof(or_fe) * (cnt ? cnt : 1));
    t[0].real = 1; t[0].img = 0;
    for (uint64_t k = 1; k < cnt; ++k) t[k] = or_mul(t[k - 1], root);
    return t;
}

/* fft(arr, logn, inverse) of src/utils.cpp:904-964, four-step:
 * j = j1 + n1 j2, k = k2 + n2 k1; column transforms of length n2 over j2,
 * twist by w^(j1 k2), row transforms of length n1 over j1. */
void or_fft_mt(or_fe *arr, int logn, int inverse, int threads) {
    if (logn < 12) { or_fft(arr, logn, inverse); return; }
    const int l1 = logn / 2, l2 = logn - l1;
    const uint64_t n = 1ULL << logn, n1 = 1ULL << l1, n2 = 1ULL << l2;
    or_fe w = or_root_of_unity(logn);
    if (inverse) w = or_inv(w);
    /* roots of order n2 and n1 are powers of w (root_of_unity(k-1) = root_of_unity(k)^2) */
    or_fe w2 = w, w1 = w;
    for (int i = 0; i < l1; ++i) w2 = or_mul(w2, w2);
    for (int i = 0; i < l2; ++i) w1 = or_mul(w1, w1);
    or_fe *t2 = pow_table(w2, n2 / 2), *t1 = pow_table(w1, n1 / 2);
    or_fe *wt = pow_table(w, n2);          /* w^k2, k2 < n2 */
    or_fe *Y = (or_fe *)malloc(sizeof(or_fe) * n);
    const int B = 8;                       /* columns gathered together */
    #pragma omp parallel num_threads(threads)
    {
        or_fe *col = (or_fe *)malloc(sizeof(or_fe) * n2 * B);
        #pragma omp for schedule(static)
        for (uint64_t j1 = 0; j1 < n1; j1 += B) {
            for (uint64_t j2 = 0; j2 < n2; ++j2)
                for (int b = 0; b < B; ++b) col[b * n2 + j2] = arr[j1 + b + n1 * j2];
            for (int b = 0; b < B; ++b) {
                or_fe *c = col + b * n2;
                fft_tab(c, l2, t2);
                /* twist w^(j1 k2) as a running power of w^j1 */
                const uint64_t jj = j1 + b;
                or_fe step = { 1, 0 }, tw = { 1, 0 };
                if (jj) {
                    /* w^jj from the table of w^k2 (jj < n1 <= n2) */
                    step = wt[jj];
                }
                for (uint64_t k2 = 0; k2 < n2; ++k2) {
                    Y[jj * n2 + k2] = or_mul(c[k2], tw);
                    tw = or_mul(tw, step);
                }
            }
        }
        free(col);
        or_fe *row = (or_fe *)malloc(sizeof(or_fe) * n1 * B);
        #pragma omp for schedule(static)
        for (uint64_t k2 = 0; k2 < n2; k2 += B) {
            for (uint64_t j1 = 0; j1 < n1; ++j1)
                for (int b = 0; b < B; ++b) row[b * n1 + j1] = Y[j1 * n2 + k2 + b];
            for (int b = 0; b < B; ++b) {
                fft_tab(row + b * n1, l1, t1);
                for (uint64_t k1 = 0; k1 < n1; ++k1) arr[k2 + b + n2 * k1] = row[b * n1 + k1];
            }
        }
        free(row);
    }
    if (inverse) {                          /* utils.cpp:952-963 */
        const or_fe ilen = or_inv(or_from_u64(n));
        #pragma omp parallel for num_threads(threads) schedule(static)
        for (uint64_t i = 0; i < n; ++i) arr[i] = or_mul(arr[i], ilen);
    }
    free(Y); free(wt); free(t1); free(t2);
}

/* Whir_Commit_initial's codeword (src/Whir.cpp:158-163) on all cores */
void or_rs_encode_mt(const or_fe *arr, uint64_t n, int rate, or_fe *cw, int threads) {
    const uint64_t m = n * (uint64_t)rate;
    memcpy(cw, arr, n * sizeof(or_fe));
    memset(cw + n, 0, (m - n) * sizeof(or_fe));
    int logm = 0;
    while ((1ULL << logm) < m) ++logm;
    or_fft_mt(cw, logm, 0, threads);
}

/* MT_commit_Blake + create_tree_blake (src/merkle_tree.cpp:195-299), every
 * level stored, leaves and each level's nodes hashed in parallel */
void or_merkle_commit_mt(const or_fe *elems, uint64_t n_elems, uint8_t *nodes, int dup_left, int threads) {
    const uint64_t leaves = n_elems / 4;
    #pragma omp parallel num_threads(threads)
    {
        #pragma omp for schedule(static)
        for (uint64_t i = 0; i < leaves; ++i)
            or_blake3_hash64((const uint8_t *)(elems + 4 * i), nodes + 32 * i);
        uint8_t *prev = nodes, *cur = nodes + 32 * leaves;
        for (uint64_t s = leaves / 2; s >= 1; s /= 2) {
            #pragma omp for schedule(static)
            for (uint64_t i = 0; i < s; ++i) {
                uint8_t data[64];
                memcpy(data, prev + 32 * (2 * i), 32);
                memcpy(data + 32, prev + 32 * (dup_left ? 2 * i : 2 * i + 1), 32);
                or_blake3_hash64(data, cur + 32 * i);
            }
            prev = cur;
            cur += 32 * s;
        }
    }
}

/* one round's (a, b, c) as round_poly in oracle.c, partial sums per thread */
static void round_poly_mt(const or_fe *v1, const or_fe *v2, uint64_t L, or_fe out[3], int threads) {
    or_fe a = { 0, 0 }, b = { 0, 0 }, c = { 0, 0 };
    #pragma omp parallel num_threads(threads)
    {
        or_fe pa = { 0, 0 }, pb = { 0, 0 }, pc = { 0, 0 };
        #pragma omp for schedule(static) nowait
        for (uint64_t j = 0; j < L; ++j) {
            const or_fe d1 = or_sub(v1[2 * j + 1], v1[2 * j]);
            const or_fe d2 = or_sub(v2[2 * j + 1], v2[2 * j]);
            pa = or_add(pa, or_mul(d1, d2));
            pb = or_add(pb, or_add(or_mul(d1, v2[2 * j]), or_mul(v1[2 * j], d2)));
            pc = or_add(pc, or_mul(v1[2 * j], v2[2 * j]));
        }
        #pragma omp critical
        {
            a = or_add(a, pa);
            b = or_add(b, pb);
            c = or_add(c, pc);
        }
    }
    out[0] = a; out[1] = b; out[2] = c;
}

static or_fe quad_eval_mt(const or_fe p[3], or_fe x) {   /* polynomial.cpp:111-113 */
    return or_add(or_mul(or_add(or_mul(p[0], x), p[1]), x), p[2]);
}

/* _quadratic_sumcheck at N = 1, all rounds Fiat-Shamir
 * (src/Distributed_Sumcheck.cpp:204-348), as or_product_sumcheck with
 * offset 0; the folds write fresh buffers so a round's reads and writes
 * never overlap across threads.  v1, v2 are not modified. */
int or_product_sumcheck_mt(or_fe y, const or_fe *v1, const or_fe *v2, uint64_t n, or_fe *prev, or_fe *polys,
                           or_fe *r, or_fe *final_out, int threads) {
    int m = 0;
    while ((1ULL << m) < n) ++m;
    or_fe *a1 = (or_fe *)malloc(sizeof(or_fe) * (n / 2 + 1)), *a2 = (or_fe *)malloc(sizeof(or_fe) * (n / 2 + 1));
    or_fe *b1 = (or_fe *)malloc(sizeof(or_fe) * (n / 4 + 1)), *b2 = (or_fe *)malloc(sizeof(or_fe) * (n / 4 + 1));
    const or_fe *s1 = v1, *s2 = v2;
    const or_fe zero = { 0, 0 }, one = { 1, 0 };
    int rc = 0;
    for (int i = 0; i < m; ++i) {
        const uint64_t L = 1ULL << (m - 1 - i);
        or_fe p[3];
        round_poly_mt(s1, s2, L, p, threads);
        const or_fe h = or_add(quad_eval_mt(p, zero), quad_eval_mt(p, one));
        if (h.real != y.real || h.img != y.img) { rc = -(i + 1); break; }
        const or_fe ch = or_hash_to_field(prev, p, 3);
        y = quad_eval_mt(p, ch);
        polys[3 * i] = p[0]; polys[3 * i + 1] = p[1]; polys[3 * i + 2] = p[2];
        r[i] = ch;
        or_fe *d1 = (i & 1) ? b1 : a1, *d2 = (i & 1) ? b2 : a2;
        #pragma omp parallel for num_threads(threads) schedule(static)
        for (uint64_t j = 0; j < L; ++j) {
            d1[j] = or_add(or_mul(ch, or_sub(s1[2 * j + 1], s1[2 * j])), s1[2 * j]);
            d2[j] = or_add(or_mul(ch, or_sub(s2[2 * j + 1], s2[2 * j])), s2[2 * j]);
        }
        s1 = d1; s2 = d2;
    }
    if (rc == 0) {
        final_out[0] = s1[0];
        final_out[1] = s2[0];
    }
    free(a1); free(a2); free(b1); free(b2);
    return rc;
}

/* the plain single-rank prover on all cores: the proof bytes of
 * pyoracle.plain_proof (root | m x (a, b, c) | v1[0], v2[0]) for the
 * transcript starting at *prev (updated).  Returns 0, or the failed round
 * as a negative number.  stage_ms (optional, 4): encode, merkle, sumcheck,
 * total wall ms. */
int or_plain_proof_mt(const or_fe *f, const or_fe *g, uint64_t n, int rate, or_fe *prev, uint8_t *proof,
                      double *stage_ms, int threads) {
    int m = 0;
    while ((1ULL << m) < n) ++m;
    const uint64_t cwn = n * (uint64_t)rate;
    const double t0 = omp_get_wtime();
    or_fe *cw = (or_fe *)malloc(sizeof(or_fe) * cwn);
    or_rs_encode_mt(f, n, rate, cw, threads);
    const double t1 = omp_get_wtime();
    const uint64_t cnt = or_merkle_node_count(cwn);
    uint8_t *nodes = (uint8_t *)malloc(32 * cnt);
    or_merkle_commit_mt(cw, cwn, nodes, 1, threads);
    const double t2 = omp_get_wtime();
    /* claimed sum y = sum f g */
    or_fe y = { 0, 0 };
    #pragma omp parallel num_threads(threads)
    {
        or_fe py = { 0, 0 };
        #pragma omp for schedule(static) nowait
        for (uint64_t i = 0; i < n; ++i) py = or_add(py, or_mul(f[i], g[i]));
        #pragma omp critical
        y = or_add(y, py);
    }
    or_fe *polys = (or_fe *)malloc(sizeof(or_fe) * 3 * (m ? m : 1));
    or_fe *r = (or_fe *)malloc(sizeof(or_fe) * (m ? m : 1));
    or_fe fin[2];
    const int rc = or_product_sumcheck_mt(y, f, g, n, prev, polys, r, fin, threads);
    const double t3 = omp_get_wtime();
    if (rc == 0) {
        memcpy(proof, nodes + 32 * (cnt - 1), 32);
        memcpy(proof + 32, polys, 48 * (size_t)m);
        memcpy(proof + 32 + 48 * (size_t)m, fin, 32);
    }
    if (stage_ms) {
        stage_ms[0] = 1e3 * (t1 - t0);
        stage_ms[1] = 1e3 * (t2 - t1);
        stage_ms[2] = 1e3 * (t3 - t2);
        stage_ms[3] = 1e3 * (t3 - t0);
    }
    free(cw); free(nodes); free(polys); free(r);
    return rc;
}

/* update_betas' weight update (src/Whir.cpp:325-342): b[j] += sum_i
 * ch[i] w[i]^j for j < n, threads over ranges of j (each starts its
 * running powers at w[i]^j0) */
void or_update_betas_mt(or_fe *b, uint64_t n, const or_fe *ch, const or_fe *w, int l, int threads) {
    #pragma omp parallel num_threads(threads)
    {
        const int nt = omp_get_num_threads(), id = omp_get_thread_num();
        const uint64_t lo = n * (uint64_t)id / (uint64_t)nt, hi = n * (uint64_t)(id + 1) / (uint64_t)nt;
        for (int i = 0; i < l && lo < hi; ++i) {
            or_fe pw = or_mul(ch[i], or_pow(w[i], lo));
            for (uint64_t j = lo; j < hi; ++j) {
                b[j] = or_add(b[j], pw);
                pw = or_mul(pw, w[i]);
            }
        }
    }
}

int or_max_threads(void) { return omp_get_max_threads(); }
