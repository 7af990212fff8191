/*
 * oracle.h -- CPU restatement of the reference prover's rank-local path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.  The
 * product (libcosnark.so) never links, loads or calls anything in here.
 *
 * Plain C (gcc), scalar, single-threaded unless a function says otherwise.
 * Every function cites the reference file:line whose behaviour it restates
 * (paths relative to the reference repository root).
 *
 * Element layout follows src/fieldElement.hpp:96-97: {u64 real; u64 img;},
 * 16 bytes, canonical values in [0, 2^61-1).
 */
#ifndef COSNARK_ORACLE_H
#define COSNARK_ORACLE_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct { uint64_t real, img; } or_fe;

/* ---- field GF((2^61-1)^2), src/fieldElement.cpp ---- */
or_fe or_add(or_fe a, or_fe b);                 /* :34-47 */
or_fe or_sub(or_fe a, or_fe b);                 /* :80-96 */
or_fe or_mul(or_fe a, or_fe b);                 /* :49-78 */
or_fe or_inv(or_fe a);                          /* :207-210 (x^(p^2-2)) */
or_fe or_pow(or_fe a, unsigned __int128 e);     /* :323-335 */
or_fe or_root_of_unity(int logn);               /* src/utils.cpp:706-717 */
or_fe or_from_u64(uint64_t x);                  /* F(long long), :24-27 */

/* array wrappers (for ctypes), element arrays as u64[2n] */
void or_vec_mul(const or_fe *a, const or_fe *b, or_fe *out, uint64_t n);
void or_vec_add(const or_fe *a, const or_fe *b, or_fe *out, uint64_t n);
void or_vec_sub(const or_fe *a, const or_fe *b, or_fe *out, uint64_t n);
void or_vec_inv(const or_fe *a, or_fe *out, uint64_t n);

/* ---- BLAKE3 of one 64-byte block, src/Blake3_hash.cpp:5-9 ---- */
void or_blake3_hash64(const uint8_t src[64], uint8_t dst[32]);
/* general single-block BLAKE3 (len <= 64), for known-answer checks */
void or_blake3_small(const uint8_t *src, size_t len, uint8_t dst[32]);
void or_blake3_hash64_many(const uint8_t *src, uint8_t *dst, uint64_t n);

/* ---- Fiat-Shamir, src/Fiat_Shamir.cpp ---- */
or_fe or_hash_F(or_fe x, or_fe y);                              /* :8-27 */
/* absorbs elems[0..n) into *prev (the reference's prev_challenge) */
or_fe or_hash_to_field(or_fe *prev, const or_fe *elems, uint64_t n); /* :29-34 */
/* writes l (N-index, M-index) pairs into idx[2l] */
void or_get_indexes(or_fe *prev, int l, uint32_t N, uint32_t M, uint32_t *idx); /* :37-56 */

/* ---- NTT, src/utils.cpp ---- */
void or_fft(or_fe *arr, int logn, int inverse);     /* fft :904-964 */
void or_pfft(or_fe *arr, int logn);                 /* pfft :1583-1673 (forward only) */
/* Whir_Commit_initial's codeword: (arr || 0) zero-padded to rate*n, forward
 * transform.  src/Whir.cpp:158-163 */
void or_rs_encode(const or_fe *arr, uint64_t n, int rate, or_fe *codeword);

/* ---- Merkle tree, src/merkle_tree.cpp ---- */
/* number of 32-byte nodes over all levels for a commit of n_elems elements */
uint64_t or_merkle_node_count(uint64_t n_elems);
/* MT_commit_Blake (:195-233) + create_tree_blake (:267-299); levels stored
 * back to back, level 0 first; root is the last node.  n_elems must be a
 * power of two >= 4.  dup_left=1 is the reference's behaviour (the parent
 * hashes its left child twice, :287-292); dup_left=0 hashes left||right. */
void or_merkle_commit(const or_fe *elems, uint64_t n_elems, uint8_t *nodes, int dup_left);
/* open_tree_blake (:320-335): sibling path, one digest per level below root */
int or_merkle_open(const uint8_t *nodes, uint64_t n_elems, uint64_t pos, uint8_t *path);

/* ---- multilinear helpers, src/utils.cpp ---- */
void or_precompute_beta(const or_fe *r, int m, or_fe *B);        /* :198-243 */
or_fe or_evaluate_vector(const or_fe *v, uint64_t n, const or_fe *r); /* :1201-1214 */
void or_compute_lagrange_coeff(or_fe r, int degree, or_fe *L);   /* :169-197 */

/* ---- glibc challenge stream, src/fieldElement.cpp:119-125,363-368 ---- */
or_fe or_random_element(void);   /* F::_random(): uses libc random() */

/* ---- product sumcheck, src/Distributed_Sumcheck.cpp:204-348 at one rank,
 *      tail src/sparse_eval.cpp:140-192 ----
 * v1, v2 (length n, power of two) are folded in place as the reference does.
 * offset = number of trailing rounds that draw F::_random() challenges
 * (the reference's sumcheck_offset); the leading rounds use the Fiat-Shamir
 * chain *prev.  polys[3*round + {0,1,2}] = (a,b,c) of a x^2 + b x + c;
 * r[round] = challenge.  final_out = {v1[0], v2[0]}.
 * Returns 0, or -(round+1) if H(0)+H(1) != y (the reference exits there). */
int or_product_sumcheck(or_fe y, or_fe *v1, or_fe *v2, uint64_t n, or_fe *prev,
                        int offset, or_fe *polys, or_fe *r, or_fe *final_out);

/* ---- synthetic inputs: splitmix64-seeded xoshiro256**, 61-bit masked,
 *      value p rejected (workload definition, not a reference function) ---- */
void or_synth_fill(uint64_t seed, uint64_t stream, or_fe *out, uint64_t n);

/* ---- WHIR, src/Whir.cpp ---- */
int or_whir_query_count(int rate);                                  /* :253-258 */
or_fe or_rs_multifold_segment_tensor(const or_fe *seg, int seg_n, const or_fe *r,
                                     int nr, uint64_t n, uint64_t idx); /* coPCS_utils.cpp:333-383 */
/* Whir_Open (:388-467) with the iteration loop of Whir_Iteration (:354-386),
 * whir_k = 4.  arr (n), codeword (rate*n, from or_rs_encode) and its Merkle
 * nodes describe the initial commitment; z (log2 n coordinates) is the
 * single claim point and K the claim.  Draws challenges from libc random()
 * (the caller seeds it).  Writes a flat record; see oracle.c for layout.
 * Returns the number of u64 words written, or a negative error. */
int64_t or_whir_open(const or_fe *arr, uint64_t n, const or_fe *z, or_fe K,
                     int rate, int stop_log, uint64_t *record, uint64_t record_cap);
/* get_whir_rates (src/Whir.cpp:18-137): the rate each array (sizes sorted
 * largest first) is committed at, starting from `rate`. */
void or_whir_rates(const uint64_t *sizes, int count, int rate, int whir_k, int *rates);
/* Whir_Batch_Open's fold schedule (src/Whir.cpp:736-755): kappas[] and
 * aggr[] per step; returns the step count (-1 beyond cap). */
int or_whir_kappas(const uint64_t *sizes, int count, int whir_k, int *kappas, int *aggr, int cap);
/* Whir_Batch_Open (src/Whir.cpp:706-899) with the prover's record; arrays
 * sorted largest first, each with its codeword (cw_sizes[i] elements, the
 * rate from or_whir_rates), its dup-left tree, claim evals[i] and eq
 * weights betas[i]; rate = codeword 0's.  Returns the word count or a
 * negative status. */
/* aggregate_claims (src/Whir.cpp:608-668) with arr, b_aggr empty on entry:
 * npoints claim points of log2 n elements each; beta_ch[npoints],
 * arr_ch[narrays] receive the libc-drawn challenges. */
/* _Whir_Open (src/Whir.cpp:469-545): one array with explicit weights b and
 * claim K, opened against its dup-left tree (nodes) or, nodes NULL, the
 * all-zero-leaf tree of zero_leaves leaves that aggregate_commitments
 * leaves; chal (log2 n) = total_r, final_claim = MLE(b)(total_r); the record
 * has or_whir_batch_open's layout */
int64_t or_whir_open_weights(const or_fe *arr, uint64_t n, const or_fe *b, or_fe K, const or_fe *cw, uint64_t cw_n,
                             const uint8_t *nodes, uint64_t zero_leaves, int rate, int whir_k, uint64_t *record,
                             uint64_t record_cap, or_fe *chal, or_fe *final_claim);
void or_whir_aggregate_claims(int narrays, const or_fe *const *arrays, uint64_t n, int npoints, const or_fe *points,
                              or_fe *arr, or_fe *b_aggr, or_fe *K, or_fe *beta_ch, or_fe *arr_ch);
int64_t or_whir_batch_open(int count, const or_fe *const *arrays, const uint64_t *sizes,
                           const or_fe *const *codewords, const uint64_t *cw_sizes, const uint8_t *const *trees,
                           const or_fe *evals, const or_fe *const *betas, int rate, int whir_k,
                           uint64_t *record, uint64_t record_cap);

/* ---- coPCS rank-local kernels ---- */
/* Encode (src/coPCS.cpp:131-147) of nrows rows of L elements each: iNTT_L,
 * times w^-j with w = root(1 + log_rate_offset + log2 L), zero-pad to
 * rate*L, NTT.  out: nrows x rate*L, row major. */
void or_copcs_encode(const or_fe *rows, int nrows, uint64_t L, int rate, int log_rate_offset, or_fe *out);
/* distributed_MT at one rank (src/coPCS.cpp:152-183): the codeword matrix
 * has 32 rows (BATCH_FACTOR, src/New_PCS.cpp:9) of Lc elements, rows >=
 * nrows zero; column j -> MT_commit_Blake of its 32 elements (8 leaves) ->
 * the column root; create_tree_blake over the Lc column roots.  base_nodes
 * receives Lc + Lc/2 + ... + 1 digests (Com.Base_MT, level 0 first). */
void or_distributed_mt(const or_fe *cw, int nrows, uint64_t Lc, uint8_t *base_nodes);
/* pcs_sumcheck_phase1 before aggregation (src/coPCS.cpp:547-569): poly (2d)
 * = sum over blocks j of NTT_2d(pad(iNTT_d(data block))) * same for B. */
void or_pcs_sumcheck_phase1(const or_fe *data, const or_fe *B, uint64_t M, int d, or_fe *poly);
/* pcs_sumcheck_phase2 (src/coPCS.cpp:571-587): data[j] = sum_k L[k]
 * data[jd+k] (same for B), L = compute_lagrange_coeff(r, d); the first M/d
 * entries of data and B are overwritten. */
void or_pcs_sumcheck_phase2(or_fe *data, or_fe *B, uint64_t M, int d, or_fe r);
/* RS_multifold5 (src/coPCS_utils.cpp:386-437): out has n/fd elements. */
void or_rs_multifold5(const or_fe *cw, uint64_t n, const or_fe *a, int fd, or_fe *out);
/* encode_concatenated (src/coPCS.cpp:653-677): cprime = (data || 0) of
 * length rate0*n, forward NTT; with S = rate0*n/degree, segment j < S takes
 * cprime[j + k*S] for k < degree, zero-pads to degree*rate1 and takes the
 * forward NTT.  C: S rows of degree*rate1 elements, row-major. */
void or_encode_concatenated(const or_fe *data, uint64_t n, int rate0, int rate1, int degree, or_fe *cprime, or_fe *C);
/* RS_multifold_segment (src/coPCS_utils.cpp:172-230): the fold of one
 * fd-element segment cw[idx + s*(n/fd)], s < fd, by the weights L. */
or_fe or_rs_multifold_segment(const or_fe *segment, const or_fe *L, int fd, uint64_t n, uint64_t idx);
/* The batching step of the opening (src/coPCS.cpp:1349-1353):
 * out[k] = sum_{i < nrows} c[i] * rows[i][k] (rows row-major, len each). */
void or_row_combine(const or_fe *rows, int nrows, uint64_t len, const or_fe *c, or_fe *out);
/* cubic_coSumcheck_phase1's block sum (src/coSumcheck.cpp:9-45), before the
 * mask and sumcheck_aggr: H (4d elements). */
void or_cosumcheck3_phase1(const or_fe *p1, const or_fe *p2, const or_fe *p3, const or_fe *b1, uint64_t M, int d,
                           or_fe *H);
/* cubic_coSumcheck_phase2 (src/coSumcheck.cpp:64-148), in place, weights L. */
void or_cosumcheck3_phase2(or_fe *p1, or_fe *p2, or_fe *p3, or_fe *b1, uint64_t M, int d, const or_fe *L);
/* unpack_secret_share (src/MPI_utils.cpp:2769-2785): secret (n shares) is
 * transformed in place as the reference leaves it; v receives kk values. */
void or_unpack_secret_share(or_fe *secret, int n, int kk, or_fe *v);
/* compute_secret_shares (src/MPI_utils.cpp:2738-2767) for M values, k per
 * group, buffers of kk, N parties: the pads (entries k..kk-1 of each
 * group's buffer, (M/k) * (kk - k) elements, group-major) come from `pads`
 * (the reference draws random(); NULL gives zeros, as its non-private
 * branch).  shares: N rows of M/k elements. */
void or_compute_secret_shares(const or_fe *v, uint64_t M, int N, int k, int kk, const or_fe *pads, or_fe *shares);
/* rank 0's combine in sumcheck_aggr (src/MPI_utils.cpp:2604-2645): polys
 * N rows of len (rank-major), b N/4 weights; out len elements. */
void or_sumcheck_aggr(const or_fe *polys, int N, int len, const or_fe *b, or_fe *out);

/* ---- coSumcheck drivers (src/coSumcheck.cpp:233-448), all N ranks in
 * lockstep: rank m's shares p1[m], p2[m] (, p3[m]) of M elements are folded
 * in place, R[m][i] (R_sizes[i] elements: 2 d_i quadratic, 4 d_i cubic) are
 * rank m's masks (modified as the reference modifies its copy); the
 * gathers of sumcheck_aggr and send_final_sumcheck_poly read the other
 * ranks' values; prev is the transcript every rank shares.  Outputs: the
 * claims (5 quadratic, 4 cubic), r_out = sumcheck_r (ndeg + 1: the closing
 * challenge first), ip_vec (sum of the R sizes) and the reference's
 * printed (not fatal) check failures as a bit mask: bit i round i's
 * eval != K, bit ndeg the final check, bit ndeg + 1 the cubic's
 * verification 2.  N >= 4 a power of two; 0 or -1. */
int or_quadratic_cosumcheck(int N, or_fe **p1, or_fe **p2, uint64_t M, or_fe ***R, const uint64_t *R_sizes, int nR,
                            or_fe K, or_fe K_r, const int *degrees, int ndeg, or_fe *prev,
                            or_fe *claims, or_fe *r_out, or_fe *ip_vec, int *failures);
int or_cubic_cosumcheck(int N, or_fe **p1, or_fe **p2, or_fe **p3, uint64_t M, const or_fe *r1, int nr1,
                        const or_fe *r2, int nr2, or_fe ***R, const uint64_t *R_sizes, int nR, or_fe K, or_fe K_r,
                        const int *degrees, int ndeg, or_fe *prev, or_fe *claims, or_fe *r_out, or_fe *ip_vec,
                        int *failures);

/* prepare_data (src/coPCS.cpp:76-128), BATCH_FACTOR 32, _queries[0] = q0:
 * rows (32 x n/16) as the reference fills them; the non-empty row count or
 * -1 where it exits or reads past its inputs */
int or_prepare_data(const or_fe *shares, uint64_t n, const or_fe *rnd, uint64_t nr, int q0, or_fe *rows);

/* ---- all-cores variants (oracle_mt.c, OpenMP): the same bytes as the
 * single-thread functions above, `threads` OpenMP threads ---- */
void or_fft_mt(or_fe *arr, int logn, int inverse, int threads);            /* utils.cpp:904-964, 1583-1673 */
void or_rs_encode_mt(const or_fe *arr, uint64_t n, int rate, or_fe *cw, int threads);  /* Whir.cpp:158-163 */
void or_merkle_commit_mt(const or_fe *elems, uint64_t n_elems, uint8_t *nodes, int dup_left, int threads);
int or_product_sumcheck_mt(or_fe y, const or_fe *v1, const or_fe *v2, uint64_t n, or_fe *prev, or_fe *polys,
                           or_fe *r, or_fe *final_out, int threads);
int or_plain_proof_mt(const or_fe *f, const or_fe *g, uint64_t n, int rate, or_fe *prev, uint8_t *proof,
                      double *stage_ms, int threads);
void or_update_betas_mt(or_fe *b, uint64_t n, const or_fe *ch, const or_fe *w, int l, int threads);
int or_max_threads(void);
/* thread count of or_whir_open / or_whir_batch_open's encodes, trees and
 * weight updates (1, the default: single-thread) */
void or_whir_set_threads(int threads);

#ifdef __cplusplus
}
#endif
#endif
