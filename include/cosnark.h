/*
 * cosnark.h -- C ABI of libcosnark.so, the gfx950 (MI355X) implementation of
 * one prover rank's local work in the code-based collaborative SNARK.
 *
 * Plain C, POD arguments only: device buffers are caller-owned device
 * pointers (hipMalloc'd or cosnark_dev_alloc'd), host buffers are ordinary
 * pointers.  Field elements are 16 bytes {u64 real; u64 img;} exactly as the
 * reference's virgo::fieldElement (src/fieldElement.hpp:96-97), so a
 * std::vector<F> is passed as (uint64_t*)v.data() with no copy.  Digests are
 * 32-byte BLAKE3 outputs (struct _hash, src/Blake3_hash.h:3-5).
 *
 * Every entry point returns 0 on success or a negative COSNARK_E_* code; no
 * exception crosses the boundary.  cosnark_last_error() gives the message.
 * Work is issued on the context's HIP stream; unless stated otherwise a call
 * returns after its results are complete.  One context per host thread.
 *
 * Citations name the reference function each entry point replaces
 * (paths relative to the reference repository root).
 */
#ifndef COSNARK_H
#define COSNARK_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define COSNARK_OK        0
#define COSNARK_E_ARG    -1   /* bad argument (size not a power of two, ...) */
#define COSNARK_E_HIP    -2   /* HIP runtime error */
#define COSNARK_E_NOMEM  -3   /* device allocation failed */
#define COSNARK_E_CHECK  -4   /* prover self-check failed where the
                                 reference prints an error and exits */
#define COSNARK_E_RANGE  -5   /* size beyond what this build supports */

typedef struct cosnark_ctx cosnark_ctx;

/* ---- context ------------------------------------------------------- */
int  cosnark_create(int device, cosnark_ctx **out);
void cosnark_destroy(cosnark_ctx *ctx);
const char *cosnark_last_error(const cosnark_ctx *ctx);
int  cosnark_sync(cosnark_ctx *ctx);
/* the hipStream_t the context issues work on */
void *cosnark_stream(cosnark_ctx *ctx);
/* plain prover scheduling: 1 (default) runs the sumcheck on a second stream
 * beside the commit, 0 runs the stages one after another (for profiling) */
int  cosnark_set_overlap(cosnark_ctx *ctx, int on);
/* library build string (arch, version) */
const char *cosnark_version(void);

/* ---- device memory helpers (for callers with no allocator of their own) */
int cosnark_dev_alloc(cosnark_ctx *ctx, uint64_t bytes, void **out);
int cosnark_dev_free(cosnark_ctx *ctx, void *p);
int cosnark_h2d(cosnark_ctx *ctx, void *d_dst, const void *h_src, uint64_t bytes);
int cosnark_d2h(cosnark_ctx *ctx, void *h_dst, const void *d_src, uint64_t bytes);
int cosnark_d2d(cosnark_ctx *ctx, void *d_dst, const void *d_src, uint64_t bytes);

/* ---- host field operations (GF((2^61-1)^2), canonical in and out) -------
 * operator* and operator+ of virgo::fieldElement, src/fieldElement.cpp:34-78;
 * for the few host-side combinations around the kernels (sumcheck masks). */
void cosnark_field_mul(const uint64_t a[2], const uint64_t b[2], uint64_t out[2]);
void cosnark_field_add(const uint64_t a[2], const uint64_t b[2], uint64_t out[2]);
/* blake3_hash(src, dst) (src/Blake3_hash.cpp:5-9): BLAKE3 of one 64-byte
 * block, 32-byte digest (host) */
void cosnark_blake3_64(const uint8_t in[64], uint8_t out[32]);
/* F::_random() (src/fieldElement.cpp:119-125, 363-368): two 20-digit
 * decimal draws from libc random() on the calling thread, reduced mod p */
void cosnark_field_random(uint64_t out[2]);

/* ---- synthetic workload inputs ---------------------------------------
 * n elements uniform in GF(p^2): splitmix64-seeded xoshiro256**, each
 * coordinate masked to 61 bits with the value p rejected.  Host buffer. */
void cosnark_synth_fill(uint64_t seed, uint64_t stream, uint64_t *h_out, uint64_t n);

/* ---- NTT: fft(arr, logn, inverse), src/utils.cpp:904-964 --------------
 * In place on d_arr (2^logn elements), natural order in and out.  Forward
 * uses getRootOfUnity(logn) (src/utils.cpp:706-717); inverse uses its
 * inverse and scales by 2^-logn.  Also covers pfft (src/utils.cpp:1583),
 * which computes the same forward transform.  1 <= logn <= 28. */
int cosnark_ntt(cosnark_ctx *ctx, uint64_t *d_arr, int logn, int inverse);

/* ---- Reed-Solomon encode, the codeword of Whir_Commit_initial ----------
 * src/Whir.cpp:158-163: codeword = pfft(coeffs || 0) of length rate*n.
 * d_coeffs (n elements) is read only; d_codeword holds rate*n elements. */
int cosnark_rs_encode(cosnark_ctx *ctx, const uint64_t *d_coeffs, uint64_t n,
                      int rate, uint64_t *d_codeword);

/* ---- Merkle tree: MT_commit_Blake + create_tree_blake -----------------
 * src/merkle_tree.cpp:195-233 and :267-299.  Leaf i = BLAKE3 of elements
 * 4i..4i+3 (64 raw bytes).  All levels are written back to back into
 * d_nodes (level 0 first, root last), cosnark_merkle_node_count(n) digests.
 * dup_left = 1 reproduces the reference (a parent hashes its left child
 * twice, :287-292); dup_left = 0 hashes left||right.  n_elems: power of
 * two, 8 <= n_elems <= 2^30. */
uint64_t cosnark_merkle_node_count(uint64_t n_elems);
int cosnark_merkle_commit(cosnark_ctx *ctx, const uint64_t *d_elems, uint64_t n_elems,
                          uint8_t *d_nodes, int dup_left);
/* create_tree_blake (src/merkle_tree.cpp:267-299) over given level-0
 * digests: d_nodes holds n_leaves digests (level 0, a power of two); levels
 * 1.. are written after them, root last (2 n_leaves - 1 digests). */
int cosnark_merkle_tree_from_leaves(cosnark_ctx *ctx, uint64_t n_leaves, uint8_t *d_nodes, int dup_left);
/* open_tree_blake, src/merkle_tree.cpp:320-335: sibling digest per level
 * below the root into h_path (32 * depth bytes); *depth = log2(n_elems/4). */
int cosnark_merkle_open(cosnark_ctx *ctx, const uint8_t *d_nodes, uint64_t n_elems,
                        uint64_t pos, uint8_t *h_path, int *depth);

/* Authentication-path verification, q paths in one launch: what
 * verify_claim_opt_blake (src/merkle_tree.cpp:337-383) stands for.  The
 * reference's version hashes with my_hhash, reads siblings from the prover's
 * tree and returns true without comparing against the root; this one checks
 * each path against the root.  Path j: the leaf's 4 elements at
 * d_leaves[4j..4j+3] (BLAKE3 of their 64 bytes), its position d_pos[j]
 * (< 2^depth), depth sibling digests at d_paths + 32*depth*j (level 0 first,
 * as cosnark_merkle_open gives them), parents H(left || left) when dup_left
 * (the reference's trees: there the root depends on leaf 0 alone, so a
 * path at position p checks its leaf only when p = 0 and its level-L
 * sibling only when p >> L = 1) or H(left || right).  d_ok[j] = 1 when the path reaches d_root (32
 * device bytes); *h_failures counts the others. */
int cosnark_merkle_verify(cosnark_ctx *ctx, const uint64_t *d_leaves, const uint32_t *d_pos, const uint8_t *d_paths,
                          int depth, uint64_t q, const uint8_t *d_root, int dup_left, uint8_t *d_ok,
                          uint64_t *h_failures);

/* ---- Fiat-Shamir transcript, src/Fiat_Shamir.cpp ----------------------
 * The reference keeps one global running value prev_challenge (starts at 0,
 * Fiat_Shamir.cpp:5); here it is explicit state owned by the caller. */
typedef struct { uint64_t prev[2]; } cosnark_transcript;
/* hash_to_field (:29-34): absorbs elems in order, returns the new state */
int cosnark_hash_to_field(cosnark_transcript *ts, const uint64_t *h_elems, uint64_t n,
                          uint64_t out[2]);
/* get_indexes (:37-56): l pairs (prev.real % N, prev.img % M) into idx[2l] */
int cosnark_get_indexes(cosnark_transcript *ts, int l, uint32_t N, uint32_t M, uint32_t *idx);

/* ---- product sumcheck ---------------------------------------------------
 * _quadratic_sumcheck(y, v1, v2, N = 1), src/Distributed_Sumcheck.cpp:204-348
 * with the tail of quadratic_sumcheck, src/sparse_eval.cpp:140-192.
 * n = 2^m elements per vector (read only).  Round i (i < m - offset) takes
 * its challenge from the transcript (hash_to_field({a,b,c})); the last
 * `offset` rounds draw F::_random() four times each from libc random()
 * (src/fieldElement.cpp:119-125,363-368), as the reference's
 * sumcheck_offset rounds do.  Outputs: h_polys[6*i..6*i+5] = (a,b,c) of
 * round i's a x^2 + b x + c; h_r[2*i..] = challenge i; h_final = {v1[0],
 * v2[0]} after all folds.  Returns COSNARK_E_CHECK if some round has
 * H(0) + H(1) != y (the reference exits there, :243-247). */
int cosnark_product_sumcheck(cosnark_ctx *ctx, const uint64_t *d_v1, const uint64_t *d_v2,
                             uint64_t n, const uint64_t y[2], cosnark_transcript *ts,
                             int offset, uint64_t *h_polys, uint64_t *h_r,
                             uint64_t h_final[4]);

/* ---- plain single-rank prover -------------------------------------------
 * The rank-local pipeline of one proof: RS-encode f at blowup `rate`
 * (Whir_Commit_initial form), Merkle-commit the codeword (MT_commit_Blake,
 * dup-left), then the all-Fiat-Shamir product sumcheck of f*g with claimed
 * sum y = sum_x f(x) g(x).  Roots are not absorbed into the transcript
 * (the reference never absorbs them).  Proof bytes (cosnark_plain_proof_bytes
 * of them): root (32) | m rounds x (a,b,c) (48 each) | v1[0], v2[0] (32).
 * times (optional, 4 doubles): encode, merkle, sumcheck, total ms measured
 * with HIP events on the context stream. */
uint64_t cosnark_plain_proof_bytes(uint64_t n);
int cosnark_prove_plain(cosnark_ctx *ctx, const uint64_t *d_f, const uint64_t *d_g,
                        uint64_t n, int rate, cosnark_transcript *ts,
                        uint8_t *h_proof, double *times);
/* The same proof, keeping the commitment: the codeword (rate*n elements)
 * is written to d_codeword and every tree level (cosnark_merkle_node_count(
 * rate*n) digests, level 0 first, root last) to d_nodes, both caller-owned
 * device buffers, as Whir_Commit_initial (src/Whir.cpp:158-190) leaves the
 * codeword and MT_commit_Blake (src/merkle_tree.cpp:195-233) the tree for
 * the openings that follow.  Either may be null (that part then stays in
 * the context's scratch).  Proof bytes equal cosnark_prove_plain's. */
int cosnark_prove_plain_ex(cosnark_ctx *ctx, const uint64_t *d_f, const uint64_t *d_g,
                           uint64_t n, int rate, cosnark_transcript *ts,
                           uint8_t *h_proof, uint64_t *d_codeword, uint8_t *d_nodes,
                           double *times);
/* The plain prover of cosnark_prove_plain_ex with no device allocation.
 * No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  The caller passes the inputs and both outputs
 * (d_codeword: rate*n elements, d_nodes: cosnark_merkle_node_count(rate*n)
 * digests, both required); the NTT temporary and the sumcheck's fold
 * buffers come from a fixed scratch array of
 * cosnark_prove_plain_noalloc_max_bytes() bytes (96 MiB) that each GPU holds
 * until the process ends.  Sizes needing more (rate*n + 1.5 n elements of 16
 * bytes, rounded up to 256 bytes between the two parts: n = 2^20 at rate 4)
 * return COSNARK_E_RANGE.  All calls on one GPU share that GPU's scratch:
 * calls that overlap on one GPU (several host threads or contexts) race.
 * Proof, codeword and tree equal cosnark_prove_plain_ex's. */
uint64_t cosnark_prove_plain_noalloc_max_bytes(void);
int cosnark_prove_plain_noalloc(cosnark_ctx *ctx, const uint64_t *d_f, const uint64_t *d_g,
                                uint64_t n, int rate, cosnark_transcript *ts, uint8_t *h_proof,
                                uint64_t *d_codeword, uint8_t *d_nodes);
/* The verifier of a plain proof's sumcheck (the checks the reference's
 * prover makes, src/Distributed_Sumcheck.cpp:243-248): with y the claimed
 * sum and the transcript t as the prover started it, each round's
 * polynomial (a, b, c) must be canonical and satisfy H(0) + H(1) = y; then
 * r = hash_to_field({a, b, c}) and y = H(r); at the end y = v1[0] v2[0].
 * Host only.  h_r (optional) receives the m challenges; *bad_round (optional)
 * the failing round (m for the final product).  COSNARK_E_CHECK on failure.
 * The root is not checked: a plain proof carries no openings. */
int cosnark_verify_plain(const uint8_t *h_proof, uint64_t n, const uint64_t y[2], cosnark_transcript *t,
                         uint64_t *h_r, int *bad_round);

/* ---- WHIR (whir_k = 4), src/Whir.cpp -----------------------------------
 * Whir_Commit_initial (:158-190): d_codeword = RS encode of d_arr at `rate`
 * (rate*n elements), d_nodes = its Merkle tree (MT_commit_Blake, 4 elements
 * per leaf, cosnark_merkle_node_count(rate*n) digests). */
int cosnark_whir_commit(cosnark_ctx *ctx, const uint64_t *d_arr, uint64_t n, int rate,
                        uint64_t *d_codeword, uint8_t *d_nodes);
/* The claim of Whir_Open (:392-395): K = sum_x arr[x] * eq(z, x), with the
 * eq table of precompute_beta (src/utils.cpp:198-243); h_z holds log2(n)
 * elements. */
int cosnark_whir_claim(cosnark_ctx *ctx, const uint64_t *d_arr, uint64_t n, const uint64_t *h_z,
                       uint64_t K[2]);
/* Whir_Open (:388-467) over one claim (K, z), with the commitment from
 * cosnark_whir_commit; d_arr is read only.  Draws its challenges from libc
 * random()/rand() in the reference's order (seed with srandom).  Iterates
 * until |arr| <= max(4, 2^stop_log) (stop_log <= 2: the reference's loop).
 * Writes the opening record (u64 words, *words of them) to h_record:
 *   [iterations]; per iteration: 2 x (a,b,c), 2 challenges, new root (4
 *   words), l, l positions, l x 4 queried values, l folded values, l auth
 *   paths (log2(|cw|/4) digests each), update_betas challenge; then the final
 *   arr, K, the final commitment root and the closing challenges.
 * Returns COSNARK_E_CHECK when the claim chain fails the reference's own
 * checks (Whir.cpp:238-241, 454-457). */
int cosnark_whir_open(cosnark_ctx *ctx, const uint64_t *d_arr, uint64_t n, const uint64_t *h_z,
                      const uint64_t K[2], int rate, int stop_log, const uint64_t *d_codeword,
                      const uint8_t *d_nodes, uint64_t *h_record, uint64_t record_cap,
                      uint64_t *words);
/* get_whir_rates (src/Whir.cpp:18-137) without its printing: for arrays of
 * sizes[0..count) sorted largest first, the rate rates[i] each is committed
 * at so that Whir_Batch_Open meets it with a codeword of the running size;
 * the first is `rate`.  Host only.  COSNARK_E_ARG for unsorted sizes. */
int cosnark_whir_rates(const uint64_t *sizes, int count, int rate, int whir_k, int *rates);
/* Whir_Batch_Open (src/Whir.cpp:706-899): one opening for count arrays of
 * sizes sorted largest first.  d_arrays[i], d_betas[i] (sizes[i] elements:
 * the array and its eq weights), d_codewords[i] (cw_sizes[i] elements, the
 * encoding at cosnark_whir_rates' rate) and d_nodes[i] (its dup-left tree as
 * cosnark_whir_commit writes it) are device pointers; h_evals holds the
 * claims (only entry 0 is read: the reference recomputes the others by
 * aggregate_arrays, :670-703).  The fold schedule (:736-755) runs Whir_Iteration
 * with fold factor kappa, and where an array size is met, aggregate_arrays
 * and aggregate_commitments (:547-606) merge the arrays of that size into the
 * running one; aggregate_commitments leaves a tree over all-zero leaf digests
 * (the reference never fills its leaf vector), which the next iteration's
 * queries open.  After an aggregation at the last scheduled step the
 * reference reads past its kappa list (:826); the last kappa is used here.
 * Challenges come from libc random()/rand() in the reference's order
 * (seed first).  The record (u64 words): [steps], per step [kind]: 0 an
 * iteration: kappa, log2(kappa) x (a,b,c), the log2(kappa) challenges, new
 * root, l, l positions, l x kappa replies, l folds, l auth paths, the
 * update_betas challenge; 1 an aggregation: cnt, challenges1[cnt],
 * challenges2[cnt] (entry 0 = 1), evals[cnt x cnt] = <arr_i, b_j>, K, l,
 * l positions, l auth paths on the running tree, the new tree's root; then
 * the final arr, K, the final root (Whir_Commit over zero digests) and the
 * closing challenges.  COSNARK_E_CHECK when one of the reference's checks
 * fails (a round, the closing K = <arr, b>, the closing codeword = RS(arr))
 * or where the reference exits (a query beyond the tree, codeword sizes
 * that differ at an aggregation).  whir_k a power of two in [2, 64]. */
/* _Whir_Open (src/Whir.cpp:469-545): the prover work of one opening of
 * d_arr (n elements) with explicit weights d_b (n elements; Whir_Open,
 * :388-412, builds them from its claims) and claim K, at fold factor whir_k
 * (a power of two in [2, 64]; the reference's global, 8 by default), down to
 * |arr| <= whir_k.  d_codeword (cw_n elements, the encoding at `rate`) is
 * opened against its dup-left tree d_nodes (as cosnark_whir_commit writes
 * it), or, when d_nodes is null, against the tree over zero_leaves all-zero
 * leaf digests that aggregate_commitments (:547-606) leaves.  Challenges
 * come from libc random()/rand() in the reference's order.  The record has
 * cosnark_whir_batch_open's layout (one array: iterations only).
 * h_challenges receives total_r (log2 n elements: every round's challenge,
 * then the closing ones) and final_claim the multilinear evaluation of the
 * initial weights at total_r: the pair _Whir_Open returns (:517-518).
 * COSNARK_E_CHECK where the reference's checks fail. */
int cosnark_whir_open_weights(cosnark_ctx *ctx, const uint64_t *d_arr, uint64_t n, const uint64_t *d_b,
                              const uint64_t K[2], const uint64_t *d_codeword, uint64_t cw_n,
                              const uint8_t *d_nodes, uint64_t zero_leaves, int rate, int whir_k,
                              uint64_t *h_record, uint64_t record_cap, uint64_t *words,
                              uint64_t *h_challenges, uint64_t final_claim[2]);
/* precompute_beta(r, B), src/utils.cpp:198-243: the 2^m eq weights
 * B[x] = prod_k (bit_k(x) ? r_k : 1 - r_k) of the point h_z (m elements,
 * r_0 on bit 0) into d_out.  1 <= m <= 28. */
int cosnark_precompute_beta(cosnark_ctx *ctx, const uint64_t *h_z, int m, uint64_t *d_out);
/* aggregate_claims (src/Whir.cpp:608-668) for claims on narrays arrays of n
 * elements (device pointers): the eq weights beta_j of each of the npoints
 * claim points (h_points: npoints x log2 n elements, point j's coordinates
 * as precompute_beta takes them), Sums[i][j] = <beta_j, arrays_i>, the
 * challenges h_beta_ch[npoints] then h_arr_ch[narrays] from libc (in that
 * order), d_b_aggr = sum_j beta_ch_j beta_j, d_arr = sum_i arr_ch_i arrays_i
 * and K = sum_ij beta_ch_j arr_ch_i Sums[i][j].  The reference adds to arr
 * and b_aggr as given (its callers pass empty vectors); here they are
 * overwritten.  1..256 arrays and points. */
int cosnark_whir_aggregate_claims(cosnark_ctx *ctx, int narrays, const uint64_t *const *d_arrays, uint64_t n,
                                  int npoints, const uint64_t *h_points, uint64_t *d_arr, uint64_t *d_b_aggr,
                                  uint64_t K[2], uint64_t *h_beta_ch, uint64_t *h_arr_ch);
int cosnark_whir_batch_open(cosnark_ctx *ctx, int count, const uint64_t *const *d_arrays, const uint64_t *sizes,
                            const uint64_t *const *d_codewords, const uint64_t *cw_sizes,
                            const uint8_t *const *d_nodes, const uint64_t *h_evals, const uint64_t *const *d_betas,
                            int rate, int whir_k, uint64_t *h_record, uint64_t record_cap, uint64_t *words);

/* ---- coPCS rank-local kernels --------------------------------------------
 * Encode, src/coPCS.cpp:131-147: each of the nrows rows of L elements
 * (row-major in d_rows) -> iNTT_L, times w^-j with w = getRootOfUnity(1 +
 * log_rate_offset + log2 L), zero-padded to rate*L, NTT.  log_rate_offset is
 * log2 of the reference's global rate_offset (src/coPCS.cpp:22,46).
 * d_codewords: nrows x rate*L elements, row-major.  L >= 2 and rate are
 * powers of two; rate*L <= 2^28. */
int cosnark_copcs_encode(cosnark_ctx *ctx, const uint64_t *d_rows, int nrows, uint64_t L, int rate,
                         int log_rate_offset, uint64_t *d_codewords);
/* PCS_commit at one rank (src/coPCS.cpp:236-250): prepare_data
 * (:76-128) lays this rank's n_shares shares and its random mask values out
 * in rows of B = n_shares / 16 elements (BATCH_FACTOR / 2,
 * src/New_PCS.cpp:9): each row takes B - q0 shares and then q0 random
 * values until the shares run out (the rest of that row's share part is
 * zero), and one more row takes random values only (B of them while at
 * least B + 1 shares exceed the random values used, else B - (n_shares -
 * used)); q0 is the reference's _queries[0] (333 at rate 4).  The rows
 * (*nrows of them, at most 32) are written to d_rows, encoded with
 * cosnark_copcs_encode at `rate` and log_rate_offset into d_codewords
 * (*nrows x rate B) and committed with cosnark_distributed_mt into d_nodes
 * (2 rate B - 1 digests).  d_rows holds 32 B elements, d_codewords 32 rate B.
 * COSNARK_E_RANGE where the reference exits (more than 32 rows needed). */
int cosnark_pcs_commit(cosnark_ctx *ctx, const uint64_t *d_shares, uint64_t n_shares, const uint64_t *d_random,
                       uint64_t n_random, int q0, int rate, int log_rate_offset, uint64_t *d_rows,
                       uint64_t *d_codewords, uint8_t *d_nodes, int *nrows);
/* prepare_data's row count and row size for these sizes, without a GPU */
int cosnark_pcs_prepare_rows(uint64_t n_shares, uint64_t n_random, int q0, int *nrows, uint64_t *batch_size);
/* The segments and folds of delegate_checks (src/coPCS.cpp:740-760):
 * for each query q < l, segment q = d_cw[h_idx[q] + s * (n/fd)], s < fd,
 * and h_y[q] = RS_multifold_segment(segment q, h_L, n, h_idx[q])
 * (src/coPCS_utils.cpp:172-230), which equals RS_multifold5's output at
 * h_idx[q].  h_segments (optional): l x fd elements.  h_idx[q] < n/fd;
 * fd and n powers of two, 2 <= fd <= min(n, 4096); l <= 65536. */
int cosnark_multifold_segments(cosnark_ctx *ctx, const uint64_t *d_cw, uint64_t n, const uint64_t *h_L, int fd,
                               const uint32_t *h_idx, int l, uint64_t *h_segments, uint64_t *h_y);
/* The batching step of the opening, src/coPCS.cpp:1349-1353:
 * d_out[k] = sum_{i < nrows} h_c[i] * d_rows[i * len + k] (rows back to
 * back; the reference combines the first batches - 1 rows of the codeword
 * and of the committed data this way).  1 <= nrows <= 32. */
int cosnark_row_combine(cosnark_ctx *ctx, const uint64_t *d_rows, int nrows, uint64_t len, const uint64_t *h_c,
                        uint64_t *d_out);
/* encode_concatenated, src/coPCS.cpp:653-677: d_cprime = (data || 0) of
 * rate0*n elements, forward NTT (RS encode); with S = rate0*n/degree,
 * segment j < S takes cprime[j + k*S] for k < degree, zero-pads to
 * degree*rate1 and takes the forward NTT: row j of d_C (S rows of
 * degree*rate1 elements, row-major; commit_concatenated, :679-689, hashes
 * d_C as one vector with cosnark_merkle_commit).  n, rate0, rate1, degree
 * powers of two, degree <= rate0*n <= 2^28, 2 <= degree*rate1 <= 4096. */
int cosnark_encode_concatenated(cosnark_ctx *ctx, const uint64_t *d_data, uint64_t n, int rate0, int rate1,
                                int degree, uint64_t *d_cprime, uint64_t *d_C);
/* distributed_MT at one rank, src/coPCS.cpp:152-183: the codeword matrix has
 * 32 rows (BATCH_FACTOR, src/New_PCS.cpp:9) of Lc elements; d_cw holds the
 * first nrows (row-major), the rest are zero.  Column j is committed with
 * MT_commit_Blake (32 elements, 8 leaves) and its root is leaf j of a
 * create_tree_blake tree.  d_nodes receives Com.Base_MT: Lc + Lc/2 + ... + 1
 * digests, level 0 first (2 Lc - 1).  With the reference's dup-left parents
 * only rows 0..3 reach a column root.  1 <= nrows <= 32; Lc a power of two
 * <= 2^28.  The all-to-all and rank-0 tree of roots (:186-230) span ranks. */
int cosnark_distributed_mt(cosnark_ctx *ctx, const uint64_t *d_cw, int nrows, uint64_t Lc, uint8_t *d_nodes);
/* Block sumcheck, phase 1: pcs_sumcheck_phase1 before pcs_sumcheck_aggr
 * (src/coPCS.cpp:547-569), equal to the block sum of
 * quadratic_coSumcheck_phase1 (src/coSumcheck.cpp:150-180) before its
 * r * H_r mask and sumcheck_aggr.  h_poly (2d elements) = sum over blocks j
 * of NTT_2d(pad(iNTT_d(p1 block j))) * NTT_2d(pad(iNTT_d(p2 block j))).
 * M = |p1| = |p2|; d and M powers of two, 2 <= d <= min(M, 2048). */
int cosnark_cosumcheck_phase1(cosnark_ctx *ctx, const uint64_t *d_p1, const uint64_t *d_p2, uint64_t M, int d,
                              uint64_t *h_poly);
/* Block sumcheck, phase 2: pcs_sumcheck_phase2 (src/coPCS.cpp:571-587),
 * equal to quadratic_coSumcheck_phase2 (src/coSumcheck.cpp:201-213) with
 * L = compute_lagrange_coeff(., r, d) (src/utils.cpp:169-173):
 * p[j] = sum_k L_k p[jd + k] for j < M/d, in place on both vectors (the
 * reference then shrinks them to M/d).  2 <= d <= min(M, 8192). */
int cosnark_cosumcheck_phase2(cosnark_ctx *ctx, uint64_t *d_p1, uint64_t *d_p2, uint64_t M, int d,
                              const uint64_t r[2]);
/* The same fold with caller-given weights h_L (d elements):
 * quadratic_coSumcheck_phase2(p1, p2, L, d), src/coSumcheck.cpp:201-213. */
int cosnark_cosumcheck_fold(cosnark_ctx *ctx, uint64_t *d_p1, uint64_t *d_p2, uint64_t M, int d,
                            const uint64_t *h_L);
/* Cubic block sumcheck, phase 1: the block sum of cubic_coSumcheck_phase1
 * (src/coSumcheck.cpp:9-45) before its r * H_r mask and sumcheck_aggr.
 * With h1, h2, h3, h4 = blocks of p1, p2, b1, p3 (:17-20), each iNTT_d,
 * zero-padded to 4d and NTT_4d, h_poly (4d elements) = sum over blocks of
 * (h1 h2 - h4) h3.  M = |p1| = |p2| = |p3| = |b1|; d and M powers of two,
 * 2 <= d <= min(M, 1024). */
int cosnark_cosumcheck3_phase1(cosnark_ctx *ctx, const uint64_t *d_p1, const uint64_t *d_p2, const uint64_t *d_p3,
                               const uint64_t *d_b1, uint64_t M, int d, uint64_t *h_poly);
/* Cubic block sumcheck, phase 2: cubic_coSumcheck_phase2(p1, p2, p3, b1, L, d)
 * (src/coSumcheck.cpp:64-148): p[j] = sum_k L_k p[jd + k] on each of the four
 * vectors, in place (an all-zero block gives 0, as the reference's skip
 * does).  2 <= d <= min(M, 8192). */
int cosnark_cosumcheck3_fold(cosnark_ctx *ctx, uint64_t *d_p1, uint64_t *d_p2, uint64_t *d_p3, uint64_t *d_b1,
                             uint64_t M, int d, const uint64_t *h_L);
/* The same with L = compute_lagrange_coeff(., r, d), as cubic_coSumcheck
 * derives it (src/coSumcheck.cpp:421). */
int cosnark_cosumcheck3_phase2(cosnark_ctx *ctx, uint64_t *d_p1, uint64_t *d_p2, uint64_t *d_p3, uint64_t *d_b1,
                               uint64_t M, int d, const uint64_t r[2]);

/* ---- rank 0's share arithmetic (host only, no context) --------------------
 * unpack_secret_share(secret, _k, k), src/MPI_utils.cpp:2769-2785: the n
 * shares in h_secret (values at w_2n^(2j+1) of a polynomial of degree < n)
 * are replaced, as the reference leaves them, by the polynomial's values at
 * the n-th roots of unity, and h_out[j] = h_secret[n j / kk] for j < kk.
 * n a power of two <= 8192, 1 <= kk <= n.  Returns COSNARK_E_ARG otherwise. */
int cosnark_share_unpack(uint64_t *h_secret, int n, int kk, uint64_t *h_out);
/* Rank 0's combine in sumcheck_aggr(poly, b, N), src/MPI_utils.cpp:2604-2645,
 * given the gathered polynomials: h_polys holds nranks rows of len elements
 * (row m = rank m's poly); coordinate i's shares are unpacked
 * (cosnark_share_unpack with kk = n = nranks) and
 * h_out[i] = sum_{j < nranks/4} b_j * unpacked[2j].  h_b holds nranks/4
 * elements (the reference's ones(N/4) or b2); for nranks < 4 the sum is
 * empty and h_out is zero, as in the reference.  nranks a power of two
 * <= 8192. */
int cosnark_sumcheck_aggr(const uint64_t *h_polys, int nranks, int len, const uint64_t *h_b, uint64_t *h_out);
/* ---- coSumcheck drivers (src/coSumcheck.cpp:233-448), one rank of N ----
 * The exchanges of the reference -- sumcheck_aggr's gather to rank 0 and
 * broadcast of the combination (src/MPI_utils.cpp:2604-2670),
 * send_final_sumcheck_poly's (:2672-2736) -- are one all-gather supplied by
 * the caller (MPI, RCCL, or threads in one process): the callback receives
 * this rank's `elems` elements in `mine` and writes every rank's, rank
 * after rank (nranks * elems elements), to `all`; it returns 0 on success.
 * Every rank then forms rank 0's combination itself. */
typedef int (*cosnark_allgather_fn)(void *user, const uint64_t *mine, uint64_t elems, uint64_t *all);
/* quadratic_coSumcheck(p1, p2, R, K, K_r, degrees, N, ps), :233-331, as
 * rank `rank` of nranks (a power of two >= 4: the reference divides by
 * N/4).  d_p1, d_p2: this rank's shares, M elements each, folded in place
 * (M / prod(degrees) >= 2 elements remain; the reference resizes them).
 * h_R[i] (R_sizes[i] = 2 degrees[i] elements, i < nR = ndeg): this rank's
 * masks, read only (the reference works on its copy).  The transcript ts
 * absorbs {K, K_r} first (c = hash_to_field({K, K_r})).  Per round: block
 * sums (cosnark_cosumcheck_phase1) plus c times the mask, the all-gather and
 * combination with ones(N/4), r = hash_to_field(H), ip_vec, the sparse mask
 * fold (quadratic_coSumcheck_phase2_sparse, :215-228), the fold by
 * compute_lagrange_coeff(r, d).  Then {p1[0], p2[0], p1[1], p2[1],
 * final_eval_r} are gathered and unpacked to N/2 values each.  Outputs:
 * h_claims (5 elements: the claims' values), h_r (ndeg + 1 elements: the
 * claims' common point, the closing challenge first), h_ip_vec (optional,
 * sum of 2 degrees[i] elements: the global ip_vec, round after round), ps
 * (optional: rank 0 adds the proof size in KiB as the reference does),
 * check_failures (optional): the reference's printed, not fatal, checks as
 * a bit mask -- bit i round i's sum != K, bit ndeg the final check. */
int cosnark_quadratic_cosumcheck(cosnark_ctx *ctx, uint64_t *d_p1, uint64_t *d_p2, uint64_t M,
                                 const uint64_t *const *h_R, const uint64_t *R_sizes, int nR, const uint64_t K[2],
                                 const uint64_t K_r[2], const int *degrees, int ndeg, int nranks, int rank,
                                 cosnark_allgather_fn gather, void *user, cosnark_transcript *ts, uint64_t *h_claims,
                                 uint64_t *h_r, uint64_t *h_ip_vec, double *ps, int *check_failures);
/* cubic_coSumcheck(p1, p2, p3, r1, r2, R, K, K_r, degrees, N, ps),
 * :335-448, as rank `rank` of nranks: as above with b1 = precompute_beta(r1)
 * (M = 2^nr1 elements, folded with the shares) in the cubic block sums
 * (cosnark_cosumcheck3_phase1), the combination weighted by b2 =
 * precompute_beta(r2) (2^nr2 >= N/4), masks of 4 degrees[i] elements,
 * final values {p1[0], p2[0], p3[0], final_eval_r} (M / prod(degrees) >= 1),
 * h_claims 4 elements, h_ip_vec sum of 4 degrees[i] elements, and
 * check_failures bit ndeg + 1 for the reference's verification 2
 * (b1[0] against the rounds' Lagrange weights). */
int cosnark_cubic_cosumcheck(cosnark_ctx *ctx, uint64_t *d_p1, uint64_t *d_p2, uint64_t *d_p3, uint64_t M,
                             const uint64_t *h_r1, int nr1, const uint64_t *h_r2, int nr2, const uint64_t *const *h_R,
                             const uint64_t *R_sizes, int nR, const uint64_t K[2], const uint64_t K_r[2],
                             const int *degrees, int ndeg, int nranks, int rank, cosnark_allgather_fn gather,
                             void *user, cosnark_transcript *ts, uint64_t *h_claims, uint64_t *h_r, uint64_t *h_ip_vec,
                             double *ps, int *check_failures);
/* RS_multifold5, src/coPCS_utils.cpp:386-437: folds the n-element codeword
 * d_cw by the fd challenges h_a (fd elements) into n/fd elements at d_out
 * (the reference replaces the codeword; here d_out must not alias d_cw).
 * fd and n powers of two, 2 <= fd <= min(n, 4096), n <= 2^28. */
int cosnark_rs_multifold5(cosnark_ctx *ctx, const uint64_t *d_cw, uint64_t n, const uint64_t *h_a, int fd,
                          uint64_t *d_out);

/* ---- batched plain proofs (no reference counterpart: the reference proves
 * one polynomial at a time) ---------------------------------------------
 * A batch holds `count` plain proofs (cosnark_prove_plain) of sizes
 * sizes[i] (powers of two, n >= 4, 8 <= n*rate <= 2^28) at one blowup.
 * cosnark_batch_create plans it once: device slots for every input and
 * intermediate, and the launch tables.  The caller writes proof i's f and g
 * into the slots cosnark_batch_input returns (sizes[i] elements each).
 * cosnark_batch_prove runs every proof: the RS encodes of one size share
 * each NTT launch, all trees share each Merkle launch and all sumchecks share
 * each round's launch.  The pass is captured into a hipGraph at the first
 * call and replayed afterwards (cosnark_batch_set_graph(b, 0): plain
 * launches).  ts[i] is proof i's transcript (in and out); h_proofs receives
 * the proofs back to back, proof i at cosnark_batch_proof_offset(b, i), each
 * byte-identical to cosnark_prove_plain's for the same inputs and
 * transcript.  ms (optional): device time of the pass. */
typedef struct cosnark_batch cosnark_batch;
int cosnark_batch_create(cosnark_ctx *ctx, int count, const uint64_t *sizes, int rate, cosnark_batch **out);
void cosnark_batch_destroy(cosnark_batch *b);
int cosnark_batch_input(cosnark_batch *b, int i, uint64_t **d_f, uint64_t **d_g);
uint64_t cosnark_batch_proof_offset(const cosnark_batch *b, int i);
/* Proof i's commitment as the last cosnark_batch_prove left it: its
 * codeword (rate * sizes[i] elements) and every level of its tree
 * (cosnark_merkle_node_count(rate * sizes[i]) digests, level 0 first, root
 * last), device pointers into the batch's memory, valid until the next
 * prove or the batch's destruction.  They equal cosnark_prove_plain_ex's
 * for the same input. */
int cosnark_batch_commitment(cosnark_batch *b, int i, uint64_t **d_codeword, uint8_t **d_nodes);
int cosnark_batch_set_graph(cosnark_batch *b, int on);
int cosnark_batch_prove(cosnark_batch *b, cosnark_transcript *ts, uint8_t *h_proofs, double *ms);

/* ---- proof server (no reference counterpart: the reference proves one
 * request per process run) ---------------------------------------------------
 * Continuous batching of plain proofs under a fixed HBM budget.  The server
 * owns a context on `device`, a scheduler thread and an arena of `budget`
 * bytes of device memory, from which every request's working memory and
 * every pass's tables come.  Each tick it retires finished passes, admits
 * queued requests in arrival order while their memory fits, and launches
 * them as one batched pass on a free slot (one of `slots` stream pairs), so
 * passes overlap on the device.  Requests are n = 2^k <= 2^max_log_n
 * elements at blowup `rate`; budget must hold `slots` NTT temporaries of
 * rate * 2^max_log_n elements plus one largest request.
 *   submit: d_f, d_g are device arrays on the same device (n elements each,
 *           read during the proof: keep them until wait returns); prev is
 *           the transcript start (null: zero).  Thread-safe.
 *   wait:   blocks until the proof is ready (one wait per ticket: a
 *           second wait on a ticket already being waited for returns
 *           COSNARK_E_ARG); h_proof receives
 *           cosnark_plain_proof_bytes(n) bytes, identical to
 *           cosnark_prove_plain's for the same inputs and transcript;
 *           latency_ms = completion - submission; releases the ticket.
 *   stats:  passes, completed requests, most requests in flight at once,
 *           peak arena bytes, arena bytes, queued requests. */
typedef struct cosnark_server cosnark_server;
int cosnark_server_create(int device, uint64_t budget, int rate, int max_log_n, int slots, cosnark_server **out);
int cosnark_server_submit(cosnark_server *s, const uint64_t *d_f, const uint64_t *d_g, uint64_t n,
                          const uint64_t prev[2], uint64_t *ticket);
/* submit_ex: as submit, and the request's commitment -- its codeword
 * (rate * n elements) and every level of its tree (cosnark_merkle_node_count(
 * rate * n) digests) -- is written to the caller's device buffers d_codeword
 * and d_nodes (both or neither) instead of the arena, complete when wait
 * returns, equal to cosnark_prove_plain_ex's. */
int cosnark_server_submit_ex(cosnark_server *s, const uint64_t *d_f, const uint64_t *d_g, uint64_t n,
                             const uint64_t prev[2], uint64_t *d_codeword, uint8_t *d_nodes, uint64_t *ticket);
int cosnark_server_poll(cosnark_server *s, uint64_t ticket, int *done);
int cosnark_server_wait(cosnark_server *s, uint64_t ticket, uint8_t *h_proof, uint64_t prev_out[2],
                        double *latency_ms);
int cosnark_server_stats(cosnark_server *s, uint64_t out[6]);
void cosnark_server_destroy(cosnark_server *s);

/* ---- NTT plan autotuner (no reference counterpart) -----------------------
 * Times every radix plan (largest pass radix 2^5..2^8, LDS tile 1024..4096
 * elements) for 2^logn-point transforms, logn_lo <= logn <= logn_hi, with
 * `batch` transforms per launch, and keeps the fastest per size in ctx.
 * Plans differ only in speed; outputs are identical.  best_ms (optional,
 * logn_hi - logn_lo + 1 doubles) receives the winning times.  Run it before
 * creating batches (a captured graph keeps the plan it was captured with). */
int cosnark_autotune(cosnark_ctx *ctx, int logn_lo, int logn_hi, int batch, double *best_ms);
int cosnark_ntt_plan(cosnark_ctx *ctx, int logn, int *passbits, int *tile);
/* Sets the plan for 2^logn-point transforms (as cosnark_autotune would):
 * passes of at most `passbits` bits (2..8), tiles of `tile` elements
 * (a power of two in 256..4096).  For callers that tune end to end, e.g.
 * the encode beside an overlapped sumcheck.  passbits = tile = 0 restores
 * the context defaults for that size. */
int cosnark_set_ntt_plan(cosnark_ctx *ctx, int logn, int passbits, int tile);
/* The Merkle schedule (no reference counterpart): a workgroup takes 2^b
 * first-level nodes of a tree and climbs b levels in LDS before the next
 * launch (b in [6, 10]; 10 by default), so a tree of 2^k leaves takes about
 * k / (b + 1) launches.  Fan-in stays 2 (the reference's binary dup-left
 * tree) and every b builds the same tree.  cosnark_autotune_merkle times a
 * batch of trees of the given leaf counts (built in shared launches, as the
 * batched prover builds them) under each b and keeps the fastest in ctx;
 * run it before creating batches. */
int cosnark_autotune_merkle(cosnark_ctx *ctx, const uint64_t *leaf_counts, int count, int *best_block_log,
                            double *best_ms);
int cosnark_merkle_schedule(cosnark_ctx *ctx, int *block_log);
int cosnark_set_merkle_schedule(cosnark_ctx *ctx, int block_log);
/* Measured throughput of the library's GF(p^2) multiply on this device
 * (independent multiply chains, no memory traffic): the compute ceiling the
 * transform kernels are reported against.  No reference counterpart. */
int cosnark_fmul_rate(cosnark_ctx *ctx, double *per_s);
/* Measured throughput of the library's single-block BLAKE3 compression on
 * this device (independent chains, no memory traffic): the ceiling the
 * Merkle and leaf-hashing kernels are reported against.  No reference
 * counterpart. */
int cosnark_blake3_rate(cosnark_ctx *ctx, double *per_s);

#ifdef __cplusplus
}
#endif
#endif
