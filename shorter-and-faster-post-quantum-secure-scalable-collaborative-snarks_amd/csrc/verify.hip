// verify.hip -- verifiers for what the prover emits.
//
// Merkle authentication paths.  The reference's verify_claim_opt_blake
// (src/merkle_tree.cpp:337-383) hashes with my_hhash rather than BLAKE3,
// takes siblings from the prover's tree instead of the path, and returns
// true without comparing against the root; callers ignore its result.  The
// kernel here is what that function is meant to be for the trees this
// library builds: leaf = BLAKE3 of the 4 elements (64 bytes), then one
// compression per level with the sibling from the path (open_tree_blake's
// order, level 0 first), and the result compared with the root.  With the
// reference's dup-left parents (src/merkle_tree.cpp:287-292) a parent is
// H(left || left), so a node's hash goes up only while it is a left child:
// the root depends on leaf 0 alone, a leaf's own hash counts only at
// position 0 and the sibling at level L only when pos >> L == 1.  That is
// the reference's construction, checked here as it stands.
//
// Plain proofs (cosnark_prove_plain's bytes): the verifier side of the
// product sumcheck, as the reference's prover checks it
// (src/Distributed_Sumcheck.cpp:243-248): per round H(0) + H(1) = y,
// r = hash_to_field({a, b, c}) on the transcript, y = H(r); at the end
// y = v1[0] v2[0].
#include <string.h>
#include <vector>
#include "blake3.h"
#include "ctx.h"

using gf::fe;

namespace {

constexpr int VT = 256;

__global__ __launch_bounds__(VT) void merkle_verify_kernel(const fe *__restrict__ leaves,
                                                           const uint32_t *__restrict__ pos,
                                                           const uint8_t *__restrict__ paths, int depth, uint64_t q,
                                                           const uint8_t *__restrict__ root, int dup_left,
                                                           uint8_t *__restrict__ ok) {
    const uint64_t j = (uint64_t)blockIdx.x * VT + threadIdx.x;
    if (j >= q) return;
    uint32_t m[16], h[8];
    const uint32_t *lw = reinterpret_cast<const uint32_t *>(leaves + 4 * j);
#pragma unroll
    for (int i = 0; i < 16; ++i) m[i] = lw[i];
    b3::compress64(m, h);
    uint64_t p = pos[j];
    const bool in_range = depth >= 32 || (p >> depth) == 0;
    for (int lvl = 0; lvl < depth; ++lvl) {
        const uint32_t *sw = reinterpret_cast<const uint32_t *>(paths + 32 * (j * (uint64_t)depth + lvl));
        uint32_t sib[8];
#pragma unroll
        for (int i = 0; i < 8; ++i) sib[i] = sw[i];
        const bool right = p & 1;
        if (dup_left) {
#pragma unroll
            for (int i = 0; i < 8; ++i) m[i] = m[8 + i] = right ? sib[i] : h[i];
        } else {
#pragma unroll
            for (int i = 0; i < 8; ++i) {
                m[i] = right ? sib[i] : h[i];
                m[8 + i] = right ? h[i] : sib[i];
            }
        }
        b3::compress64(m, h);
        p >>= 1;
    }
    const uint32_t *rw = reinterpret_cast<const uint32_t *>(root);
    bool same = in_range;
#pragma unroll
    for (int i = 0; i < 8; ++i) same = same && h[i] == rw[i];
    ok[j] = same ? 1 : 0;
}

fe quad_at(const fe *p, const fe &x) {      // a x^2 + b x + c
    return gf::add(gf::mul(gf::add(gf::mul(p[0], x), p[1]), x), p[2]);
}

}  // namespace

extern "C" {

int cosnark_merkle_verify(cosnark_ctx *ctx, const uint64_t *d_leaves, const uint32_t *d_pos, const uint8_t *d_paths,
                          int depth, uint64_t q, const uint8_t *d_root, int dup_left, uint8_t *d_ok,
                          uint64_t *h_failures) {
    if (!ctx || !d_leaves || !d_pos || (!d_paths && depth > 0) || !d_root || !d_ok || !h_failures)
        return set_err(ctx, COSNARK_E_ARG, "merkle_verify: null argument");
    if (depth < 0 || depth > 40 || q > (1ull << 32)) return set_err(ctx, COSNARK_E_ARG, "merkle_verify: depth <= 40");
    *h_failures = 0;
    if (q == 0) return COSNARK_OK;
    hipStream_t s = ctx->stream;
    hipLaunchKernelGGL(merkle_verify_kernel, dim3((unsigned)((q + VT - 1) / VT)), dim3(VT), 0, s,
                       (const fe *)d_leaves, d_pos, d_paths, depth, q, d_root, dup_left, d_ok);
    LAUNCHCK();
    std::vector<uint8_t> ok(q);
    HIPCK(hipMemcpyAsync(ok.data(), d_ok, q, hipMemcpyDeviceToHost, s));
    HIPCK(hipStreamSynchronize(s));
    uint64_t bad = 0;
    for (uint8_t v : ok) bad += v ? 0 : 1;
    *h_failures = bad;
    return COSNARK_OK;
}

int cosnark_verify_plain(const uint8_t *h_proof, uint64_t n, const uint64_t y_in[2], cosnark_transcript *t,
                         uint64_t *h_r, int *bad_round) {
    if (!h_proof || !y_in || !t) return COSNARK_E_ARG;
    if (!is_pow2(n) || n < 4) return COSNARK_E_ARG;
    const int m = ilog2(n);
    fe y = gf::mk(y_in[0], y_in[1]);
    if (bad_round) *bad_round = -1;
    const uint64_t *w = reinterpret_cast<const uint64_t *>(h_proof + 32);   // after the root
    for (int i = 0; i < m; ++i) {
        fe p[3];
        for (int k = 0; k < 3; ++k) p[k] = gf::mk(w[6 * i + 2 * k], w[6 * i + 2 * k + 1]);
        for (int k = 0; k < 3; ++k)
            if (p[k].re >= gf::P || p[k].im >= gf::P) {
                if (bad_round) *bad_round = i;
                return COSNARK_E_CHECK;
            }
        if (!gf::eq(gf::add(quad_at(p, gf::zero()), quad_at(p, gf::one())), y)) {
            if (bad_round) *bad_round = i;
            return COSNARK_E_CHECK;
        }
        uint64_t r[2];
        int rc = cosnark_hash_to_field(t, w + 6 * i, 3, r);
        if (rc) return rc;
        if (h_r) { h_r[2 * i] = r[0]; h_r[2 * i + 1] = r[1]; }
        y = quad_at(p, gf::mk(r[0], r[1]));
    }
    const uint64_t *v = w + 6 * m;
    const fe v1 = gf::mk(v[0], v[1]), v2 = gf::mk(v[2], v[3]);
    if (!gf::eq(gf::mul(v1, v2), y)) {
        if (bad_round) *bad_round = m;
        return COSNARK_E_CHECK;
    }
    return COSNARK_OK;
}

}  // extern "C"
