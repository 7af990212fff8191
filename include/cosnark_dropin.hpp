// cosnark_dropin.hpp -- the reference prover's rank-local entry points, with
// their names, argument order and meaning, running on the GPU through the
// cosnark C ABI (cosnark.h).  Header only; link libcosnark.so.
//
// F is the caller's field element: the reference's virgo::fieldElement
// (src/fieldElement.hpp:96-97), 16 bytes, u64 real then u64 img, canonical.
// H is its digest type: _hash (src/Blake3_hash.h:3-5), uint8_t arr[32].
// A reference source file adopts an entry point by forwarding to it, e.g.
//     void merkle_tree::merkle_tree_prover::MT_commit_Blake(F *l, vector<vector<_hash>> &h, int N)
//     { cosnark::dropin::MT_commit_Blake(l, h, N); }
//
// Conventions kept from the reference: vectors by reference, mutated and
// resized as the reference does; the Fiat-Shamir running value is one
// state per thread here (the reference's global prev_challenge,
// src/Fiat_Shamir.cpp:5), read and set with transcript(); the globals
// sumcheck_offset (src/main.cpp:36) and rate_offset (src/coPCS.cpp:22) are
// sumcheck_offset() and rate_offset().  Differences: where the reference
// prints a message and calls exit(-1), these throw cosnark::dropin::error;
// work runs on a per-thread context on GPU 0 (COSNARK_DEVICE selects
// another), and each call copies its inputs to the device and its results
// back.  Cross-rank steps (MPI aggregation, all-to-all, the rank-0 tree of
// roots for N > 1) are outside this layer: the functions take N and accept
// N = 1, where those steps are the identity or a single root.
#pragma once
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <algorithm>
#include <cmath>
#include <type_traits>
#include <unordered_map>
#include <utility>
#include <vector>
#include "cosnark.h"

namespace cosnark {
namespace dropin {

struct error : std::runtime_error {
    int code;
    error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

// ---- process state mirroring the reference's globals ----------------------
inline int &sumcheck_offset() { static int v = 0; return v; }          // src/main.cpp:36
inline unsigned &rate_offset() { static unsigned v = 1; return v; }     // src/coPCS.cpp:22,46
inline cosnark_transcript &transcript() {                               // src/Fiat_Shamir.cpp:5
    thread_local cosnark_transcript t{{0, 0}};
    return t;
}
// This rank and the exchange the coSumcheck drivers use in place of the
// reference's MPI gathers and broadcasts (MPI_Comm_rank and
// sumcheck_aggr / send_final_sumcheck_poly, src/MPI_utils.cpp:2604-2736):
// an all-gather over the N ranks, e.g. MPI_Allgather of 2 * elems
// MPI_UINT64_T.  Per thread, so threads of one process can act as ranks.
struct Comm {
    int rank = 0;
    cosnark_allgather_fn gather = nullptr;
    void *user = nullptr;
};
inline Comm &comm() {
    thread_local Comm c;
    return c;
}
// ip_vec (src/coSumcheck.cpp:231), written by the coSumcheck drivers
template <class F>
std::vector<std::vector<F>> &ip_vec() {
    thread_local std::vector<std::vector<F>> v;
    return v;
}

namespace detail {

struct Ctx {
    cosnark_ctx *c = nullptr;
    ~Ctx() { if (c) cosnark_destroy(c); }
};

inline cosnark_ctx *ctx() {
    thread_local Ctx h;
    if (!h.c) {
        const char *d = std::getenv("COSNARK_DEVICE");
        const int rc = cosnark_create(d ? std::atoi(d) : 0, &h.c);
        if (rc) throw error(rc, "cosnark: no GPU context (is the HIP device visible?)");
    }
    return h.c;
}

inline void check(int rc, const char *what) {
    if (rc) throw error(rc, std::string(what) + ": " + cosnark_last_error(ctx()));
}

template <class F>
void elem_type() {
    static_assert(sizeof(F) == 16 && std::is_trivially_copyable<F>::value,
                  "field element must be 16 bytes: u64 real, u64 img");
}

template <class H>
void digest_type() {
    static_assert(sizeof(H) == 32 && std::is_trivially_copyable<H>::value, "digest must be 32 bytes");
}

// device buffer, freed on scope exit
class Dev {
  public:
    explicit Dev(uint64_t bytes) : n_(bytes) {
        check(cosnark_dev_alloc(ctx(), bytes ? bytes : 16, &p_), "device allocation");
    }
    Dev(const void *host, uint64_t bytes) : Dev(bytes) { put(host, bytes); }
    ~Dev() { if (p_) cosnark_dev_free(ctx(), p_); }
    Dev(const Dev &) = delete;
    Dev &operator=(const Dev &) = delete;
    void put(const void *host, uint64_t bytes) { if (bytes) check(cosnark_h2d(ctx(), p_, host, bytes), "copy in"); }
    void get(void *host, uint64_t bytes) const { if (bytes) check(cosnark_d2h(ctx(), host, p_, bytes), "copy out"); }
    uint64_t *u64() const { return static_cast<uint64_t *>(p_); }
    uint8_t *u8() const { return static_cast<uint8_t *>(p_); }

  private:
    void *p_ = nullptr;
    uint64_t n_;
};

inline int ilog2(uint64_t x) { int l = 0; while ((1ull << l) < x) ++l; return l; }
inline bool pow2(uint64_t x) { return x && !(x & (x - 1)); }

// all levels of a tree, level 0 first, into vector<vector<H>>
template <class H>
void unpack_levels(const std::vector<uint8_t> &nodes, uint64_t leaves, std::vector<std::vector<H>> &hashes) {
    const int depth = ilog2(leaves);
    if ((int)hashes.size() < depth + 1) hashes.resize(depth + 1);
    uint64_t off = 0;
    for (int l = 0; l <= depth; ++l) {
        const uint64_t cnt = leaves >> l;
        hashes[l].resize(cnt);
        std::memcpy(hashes[l].data(), nodes.data() + 32 * off, 32 * cnt);
        off += cnt;
    }
}

template <class F>
F make(uint64_t re, uint64_t im) {
    F x;
    uint64_t w[2] = {re, im};
    std::memcpy(&x, w, 16);
    return x;
}

}  // namespace detail

// ---- NTT -------------------------------------------------------------------

// fft(arr, logn, flag), src/utils.cpp:904-964: in place, natural order;
// flag = inverse (scaled by 2^-logn).  arr.size() must be 2^logn.
template <class F>
void fft(std::vector<F> &arr, int logn, bool flag) {
    detail::elem_type<F>();
    if (arr.size() != (1ull << logn)) throw error(COSNARK_E_ARG, "fft: arr.size() != 2^logn");
    if (logn == 0) return;
    detail::Dev d(arr.data(), 16 * arr.size());
    detail::check(cosnark_ntt(detail::ctx(), d.u64(), logn, flag ? 1 : 0), "fft");
    d.get(arr.data(), 16 * arr.size());
}

// pfft(arr, logn, flag), src/utils.cpp:1583-1673: the forward transform
// (the reference ignores flag and does not scale).
template <class F>
void pfft(std::vector<F> &arr, int logn, bool) { fft(arr, logn, false); }

// ---- Fiat-Shamir, src/Fiat_Shamir.cpp ----------------------------------------

// hash_to_field(elements), :29-34: absorbs the elements into the thread's
// transcript and returns its new value
template <class F>
F hash_to_field(const std::vector<F> &elements) {
    detail::elem_type<F>();
    uint64_t out[2];
    detail::check(cosnark_hash_to_field(&transcript(), reinterpret_cast<const uint64_t *>(elements.data()),
                                        elements.size(), out), "hash_to_field");
    return detail::make<F>(out[0], out[1]);
}

// get_indexes(l, N, M, pos), :37-56 (pos = -1: all l pairs)
inline std::vector<std::vector<uint32_t>> get_indexes(int l, int N, int M, int pos) {
    std::vector<uint32_t> flat(2 * (size_t)(l > 0 ? l : 0));
    detail::check(cosnark_get_indexes(&transcript(), l, (uint32_t)N, (uint32_t)M, flat.data()), "get_indexes");
    std::vector<std::vector<uint32_t>> idx;
    for (int i = 0; i < l; ++i) {
        if (pos != -1 && flat[2 * i] != (uint32_t)pos) continue;
        idx.push_back({flat[2 * i], flat[2 * i + 1]});
    }
    return idx;
}

// ---- Merkle tree, src/merkle_tree.cpp ----------------------------------------

// MT_commit_Blake(leafs, hashes, N), :195-233: leaf i = BLAKE3 of elements
// 4i..4i+3; every level into hashes (level 0 first).  N: power of two >= 8
// (the reference's N <= 2 branch hashes 32 uninitialised bytes).
template <class F, class H>
void MT_commit_Blake(F *leafs, std::vector<std::vector<H>> &hashes, int N) {
    detail::elem_type<F>();
    detail::digest_type<H>();
    if (N < 8 || !detail::pow2((uint64_t)N)) throw error(COSNARK_E_ARG, "MT_commit_Blake: N must be a power of two >= 8");
    const uint64_t cnt = cosnark_merkle_node_count((uint64_t)N);
    detail::Dev e(leafs, 16ull * N), nodes(32 * cnt);
    detail::check(cosnark_merkle_commit(detail::ctx(), e.u64(), (uint64_t)N, nodes.u8(), 1), "MT_commit_Blake");
    std::vector<uint8_t> h(32 * cnt);
    nodes.get(h.data(), h.size());
    detail::unpack_levels(h, (uint64_t)N / 4, hashes);
}

// create_tree_blake(ele_num, hashes, element_size, alloc_required),
// :267-299: levels 1.. over hashes[0] (ele_num digests, a power of two),
// each parent the hash of its left child twice.
template <class H>
void create_tree_blake(int ele_num, std::vector<std::vector<H>> &hashes, int = 32, bool = false) {
    detail::digest_type<H>();
    if (ele_num < 1 || !detail::pow2((uint64_t)ele_num) || hashes.empty() || (int)hashes[0].size() < ele_num)
        throw error(COSNARK_E_ARG, "create_tree_blake: need hashes[0] with ele_num (a power of two) digests");
    const uint64_t total = 2ull * ele_num - 1;
    detail::Dev nodes(32 * total);
    nodes.put(hashes[0].data(), 32ull * ele_num);
    detail::check(cosnark_merkle_tree_from_leaves(detail::ctx(), (uint64_t)ele_num, nodes.u8(), 1), "create_tree_blake");
    std::vector<uint8_t> h(32 * total);
    nodes.get(h.data(), h.size());
    detail::unpack_levels(h, (uint64_t)ele_num, hashes);
}

// open_tree_blake(MT_hashes, pos), :320-335: the sibling of pos on every
// level below the root (a lookup in the host copy of the tree)
template <class H>
std::vector<H> open_tree_blake(std::vector<std::vector<H>> &MT_hashes, int pos) {
    if (MT_hashes.empty() || pos < 0 || pos >= (int)MT_hashes[0].size())
        throw error(COSNARK_E_ARG, "open_tree_blake: position out of range");
    std::vector<H> path;
    for (size_t i = 0; i + 1 < MT_hashes.size(); ++i) {
        const size_t sib = 2 * (size_t)(pos / 2) + (1 - pos % 2);
        if (sib >= MT_hashes[i].size()) throw error(COSNARK_E_ARG, "open_tree_blake: tree level too short");
        path.push_back(MT_hashes[i][sib]);
        pos /= 2;
    }
    return path;
}

// verify_claim_opt_blake(MT, path, pos, N, visited, ps), :337-383: the
// reference's verifier walks from pos towards the root until it meets a
// node an earlier path already visited, marking visited nodes (indices
// N + pos, halved per level) and adding 32/1024 to the proof-size counter
// ps per level; it hashes with my_hhash and never compares the result with
// the root, so it returns true.  This does the same bookkeeping and exits
// (throws) where the reference does; cosnark_merkle_verify checks paths
// against the root on the GPU.
template <class H>
bool verify_claim_opt_blake(std::vector<std::vector<H>> &MT, const H *path, int pos_element_arr, int N,
                            std::vector<bool> &visited, double &ps) {
    (void)path;
    if (N <= 0 || (N & (-N)) != N) throw error(COSNARK_E_ARG, "verify_claim_opt_blake: N not a power of two");
    if (MT.empty() || (int)MT[0].size() != N) throw error(COSNARK_E_CHECK, "verify_claim_opt_blake: error in sizes");
    int pos = pos_element_arr;
    if (pos < 0 || pos >= (int)MT[0].size()) throw error(COSNARK_E_CHECK, "verify_claim_opt_blake: position out of range");
    if (visited.size() < 2 * (size_t)N) throw error(COSNARK_E_ARG, "verify_claim_opt_blake: visited shorter than 2N");
    int pos_element = N + pos_element_arr;
    for (size_t i = 0; i + 1 < MT.size(); ++i) {
        if ((pos_element ^ 1) >= 2 * N) throw error(COSNARK_E_CHECK, "verify_claim_opt_blake: node index out of range");
        if (visited[pos_element ^ 1]) return true;
        if ((size_t)(pos ^ 1) >= MT[i].size()) throw error(COSNARK_E_CHECK, "verify_claim_opt_blake: error in verification");
        visited[pos_element ^ 1] = true;
        pos_element /= 2;
        pos /= 2;
        visited[pos_element] = true;
        ps += 32.0 / 1024.0;
    }
    return true;
}

// ---- product sumcheck ----------------------------------------------------------

// _quadratic_sumcheck(y, v1, v2, N), src/Distributed_Sumcheck.cpp:204-348 at
// N = 1 with the tail of src/sparse_eval.cpp:140-192: the first
// log2|v| - sumcheck_offset() rounds take Fiat-Shamir challenges from the
// thread's transcript, the rest draw F::_random() from libc random().
// Returns {(v1 folded, r), (v2 folded, r)}, r the challenges in round order.
// Throws where the reference exits (H(0) + H(1) != y).  v1, v2 are not
// modified (the reference leaves them partly folded).
template <class F>
std::vector<std::pair<F, std::vector<F>>> _quadratic_sumcheck(F y, std::vector<F> &v1, std::vector<F> &v2, int N) {
    detail::elem_type<F>();
    if (N != 1) throw error(COSNARK_E_ARG, "_quadratic_sumcheck: one rank (N = 1) only");
    const uint64_t n = v1.size();
    if (v2.size() != n || !detail::pow2(n) || n < 2) throw error(COSNARK_E_ARG, "_quadratic_sumcheck: sizes");
    const int m = detail::ilog2(n);
    detail::Dev a(v1.data(), 16 * n), b(v2.data(), 16 * n);
    std::vector<uint64_t> polys(6 * (size_t)m), r(2 * (size_t)m), fin(4);
    uint64_t yy[2];
    std::memcpy(yy, &y, 16);
    detail::check(cosnark_product_sumcheck(detail::ctx(), a.u64(), b.u64(), n, yy, &transcript(), sumcheck_offset(),
                                           polys.data(), r.data(), fin.data()),
                  "_quadratic_sumcheck");
    std::vector<F> rr(m);
    for (int i = 0; i < m; ++i) rr[i] = detail::make<F>(r[2 * i], r[2 * i + 1]);
    return {std::make_pair(detail::make<F>(fin[0], fin[1]), rr), std::make_pair(detail::make<F>(fin[2], fin[3]), rr)};
}

// ---- WHIR, src/Whir.cpp ------------------------------------------------------

// The fold factor, the reference's global `int whir_k = 8;` (src/Whir.cpp:6).
inline int &whir_k() { static int v = 8; return v; }

// get_whir_rates' globals (declared extern at src/Whir.cpp:8-16): the
// sizes it reads and the schedules it writes.
inline std::vector<int> &whir_sizes() { static std::vector<int> v; return v; }
inline std::vector<int> &whir_rates() { static std::vector<int> v; return v; }
inline std::vector<int> &sorted_indexes() { static std::vector<int> v; return v; }
inline std::vector<int> &sorted_whir_rates() { static std::vector<int> v; return v; }
inline std::unordered_map<int, int> &whir_kappas() { static std::unordered_map<int, int> v; return v; }

namespace detail {
// every level of a dup-left tree over `leaves` all-zero digests: what
// Whir_Commit, Whir_Commit_Codeword and aggregate_commitments leave (their
// leaf vector is never filled), level l holding copies of the l-fold
// self-hash of the zero digest
template <class H>
void zero_tree(uint64_t leaves, std::vector<std::vector<H>> &MT) {
    if (leaves < 1 || !pow2(leaves)) throw error(COSNARK_E_ARG, "zero tree: leaf count not a power of two");
    const int depth = ilog2(leaves);
    MT.assign(depth + 1, {});
    uint8_t cur[32] = {0}, blk[64];
    for (int l = 0; l <= depth; ++l) {
        if (l > 0) {
            std::memcpy(blk, cur, 32);
            std::memcpy(blk + 32, cur, 32);
            cosnark_blake3_64(blk, cur);
        }
        H h;
        std::memcpy(&h, cur, 32);
        MT[l].assign(leaves >> l, h);
    }
}

// the query count of whir_query and aggregate_commitments without the
// conjectured-security option (src/Whir.cpp:253-258, :567-572)
inline int whir_query_count(int rate) { return (int)(-100.0 / std::log2(std::sqrt(1.0 / (double)rate))); }

// a tree to open on the device: the dup-left tree of the codeword
// (MT_commit_Blake's levels) or an all-zero-leaf tree, told apart by its
// leaf digests; returns the zero tree's leaf count (0: nodes holds the tree)
template <class H>
uint64_t tree_or_zero(const std::vector<std::vector<H>> &MT, uint64_t cw, std::unique_ptr<Dev> &nodes,
                      const char *what) {
    if (MT.empty() || MT[0].empty()) throw error(COSNARK_E_ARG, std::string(what) + ": empty tree");
    static const uint8_t zero[32] = {0};
    bool all_zero = true;
    for (const H &h : MT[0])
        if (std::memcmp(&h, zero, 32) != 0) { all_zero = false; break; }
    if (all_zero) return MT[0].size();
    nodes.reset(new Dev(32 * cosnark_merkle_node_count(cw)));
    const uint64_t cnt = cosnark_merkle_node_count(cw);
    if (MT[0].size() != cw / 4) throw error(COSNARK_E_ARG, std::string(what) + ": tree does not match the codeword");
    uint64_t off = 0;
    for (size_t l = 0; l < MT.size(); ++l) {
        if (off + MT[l].size() > cnt) throw error(COSNARK_E_ARG, std::string(what) + ": tree larger than expected");
        check(cosnark_h2d(ctx(), nodes->u8() + 32 * off, MT[l].data(), 32 * MT[l].size()), "copy in");
        off += MT[l].size();
    }
    if (off != cnt) throw error(COSNARK_E_ARG, std::string(what) + ": tree smaller than expected");
    return 0;
}
}  // namespace detail

// Whir_Commit_initial(arr, codeword, MT, rate, kappa), :158-190: codeword =
// the rate*|arr| evaluations of arr's coefficients, MT its tree (any kappa)
template <class F, class H>
void Whir_Commit_initial(std::vector<F> &arr, std::vector<F> &codeword, std::vector<std::vector<H>> &MT, int rate,
                         int kappa) {
    detail::elem_type<F>();
    detail::digest_type<H>();
    (void)kappa;   // the reference only fills an unused buffer with it (:165-175)
    const uint64_t n = arr.size(), cw = n * (uint64_t)rate, cnt = cosnark_merkle_node_count(cw);
    detail::Dev a(arr.data(), 16 * n), c(16 * cw), nodes(32 * cnt);
    detail::check(cosnark_whir_commit(detail::ctx(), a.u64(), n, rate, c.u64(), nodes.u8()), "Whir_Commit_initial");
    codeword.resize(cw);
    c.get(codeword.data(), 16 * cw);
    std::vector<uint8_t> h(32 * cnt);
    nodes.get(h.data(), h.size());
    detail::unpack_levels(h, cw / 4, MT);
}

// Whir_Commit(arr, codeword, MT, rate, kappa = whir_k), :192-213: the same
// codeword; MT is the tree the reference leaves, over |codeword| / kappa
// leaf digests it never fills (all zero).
template <class F, class H>
void Whir_Commit(std::vector<F> &arr, std::vector<F> &codeword, std::vector<std::vector<H>> &MT, int rate,
                 int kappa = whir_k()) {
    detail::elem_type<F>();
    detail::digest_type<H>();
    const uint64_t n = arr.size(), cw = n * (uint64_t)rate;
    if (!detail::pow2(n) || rate < 1 || !detail::pow2((uint64_t)rate) || kappa < 1 || cw < (uint64_t)kappa)
        throw error(COSNARK_E_ARG, "Whir_Commit: sizes");
    detail::Dev a(arr.data(), 16 * n), c(16 * cw);
    detail::check(cosnark_rs_encode(detail::ctx(), a.u64(), n, rate, c.u64()), "Whir_Commit");
    codeword.resize(cw);
    c.get(codeword.data(), 16 * cw);
    detail::zero_tree(cw / (uint64_t)kappa, MT);
}

// Whir_Commit_Codeword(codeword, MT, rate), :139-156: MT over
// |codeword| / whir_k all-zero leaf digests (the codeword is not read).
template <class F, class H>
void Whir_Commit_Codeword(std::vector<F> &codeword, std::vector<std::vector<H>> &MT, int) {
    detail::elem_type<F>();
    detail::digest_type<H>();
    detail::zero_tree(codeword.size() / (uint64_t)whir_k(), MT);
}

// The opening record of the thread's last Whir_Open / _Whir_Open /
// Whir_Batch_Open (cosnark_whir_batch_open's layout), for callers that
// serialise or verify it.
inline std::vector<uint64_t> &last_whir_record() {
    thread_local std::vector<uint64_t> rec;
    return rec;
}

namespace detail {
template <class F, class H>
std::pair<F, std::vector<F>> whir_open_weights(std::vector<F> &arr, std::vector<F> &codeword,
                                               std::vector<std::vector<H>> &MT, const Dev &b, const F &K, int rate,
                                               const char *what) {
    const uint64_t n = arr.size(), cw = codeword.size();
    if (!pow2(n) || n < 2 || !pow2(cw) || cw < 2 * n) throw error(COSNARK_E_ARG, std::string(what) + ": sizes");
    std::unique_ptr<Dev> nodes;
    const uint64_t zl = tree_or_zero(MT, cw, nodes, what);
    Dev a(arr.data(), 16 * n), c(codeword.data(), 16 * cw);
    uint64_t k[2], fin[2];
    std::memcpy(k, &K, 16);
    auto &rec = last_whir_record();
    rec.assign(1u << 22, 0);
    uint64_t words = 0;
    std::vector<F> r((size_t)ilog2(n));
    check(cosnark_whir_open_weights(ctx(), a.u64(), n, b.u64(), k, c.u64(), cw, nodes ? nodes->u8() : nullptr, zl,
                                    rate, whir_k(), rec.data(), rec.size(), &words,
                                    reinterpret_cast<uint64_t *>(r.data()), fin),
          what);
    rec.resize(words);
    return {make<F>(fin[0], fin[1]), r};
}
}  // namespace detail

// Whir_Open(arr, codeword, MT, claims, rate), :388-467, at the global
// whir_k: the weights are eq(claims[0]'s point); with more claims one
// F::_random() r batches them, K += r^i K_i and b += r^i eq(claims[1]'s
// point) for every i >= 1 -- the reference reads claims[1] there (:403), so
// three or more claims at distinct points fail the first sumcheck check,
// where the reference exits.  Iterations down to |arr| <= whir_k,
// challenges from libc random()/rand() in the reference's order.  Throws
// where the reference's checks exit.  arr, codeword and MT are left as
// given (the reference folds them).
template <class F, class H>
void Whir_Open(std::vector<F> &arr, std::vector<F> &codeword, std::vector<std::vector<H>> &MT,
               std::vector<std::pair<F, std::vector<F>>> claims, int rate) {
    detail::elem_type<F>();
    detail::digest_type<H>();
    if (claims.empty()) throw error(COSNARK_E_ARG, "Whir_Open: no claim");
    const uint64_t n = arr.size();
    const int m = detail::ilog2(n);
    for (auto &cl : claims)
        if ((int)cl.second.size() != m) throw error(COSNARK_E_ARG, "Whir_Open: claim point size != log2 |arr|");
    F K = claims[0].first;
    detail::Dev b(16 * n);
    detail::check(cosnark_precompute_beta(detail::ctx(), reinterpret_cast<const uint64_t *>(claims[0].second.data()),
                                          m, b.u64()),
                  "Whir_Open");
    if (claims.size() > 1) {
        uint64_t r[2], p[2], psum[2] = {0, 0};
        cosnark_field_random(r);                                    // :399
        std::memcpy(p, r, 16);
        for (size_t i = 1; i < claims.size(); ++i) {
            uint64_t t[2], k[2];
            std::memcpy(k, &K, 16);
            cosnark_field_mul(p, reinterpret_cast<const uint64_t *>(&claims[i].first), t);
            cosnark_field_add(k, t, k);
            std::memcpy(&K, k, 16);
            cosnark_field_add(psum, p, psum);
            cosnark_field_mul(p, r, p);
        }
        // b = eq(z_0) + (sum of the powers) eq(z_1)
        detail::Dev both(32 * n);
        detail::check(cosnark_d2d(detail::ctx(), both.u8(), b.u8(), 16 * n), "Whir_Open");
        detail::check(cosnark_precompute_beta(detail::ctx(), reinterpret_cast<const uint64_t *>(claims[1].second.data()),
                                              m, both.u64() + 2 * n),
                      "Whir_Open");
        const uint64_t c[4] = {1, 0, psum[0], psum[1]};
        detail::check(cosnark_row_combine(detail::ctx(), both.u64(), 2, n, c, b.u64()), "Whir_Open");
    }
    detail::whir_open_weights(arr, codeword, MT, b, K, rate, "Whir_Open");
}

// _Whir_Open(arr, codeword, MT, b, claims, challenges, K, rate, Verify),
// :469-545, at the global whir_k: returns the final claim, the initial
// weights' multilinear evaluation at every challenge, with those challenges
// (total_r).  claims and challenges feed only the reference's verifier-side
// step 3, which does not exit.  MT may be MT_commit_Blake's tree of the
// codeword or the all-zero-leaf tree aggregate_commitments leaves.  A failed
// check throws for either Verify (the reference continues when Verify is
// false).  arr, codeword, MT and b are left as given (the reference folds
// them).
template <class F, class H>
std::pair<F, std::vector<F>> _Whir_Open(std::vector<F> &arr, std::vector<F> &codeword, std::vector<std::vector<H>> &MT,
                                        std::vector<F> &b, std::vector<std::vector<std::pair<F, std::vector<F>>>>,
                                        std::vector<F>, F K, int rate, bool = true) {
    detail::elem_type<F>();
    detail::digest_type<H>();
    if (b.size() != arr.size()) throw error(COSNARK_E_ARG, "_Whir_Open: weights and array sizes differ");
    detail::Dev db(b.data(), 16 * b.size());
    return detail::whir_open_weights(arr, codeword, MT, db, K, rate, "_Whir_Open");
}

// aggregate_commitments(codeword, MT, codewords_aggr, MT_aggr, r, rate),
// :547-606: codeword = sum_i r[i] codewords_aggr[i]; l = the query count at
// `rate` positions rand() % (|codeword| / 4) are opened on MT_aggr[0] (the
// paths feed only the proof-size counter); MT becomes the tree over
// |codeword| / whir_k all-zero leaf digests.  Throws where the reference
// exits (codewords of different sizes, a position beyond MT_aggr[0]).
template <class F, class H>
void aggregate_commitments(std::vector<F> &codeword, std::vector<std::vector<H>> &MT,
                           std::vector<std::vector<F>> &codewords_aggr, std::vector<std::vector<std::vector<H>>> &MT_aggr,
                           std::vector<F> r, int rate) {
    detail::elem_type<F>();
    detail::digest_type<H>();
    const size_t c = codewords_aggr.size();
    if (!c || r.size() < c || MT_aggr.empty()) throw error(COSNARK_E_ARG, "aggregate_commitments: arguments");
    const uint64_t L = codewords_aggr[0].size();
    for (size_t i = 0; i < c; ++i)
        if (codewords_aggr[i].size() != L) throw error(COSNARK_E_CHECK, "aggregate_commitments: incorrect codeword size");
    // codeword = sum r_i C_i, 32 rows per combine (the partial sum carried on)
    detail::Dev acc(16 * L), rows(16 * L * (c < 32 ? c : 32));
    size_t done = 0;
    bool first = true;
    while (done < c) {
        const size_t take = std::min<size_t>(c - done, first ? 32 : 31);
        std::vector<uint64_t> coef;
        size_t row = 0;
        if (!first) {
            detail::check(cosnark_d2d(detail::ctx(), rows.u8(), acc.u8(), 16 * L), "aggregate_commitments");
            coef.push_back(1);
            coef.push_back(0);
            ++row;
        }
        for (size_t i = 0; i < take; ++i, ++row) {
            detail::check(cosnark_h2d(detail::ctx(), rows.u8() + 16 * L * row, codewords_aggr[done + i].data(), 16 * L),
                          "copy in");
            const uint64_t *ri = reinterpret_cast<const uint64_t *>(&r[done + i]);
            coef.push_back(ri[0]);
            coef.push_back(ri[1]);
        }
        detail::check(cosnark_row_combine(detail::ctx(), rows.u64(), (int)row, L, coef.data(), acc.u64()),
                      "aggregate_commitments");
        done += take;
        first = false;
    }
    codeword.resize(L);
    acc.get(codeword.data(), 16 * L);
    const int l = detail::whir_query_count(rate);
    for (int i = 0; i < l; ++i) {
        const int q = rand() % (int)(L / 4);                        // :574-575
        (void)open_tree_blake(MT_aggr[0], q);                       // exits (throws) beyond the tree
    }
    detail::zero_tree(L / (uint64_t)whir_k(), MT);
}

// get_whir_rates(rate), :18-137, on the globals: reads whir_sizes(); writes
// sorted_indexes() (array i's place in the largest-first order, equal sizes
// in reverse index order, as the reference's sort gives them), appends the
// commitment rates in that order to sorted_whir_rates() (the reference
// never clears it) and copies it to whir_rates(), and sets whir_kappas()
// (size -> fold factor where the schedule meets that size).  It does not
// print the schedule as the reference's rank 0 does.
inline void get_whir_rates(int rate) {
    const std::vector<int> &ws = whir_sizes();
    if (ws.empty()) throw error(COSNARK_E_ARG, "get_whir_rates: whir_sizes is empty");
    const int wk = whir_k();
    std::vector<std::pair<int, int>> sp(ws.size());
    for (size_t i = 0; i < ws.size(); ++i) sp[i] = {ws[i], (int)i};
    std::sort(sp.begin(), sp.end());                                 // :23-25
    std::vector<int> order;
    sorted_indexes().assign(ws.size(), 0);
    for (int i = (int)ws.size() - 1; i >= 0; --i) {
        order.push_back(sp[i].second);
        sorted_indexes()[sp[i].second] = (int)ws.size() - 1 - i;
    }
    std::vector<uint64_t> sorted;
    for (int idx : order) sorted.push_back((uint64_t)ws[idx]);
    std::vector<int> rates(sorted.size());
    if (cosnark_whir_rates(sorted.data(), (int)sorted.size(), rate, wk, rates.data()))
        throw error(COSNARK_E_ARG, "get_whir_rates: bad sizes");
    for (int x : rates) sorted_whir_rates().push_back(x);
    // the fold factor at each size the schedule meets (:76-123)
    std::vector<int> s2(sorted.begin(), sorted.end());
    int size = s2[0];
    s2.erase(s2.begin());
    while (!s2.empty() && s2[0] == size) s2.erase(s2.begin());
    if (s2.empty()) whir_kappas()[size] = wk;
    bool flag = true;
    while (!s2.empty()) {
        if (size / wk > s2[0]) {
            if (flag) {
                whir_kappas()[size] = wk;
                flag = false;
            }
            size /= wk;
        } else {
            if (flag) whir_kappas()[size] = size / s2[0];
            size = s2[0];
            flag = true;
            const int s = s2[0];
            while (!s2.empty() && s2[0] == s) s2.erase(s2.begin());
            if (s2.empty()) whir_kappas()[size] = wk;
        }
        if (size <= wk) break;
    }
    whir_rates() = sorted_whir_rates();
}

// ---- coPCS, src/coPCS.cpp and src/coPCS_utils.cpp ----------------------------------

// Encode(Codeword, N, rate), src/coPCS.cpp:131-147: every non-empty row
// (before the first empty one, at most 32) becomes its coset encoding at
// blowup rate[0], with w = root(1 + log2 rate_offset() + log2 |row|).
template <class F>
void Encode(std::vector<std::vector<F>> &Codeword, int, std::vector<int> rate) {
    detail::elem_type<F>();
    size_t rows = 0;
    while (rows < Codeword.size() && rows < 32 && !Codeword[rows].empty()) ++rows;
    if (rows == 0) return;
    const uint64_t L = Codeword[0].size(), M = L * (uint64_t)rate.at(0);
    for (size_t i = 1; i < rows; ++i)
        if (Codeword[i].size() != L) throw error(COSNARK_E_ARG, "Encode: rows of different sizes");
    detail::Dev in(16 * L * rows), out(16 * M * rows);
    for (size_t i = 0; i < rows; ++i)
        detail::check(cosnark_h2d(detail::ctx(), in.u8() + 16 * L * i, Codeword[i].data(), 16 * L), "copy in");
    detail::check(cosnark_copcs_encode(detail::ctx(), in.u64(), (int)rows, L, rate[0], detail::ilog2(rate_offset()),
                                       out.u64()),
                  "Encode");
    for (size_t i = 0; i < rows; ++i) {
        Codeword[i].resize(M);
        detail::check(cosnark_d2h(detail::ctx(), Codeword[i].data(), out.u8() + 16 * M * i, 16 * M), "copy out");
    }
}

// prepare_parameters' globals (src/coPCS.cpp:18-24, src/New_PCS.cpp:3-16)
inline std::vector<int> &_degrees() { static std::vector<int> v; return v; }
inline std::vector<int> &_queries() { static std::vector<int> v; return v; }
inline std::vector<std::vector<int>> &_rates() { static std::vector<std::vector<int>> v; return v; }
inline std::vector<int> &delegation_size() { static std::vector<int> v; return v; }
inline int &randomness() { static int v = 0; return v; }
inline int &max_delegation_size() { static int v = 0; return v; }
inline int &Security_Level() { static int v = 100; return v; }     // src/New_PCS.cpp:3
constexpr int BATCH_FACTOR = 32;                                   // src/New_PCS.cpp:9

// prepare_parameters(rates, rounds, N, M), src/coPCS.cpp:38-70: get_degrees
// (rounds, 2 (4 M / N)) and get_queries(rounds, rates) (src/New_PCS.cpp:
// 23-76) append to _degrees() and _queries(); rate_offset() multiplies by
// rates[i][0] for i >= 1; delegation_size() gets next_pow2(d_i q_i) for
// i >= 2; randomness() adds q_i d_i; max_delegation_size() the largest
// delegation size.  The reference's globals accumulate across calls, as
// here; its rank 0 prints them, this does not.
inline void prepare_parameters(std::vector<std::vector<int>> rates, int rounds, int N, int M) {
    if (rounds < 3 || (int)rates.size() < rounds || N < 1) throw error(COSNARK_E_ARG, "prepare_parameters: arguments");
    // get_degrees (New_PCS.cpp:53-76)
    int log_size = (int)std::log2((double)(2 * (4 * (long long)M / N)));
    _degrees().push_back(BATCH_FACTOR);
    log_size -= (int)std::log2((double)BATCH_FACTOR);
    _degrees().push_back(8);
    log_size -= 3;
    const int deg = log_size / (rounds - 2);
    for (int i = 1; i < rounds - 2; ++i) _degrees().push_back(1 << deg);
    _degrees().push_back(1 << (log_size - deg * (rounds - 3)));
    // get_queries (New_PCS.cpp:23-47)
    const double SL = Security_Level();
    for (int i = 0; i < rounds; ++i) {
        if (i <= 1) {
            _queries().push_back((int)(SL / (-std::log2(1 - 0.25 * (1.0 - 1.0 / (double)rates[0][0])))));
        } else {
            const double r0 = rates[i][0], r1 = rates[i][1];
            const int l1 = (int)(SL / (-std::log2(1 - 0.5 * 0.5 * (1.0 - 1.0 / r0) * (1 - std::sqrt(1.0 / r1)))));
            const int l2 = (int)(SL / (-std::log2(1 - 0.5 * 0.5 * (1.0 - 1.0 / r0))) +
                                 SL / (-std::log2(std::sqrt(1.0 / r1))));
            _queries().push_back(l1 > l2 ? l2 : l1);
        }
    }
    _rates() = rates;
    for (size_t i = 1; i < rates.size(); ++i) rate_offset() *= (unsigned)rates[i][0];   // :44-46
    auto next_pow2 = [](size_t x) -> int {                                               // src/utils.cpp:20-25
        if ((1 << (int)std::log2((double)x)) == (int)x) return (int)x;
        return 1 << (1 + (int)std::log2((double)x));
    };
    for (size_t i = 2; i < _degrees().size(); ++i)
        delegation_size().push_back(next_pow2((size_t)_degrees()[i] * _queries()[i]));
    for (size_t i = 0; i < _degrees().size(); ++i) randomness() += _queries()[i] * _degrees()[i];
    if (delegation_size().empty()) throw error(COSNARK_E_ARG, "prepare_parameters: fewer than three rounds");
    max_delegation_size() = delegation_size()[0];
    for (int d : delegation_size())
        if (max_delegation_size() <= d) max_delegation_size() = d;
}

// PCS_commit(C, Codeword, Commit_data, Shares, Random_Shares, rate, N),
// src/coPCS.cpp:236-250, one rank: prepare_data's rows (_queries()[0]
// random values closing each, one random-only row) into Commit_data, their
// encodings at rate[0][0] (with rate_offset()) into Codeword, distributed_MT
// into C; Shares and Random_Shares are cleared as the reference clears them.
template <class F, class MTT>
void PCS_commit(MTT &C, std::vector<std::vector<F>> &Codeword, std::vector<std::vector<F>> &Commit_data,
                std::vector<F> &Shares, std::vector<F> &Random_Shares, std::vector<std::vector<int>> rate, int N) {
    detail::elem_type<F>();
    if (_queries().empty() || rate.empty() || rate[0].empty())
        throw error(COSNARK_E_ARG, "PCS_commit: call prepare_parameters first");
    const uint64_t n = Shares.size(), nr = Random_Shares.size(), B = n / (BATCH_FACTOR / 2);
    const int r = rate[0][0];
    detail::Dev s(Shares.data(), 16 * n), rn(Random_Shares.data(), 16 * (nr ? nr : 1));
    detail::Dev rows(16 * 32 * B), cws(16 * 32 * B * r), nodes(32 * (2 * B * r - 1));
    int c = 0;
    detail::check(cosnark_pcs_commit(detail::ctx(), s.u64(), n, rn.u64(), nr, _queries()[0], r,
                                     detail::ilog2(rate_offset()), rows.u64(), cws.u64(), nodes.u8(), &c),
                  "PCS_commit");
    Commit_data.resize(BATCH_FACTOR);
    Codeword.assign(BATCH_FACTOR, {});
    for (int i = 0; i < c; ++i) {
        Commit_data[i].resize(B);
        detail::check(cosnark_d2h(detail::ctx(), Commit_data[i].data(), rows.u8() + 16 * B * i, 16 * B), "copy out");
        Codeword[i].resize(B * r);
        detail::check(cosnark_d2h(detail::ctx(), Codeword[i].data(), cws.u8() + 16 * B * r * i, 16 * B * r), "copy out");
    }
    std::vector<uint8_t> h(32 * (2 * B * r - 1));
    nodes.get(h.data(), h.size());
    C.Base_MT.clear();
    detail::unpack_levels(h, B * r, C.Base_MT);
    if (N == 1) C.Root_MT.assign(1, {C.Base_MT.back()[0]});
    Shares.clear();
    Random_Shares.clear();
}

// distributed_MT(Codeword, Com, N), src/coPCS.cpp:152-230, one rank: Base_MT
// is the tree over the column roots; at N = 1 Root_MT is that one root.
// MTT is the reference's MT struct (Base_MT, Root_MT of vector<vector<_hash>>).
template <class F, class MTT>
void distributed_MT(std::vector<std::vector<F>> &Codeword, MTT &Com, int N) {
    detail::elem_type<F>();
    int rows = 0;
    while (rows < (int)Codeword.size() && !Codeword[rows].empty()) ++rows;
    if (rows == 0) throw error(COSNARK_E_ARG, "distributed_MT: empty codeword");
    if (rows > 32) throw error(COSNARK_E_ARG, "distributed_MT: more than 32 rows");
    const uint64_t Lc = Codeword[0].size();
    detail::Dev cw(16 * Lc * rows), nodes(32 * (2 * Lc - 1));
    for (int i = 0; i < rows; ++i)
        detail::check(cosnark_h2d(detail::ctx(), cw.u8() + 16 * Lc * i, Codeword[i].data(), 16 * Lc), "copy in");
    detail::check(cosnark_distributed_mt(detail::ctx(), cw.u64(), rows, Lc, nodes.u8()), "distributed_MT");
    std::vector<uint8_t> h(32 * (2 * Lc - 1));
    nodes.get(h.data(), h.size());
    Com.Base_MT.clear();
    detail::unpack_levels(h, Lc, Com.Base_MT);
    if (N == 1) {
        Com.Root_MT.assign(1, {Com.Base_MT.back()[0]});
    }
}

// encode_concatenated(Commit_Data, C_prime, C, rate, degree),
// src/coPCS.cpp:653-677: C_prime = RS encode of Commit_Data at rate[0];
// C[j] = the rate[1]-blowup encoding of segment j (C_prime[j + k S], k < degree)
template <class F>
void encode_concatenated(std::vector<F> &Commit_Data, std::vector<F> &C_prime, std::vector<std::vector<F>> &C,
                         std::vector<int> rate, int degree) {
    detail::elem_type<F>();
    const uint64_t n = Commit_Data.size(), m = n * (uint64_t)rate.at(0);
    const uint64_t S = m / (uint64_t)degree, DP = (uint64_t)degree * (uint64_t)rate.at(1);
    detail::Dev d(Commit_Data.data(), 16 * n), cp(16 * m), c(16 * S * DP);
    detail::check(cosnark_encode_concatenated(detail::ctx(), d.u64(), n, rate[0], rate[1], degree, cp.u64(), c.u64()),
                  "encode_concatenated");
    C_prime.resize(m);
    cp.get(C_prime.data(), 16 * m);
    C.resize(S);
    for (uint64_t j = 0; j < S; ++j) {
        C[j].resize(DP);
        detail::check(cosnark_d2h(detail::ctx(), C[j].data(), c.u8() + 16 * DP * j, 16 * DP), "copy out");
    }
}

namespace detail {
// H[i] += r * H_r[i] with H_r = R[iter] and H_r[0] += R[j][stride * k] for
// j > iter, k < |R[j]| / stride (the coSumcheck masks: stride 2 at
// src/coSumcheck.cpp:182-193, stride 4 at :48-58)
template <class F>
void add_mask(std::vector<F> &H, const std::vector<std::vector<F>> &R, int iter, const F &r, size_t stride,
              const char *what) {
    std::vector<F> Hr = R.at(iter);
    if (Hr.size() < H.size()) throw error(COSNARK_E_ARG, std::string(what) + ": R[iter] shorter than the polynomial");
    auto w = [](const F &x) { return reinterpret_cast<const uint64_t *>(&x); };
    uint64_t h0[2] = {w(Hr[0])[0], w(Hr[0])[1]};
    for (size_t j = iter + 1; j < R.size(); ++j)
        for (size_t k = 0; k < R[j].size() / stride; ++k) cosnark_field_add(h0, w(R[j][stride * k]), h0);
    std::memcpy(&Hr[0], h0, 16);
    uint64_t rr[2];
    std::memcpy(rr, &r, 16);
    for (size_t i = 0; i < H.size(); ++i) {
        uint64_t t[2], h[2];
        cosnark_field_mul(rr, w(Hr[i]), t);
        cosnark_field_add(w(H[i]), t, h);
        std::memcpy(&H[i], h, 16);
    }
}
}  // namespace detail

// commit_concatenated(C, Com, N), src/coPCS.cpp:679-689 (one rank):
// MT_commit_Blake over the rows of C back to back; at N = 1 Root_MT is its root
template <class F, class MTT>
void commit_concatenated(std::vector<std::vector<F>> &C, MTT &Com, int N) {
    detail::elem_type<F>();
    std::vector<F> leaves;
    for (auto &row : C) leaves.insert(leaves.end(), row.begin(), row.end());
    Com.Base_MT.clear();
    Com.Root_MT.clear();
    MT_commit_Blake(leaves.data(), Com.Base_MT, (int)leaves.size());
    if (N == 1) Com.Root_MT.assign(1, {Com.Base_MT.back()[0]});
}

// pcs_sumcheck_phase1(Commit_data, B, degree, N), src/coPCS.cpp:547-569:
// this rank's 2*degree-element polynomial (pcs_sumcheck_aggr, the MPI
// aggregation, is the caller's).
template <class F>
std::vector<F> pcs_sumcheck_phase1(std::vector<F> &Commit_data, std::vector<F> &B, int degree, int) {
    detail::elem_type<F>();
    const uint64_t M = Commit_data.size();
    if (B.size() != M) throw error(COSNARK_E_ARG, "pcs_sumcheck_phase1: sizes");
    detail::Dev a(Commit_data.data(), 16 * M), b(B.data(), 16 * M);
    std::vector<F> poly(2 * (size_t)degree);
    detail::check(cosnark_cosumcheck_phase1(detail::ctx(), a.u64(), b.u64(), M, degree,
                                            reinterpret_cast<uint64_t *>(poly.data())),
                  "pcs_sumcheck_phase1");
    return poly;
}

// pcs_sumcheck_phase2(Commit_data, B, degree, r), src/coPCS.cpp:571-587
template <class F>
void pcs_sumcheck_phase2(std::vector<F> &Commit_data, std::vector<F> &B, int degree, F r) {
    detail::elem_type<F>();
    const uint64_t M = Commit_data.size();
    if (B.size() != M) throw error(COSNARK_E_ARG, "pcs_sumcheck_phase2: sizes");
    detail::Dev a(Commit_data.data(), 16 * M), b(B.data(), 16 * M);
    uint64_t rr[2];
    std::memcpy(rr, &r, 16);
    detail::check(cosnark_cosumcheck_phase2(detail::ctx(), a.u64(), b.u64(), M, degree, rr), "pcs_sumcheck_phase2");
    Commit_data.resize(M / degree);
    B.resize(M / degree);
    a.get(Commit_data.data(), 16 * (M / degree));
    b.get(B.data(), 16 * (M / degree));
}

// quadratic_coSumcheck_phase1(p1, p2, R, r, d, iter, N),
// src/coSumcheck.cpp:150-198, this rank's part: the block sum plus r * H_r
// with H_r = R[iter] and H_r[0] += the even entries of R[iter+1..]
// (sumcheck_aggr, the MPI aggregation, is the caller's).
template <class F>
std::vector<F> quadratic_coSumcheck_phase1(std::vector<F> &p1, std::vector<F> &p2, std::vector<std::vector<F>> R, F r,
                                           int d, int iter, int) {
    detail::elem_type<F>();
    const uint64_t M = p1.size();
    if (p2.size() != M) throw error(COSNARK_E_ARG, "quadratic_coSumcheck_phase1: sizes");
    detail::Dev a(p1.data(), 16 * M), b(p2.data(), 16 * M);
    std::vector<F> H(2 * (size_t)d);
    detail::check(cosnark_cosumcheck_phase1(detail::ctx(), a.u64(), b.u64(), M, d, reinterpret_cast<uint64_t *>(H.data())),
                  "quadratic_coSumcheck_phase1");
    // the mask (src/coSumcheck.cpp:182-193)
    detail::add_mask(H, R, iter, r, 2, "quadratic_coSumcheck_phase1");
    return H;
}

// quadratic_coSumcheck_phase2(p1, p2, L, d), src/coSumcheck.cpp:201-213
template <class F>
void quadratic_coSumcheck_phase2(std::vector<F> &p1, std::vector<F> &p2, std::vector<F> L, int d) {
    detail::elem_type<F>();
    const uint64_t M = p1.size();
    if (p2.size() != M || (int)L.size() < d) throw error(COSNARK_E_ARG, "quadratic_coSumcheck_phase2: sizes");
    detail::Dev a(p1.data(), 16 * M), b(p2.data(), 16 * M);
    detail::check(cosnark_cosumcheck_fold(detail::ctx(), a.u64(), b.u64(), M, d, reinterpret_cast<const uint64_t *>(L.data())),
                  "quadratic_coSumcheck_phase2");
    p1.resize(M / d);
    p2.resize(M / d);
    a.get(p1.data(), 16 * (M / d));
    b.get(p2.data(), 16 * (M / d));
}

// cubic_coSumcheck_phase1(p1, p2, p3, b1, b2, R, r, d, iter, N),
// src/coSumcheck.cpp:9-62, this rank's part: the cubic block sum (blocks
// whose p1 part is zero skipped) plus r * H_r with H_r = R[iter] and
// H_r[0] += the entries 4k of R[iter+1..].  sumcheck_aggr(H, b2, N), the
// MPI aggregation, is the caller's: rank 0 combines the gathered parts
// with sumcheck_aggr_gathered.
template <class F>
std::vector<F> cubic_coSumcheck_phase1(std::vector<F> &p1, std::vector<F> &p2, std::vector<F> &p3, std::vector<F> &b1,
                                       std::vector<F> &, std::vector<std::vector<F>> R, F r, int d, int iter, int) {
    detail::elem_type<F>();
    const uint64_t M = p1.size();
    if (p2.size() != M || p3.size() != M || b1.size() != M) throw error(COSNARK_E_ARG, "cubic_coSumcheck_phase1: sizes");
    detail::Dev a(p1.data(), 16 * M), b(p2.data(), 16 * M), c(p3.data(), 16 * M), e(b1.data(), 16 * M);
    std::vector<F> H(4 * (size_t)d);
    detail::check(cosnark_cosumcheck3_phase1(detail::ctx(), a.u64(), b.u64(), c.u64(), e.u64(), M, d,
                                             reinterpret_cast<uint64_t *>(H.data())),
                  "cubic_coSumcheck_phase1");
    detail::add_mask(H, R, iter, r, 4, "cubic_coSumcheck_phase1");
    return H;
}

// cubic_coSumcheck_phase2(p1, p2, p3, b1, L, d), src/coSumcheck.cpp:64-148
template <class F>
void cubic_coSumcheck_phase2(std::vector<F> &p1, std::vector<F> &p2, std::vector<F> &p3, std::vector<F> &b1,
                             std::vector<F> L, int d) {
    detail::elem_type<F>();
    const uint64_t M = p1.size();
    if (p2.size() != M || p3.size() != M || b1.size() != M || (int)L.size() < d)
        throw error(COSNARK_E_ARG, "cubic_coSumcheck_phase2: sizes");
    detail::Dev a(p1.data(), 16 * M), b(p2.data(), 16 * M), c(p3.data(), 16 * M), e(b1.data(), 16 * M);
    detail::check(cosnark_cosumcheck3_fold(detail::ctx(), a.u64(), b.u64(), c.u64(), e.u64(), M, d,
                                           reinterpret_cast<const uint64_t *>(L.data())),
                  "cubic_coSumcheck_phase2");
    for (auto *v : {&p1, &p2, &p3, &b1}) v->resize(M / d);
    a.get(p1.data(), 16 * (M / d));
    b.get(p2.data(), 16 * (M / d));
    c.get(p3.data(), 16 * (M / d));
    e.get(b1.data(), 16 * (M / d));
}

namespace detail {
template <class F>
void split_ip_vec(const std::vector<F> &flat, const std::vector<int> &degrees, int mult) {
    auto &ip = ip_vec<F>();
    ip.assign(degrees.size(), {});
    size_t off = 0;
    for (size_t i = 0; i < degrees.size(); ++i) {
        ip[i].assign(flat.begin() + off, flat.begin() + off + (size_t)mult * degrees[i]);
        off += (size_t)mult * degrees[i];
    }
}
template <class F>
std::vector<const uint64_t *> mask_ptrs(const std::vector<std::vector<F>> &R, std::vector<uint64_t> &sizes) {
    std::vector<const uint64_t *> p;
    for (auto &x : R) {
        p.push_back(reinterpret_cast<const uint64_t *>(x.data()));
        sizes.push_back(x.size());
    }
    return p;
}
}  // namespace detail

// quadratic_coSumcheck(p1, p2, R, K, K_r, degrees, N, ps), src/coSumcheck.cpp:
// 233-331, as comm().rank of N (N >= 4) exchanging through comm().gather:
// returns the five claims with their common point; p1 and p2 are folded
// and resized as the reference leaves them, ip_vec<F>() is set, rank 0
// adds the proof size to ps.  The reference's printed check failures are
// not errors there; here they are counted in last_check_failures().
inline int &last_check_failures() { thread_local int v = 0; return v; }
template <class F>
std::vector<std::pair<F, std::vector<F>>> quadratic_coSumcheck(std::vector<F> &p1, std::vector<F> &p2,
                                                               std::vector<std::vector<F>> R, F K, F K_r,
                                                               std::vector<int> degrees, int N, double &ps) {
    detail::elem_type<F>();
    const uint64_t M = p1.size();
    if (p2.size() != M || !comm().gather) throw error(COSNARK_E_ARG, "quadratic_coSumcheck: sizes or no comm().gather");
    detail::Dev a(p1.data(), 16 * M), b(p2.data(), 16 * M);
    std::vector<uint64_t> sizes;
    auto ptrs = detail::mask_ptrs(R, sizes);
    uint64_t k[2], kr[2];
    std::memcpy(k, &K, 16);
    std::memcpy(kr, &K_r, 16);
    uint64_t total = 0;
    for (int d : degrees) total += 2 * (uint64_t)d;
    std::vector<F> claims(5), r(degrees.size() + 1), ip(total);
    int fails = 0;
    detail::check(cosnark_quadratic_cosumcheck(detail::ctx(), a.u64(), b.u64(), M, ptrs.data(), sizes.data(),
                                               (int)R.size(), k, kr, degrees.data(), (int)degrees.size(), N,
                                               comm().rank, comm().gather, comm().user, &transcript(),
                                               reinterpret_cast<uint64_t *>(claims.data()),
                                               reinterpret_cast<uint64_t *>(r.data()),
                                               reinterpret_cast<uint64_t *>(ip.data()), &ps, &fails),
                  "quadratic_coSumcheck");
    last_check_failures() = fails;
    uint64_t Mc = M;
    for (int d : degrees) Mc /= (uint64_t)d;
    p1.resize(Mc);
    p2.resize(Mc);
    a.get(p1.data(), 16 * Mc);
    b.get(p2.data(), 16 * Mc);
    detail::split_ip_vec(ip, degrees, 2);
    std::vector<std::pair<F, std::vector<F>>> out;
    for (auto &c : claims) out.push_back({c, r});
    return out;
}

// cubic_coSumcheck(p1, p2, p3, r1, r2, R, K, K_r, degrees, N, ps),
// src/coSumcheck.cpp:335-448, as comm().rank of N: the four claims with
// their common point; p1, p2, p3 folded and resized; ip_vec<F>() set.
template <class F>
std::vector<std::pair<F, std::vector<F>>> cubic_coSumcheck(std::vector<F> &p1, std::vector<F> &p2, std::vector<F> &p3,
                                                           std::vector<F> &r1, std::vector<F> &r2,
                                                           std::vector<std::vector<F>> R, F K, F K_r,
                                                           std::vector<int> degrees, int N, double &ps) {
    detail::elem_type<F>();
    const uint64_t M = p1.size();
    if (p2.size() != M || p3.size() != M || !comm().gather)
        throw error(COSNARK_E_ARG, "cubic_coSumcheck: sizes or no comm().gather");
    detail::Dev a(p1.data(), 16 * M), b(p2.data(), 16 * M), c(p3.data(), 16 * M);
    std::vector<uint64_t> sizes;
    auto ptrs = detail::mask_ptrs(R, sizes);
    uint64_t k[2], kr[2];
    std::memcpy(k, &K, 16);
    std::memcpy(kr, &K_r, 16);
    uint64_t total = 0;
    for (int d : degrees) total += 4 * (uint64_t)d;
    std::vector<F> claims(4), r(degrees.size() + 1), ip(total);
    int fails = 0;
    detail::check(cosnark_cubic_cosumcheck(detail::ctx(), a.u64(), b.u64(), c.u64(), M,
                                           reinterpret_cast<const uint64_t *>(r1.data()), (int)r1.size(),
                                           reinterpret_cast<const uint64_t *>(r2.data()), (int)r2.size(), ptrs.data(),
                                           sizes.data(), (int)R.size(), k, kr, degrees.data(), (int)degrees.size(), N,
                                           comm().rank, comm().gather, comm().user, &transcript(),
                                           reinterpret_cast<uint64_t *>(claims.data()),
                                           reinterpret_cast<uint64_t *>(r.data()),
                                           reinterpret_cast<uint64_t *>(ip.data()), &ps, &fails),
                  "cubic_coSumcheck");
    last_check_failures() = fails;
    uint64_t Mc = M;
    for (int d : degrees) Mc /= (uint64_t)d;
    for (auto *v : {&p1, &p2, &p3}) v->resize(Mc);
    a.get(p1.data(), 16 * Mc);
    b.get(p2.data(), 16 * Mc);
    c.get(p3.data(), 16 * Mc);
    detail::split_ip_vec(ip, degrees, 4);
    std::vector<std::pair<F, std::vector<F>>> out;
    for (auto &x : claims) out.push_back({x, r});
    return out;
}

// unpack_secret_share(secret, _k, k), src/MPI_utils.cpp:2769-2785 (host):
// secret is left transformed as the reference leaves it
template <class F>
std::vector<F> unpack_secret_share(std::vector<F> &secret, int _k, int) {
    detail::elem_type<F>();
    std::vector<F> v((size_t)_k);
    const int rc = cosnark_share_unpack(reinterpret_cast<uint64_t *>(secret.data()), (int)secret.size(), _k,
                                        reinterpret_cast<uint64_t *>(v.data()));
    if (rc) throw error(rc, "unpack_secret_share: sizes");
    return v;
}

// rank 0's combine in sumcheck_aggr(poly, b, N), src/MPI_utils.cpp:2604-2645,
// over the polynomials gathered from the N = |P| ranks (P[m] = rank m's)
template <class F>
std::vector<F> sumcheck_aggr_gathered(const std::vector<std::vector<F>> &P, const std::vector<F> &b) {
    detail::elem_type<F>();
    const int N = (int)P.size();
    const size_t len = N ? P[0].size() : 0;
    std::vector<F> flat;
    for (auto &row : P) {
        if (row.size() != len) throw error(COSNARK_E_ARG, "sumcheck_aggr: ranks' polynomials differ in size");
        flat.insert(flat.end(), row.begin(), row.end());
    }
    if ((int)b.size() < N / 4) throw error(COSNARK_E_ARG, "sumcheck_aggr: b shorter than N/4");
    std::vector<F> out(len);
    const int rc = cosnark_sumcheck_aggr(reinterpret_cast<const uint64_t *>(flat.data()), N, (int)len,
                                         reinterpret_cast<const uint64_t *>(b.data()),
                                         reinterpret_cast<uint64_t *>(out.data()));
    if (rc) throw error(rc, "sumcheck_aggr: sizes");
    return out;
}

namespace detail {
// a tree as MT_commit_Blake leaves it (levels, leaves first) onto the device
template <class H>
void tree_in(const std::vector<std::vector<H>> &MT, uint64_t cw, Dev &nodes, const char *what) {
    const uint64_t cnt = cosnark_merkle_node_count(cw);
    if (MT.empty() || MT[0].size() != cw / 4) throw error(COSNARK_E_ARG, std::string(what) + ": tree does not match");
    uint64_t off = 0;
    for (size_t l = 0; l < MT.size(); ++l) {
        if (off + MT[l].size() > cnt) throw error(COSNARK_E_ARG, std::string(what) + ": tree larger than expected");
        check(cosnark_h2d(ctx(), nodes.u8() + 32 * off, MT[l].data(), 32 * MT[l].size()), "copy in");
        off += MT[l].size();
    }
    if (off != cnt) throw error(COSNARK_E_ARG, std::string(what) + ": tree smaller than expected");
}
}  // namespace detail

// get_whir_rates with the sizes (sorted largest first) passed in and the
// rates returned, for callers that keep no globals: the rates
// get_whir_rates(rate) appends for the same sizes.
inline std::vector<int> get_whir_rates(const std::vector<int> &sizes, int rate, int whir_k) {
    std::vector<uint64_t> sz(sizes.begin(), sizes.end());
    std::vector<int> rates(sizes.size());
    if (sizes.empty() || cosnark_whir_rates(sz.data(), (int)sz.size(), rate, whir_k, rates.data()))
        throw error(COSNARK_E_ARG, "get_whir_rates: sizes must be sorted largest first");
    return rates;
}

// Whir_Batch_Open(arrays, codewords, MTs, evals, betas, rate, Verify),
// src/Whir.cpp:706-899, with the global whir_k (or the last argument when
// it is given).  The arrays are
// sorted largest first, each committed at get_whir_rates' rate.  The
// reference consumes its vector arguments as it aggregates them; they are
// left empty here.  Its record is in last_whir_record().  A failed check
// throws whatever Verify says (the reference prints for Verify = false).
template <class F, class H>
void Whir_Batch_Open(std::vector<std::vector<F>> &arrays, std::vector<std::vector<F>> &codewords,
                     std::vector<std::vector<std::vector<H>>> &MTs, std::vector<F> evals,
                     std::vector<std::vector<F>> &betas, int rate, bool, int wk = 0) {
    const int whir_k = wk ? wk : cosnark::dropin::whir_k();
    detail::elem_type<F>();
    detail::digest_type<H>();
    const size_t c = arrays.size();
    if (!c || codewords.size() != c || MTs.size() != c || betas.size() != c || evals.empty())
        throw error(COSNARK_E_ARG, "Whir_Batch_Open: one codeword, tree and weight vector per array");
    std::vector<std::unique_ptr<detail::Dev>> keep;
    std::vector<const uint64_t *> pa(c), pc(c), pb(c);
    std::vector<const uint8_t *> pt(c);
    std::vector<uint64_t> sizes(c), cws(c);
    for (size_t i = 0; i < c; ++i) {
        sizes[i] = arrays[i].size();
        cws[i] = codewords[i].size();
        if (betas[i].size() != sizes[i]) throw error(COSNARK_E_ARG, "Whir_Batch_Open: weights and array sizes differ");
        keep.emplace_back(new detail::Dev(arrays[i].data(), 16 * sizes[i]));
        pa[i] = keep.back()->u64();
        keep.emplace_back(new detail::Dev(codewords[i].data(), 16 * cws[i]));
        pc[i] = keep.back()->u64();
        keep.emplace_back(new detail::Dev(betas[i].data(), 16 * sizes[i]));
        pb[i] = keep.back()->u64();
        keep.emplace_back(new detail::Dev(32 * cosnark_merkle_node_count(cws[i])));
        detail::tree_in(MTs[i], cws[i], *keep.back(), "Whir_Batch_Open");
        pt[i] = keep.back()->u8();
    }
    auto &rec = last_whir_record();
    rec.assign(1u << 22, 0);
    uint64_t words = 0;
    detail::check(cosnark_whir_batch_open(detail::ctx(), (int)c, pa.data(), sizes.data(), pc.data(), cws.data(),
                                          pt.data(), reinterpret_cast<const uint64_t *>(evals.data()), pb.data(), rate,
                                          whir_k, rec.data(), rec.size(), &words),
                  "Whir_Batch_Open");
    rec.resize(words);
    arrays.clear();
    codewords.clear();
    MTs.clear();
    betas.clear();
}

// aggregate_claims(arrays, arr, b_aggr, claims, K), src/Whir.cpp:608-668:
// arr and b_aggr receive the combined array and weights (the reference adds
// to them as given; its callers pass empty vectors), K the combined claim;
// returns {weight challenges, array challenges} drawn from libc.
template <class F>
std::vector<std::vector<F>> aggregate_claims(std::vector<std::vector<F>> &arrays, std::vector<F> &arr,
                                             std::vector<F> &b_aggr,
                                             std::vector<std::vector<std::pair<F, std::vector<F>>>> claims, F &K) {
    detail::elem_type<F>();
    if (arrays.empty()) throw error(COSNARK_E_ARG, "aggregate_claims: no arrays");
    const uint64_t n = arrays[0].size();
    std::vector<F> pts;
    int np = 0;
    for (auto &ci : claims)
        for (auto &cl : ci) {
            if ((uint64_t(1) << cl.second.size()) != n) throw error(COSNARK_E_ARG, "aggregate_claims: point size");
            pts.insert(pts.end(), cl.second.begin(), cl.second.end());
            ++np;
        }
    std::vector<std::unique_ptr<detail::Dev>> keep;
    std::vector<const uint64_t *> pa;
    for (auto &a : arrays) {
        if (a.size() != n) throw error(COSNARK_E_ARG, "aggregate_claims: arrays differ in size");
        keep.emplace_back(new detail::Dev(a.data(), 16 * n));
        pa.push_back(keep.back()->u64());
    }
    detail::Dev da(16 * n), db(16 * n);
    std::vector<F> bc((size_t)np), ac(arrays.size());
    uint64_t k[2];
    detail::check(cosnark_whir_aggregate_claims(detail::ctx(), (int)arrays.size(), pa.data(), n, np,
                                                reinterpret_cast<const uint64_t *>(pts.data()), da.u64(), db.u64(), k,
                                                reinterpret_cast<uint64_t *>(bc.data()),
                                                reinterpret_cast<uint64_t *>(ac.data())),
                  "aggregate_claims");
    arr.resize(n);
    b_aggr.resize(n);
    da.get(arr.data(), 16 * n);
    db.get(b_aggr.data(), 16 * n);
    std::memcpy(&K, k, 16);
    return {bc, ac};
}

// RS_multifold5(codeword, a, fold_degree), src/coPCS_utils.cpp:386-437:
// the codeword is replaced by its fold (|codeword| / fold_degree elements)
template <class F>
void RS_multifold5(std::vector<F> &codeword, std::vector<F> a, int fold_degree) {
    detail::elem_type<F>();
    if ((int)a.size() != fold_degree) throw error(COSNARK_E_ARG, "RS_multifold5: a.size() != fold_degree");
    const uint64_t n = codeword.size();
    detail::Dev c(codeword.data(), 16 * n), out(16 * (n / fold_degree));
    detail::check(cosnark_rs_multifold5(detail::ctx(), c.u64(), n, reinterpret_cast<const uint64_t *>(a.data()),
                                        fold_degree, out.u64()),
                  "RS_multifold5");
    codeword.resize(n / fold_degree);
    out.get(codeword.data(), 16 * codeword.size());
}

}  // namespace dropin
}  // namespace cosnark
