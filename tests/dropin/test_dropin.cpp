// The drop-in layer (include/cosnark_dropin.hpp) driven as the reference's
// callers drive its functions, with the reference's type shapes (F =
// {u64 real, img}, _hash = {u8 arr[32]}, MT = {Base_MT, Root_MT}), checked
// against the oracle (oracle/oracle.h).  Exit status 0 when every check
// passes; prints the first failure otherwise.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <algorithm>
#include <condition_variable>
#include <mutex>
#include <thread>
#include <vector>
#include "cosnark_dropin.hpp"
#include "../../oracle/oracle.h"

using std::vector;
namespace D = cosnark::dropin;

struct F { uint64_t real, img; };
struct _hash { uint8_t arr[32]; };
struct MT { vector<vector<_hash>> Base_MT, Root_MT; };

static int failures = 0;
#define CHECK(cond, what)                                                  \
    do {                                                                   \
        if (!(cond)) { std::printf("FAIL: %s\n", what); ++failures; }      \
        else std::printf("ok: %s\n", what);                                \
        std::fflush(stdout);                                               \
    } while (0)

static vector<F> synth(uint64_t seed, uint64_t stream, uint64_t n) {
    vector<F> v(n);
    or_synth_fill(seed, stream, reinterpret_cast<or_fe *>(v.data()), n);
    return v;
}
static bool same(const void *a, const void *b, size_t bytes) { return std::memcmp(a, b, bytes) == 0; }
static or_fe O(const F &x) { return or_fe{x.real, x.img}; }

// an all-gather among the threads of this process acting as ranks
struct ThreadGather {
    int n;
    std::mutex mu;
    std::condition_variable cv;
    int arrived = 0;
    long gen = 0;
    std::vector<std::vector<uint64_t>> parts;
    explicit ThreadGather(int n_) : n(n_), parts(n_) {}
    void barrier() {
        std::unique_lock<std::mutex> lk(mu);
        const long g = gen;
        if (++arrived == n) { arrived = 0; ++gen; cv.notify_all(); }
        else cv.wait(lk, [&] { return gen != g; });
    }
    static int fn(void *user, const uint64_t *mine, uint64_t elems, uint64_t *all) {
        auto *T = static_cast<std::pair<ThreadGather *, int> *>(user);
        ThreadGather &G = *T->first;
        G.parts[T->second].assign(mine, mine + 2 * elems);
        G.barrier();
        for (int m = 0; m < G.n; ++m) std::memcpy(all + 2 * elems * m, G.parts[m].data(), 16 * elems);
        G.barrier();
        return 0;
    }
};

int main() {
    const uint64_t seed = 0xC05A0040;
    // fft / inverse
    {
        vector<F> a = synth(seed, 0, 1 << 12), ref = a;
        D::fft(a, 12, false);
        or_fft(reinterpret_cast<or_fe *>(ref.data()), 12, 0);
        CHECK(same(a.data(), ref.data(), 16 * a.size()), "fft forward 2^12");
        D::fft(a, 12, true);
        or_fft(reinterpret_cast<or_fe *>(ref.data()), 12, 1);
        CHECK(same(a.data(), ref.data(), 16 * a.size()), "fft inverse 2^12");
    }
    // MT_commit_Blake, create_tree_blake, open_tree_blake
    {
        const int N = 1 << 12;
        vector<F> e = synth(seed, 1, N);
        vector<vector<_hash>> H;
        D::MT_commit_Blake(e.data(), H, N);
        vector<uint8_t> nodes(32 * or_merkle_node_count(N));
        or_merkle_commit(reinterpret_cast<or_fe *>(e.data()), N, nodes.data(), 1);
        bool ok = H.size() == 11;
        uint64_t off = 0;
        for (size_t l = 0; ok && l < H.size(); ++l) {
            ok = same(H[l].data(), nodes.data() + 32 * off, 32 * H[l].size());
            off += H[l].size();
        }
        CHECK(ok, "MT_commit_Blake 2^12 elements, every level");
        vector<vector<_hash>> T(H.size());
        T[0] = H[0];
        D::create_tree_blake((int)H[0].size(), T, 32, true);
        ok = true;
        for (size_t l = 0; l < H.size(); ++l) ok = ok && same(T[l].data(), H[l].data(), 32 * H[l].size());
        CHECK(ok, "create_tree_blake over the leaf digests");
        vector<_hash> path = D::open_tree_blake(H, 77);
        vector<uint8_t> opath(32 * 64);
        const int depth = or_merkle_open(nodes.data(), N, 77, opath.data());
        CHECK((int)path.size() == depth && same(path.data(), opath.data(), 32 * depth), "open_tree_blake");
    }
    // Fiat-Shamir
    {
        D::transcript() = cosnark_transcript{{0, 0}};
        vector<F> el = synth(seed, 2, 5);
        F r = D::hash_to_field(el);
        or_fe prev{0, 0};
        or_fe ro = or_hash_to_field(&prev, reinterpret_cast<or_fe *>(el.data()), 5);
        CHECK(r.real == ro.real && r.img == ro.img, "hash_to_field");
        auto idx = D::get_indexes(6, 1 << 10, 1 << 8, -1);
        vector<uint32_t> oi(12);
        or_get_indexes(&prev, 6, 1 << 10, 1 << 8, oi.data());
        bool ok = idx.size() == 6;
        for (int i = 0; ok && i < 6; ++i) ok = idx[i][0] == oi[2 * i] && idx[i][1] == oi[2 * i + 1];
        CHECK(ok, "get_indexes");
    }
    // _quadratic_sumcheck, all Fiat-Shamir
    {
        D::transcript() = cosnark_transcript{{0, 0}};
        D::sumcheck_offset() = 0;
        const uint64_t n = 1 << 14;
        vector<F> v1 = synth(seed, 3, n), v2 = synth(seed, 4, n);
        or_fe y{0, 0};
        for (uint64_t i = 0; i < n; ++i) y = or_add(y, or_mul(O(v1[i]), O(v2[i])));
        auto claims = D::_quadratic_sumcheck(F{y.real, y.img}, v1, v2, 1);
        vector<F> a = v1, b = v2, polys(3 * 14), r(14), fin(2);
        or_fe prev{0, 0};
        const int rc = or_product_sumcheck(y, reinterpret_cast<or_fe *>(a.data()), reinterpret_cast<or_fe *>(b.data()),
                                           n, &prev, 0, reinterpret_cast<or_fe *>(polys.data()),
                                           reinterpret_cast<or_fe *>(r.data()), reinterpret_cast<or_fe *>(fin.data()));
        CHECK(rc == 0 && claims.size() == 2 && same(&claims[0].first, &fin[0], 16) && same(&claims[1].first, &fin[1], 16) &&
                  claims[0].second.size() == 14 && same(claims[0].second.data(), r.data(), 16 * 14),
              "_quadratic_sumcheck 2^14");
        CHECK(D::transcript().prev[0] == prev.real && D::transcript().prev[1] == prev.img, "transcript after sumcheck");
    }
    // WHIR commit + open against the oracle's record
    {
        const uint64_t n = 1 << 12;
        const int rate = 4;
        vector<F> arr = synth(seed, 5, n), z = synth(seed, 6, 12), cw;
        vector<vector<_hash>> tree;
        D::Whir_Commit_initial(arr, cw, tree, rate, 4);
        vector<F> ocw(n * rate);
        or_rs_encode(reinterpret_cast<or_fe *>(arr.data()), n, rate, reinterpret_cast<or_fe *>(ocw.data()));
        CHECK(same(cw.data(), ocw.data(), 16 * cw.size()), "Whir_Commit_initial codeword");
        vector<F> beta(n);
        or_precompute_beta(reinterpret_cast<or_fe *>(z.data()), 12, reinterpret_cast<or_fe *>(beta.data()));
        or_fe K{0, 0};
        for (uint64_t i = 0; i < n; ++i) K = or_add(K, or_mul(O(arr[i]), O(beta[i])));
        vector<uint8_t> onodes(32 * or_merkle_node_count(n * rate));
        or_merkle_commit(reinterpret_cast<or_fe *>(ocw.data()), n * rate, onodes.data(), 1);
        // at the reference's default whir_k = 8 and at 4
        for (int wk : {8, 4}) {
            D::whir_k() = wk;
            srandom(1234);
            D::Whir_Open(arr, cw, tree, {{F{K.real, K.img}, z}}, rate);
            vector<uint64_t> orec(1 << 20);
            vector<F> ochal(12);
            srandom(1234);
            const int64_t w = or_whir_open_weights(reinterpret_cast<or_fe *>(arr.data()), n,
                                                   reinterpret_cast<or_fe *>(beta.data()), K,
                                                   reinterpret_cast<or_fe *>(ocw.data()), n * rate, onodes.data(), 0,
                                                   rate, wk, orec.data(), orec.size(),
                                                   reinterpret_cast<or_fe *>(ochal.data()), nullptr);
            const auto &rec = D::last_whir_record();
            CHECK(w > 0 && (int64_t)rec.size() == w && same(rec.data(), orec.data(), 8 * w),
                  wk == 8 ? "Whir_Open record 2^12, whir_k 8" : "Whir_Open record 2^12, whir_k 4");
        }
        // whir_k = 4 with one claim: the same record as the one-array batch open
        {
            vector<uint64_t> orec(1 << 20), orec2(1 << 20);
            const or_fe *pa = reinterpret_cast<or_fe *>(arr.data()), *pc = reinterpret_cast<or_fe *>(ocw.data()),
                        *pb = reinterpret_cast<or_fe *>(beta.data());
            const uint8_t *pt = onodes.data();
            const uint64_t sz = n, cs = n * rate;
            srandom(77);
            const int64_t w1 = or_whir_batch_open(1, &pa, &sz, &pc, &cs, &pt, &K, &pb, rate, 4, orec.data(), orec.size());
            srandom(77);
            D::whir_k() = 4;
            D::Whir_Open(arr, cw, tree, {{F{K.real, K.img}, z}}, rate);
            const auto &rec = D::last_whir_record();
            CHECK(w1 > 0 && (int64_t)rec.size() == w1 && same(rec.data(), orec.data(), 8 * w1),
                  "Whir_Open = one-array Whir_Batch_Open at whir_k 4");
        }
        // two claims batched with one F::_random() power (Whir.cpp:398-411)
        {
            vector<F> z2 = synth(seed, 7, 12), beta2(n);
            or_precompute_beta(reinterpret_cast<or_fe *>(z2.data()), 12, reinterpret_cast<or_fe *>(beta2.data()));
            or_fe K2{0, 0};
            for (uint64_t i = 0; i < n; ++i) K2 = or_add(K2, or_mul(O(arr[i]), O(beta2[i])));
            D::whir_k() = 8;
            srandom(555);
            D::Whir_Open(arr, cw, tree, {{F{K.real, K.img}, z}, {F{K2.real, K2.img}, z2}}, rate);
            const auto rec = D::last_whir_record();
            srandom(555);
            const or_fe r = or_random_element();
            vector<F> bb(n);
            for (uint64_t i = 0; i < n; ++i) {
                const or_fe t = or_add(O(beta[i]), or_mul(r, O(beta2[i])));
                bb[i] = F{t.real, t.img};
            }
            const or_fe KK = or_add(K, or_mul(r, K2));
            vector<uint64_t> orec(1 << 20);
            const int64_t w = or_whir_open_weights(reinterpret_cast<or_fe *>(arr.data()), n,
                                                   reinterpret_cast<or_fe *>(bb.data()), KK,
                                                   reinterpret_cast<or_fe *>(ocw.data()), n * rate, onodes.data(), 0,
                                                   rate, 8, orec.data(), orec.size(), nullptr, nullptr);
            CHECK(w > 0 && (int64_t)rec.size() == w && same(rec.data(), orec.data(), 8 * w), "Whir_Open two claims");
            // three claims at distinct points: the reference batches claims[1]'s
            // weights for every i >= 1 (:403), so the first sumcheck check fails
            vector<F> z3 = synth(seed, 8, 12);
            bool threw = false;
            try {
                D::Whir_Open(arr, cw, tree, {{F{K.real, K.img}, z}, {F{K2.real, K2.img}, z2}, {F{3, 4}, z3}}, rate);
            } catch (const D::error &e) {
                threw = e.code == COSNARK_E_CHECK;
            }
            CHECK(threw, "Whir_Open three claims: the reference's claims[1] batching fails its check");
        }
        // _Whir_Open on the codeword's tree: record, total_r and the final claim
        {
            D::whir_k() = 8;
            srandom(2024);
            vector<vector<std::pair<F, vector<F>>>> cl;
            auto fc = D::_Whir_Open(arr, cw, tree, beta, cl, {}, F{K.real, K.img}, rate, true);
            const auto rec = D::last_whir_record();
            vector<uint64_t> orec(1 << 20);
            vector<F> ochal(12);
            or_fe ofin;
            srandom(2024);
            const int64_t w = or_whir_open_weights(reinterpret_cast<or_fe *>(arr.data()), n,
                                                   reinterpret_cast<or_fe *>(beta.data()), K,
                                                   reinterpret_cast<or_fe *>(ocw.data()), n * rate, onodes.data(), 0,
                                                   rate, 8, orec.data(), orec.size(),
                                                   reinterpret_cast<or_fe *>(ochal.data()), &ofin);
            CHECK(w > 0 && (int64_t)rec.size() == w && same(rec.data(), orec.data(), 8 * w) && fc.second.size() == 12 &&
                      same(fc.second.data(), ochal.data(), 16 * 12) && fc.first.real == ofin.real &&
                      fc.first.img == ofin.img,
                  "_Whir_Open 2^12 whir_k 8: record, total_r, final claim");
        }
    }
    // Whir_Commit, Whir_Commit_Codeword: codeword and the all-zero-leaf tree
    {
        D::whir_k() = 8;
        const uint64_t n = 1 << 10;
        vector<F> arr = synth(seed, 9, n), cw;
        vector<vector<_hash>> mt;
        D::Whir_Commit(arr, cw, mt, 4);
        vector<F> ocw(4 * n);
        or_rs_encode(reinterpret_cast<or_fe *>(arr.data()), n, 4, reinterpret_cast<or_fe *>(ocw.data()));
        // level l of the zero tree: the l-fold self-hash of 32 zero bytes
        auto zero_ok = [&](const vector<vector<_hash>> &t, uint64_t leaves) {
            uint8_t cur[32] = {0}, blk[64];
            int depth = 0;
            while ((1ull << depth) < leaves) ++depth;
            if ((int)t.size() != depth + 1) return false;
            for (int l = 0; l <= depth; ++l) {
                if (l) { memcpy(blk, cur, 32); memcpy(blk + 32, cur, 32); or_blake3_hash64(blk, cur); }
                if (t[l].size() != (leaves >> l)) return false;
                for (auto &h : t[l]) if (!same(h.arr, cur, 32)) return false;
            }
            return true;
        };
        CHECK(same(cw.data(), ocw.data(), 16 * cw.size()) && zero_ok(mt, 4 * n / 8), "Whir_Commit codeword and tree");
        vector<vector<_hash>> mt4;
        D::Whir_Commit(arr, cw, mt4, 4, 4);
        CHECK(zero_ok(mt4, n), "Whir_Commit kappa 4 tree");
        vector<vector<_hash>> mt2;
        D::Whir_Commit_Codeword(cw, mt2, 4);
        CHECK(zero_ok(mt2, 4 * n / 8), "Whir_Commit_Codeword tree");
        // aggregate_commitments over three codewords, then _Whir_Open on its tree
        vector<vector<F>> arrs, cws;
        vector<vector<vector<_hash>>> mts;
        for (int i = 0; i < 3; ++i) {
            arrs.push_back(synth(seed, 120 + i, n));
            vector<F> c;
            vector<vector<_hash>> t;
            D::Whir_Commit_initial(arrs[i], c, t, 4, 4);
            cws.push_back(c);
            mts.push_back(t);
        }
        vector<F> rr = synth(seed, 130, 3), acw;
        vector<vector<_hash>> amt;
        srandom(99);
        D::aggregate_commitments(acw, amt, cws, mts, rr, 4);
        const int after = rand();
        srandom(99);
        for (int i = 0; i < 100; ++i) (void)rand();                  // l = 100 at rate 4
        vector<F> expect(4 * n, F{0, 0});
        for (int i = 0; i < 3; ++i)
            for (uint64_t j = 0; j < 4 * n; ++j) {
                or_fe t = or_add(O(expect[j]), or_mul(O(rr[i]), O(cws[i][j])));
                expect[j] = F{t.real, t.img};
            }
        CHECK(same(acw.data(), expect.data(), 16 * expect.size()) && zero_ok(amt, 4 * n / 8) && after == rand(),
              "aggregate_commitments codeword, tree, libc draws");
        // the aggregated array with combined weights, opened against that tree
        vector<F> aarr(n, F{0, 0}), b = synth(seed, 131, n);
        for (int i = 0; i < 3; ++i)
            for (uint64_t j = 0; j < n; ++j) {
                or_fe t = or_add(O(aarr[j]), or_mul(O(rr[i]), O(arrs[i][j])));
                aarr[j] = F{t.real, t.img};
            }
        or_fe K{0, 0};
        for (uint64_t i = 0; i < n; ++i) K = or_add(K, or_mul(O(aarr[i]), O(b[i])));
        srandom(7);
        vector<vector<std::pair<F, vector<F>>>> cl;
        auto fc = D::_Whir_Open(aarr, acw, amt, b, cl, {}, F{K.real, K.img}, 4, true);
        const auto rec = D::last_whir_record();
        vector<uint64_t> orec(1 << 20);
        vector<F> ochal(10);
        or_fe ofin;
        srandom(7);
        const int64_t w = or_whir_open_weights(reinterpret_cast<or_fe *>(aarr.data()), n, reinterpret_cast<or_fe *>(b.data()),
                                               K, reinterpret_cast<or_fe *>(acw.data()), 4 * n, nullptr, 4 * n / 8, 4, 8,
                                               orec.data(), orec.size(), reinterpret_cast<or_fe *>(ochal.data()), &ofin);
        CHECK(w > 0 && (int64_t)rec.size() == w && same(rec.data(), orec.data(), 8 * w) &&
                  same(fc.second.data(), ochal.data(), 16 * 10) && fc.first.real == ofin.real && fc.first.img == ofin.img,
              "_Whir_Open on aggregate_commitments' tree (coPCS.cpp:1174-1175 pattern)");
    }
    // get_whir_rates(rate) on the globals
    {
        D::whir_k() = 8;
        D::whir_sizes() = {1 << 10, 1 << 12, 1 << 8, 1 << 12};
        D::sorted_whir_rates().clear();
        D::whir_kappas().clear();
        D::get_whir_rates(4);
        const vector<int> want_idx = {2, 1, 3, 0}, want_rates = {4, 4, 8, 16};
        std::vector<uint64_t> srt = {1 << 12, 1 << 12, 1 << 10, 1 << 8};
        vector<int> orates(4);
        or_whir_rates(srt.data(), 4, 4, 8, orates.data());
        CHECK(D::sorted_indexes() == want_idx && D::sorted_whir_rates() == want_rates && D::whir_rates() == want_rates &&
                  orates == want_rates && D::whir_kappas().size() == 3 && D::whir_kappas()[1 << 12] == 4 &&
                  D::whir_kappas()[1 << 10] == 4 && D::whir_kappas()[1 << 8] == 8,
              "get_whir_rates(rate): sorted_indexes, rates, kappas");
        D::whir_sizes() = {1 << 16};
        D::whir_kappas().clear();
        D::get_whir_rates(4);
        CHECK(D::sorted_whir_rates().size() == 5 && D::sorted_whir_rates()[4] == 4 && D::whir_kappas()[1 << 16] == 8,
              "get_whir_rates appends, one size");
    }
    // verify_claim_opt_blake: the reference's bookkeeping
    {
        vector<F> e = synth(seed, 140, 32);
        vector<vector<_hash>> H;
        D::MT_commit_Blake(e.data(), H, 32);                          // 8 leaves, 4 levels
        vector<bool> visited(16, false);
        double ps = 0;
        auto p5 = D::open_tree_blake(H, 5);
        bool ok = D::verify_claim_opt_blake(H, p5.data(), 5, 8, visited, ps);
        // 13 -> siblings 12, 7, 2 marked with the parents 6, 3, 1
        const vector<int> marked = {1, 2, 3, 6, 7, 12};
        bool vis_ok = true;
        for (int i = 0; i < 16; ++i)
            vis_ok = vis_ok && visited[i] == (std::find(marked.begin(), marked.end(), i) != marked.end());
        ok = ok && vis_ok && ps == 3 * 32.0 / 1024.0;
        auto p4 = D::open_tree_blake(H, 4);
        // 12's sibling 13 is new: one level, then its parent's sibling 7 was visited
        ok = ok && D::verify_claim_opt_blake(H, p4.data(), 4, 8, visited, ps) && ps == 4 * 32.0 / 1024.0 && visited[13];
        bool threw = false;
        try { D::verify_claim_opt_blake(H, p4.data(), 4, 4, visited, ps); } catch (const D::error &) { threw = true; }
        CHECK(ok && threw, "verify_claim_opt_blake visited / proof size / exits");
    }
    // coPCS commit: Encode + distributed_MT
    {
        D::rate_offset() = 64;
        const uint64_t L = 1 << 10;
        vector<vector<F>> Codeword(32);
        vector<F> rows;
        for (int i = 0; i < 17; ++i) {
            Codeword[i] = synth(seed, 10 + i, L);
            rows.insert(rows.end(), Codeword[i].begin(), Codeword[i].end());
        }
        D::Encode(Codeword, 1, {4});
        vector<F> ocw(17 * 4 * L);
        or_copcs_encode(reinterpret_cast<or_fe *>(rows.data()), 17, L, 4, 6, reinterpret_cast<or_fe *>(ocw.data()));
        bool ok = Codeword[17].empty();
        for (int i = 0; ok && i < 17; ++i) ok = Codeword[i].size() == 4 * L && same(Codeword[i].data(), &ocw[i * 4 * L], 64 * L);
        CHECK(ok, "Encode 17 rows of 2^10, rate 4, rate_offset 64");
        MT Com;
        D::distributed_MT(Codeword, Com, 1);
        vector<uint8_t> base(32 * (2 * 4 * L - 1));
        or_distributed_mt(reinterpret_cast<or_fe *>(ocw.data()), 17, 4 * L, base.data());
        ok = Com.Base_MT.size() == 13 && Com.Root_MT.size() == 1 && Com.Root_MT[0].size() == 1;
        uint64_t off = 0;
        for (size_t l = 0; ok && l < Com.Base_MT.size(); ++l) {
            ok = same(Com.Base_MT[l].data(), base.data() + 32 * off, 32 * Com.Base_MT[l].size());
            off += Com.Base_MT[l].size();
        }
        ok = ok && same(Com.Root_MT[0][0].arr, Com.Base_MT.back()[0].arr, 32);
        CHECK(ok, "distributed_MT Base_MT and Root_MT");
    }
    // encode_concatenated + commit_concatenated
    {
        vector<F> data = synth(seed, 20, 1 << 10), cprime;
        vector<vector<F>> C;
        D::encode_concatenated(data, cprime, C, {4, 4}, 8);
        vector<F> ocp(4 << 10), oC((4 << 10) / 8 * 32);
        or_encode_concatenated(reinterpret_cast<or_fe *>(data.data()), 1 << 10, 4, 4, 8,
                               reinterpret_cast<or_fe *>(ocp.data()), reinterpret_cast<or_fe *>(oC.data()));
        bool ok = cprime.size() == ocp.size() && same(cprime.data(), ocp.data(), 16 * ocp.size()) && C.size() == 512;
        for (size_t j = 0; ok && j < C.size(); ++j) ok = C[j].size() == 32 && same(C[j].data(), &oC[32 * j], 16 * 32);
        CHECK(ok, "encode_concatenated 2^10, rates {4,4}, degree 8");
        MT Com;
        D::commit_concatenated(C, Com, 1);
        vector<uint8_t> nodes(32 * or_merkle_node_count(oC.size()));
        or_merkle_commit(reinterpret_cast<or_fe *>(oC.data()), oC.size(), nodes.data(), 1);
        CHECK(Com.Root_MT.size() == 1 && same(Com.Root_MT[0][0].arr, nodes.data() + nodes.size() - 32, 32),
              "commit_concatenated root");
    }
    // block sumcheck: pcs phases, coSumcheck phases with the mask
    {
        const uint64_t M = 1 << 12;
        vector<F> p1 = synth(seed, 30, M), p2 = synth(seed, 31, M);
        vector<F> poly = D::pcs_sumcheck_phase1(p1, p2, 8, 1), opoly(16);
        or_pcs_sumcheck_phase1(reinterpret_cast<or_fe *>(p1.data()), reinterpret_cast<or_fe *>(p2.data()), M, 8,
                               reinterpret_cast<or_fe *>(opoly.data()));
        CHECK(poly.size() == 16 && same(poly.data(), opoly.data(), 16 * 16), "pcs_sumcheck_phase1 d=8");
        vector<F> q1 = p1, q2 = p2;
        F r = synth(seed, 32, 1)[0];
        D::pcs_sumcheck_phase2(q1, q2, 8, r);
        vector<F> o1 = p1, o2 = p2;
        or_pcs_sumcheck_phase2(reinterpret_cast<or_fe *>(o1.data()), reinterpret_cast<or_fe *>(o2.data()), M, 8, O(r));
        CHECK(q1.size() == M / 8 && same(q1.data(), o1.data(), 16 * (M / 8)) && same(q2.data(), o2.data(), 16 * (M / 8)),
              "pcs_sumcheck_phase2 d=8");
        // quadratic_coSumcheck: mask from R, fold with given weights
        vector<vector<F>> R = {synth(seed, 33, 64), synth(seed, 34, 16), synth(seed, 35, 8)};
        vector<F> H = D::quadratic_coSumcheck_phase1(p1, p2, R, r, 32, 0, 1);
        vector<F> oh(64);
        or_pcs_sumcheck_phase1(reinterpret_cast<or_fe *>(p1.data()), reinterpret_cast<or_fe *>(p2.data()), M, 32,
                               reinterpret_cast<or_fe *>(oh.data()));
        or_fe h0 = O(R[0][0]);
        for (int j = 1; j < 3; ++j)
            for (size_t k = 0; k < R[j].size() / 2; ++k) h0 = or_add(h0, O(R[j][2 * k]));
        bool ok = H.size() == 64;
        for (int i = 0; ok && i < 64; ++i) {
            const or_fe hr = i == 0 ? h0 : O(R[0][i]);
            const or_fe want = or_add(O(oh[i]), or_mul(O(r), hr));
            ok = H[i].real == want.real && H[i].img == want.img;
        }
        CHECK(ok, "quadratic_coSumcheck_phase1 d=32 with the R mask");
        vector<F> L(32);
        or_compute_lagrange_coeff(O(r), 32, reinterpret_cast<or_fe *>(L.data()));
        vector<F> c1 = p1, c2 = p2, e1 = p1, e2 = p2;
        D::quadratic_coSumcheck_phase2(c1, c2, L, 32);
        or_pcs_sumcheck_phase2(reinterpret_cast<or_fe *>(e1.data()), reinterpret_cast<or_fe *>(e2.data()), M, 32, O(r));
        CHECK(c1.size() == M / 32 && same(c1.data(), e1.data(), 16 * (M / 32)) && same(c2.data(), e2.data(), 16 * (M / 32)),
              "quadratic_coSumcheck_phase2 d=32");
    }
    // cubic coSumcheck phases with the mask; rank 0's share arithmetic
    {
        const uint64_t M = 1 << 12;
        vector<F> p1 = synth(seed, 60, M), p2 = synth(seed, 61, M), p3 = synth(seed, 62, M), b1 = synth(seed, 63, M);
        for (int k = 0; k < 16; ++k) p1[16 * 5 + k] = F{0, 0};           // a skipped block
        vector<F> b2 = synth(seed, 64, 4);
        F r = synth(seed, 65, 1)[0];
        vector<vector<F>> R = {synth(seed, 66, 64), synth(seed, 67, 32), synth(seed, 68, 12)};
        vector<F> H = D::cubic_coSumcheck_phase1(p1, p2, p3, b1, b2, R, r, 16, 0, 1), oh(64);
        or_cosumcheck3_phase1(reinterpret_cast<or_fe *>(p1.data()), reinterpret_cast<or_fe *>(p2.data()),
                              reinterpret_cast<or_fe *>(p3.data()), reinterpret_cast<or_fe *>(b1.data()), M, 16,
                              reinterpret_cast<or_fe *>(oh.data()));
        or_fe h0 = O(R[0][0]);
        for (int j = 1; j < 3; ++j)
            for (size_t k = 0; k < R[j].size() / 4; ++k) h0 = or_add(h0, O(R[j][4 * k]));
        bool ok = H.size() == 64;
        for (int i = 0; ok && i < 64; ++i) {
            const or_fe hr = i == 0 ? h0 : O(R[0][i]);
            const or_fe want = or_add(O(oh[i]), or_mul(O(r), hr));
            ok = H[i].real == want.real && H[i].img == want.img;
        }
        CHECK(ok, "cubic_coSumcheck_phase1 d=16 with the R mask and a zero block");
        vector<F> L(16);
        or_compute_lagrange_coeff(O(r), 16, reinterpret_cast<or_fe *>(L.data()));
        vector<F> c[4] = {p1, p2, p3, b1}, e[4] = {p1, p2, p3, b1};
        D::cubic_coSumcheck_phase2(c[0], c[1], c[2], c[3], L, 16);
        or_cosumcheck3_phase2(reinterpret_cast<or_fe *>(e[0].data()), reinterpret_cast<or_fe *>(e[1].data()),
                              reinterpret_cast<or_fe *>(e[2].data()), reinterpret_cast<or_fe *>(e[3].data()), M, 16,
                              reinterpret_cast<or_fe *>(L.data()));
        ok = true;
        for (int a = 0; a < 4; ++a) ok = ok && c[a].size() == M / 16 && same(c[a].data(), e[a].data(), 16 * (M / 16));
        CHECK(ok, "cubic_coSumcheck_phase2 d=16");
        // unpack_secret_share leaves the shares transformed, returns _k values
        vector<F> sh = synth(seed, 69, 16), osh = sh, v, ov(8);
        v = D::unpack_secret_share(sh, 8, 4);
        or_unpack_secret_share(reinterpret_cast<or_fe *>(osh.data()), 16, 8, reinterpret_cast<or_fe *>(ov.data()));
        CHECK(v.size() == 8 && same(v.data(), ov.data(), 16 * 8) && same(sh.data(), osh.data(), 16 * 16),
              "unpack_secret_share 16 shares, _k = 8");
        // rank 0 combining 8 ranks' polynomials with b2
        vector<vector<F>> P;
        vector<F> flat;
        for (int m = 0; m < 8; ++m) {
            P.push_back(synth(seed, 70 + m, 64));
            flat.insert(flat.end(), P.back().begin(), P.back().end());
        }
        vector<F> agg = D::sumcheck_aggr_gathered(P, b2), oagg(64);
        or_sumcheck_aggr(reinterpret_cast<or_fe *>(flat.data()), 8, 64, reinterpret_cast<or_fe *>(b2.data()),
                         reinterpret_cast<or_fe *>(oagg.data()));
        CHECK(agg.size() == 64 && same(agg.data(), oagg.data(), 16 * 64), "sumcheck_aggr over 8 ranks' polynomials");
    }
    // batched WHIR: rates, Whir_Batch_Open, aggregate_claims
    {
        const std::vector<int> logs = {12, 10, 8, 8};
        std::vector<int> sizes;
        for (int lg : logs) sizes.push_back(1 << lg);
        std::vector<int> rates = D::get_whir_rates(sizes, 4, 4);
        std::vector<uint64_t> usz(sizes.begin(), sizes.end());
        std::vector<int> orates(sizes.size());
        or_whir_rates(usz.data(), (int)usz.size(), 4, 4, orates.data());
        CHECK(rates == orates, "get_whir_rates");
        vector<vector<F>> arrays, codewords, betas;
        vector<vector<vector<_hash>>> MTs;
        vector<F> evals;
        vector<vector<uint8_t>> flat;
        for (size_t i = 0; i < logs.size(); ++i) {
            arrays.push_back(synth(seed, 80 + i, sizes[i]));
            vector<F> z = synth(seed, 90 + i, logs[i]), b(sizes[i]), cw;
            or_precompute_beta(reinterpret_cast<or_fe *>(z.data()), logs[i], reinterpret_cast<or_fe *>(b.data()));
            or_fe K = {0, 0};
            for (int k = 0; k < sizes[i]; ++k) K = or_add(K, or_mul(O(arrays[i][k]), O(b[k])));
            evals.push_back(F{K.real, K.img});
            betas.push_back(b);
            vector<vector<_hash>> mt;
            D::Whir_Commit_initial(arrays[i], cw, mt, rates[i], 4);
            codewords.push_back(cw);
            MTs.push_back(mt);
            vector<uint8_t> nodes(32 * or_merkle_node_count(cw.size()));
            or_merkle_commit(reinterpret_cast<const or_fe *>(cw.data()), cw.size(), nodes.data(), 1);
            flat.push_back(nodes);
        }
        vector<const or_fe *> oa, oc, ob;
        vector<const uint8_t *> ot;
        vector<uint64_t> ocs;
        for (size_t i = 0; i < logs.size(); ++i) {
            oa.push_back(reinterpret_cast<const or_fe *>(arrays[i].data()));
            oc.push_back(reinterpret_cast<const or_fe *>(codewords[i].data()));
            ob.push_back(reinterpret_cast<const or_fe *>(betas[i].data()));
            ot.push_back(flat[i].data());
            ocs.push_back(codewords[i].size());
        }
        vector<uint64_t> orec(1 << 20);
        srandom(41);
        const int64_t ow = or_whir_batch_open((int)logs.size(), oa.data(), usz.data(), oc.data(), ocs.data(), ot.data(),
                                              reinterpret_cast<const or_fe *>(evals.data()), ob.data(), 4, 4,
                                              orec.data(), orec.size());
        srandom(41);
        D::whir_k() = 4;
        D::Whir_Batch_Open(arrays, codewords, MTs, evals, betas, 4, true);
        const auto &rec = D::last_whir_record();
        CHECK(ow > 0 && (int64_t)rec.size() == ow && same(rec.data(), orec.data(), 8 * ow) && arrays.empty(),
              "Whir_Batch_Open 2^12, 2^10, 2^8, 2^8");
        // aggregate_claims over three arrays of 2^8 with two points each
        vector<vector<F>> arrs = {synth(seed, 100, 256), synth(seed, 101, 256), synth(seed, 102, 256)};
        vector<vector<std::pair<F, vector<F>>>> claims(3);
        vector<F> pts;
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 2; ++j) {
                vector<F> z = synth(seed, 110 + 2 * i + j, 8);
                claims[i].push_back({F{0, 0}, z});
                pts.insert(pts.end(), z.begin(), z.end());
            }
        vector<F> arr, b_aggr, oarr(256), ob2(256), obc(6), oac(3);
        F K;
        srandom(43);
        auto ch = D::aggregate_claims(arrs, arr, b_aggr, claims, K);
        vector<const or_fe *> oas = {reinterpret_cast<const or_fe *>(arrs[0].data()),
                                     reinterpret_cast<const or_fe *>(arrs[1].data()),
                                     reinterpret_cast<const or_fe *>(arrs[2].data())};
        or_fe oK;
        srandom(43);
        or_whir_aggregate_claims(3, oas.data(), 256, 6, reinterpret_cast<const or_fe *>(pts.data()),
                                 reinterpret_cast<or_fe *>(oarr.data()), reinterpret_cast<or_fe *>(ob2.data()), &oK,
                                 reinterpret_cast<or_fe *>(obc.data()), reinterpret_cast<or_fe *>(oac.data()));
        CHECK(ch.size() == 2 && same(ch[0].data(), obc.data(), 16 * 6) && same(ch[1].data(), oac.data(), 16 * 3) &&
                  same(arr.data(), oarr.data(), 16 * 256) && same(b_aggr.data(), ob2.data(), 16 * 256) &&
                  K.real == oK.real && K.img == oK.img,
              "aggregate_claims 3 arrays x 2 points");
    }
    // prepare_parameters with the reference's test rates (src/main.cpp:54)
    {
        D::rate_offset() = 1;                  // the reference's initial value (src/coPCS.cpp:22)
        D::prepare_parameters({{4}, {4}, {4, 4}, {4, 4}}, 4, 4, 1 << 20);
        const vector<int> wd = {32, 8, 64, 128}, wq = {333, 333, 433, 433}, wdel = {32768, 65536};
        CHECK(D::_degrees() == wd && D::_queries() == wq && D::rate_offset() == 64 && D::delegation_size() == wdel &&
                  D::randomness() == 96456 && D::max_delegation_size() == 65536,
              "prepare_parameters degrees, queries, rate_offset, delegation sizes");
    }
    // PCS_commit at one rank: prepare_data + Encode + distributed_MT
    {
        const uint64_t n = 1 << 14, B = n / 16;
        vector<F> shares = synth(seed, 150, n), rnd = synth(seed, 151, 9016);
        vector<F> sh0 = shares, rn0 = rnd;
        MT C;
        vector<vector<F>> cw, cd;
        D::PCS_commit(C, cw, cd, shares, rnd, {{4}, {4}, {4, 4}, {4, 4}}, 1);
        vector<F> orows(32 * B);
        const int c = or_prepare_data(reinterpret_cast<or_fe *>(sh0.data()), n, reinterpret_cast<or_fe *>(rn0.data()),
                                      rn0.size(), 333, reinterpret_cast<or_fe *>(orows.data()));
        vector<F> ocw((size_t)c * 4 * B);
        or_copcs_encode(reinterpret_cast<or_fe *>(orows.data()), c, B, 4, 6, reinterpret_cast<or_fe *>(ocw.data()));
        vector<uint8_t> base(32 * (2 * 4 * B - 1));
        or_distributed_mt(reinterpret_cast<or_fe *>(ocw.data()), c, 4 * B, base.data());
        bool ok = c == 25 && shares.empty() && rnd.empty() && cd.size() == 32 && cd[c].empty() && cw[c].empty();
        for (int i = 0; ok && i < c; ++i)
            ok = same(cd[i].data(), &orows[i * B], 16 * B) && same(cw[i].data(), &ocw[(size_t)i * 4 * B], 64 * B);
        uint64_t off = 0;
        for (size_t l = 0; ok && l < C.Base_MT.size(); ++l) {
            ok = same(C.Base_MT[l].data(), base.data() + 32 * off, 32 * C.Base_MT[l].size());
            off += C.Base_MT[l].size();
        }
        CHECK(ok && off == 2 * 4 * B - 1, "PCS_commit 2^14 shares: rows, codewords, Base_MT");
    }
    // quadratic / cubic coSumcheck over 4 ranks (threads), against the oracle
    {
        const int N = 4;
        const uint64_t M = 1 << 10;
        const std::vector<int> degrees = {16, 8, 4};
        vector<F> s1 = synth(seed, 160, M), s2 = synth(seed, 161, M), s3 = synth(seed, 162, M);
        vector<F> pads = synth(seed, 163, 3 * M);
        vector<vector<F>> p1(N, vector<F>(M)), p2(N, vector<F>(M)), p3(N, vector<F>(M));
        vector<F> sh(N * M);
        auto share = [&](const vector<F> &sec, const F *pd, vector<vector<F>> &out) {
            or_compute_secret_shares(reinterpret_cast<const or_fe *>(sec.data()), M, N, 1, 2,
                                     reinterpret_cast<const or_fe *>(pd), reinterpret_cast<or_fe *>(sh.data()));
            for (int m = 0; m < N; ++m) std::copy(sh.begin() + m * M, sh.begin() + (m + 1) * M, out[m].begin());
        };
        share(s1, pads.data(), p1);
        share(s2, pads.data() + M, p2);
        share(s3, pads.data() + 2 * M, p3);
        or_fe K{0, 0};
        for (uint64_t i = 0; i < M; ++i) K = or_add(K, or_mul(O(s1[i]), O(s2[i])));
        vector<vector<vector<F>>> R(N);
        for (int m = 0; m < N; ++m)
            for (int d : degrees) R[m].push_back(vector<F>(2 * d, F{0, 0}));
        // the oracle over all ranks
        vector<vector<F>> q1 = p1, q2 = p2;
        vector<or_fe *> a1, a2;
        vector<vector<or_fe *>> rr(N);
        vector<or_fe **> rp;
        vector<vector<vector<F>>> Rc = R;
        for (int m = 0; m < N; ++m) {
            a1.push_back(reinterpret_cast<or_fe *>(q1[m].data()));
            a2.push_back(reinterpret_cast<or_fe *>(q2[m].data()));
            for (auto &x : Rc[m]) rr[m].push_back(reinterpret_cast<or_fe *>(x.data()));
        }
        for (int m = 0; m < N; ++m) rp.push_back(rr[m].data());
        std::vector<uint64_t> rs = {32, 16, 8};
        vector<F> oclaims(5), or_(4), oip(56);
        or_fe prev{0, 0};
        int ofails = -1;
        const int orc = or_quadratic_cosumcheck(N, a1.data(), a2.data(), M, rp.data(), rs.data(), 3, K, or_fe{0, 0},
                                                degrees.data(), 3, &prev, reinterpret_cast<or_fe *>(oclaims.data()),
                                                reinterpret_cast<or_fe *>(or_.data()),
                                                reinterpret_cast<or_fe *>(oip.data()), &ofails);
        ThreadGather G(N);
        std::vector<std::pair<ThreadGather *, int>> who;
        for (int m = 0; m < N; ++m) who.push_back({&G, m});
        std::vector<std::vector<std::pair<F, vector<F>>>> got(N);
        std::vector<int> fails(N, -1);
        std::vector<double> ps(N, 0);
        std::vector<bool> ipok(N, false);
        std::vector<std::thread> th;
        for (int m = 0; m < N; ++m)
            th.emplace_back([&, m] {
                D::comm() = D::Comm{m, &ThreadGather::fn, &who[m]};
                D::transcript() = cosnark_transcript{{0, 0}};
                vector<F> x1 = p1[m], x2 = p2[m];
                got[m] = D::quadratic_coSumcheck(x1, x2, R[m], F{K.real, K.img}, F{0, 0}, degrees, N, ps[m]);
                fails[m] = D::last_check_failures();
                auto &ip = D::ip_vec<F>();
                ipok[m] = ip.size() == 3 && ip[0].size() == 32 && same(ip[0].data(), oip.data(), 16 * 32) &&
                          same(ip[2].data(), &oip[48], 16 * 8) && x1.size() == 2;
            });
        for (auto &t : th) t.join();
        bool ok = orc == 0 && ofails == 0;
        for (int m = 0; ok && m < N; ++m) {
            ok = fails[m] == 0 && ipok[m] && got[m].size() == 5 && (ps[m] > 0) == (m == 0);
            for (int i = 0; ok && i < 5; ++i)
                ok = same(&got[m][i].first, &oclaims[i], 16) && got[m][i].second.size() == 4 &&
                     same(got[m][i].second.data(), or_.data(), 16 * 4);
        }
        CHECK(ok, "quadratic_coSumcheck over 4 ranks (threads) = the oracle, checks pass");
        // cubic: b1 = eq(r1), b2 = eq(r2) (one slot at N = 4)
        vector<F> r1 = synth(seed, 170, 10), r2 = synth(seed, 171, 1);
        vector<F> b1(M);
        or_precompute_beta(reinterpret_cast<or_fe *>(r1.data()), 10, reinterpret_cast<or_fe *>(b1.data()));
        or_fe b20 = or_sub(or_fe{1, 0}, O(r2[0]));
        or_fe K3{0, 0};
        for (uint64_t i = 0; i < M; ++i)
            K3 = or_add(K3, or_mul(or_mul(or_sub(or_mul(O(s1[i]), O(s2[i])), O(s3[i])), O(b1[i])), b20));
        vector<vector<vector<F>>> R3(N);
        for (int m = 0; m < N; ++m)
            for (int d : {16, 16, 4}) R3[m].push_back(vector<F>(4 * d, F{0, 0}));
        const std::vector<int> deg3 = {16, 16, 4};
        vector<vector<F>> c1 = p1, c2 = p2, c3 = p3;
        vector<or_fe *> e1, e2, e3;
        vector<vector<or_fe *>> rr3(N);
        vector<or_fe **> rp3;
        vector<vector<vector<F>>> Rc3 = R3;
        for (int m = 0; m < N; ++m) {
            e1.push_back(reinterpret_cast<or_fe *>(c1[m].data()));
            e2.push_back(reinterpret_cast<or_fe *>(c2[m].data()));
            e3.push_back(reinterpret_cast<or_fe *>(c3[m].data()));
            for (auto &x : Rc3[m]) rr3[m].push_back(reinterpret_cast<or_fe *>(x.data()));
        }
        for (int m = 0; m < N; ++m) rp3.push_back(rr3[m].data());
        std::vector<uint64_t> rs3 = {64, 64, 16};
        vector<F> oc3(4), or3(4), oip3(144);
        or_fe prev3{0, 0};
        int of3 = -1;
        const int orc3 = or_cubic_cosumcheck(N, e1.data(), e2.data(), e3.data(), M,
                                             reinterpret_cast<or_fe *>(r1.data()), 10,
                                             reinterpret_cast<or_fe *>(r2.data()), 1, rp3.data(), rs3.data(), 3, K3,
                                             or_fe{0, 0}, deg3.data(), 3, &prev3,
                                             reinterpret_cast<or_fe *>(oc3.data()),
                                             reinterpret_cast<or_fe *>(or3.data()),
                                             reinterpret_cast<or_fe *>(oip3.data()), &of3);
        std::vector<std::vector<std::pair<F, vector<F>>>> got3(N);
        std::vector<int> f3(N, -1);
        std::vector<std::thread> th3;
        for (int m = 0; m < N; ++m)
            th3.emplace_back([&, m] {
                D::comm() = D::Comm{m, &ThreadGather::fn, &who[m]};
                D::transcript() = cosnark_transcript{{0, 0}};
                vector<F> x1 = p1[m], x2 = p2[m], x3 = p3[m], rr1 = r1, rr2 = r2;
                double pp = 0;
                got3[m] = D::cubic_coSumcheck(x1, x2, x3, rr1, rr2, R3[m], F{K3.real, K3.img}, F{0, 0}, deg3, N, pp);
                f3[m] = D::last_check_failures();
            });
        for (auto &t : th3) t.join();
        ok = orc3 == 0 && of3 == 0;
        for (int m = 0; ok && m < N; ++m) {
            ok = f3[m] == 0 && got3[m].size() == 4;
            for (int i = 0; ok && i < 4; ++i)
                ok = same(&got3[m][i].first, &oc3[i], 16) && same(got3[m][i].second.data(), or3.data(), 16 * 4);
        }
        CHECK(ok, "cubic_coSumcheck over 4 ranks (threads) = the oracle, checks pass");
    }
    // RS_multifold5
    {
        vector<F> cw = synth(seed, 40, 1 << 14), a = synth(seed, 41, 32), ref(cw.size() / 32);
        or_rs_multifold5(reinterpret_cast<or_fe *>(cw.data()), cw.size(), reinterpret_cast<or_fe *>(a.data()), 32,
                         reinterpret_cast<or_fe *>(ref.data()));
        D::RS_multifold5(cw, a, 32);
        CHECK(cw.size() == ref.size() && same(cw.data(), ref.data(), 16 * ref.size()), "RS_multifold5 fd=32");
    }
    // errors where the reference exits
    {
        bool threw = false;
        try {
            vector<F> v1 = synth(seed, 50, 64), v2 = synth(seed, 51, 64);
            D::_quadratic_sumcheck(F{1, 2}, v1, v2, 1);
        } catch (const D::error &e) {
            threw = e.code == COSNARK_E_CHECK;
        }
        CHECK(threw, "_quadratic_sumcheck with a wrong claim throws");
    }
    std::printf("%s (%d failures)\n", failures ? "FAILED" : "PASSED", failures);
    return failures ? 1 : 0;
}
