// ctx.h -- library-internal context and launch helpers.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>
#include <vector>
#include "gf.h"
#include "cosnark.h"

// Largest transform the twiddle tables cover: 2^28 points.
constexpr int NTT_LOG_MAX = 28;
constexpr int NTT_TW_BITS = 14;                 // two-level table split
constexpr int NTT_TW_SIZE = 1 << NTT_TW_BITS;
constexpr int NTT_LDS_ELEMS = 4096;             // elements per workgroup tile
constexpr int NTT_THREADS = 256;
constexpr int NTT_RADIX_LOG_MAX = 10;           // radix twiddle table: w_1024^j, j < 1024
constexpr int NTT_PASS_LOG_MAX = 8;             // radix of one LDS pass <= 256
constexpr int NTT_FAN_LOG_MAX = 3;              // first-pass fan-out over zero padding up to 8x

// Product-sumcheck device state (one per context).  Transcript value and
// claim are kept per round: tr_prev[k], tr_y[k] = state before round k.
struct alignas(16) ScState {
    gf::fe tr_prev[65];
    gf::fe tr_y[65];
    gf::fe polys[64 * 3];
    gf::fe r[64];
    gf::fe ext_r[64];     // externally drawn challenges (libc stream rounds)
    gf::fe final_v[2];
    uint8_t root[32];     // plain prover: Merkle root, gathered for one copy-out
    int32_t err;          // 0, or round+1 of the first failed H(0)+H(1)=y check
    int32_t derive_y;     // claim mode: SC_CLAIM_CHECK / _DERIVE / _UNCHECKED
};

struct cosnark_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipStream_t stream2 = nullptr;    // concurrent stage (sumcheck beside the commit)
    hipStream_t stream3 = nullptr;    // batched prover: the second lane of commits
    std::string err;
    // twiddle tables (device): omega_{2^28}^(e & mask), omega_{2^28}^((e >> 14) << 14),
    // and the same for the inverse root; radix tables omega_1024^j (j < 512).
    gf::fe *tw_lo = nullptr, *tw_hi = nullptr, *itw_lo = nullptr, *itw_hi = nullptr;
    gf::fe *rtw = nullptr, *irtw = nullptr;
    // growable scratch arena
    void *scratch = nullptr;
    size_t scratch_bytes = 0;
    void *ntt_tmp = nullptr;          // NTT passes before the natural-order one
    size_t ntt_tmp_bytes = 0;
    ScState *sc = nullptr;            // device
    ScState *sc_host = nullptr;       // pinned host mirror for copy-in / copy-out
    ScState *sc_host_dev = nullptr;   // the same memory as a device address
    gf::fe *co_host = nullptr;        // pinned staging, CO_HOST_ELEMS elements (small results, weights)
    void *stage_host = nullptr;       // pinned staging for multi-part downloads (grows on demand)
    size_t stage_bytes = 0;
    gf::fe *sc_partials = nullptr;    // device, 2 x 3 * SC_MAX_BLOCKS elements (ping-pong)
    hipEvent_t ev[8] = {};
    int ntt_tile = NTT_LDS_ELEMS;     // elements per NTT workgroup tile (power of two, <= 4096)
    int ntt_passbits = NTT_PASS_LOG_MAX;  // largest radix (bits) of one NTT pass
    int ntt_fan_min_blocks = 512;     // fan-out first pass only with at least this many workgroups
    int ntt_split = 1;                // NTT passes exchange real and imaginary words in turn where they can
    int ntt_defer = 1;                // NTT plans of 3+ passes leave inter-pass twiddles to the consuming pass
    int ntt_leaf_levels = 64;         // most tree levels a leaf-hashing last NTT pass writes (1, 3, or > 3: all)
    int batch_leaf_levels = 3;        // the same for the batched prover's encodes (0: none)
    int mk_block_log = 10;            // Merkle schedule: 2^this first-level nodes per workgroup (6..10)
    int mk_reg_log = 2;               // Merkle: levels a lane climbs in registers (0..2; 2^this nodes per lane)
    // per-size plan (0 = the defaults above); set by the autotuner
    struct { int passbits, tile; } ntt_plan[NTT_LOG_MAX + 1] = {};
    bool overlap = true;              // plain prover: sumcheck beside the commit
    void *noalloc_scratch = nullptr;  // this device's copy of the file-scope scratch array (capi.hip)
    int sc_blocks = 512;              // workgroups of a multi-workgroup sumcheck round (<= SC_MAX_BLOCKS)
    int sc_batch_blocks = 256;        // the same per proof in a batched round (many proofs share the launch)
    uint64_t sc_tail_elems = 1024;    // vectors this short finish in the one-workgroup tail (power of two, >= 4)
    int sc_wide_blocks = 256;         // workgroups of the two-round launches on long vectors (<= SC_MAX_BLOCKS / 3)
    int sc_wide_shared = 128;         // the same beside the commit (sumcheck on stream2): leaves the NTT its CUs
};

constexpr int CO_HOST_ELEMS = 8192;     // pinned staging for block-sumcheck results and weights
constexpr int SC_MAX_BLOCKS = 4096;    // partial slots per round: the most workgroups a sumcheck round may use

int set_err(cosnark_ctx *ctx, int code, const std::string &msg);
int hip_err(cosnark_ctx *ctx, hipError_t e, const char *what);
// ensure the scratch arena holds at least `bytes`; returns base pointer
int scratch_reserve(cosnark_ctx *ctx, size_t bytes, void **base);
int ntt_tmp_reserve(cosnark_ctx *ctx, size_t bytes, void **base);
// pinned host staging of at least `bytes` (kept; replaced only when larger)
int stage_reserve(cosnark_ctx *ctx, size_t bytes, void **base);

#define HIPCK(expr)                                              \
    do {                                                         \
        hipError_t e_ = (expr);                                  \
        if (e_ != hipSuccess) return hip_err(ctx, e_, #expr);    \
    } while (0)

#define LAUNCHCK()                                                         \
    do {                                                                   \
        hipError_t e_ = hipGetLastError();                                 \
        if (e_ != hipSuccess) return hip_err(ctx, e_, "kernel launch");    \
    } while (0)

// F::_random() of the reference (src/fieldElement.cpp:119-125, 363-368):
// two 20-digit decimal draws from libc random(), consumed on the calling thread
gf::fe libc_random_element();

inline bool is_pow2(uint64_t x) { return x && !(x & (x - 1)); }
inline int ilog2(uint64_t x) { int l = 0; while ((1ull << l) < x) ++l; return l; }

// ---- internal asynchronous launchers (work on the given stream) ----
// NTT of 2^logn points: out = DFT(in || 0) where in has in_len <= 2^logn
// elements (in may equal out when in_len == 2^logn).
int ntt_launch(cosnark_ctx *ctx, const gf::fe *in, uint64_t in_len, gf::fe *out, int logn,
               bool inverse, hipStream_t s);
// `batch` transforms of the same size in one launch per pass: item y reads
// in_tab[y] (device array; or in + y * in_len when in_tab is null) and
// writes out_tab[y] (or out + y * 2^logn).  Passes before the last use
// tmp (batch * 2^logn elements) or, when null, the context's buffer.
// leaves (optional): the last pass also writes item 0's Merkle leaf digests
// (BLAKE3 of each 4 consecutive outputs, 2^logn / 4 digests: level 0 of a
// node array laid out level after level) when the plan allows it;
// *levels_done receives the number of tree levels written: 0, 1, 3 or all
// logn - 1 of them (levels above the leaves only for a tree that hashes left
// children twice), at most max_levels (and ctx->ntt_leaf_levels).
// leaves_tab (a device array of batch node arrays) gives every item its
// tree in place of `leaves`.  dry: only *levels_done, nothing launched.
int ntt_launch_batch(cosnark_ctx *ctx, const gf::fe *in, const gf::fe *const *in_tab, uint64_t in_len,
                     gf::fe *out, gf::fe *const *out_tab, int batch, int logn, bool inverse, hipStream_t s,
                     gf::fe *tmp, uint8_t *leaves = nullptr, int *levels_done = nullptr, int max_levels = 64,
                     uint8_t *const *leaves_tab = nullptr, bool dry = false);
// RS encode (forward NTT of in || 0) of 2^logn points with the codeword's
// Merkle leaf digests written by the last pass where the plan allows it
// (*levels_done, as above); the caller builds the rest of the tree.
int ntt_launch_leaves(cosnark_ctx *ctx, const gf::fe *in, uint64_t in_len, gf::fe *out, int logn, uint8_t *nodes,
                      int *levels_done, hipStream_t s, gf::fe *tmp = nullptr, int max_levels = 64);
int ntt_plan_bits(const cosnark_ctx *ctx, int logn);
int ntt_plan_tile(const cosnark_ctx *ctx, int logn);
// Merkle commit of n_elems elements into nodes (all levels).
int merkle_launch(cosnark_ctx *ctx, const gf::fe *elems, uint64_t n_elems, uint8_t *nodes,
                  bool dup_left, hipStream_t s);
// levels `have`.. of a tree whose levels 0 .. have-1 (`leaves` level-0
// digests first) are already in nodes
int merkle_from_leaves_launch(cosnark_ctx *ctx, uint64_t leaves, uint8_t *nodes, bool dup_left, hipStream_t s,
                              int have = 1);
// distributed_MT tree: first level = column roots of a row-major nrows x Lc
// matrix (rows nrows..31 zero), then the dup-left tree (2 Lc - 1 digests).
int merkle_columns_launch(cosnark_ctx *ctx, const gf::fe *cw, int nrows, uint64_t Lc, uint8_t *nodes,
                          hipStream_t s);
// Launch tables of a batch, staged on the host: reserve() hands out
// 256-byte aligned offsets; the bytes are copied to one device block and
// launches address their tables as block + offset.
struct TableBuf {
    std::vector<uint8_t> host;
    size_t reserve(size_t bytes) {
        const size_t o = (host.size() + 255) & ~(size_t)255;
        host.resize(o + bytes);
        return o;
    }
    template <class T>
    T *at(size_t off) { return reinterpret_cast<T *>(host.data() + off); }
};

// Batches of trees: one launch per stage covers every tree of the batch.
struct MerkleBatchPlan {
    struct Launch {
        int mode;
        uint32_t blocks;
        size_t jobs, blk_job;     // table offsets
    };
    std::vector<Launch> launches;
};
// have (optional): per tree, the levels already in nodes (0: from the elements)
int merkle_batch_plan(cosnark_ctx *ctx, int count, const gf::fe *const *elems, const uint64_t *n_elems,
                      uint8_t *const *nodes, bool dup_left, TableBuf &tb, MerkleBatchPlan &plan,
                      const int *have = nullptr);
int merkle_batch_run(cosnark_ctx *ctx, const MerkleBatchPlan &plan, const uint8_t *tables, hipStream_t s);
int sumcheck_init(cosnark_ctx *ctx);
int copcs_init(cosnark_ctx *ctx);
int whir_init(cosnark_ctx *ctx);
// Batches of all-Fiat-Shamir product sumchecks (the plain prover's): proof
// i folds v1[i], v2[i] (n[i] elements) in buf[i] (1.5 n[i] elements) with
// partials at parts[i] (6 * SC_MAX_BLOCKS elements) and state st[i], whose
// transcript starts from prev[i] and whose claim comes from round 0.
struct ScBatchPlan {
    struct Launch {
        int kind;                 // 0: one round (fold: FOLD form); 1: rounds 0-1; 2, 3: two rounds (direct, quadratic)
        bool fold;
        uint32_t blocks;
        size_t jobs, blk_job;     // table offsets
    };
    std::vector<Launch> rounds;
    size_t tail = 0;              // table offset: one tail job per proof
    int count = 0;
    const gf::fe *prev = nullptr; // device: transcript start of each proof
};
int sumcheck_batch_plan(cosnark_ctx *ctx, int count, const gf::fe *const *v1, const gf::fe *const *v2,
                        const uint64_t *n, gf::fe *const *buf, gf::fe *const *parts, ScState *const *st,
                        const gf::fe *prev, TableBuf &tb, ScBatchPlan &plan);
int sumcheck_batch_run(cosnark_ctx *ctx, const ScBatchPlan &plan, const uint8_t *tables, hipStream_t s);
// claim handling of round 0 / every round: check against y, derive y from
// round 0, or leave unchecked (the caller verifies the claim chain)
constexpr int SC_CLAIM_CHECK = 0, SC_CLAIM_DERIVE = 1, SC_CLAIM_UNCHECKED = 2;
int sumcheck_state_init(cosnark_ctx *ctx, const gf::fe &prev, const gf::fe &y, int claim_mode, hipStream_t s);
// copy the device state (with root_src's 32 bytes as its root, if set) to
// the pinned host mirror from a kernel
int sumcheck_export(cosnark_ctx *ctx, const uint8_t *root_src, hipStream_t s);
// one multi-workgroup round (fold by the previous round's challenge when
// `fold`); *nb_out = workgroups whose partials the next launch reduces
int sumcheck_round_launch(cosnark_ctx *ctx, const gf::fe *s1, const gf::fe *s2, gf::fe *d1, gf::fe *d2,
                          uint64_t npairs, int round, bool fold, int fs_rounds, int nb_prev, hipStream_t s,
                          int *nb_out);
// finalize `round` and fold both vectors by its challenge into nout elements
// Rounds `round` and round + 1 with host-drawn challenges (ext_r), fused with
// both folds: s hold 4 nq elements, d receive nq; finalized on the device.
int sumcheck_round2_launch(cosnark_ctx *ctx, const gf::fe *s1, const gf::fe *s2, gf::fe *d1, gf::fe *d2, uint64_t nq,
                           int round, hipStream_t stream);
int sumcheck_fold_launch(cosnark_ctx *ctx, const gf::fe *s1, const gf::fe *s2, gf::fe *d1, gf::fe *d2, uint64_t nout,
                         int round, int fs_rounds, int nb_prev, hipStream_t s);
int ntt_init(cosnark_ctx *ctx);
// Product sumcheck rounds; state in ctx->sc (initialised by the caller).
// buf: n + n/2 elements (ping n/2 per vector, pong n/4 per vector).
// fuse01: two rounds per launch while the vectors are long (sc_pair01_kernel,
// then sc_fold2_kernel), as the batch plan also does
int sumcheck_launch(cosnark_ctx *ctx, const gf::fe *v1, const gf::fe *v2, uint64_t n, int fs_rounds,
                    gf::fe *buf, hipStream_t s, bool fuse01 = false);

// ---- batches of plain proofs (shared by the batch API and the server) ----
// Proof i reads f, g (n elements each) and works in `work`
// (batch_work_bytes(n, rate) bytes: codeword, tree, sumcheck buffers and
// partials, state).
struct BatchProof {
    const gf::fe *f, *g;
    uint64_t n;
    void *work;
    // where the codeword and the tree go when the caller keeps them (both
    // or neither); work then holds only the sumcheck buffers and state
    gf::fe *cw = nullptr;
    uint8_t *nodes = nullptr;
};
struct BatchCore {
    int count = 0, rate = 0;
    std::vector<uint64_t> n;
    std::vector<size_t> out_off;     // proof i's record in the output block
    std::vector<gf::fe *> cw;        // proof i's codeword (rate * n elements)
    std::vector<uint8_t *> nodes;    // proof i's tree, every level
    size_t out_bytes = 0;
    struct Group {
        int logn, B;
        uint64_t in_len;
        size_t in_tab, out_tab;      // table offsets: input and codeword pointers
        size_t lv_tab;               // table offset: node array pointers (leaf digests from the last pass)
        int levels;                  // tree levels the group's last NTT pass writes
    };
    std::vector<Group> groups;
    std::vector<int> glane;          // per group: commit lane 0 or 1
    bool two_lanes = false;          // lane 1 holds some groups (runs on its own stream)
    MerkleBatchPlan mk;              // lane 0's trees (every tree with one lane)
    MerkleBatchPlan mk1;             // lane 1's trees
    ScBatchPlan sc;
    size_t pack = 0;
    TableBuf tb;
    gf::fe *prev_dev = nullptr;      // device: count transcript starts
    uint8_t *out_dev = nullptr;      // device: out_bytes of records
};
size_t align256(size_t x);
// working memory of one proof; without the codeword and the tree when the
// caller provides them (BatchProof::cw / nodes)
size_t batch_work_bytes(uint64_t n, int rate, bool with_commit = true);
size_t batch_record_bytes(uint64_t n);
int batch_core_build(cosnark_ctx *ctx, int count, const BatchProof *pr, int rate, gf::fe *prev_dev,
                     uint8_t *out_dev, BatchCore &C, bool two_lanes = false);
// NTT temporary that lane `lane`'s groups need at once (0 for an empty lane)
uint64_t batch_tmp_elems(const BatchCore &C, int lane = 0);
// The proving pass: prev_host (pinned) -> prev_dev; sumchecks on s2; NTT
// groups (in chunks of tmp_elems) and trees on s1; join; records -> out_host
// (pinned).  tables: the device copy of C.tb.
int batch_core_enqueue(cosnark_ctx *ctx, const BatchCore &C, const uint8_t *tables, const gf::fe *prev_host,
                       uint8_t *out_host, gf::fe *tmp, uint64_t tmp_elems, hipStream_t s1, hipStream_t s2,
                       hipEvent_t fork, hipEvent_t join, uint64_t tmp_elems1 = 0, hipStream_t s3 = nullptr,
                       hipEvent_t join3 = nullptr);
// one lane's commits: its NTT groups (in chunks of tmp_elems) and its trees
int batch_enqueue_commits(cosnark_ctx *ctx, const BatchCore &C, const uint8_t *tables, gf::fe *tmp,
                          uint64_t tmp_elems, hipStream_t s, int lane);
// proof i's bytes (to proofs[i] when non-null), transcript and self-check
// flag out of the copied records; returns the first failed proof or -1
int batch_unpack(const BatchCore &C, const uint8_t *out_host, uint8_t *const *proofs, cosnark_transcript *ts,
                 int32_t *errs);
