// ctx.h -- library-internal context and launch helpers.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>
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

// Product-sumcheck device state (one per context).  Transcript value and
// claim are kept per round: tr_prev[k], tr_y[k] = state before round k.
struct ScState {
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
    gf::fe *co_host = nullptr;        // pinned staging, CO_HOST_ELEMS elements (small results, weights)
    gf::fe *sc_partials = nullptr;    // device, 2 x 3 * SC_MAX_BLOCKS elements (ping-pong)
    hipEvent_t ev[8] = {};
    int ntt_tile = NTT_LDS_ELEMS;     // elements per NTT workgroup tile (power of two, <= 4096)
    int ntt_passbits = NTT_PASS_LOG_MAX;  // largest radix (bits) of one NTT pass
    bool overlap = true;              // plain prover: sumcheck beside the commit
};

constexpr int CO_HOST_ELEMS = 8192;     // pinned staging for block-sumcheck results and weights
constexpr int SC_MAX_BLOCKS = 256;     // workgroups per multi-workgroup sumcheck round

int set_err(cosnark_ctx *ctx, int code, const std::string &msg);
int hip_err(cosnark_ctx *ctx, hipError_t e, const char *what);
// ensure the scratch arena holds at least `bytes`; returns base pointer
int scratch_reserve(cosnark_ctx *ctx, size_t bytes, void **base);
int ntt_tmp_reserve(cosnark_ctx *ctx, size_t bytes, void **base);

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
// Merkle commit of n_elems elements into nodes (all levels).
int merkle_launch(cosnark_ctx *ctx, const gf::fe *elems, uint64_t n_elems, uint8_t *nodes,
                  bool dup_left, hipStream_t s);
// distributed_MT tree: first level = column roots of a row-major nrows x Lc
// matrix (rows nrows..31 zero), then the dup-left tree (2 Lc - 1 digests).
int merkle_columns_launch(cosnark_ctx *ctx, const gf::fe *cw, int nrows, uint64_t Lc, uint8_t *nodes,
                          hipStream_t s);
int sumcheck_init(cosnark_ctx *ctx);
int copcs_init(cosnark_ctx *ctx);
// claim handling of round 0 / every round: check against y, derive y from
// round 0, or leave unchecked (the caller verifies the claim chain)
constexpr int SC_CLAIM_CHECK = 0, SC_CLAIM_DERIVE = 1, SC_CLAIM_UNCHECKED = 2;
int sumcheck_state_init(cosnark_ctx *ctx, const gf::fe &prev, const gf::fe &y, int claim_mode, hipStream_t s);
// one multi-workgroup round (fold by the previous round's challenge when
// `fold`); *nb_out = workgroups whose partials the next launch reduces
int sumcheck_round_launch(cosnark_ctx *ctx, const gf::fe *s1, const gf::fe *s2, gf::fe *d1, gf::fe *d2,
                          uint64_t npairs, int round, bool fold, int fs_rounds, int nb_prev, hipStream_t s,
                          int *nb_out);
// finalize `round` and fold both vectors by its challenge into nout elements
int sumcheck_fold_launch(cosnark_ctx *ctx, const gf::fe *s1, const gf::fe *s2, gf::fe *d1, gf::fe *d2, uint64_t nout,
                         int round, int fs_rounds, int nb_prev, hipStream_t s);
int ntt_init(cosnark_ctx *ctx);
// Product sumcheck rounds; state in ctx->sc (initialised by the caller).
// scratch buffers: a1,a2,b1,b2 each of n/2 elements.
int sumcheck_launch(cosnark_ctx *ctx, const gf::fe *v1, const gf::fe *v2, uint64_t n, int fs_rounds,
                    gf::fe *buf, hipStream_t s);
