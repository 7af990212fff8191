// capi.hip -- the extern "C" boundary of libcosnark.so (include/cosnark.h).
#include <stdlib.h>
#include <string.h>
#include <vector>
#include "ctx.h"
#include "transcript.h"
#include "blake3.h"

using gf::fe;

// ---------------------------------------------------------------------
// errors, scratch
// ---------------------------------------------------------------------

int set_err(cosnark_ctx *ctx, int code, const std::string &msg) {
    if (ctx) ctx->err = msg;
    return code;
}

int hip_err(cosnark_ctx *ctx, hipError_t e, const char *what) {
    std::string m = std::string(what) + ": " + hipGetErrorString(e);
    return set_err(ctx, e == hipErrorOutOfMemory ? COSNARK_E_NOMEM : COSNARK_E_HIP, m);
}

int scratch_reserve(cosnark_ctx *ctx, size_t bytes, void **base) {
    if (bytes > ctx->scratch_bytes) {
        HIPCK(hipDeviceSynchronize());
        if (ctx->scratch) HIPCK(hipFree(ctx->scratch));
        ctx->scratch = nullptr;
        ctx->scratch_bytes = 0;
        HIPCK(hipMalloc(&ctx->scratch, bytes));
        ctx->scratch_bytes = bytes;
    }
    *base = ctx->scratch;
    return COSNARK_OK;
}

int stage_reserve(cosnark_ctx *ctx, size_t bytes, void **base) {
    if (bytes > ctx->stage_bytes) {
        // every earlier download through it has been synchronized
        if (ctx->stage_host) HIPCK(hipHostFree(ctx->stage_host));
        ctx->stage_host = nullptr;
        ctx->stage_bytes = 0;
        const size_t want = bytes > ((size_t)1 << 20) ? bytes : ((size_t)1 << 20);
        HIPCK(hipHostMalloc(&ctx->stage_host, want, hipHostMallocDefault));
        ctx->stage_bytes = want;
    }
    *base = ctx->stage_host;
    return COSNARK_OK;
}

int ntt_tmp_reserve(cosnark_ctx *ctx, size_t bytes, void **base) {
    if (bytes > ctx->ntt_tmp_bytes) {
        // work on either stream may still use the old buffer
        HIPCK(hipDeviceSynchronize());
        if (ctx->ntt_tmp) HIPCK(hipFree(ctx->ntt_tmp));
        ctx->ntt_tmp = nullptr;
        ctx->ntt_tmp_bytes = 0;
        HIPCK(hipMalloc(&ctx->ntt_tmp, bytes));
        ctx->ntt_tmp_bytes = bytes;
    }
    *base = ctx->ntt_tmp;
    return COSNARK_OK;
}

static size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

// Scratch of cosnark_prove_plain_noalloc: a file-scope device array, one
// copy per GPU, held until the process ends.  Host code never uses the
// array's name as an address: cosnark_create looks up this device's copy
// with hipGetSymbolAddress and keeps it in the context.
constexpr size_t NOALLOC_SCRATCH_BYTES = (size_t)96 << 20;
static __device__ __align__(256) uint8_t scratch[NOALLOC_SCRATCH_BYTES];

// ---------------------------------------------------------------------
// host side of the libc challenge stream (src/fieldElement.cpp:119-125,
// 363-368): F::_random() = two 20-digit decimal numbers from random()
// ---------------------------------------------------------------------

uint64_t libc_random_number() {
    uint64_t ret = (uint64_t)(random() % 10);
    for (int i = 1; i < 20; ++i) ret = (ret * 10ull + (uint64_t)(random() % 10)) % gf::P;
    return ret;
}

fe libc_random_element() {
    fe r;
    r.re = libc_random_number() % gf::P;
    r.im = libc_random_number() % gf::P;
    return r;
}

extern "C" {

const char *cosnark_version(void) { return "cosnark-hip 0.1 gfx950"; }

void cosnark_field_mul(const uint64_t a[2], const uint64_t b[2], uint64_t out[2]) {
    const fe r = gf::mul(gf::mk(a[0], a[1]), gf::mk(b[0], b[1]));
    out[0] = r.re;
    out[1] = r.im;
}

void cosnark_blake3_64(const uint8_t in[64], uint8_t out[32]) {
    uint32_t m[16], h[8];
    memcpy(m, in, 64);                 // little-endian words, as BLAKE3 reads them
    b3::compress64(m, h);
    memcpy(out, h, 32);
}

void cosnark_field_random(uint64_t out[2]) {
    const fe r = libc_random_element();
    out[0] = r.re;
    out[1] = r.im;
}

void cosnark_field_add(const uint64_t a[2], const uint64_t b[2], uint64_t out[2]) {
    const fe r = gf::add(gf::mk(a[0], a[1]), gf::mk(b[0], b[1]));
    out[0] = r.re;
    out[1] = r.im;
}

int cosnark_create(int device, cosnark_ctx **out) {
    if (!out) return COSNARK_E_ARG;
    *out = nullptr;
    cosnark_ctx *ctx = new cosnark_ctx();
    ctx->device = device;
    auto fail = [&](int rc) { cosnark_destroy(ctx); return rc; };
    hipError_t e = hipSetDevice(device);
    if (e != hipSuccess) { delete ctx; return COSNARK_E_HIP; }
    if (const char *t = getenv("COSNARK_NTT_PASSBITS")) {
        const int v = atoi(t);
        if (v >= 2 && v <= NTT_PASS_LOG_MAX) ctx->ntt_passbits = v;
    }
    if (const char *t = getenv("COSNARK_NTT_TILE")) {
        const int v = atoi(t);
        if (v >= 256 && v <= NTT_LDS_ELEMS && is_pow2((uint64_t)v)) ctx->ntt_tile = v;
    }
    if (const char *t = getenv("COSNARK_NTT_SPLIT")) ctx->ntt_split = atoi(t) != 0;
    if (const char *t = getenv("COSNARK_NTT_DEFER")) ctx->ntt_defer = atoi(t) != 0;
    if (const char *t = getenv("COSNARK_MK_BLOCK_LOG")) {
        const int v = atoi(t);
        if (v >= 6 && v <= 10) ctx->mk_block_log = v;
    }
    if (const char *t = getenv("COSNARK_MK_REG_LOG")) {
        const int v = atoi(t);
        if (v >= 0 && v <= 2) ctx->mk_reg_log = v;
    }
    if (const char *t = getenv("COSNARK_BATCH_LEAF_LEVELS")) {
        const int v = atoi(t);
        ctx->batch_leaf_levels = v > 3 ? 64 : v == 3 ? 3 : v >= 1 ? 1 : 0;
    }
    if (const char *t = getenv("COSNARK_NTT_LEAF_LEVELS")) {
        const int v = atoi(t);
        ctx->ntt_leaf_levels = v > 3 ? 64 : v == 3 ? 3 : 1;
    }
    if (const char *t = getenv("COSNARK_NTT_FAN_MIN")) {
        const int v = atoi(t);
        if (v >= 1) ctx->ntt_fan_min_blocks = v;
    }
    if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) return fail(COSNARK_E_HIP);
    if (hipStreamCreateWithFlags(&ctx->stream2, hipStreamNonBlocking) != hipSuccess) return fail(COSNARK_E_HIP);
    if (hipStreamCreateWithFlags(&ctx->stream3, hipStreamNonBlocking) != hipSuccess) return fail(COSNARK_E_HIP);
    if (const char *e = getenv("COSNARK_SC_BLOCKS")) {
        const int b = atoi(e);
        if (b >= 1 && b <= SC_MAX_BLOCKS) ctx->sc_blocks = b;
    }
    if (const char *e = getenv("COSNARK_SC_BATCH_BLOCKS")) {
        const int b = atoi(e);
        if (b >= 1 && b <= SC_MAX_BLOCKS) ctx->sc_batch_blocks = b;
    }
    if (const char *e = getenv("COSNARK_SC_WIDE")) {
        const int b = atoi(e);
        if (b >= 1 && b <= SC_MAX_BLOCKS / 3) ctx->sc_wide_blocks = b;
    }
    if (const char *e = getenv("COSNARK_SC_WIDE_SHARED")) {
        const int b = atoi(e);
        if (b >= 1 && b <= SC_MAX_BLOCKS / 3) ctx->sc_wide_shared = b;
    }
    if (const char *e = getenv("COSNARK_SC_TAIL")) {
        const long long v = atoll(e);
        if (v >= 4 && v <= (1 << 20) && is_pow2((uint64_t)v)) ctx->sc_tail_elems = (uint64_t)v;
    }
    for (auto &ev : ctx->ev)
        if (hipEventCreate(&ev) != hipSuccess) return fail(COSNARK_E_HIP);

    // twiddle tables: omega = root of order 2^28 (and its inverse)
    std::vector<fe> lo(NTT_TW_SIZE), hi(NTT_TW_SIZE), ilo(NTT_TW_SIZE), ihi(NTT_TW_SIZE);
    std::vector<fe> rt(1024), irt(1024);
    {
        const fe w = gf::root_of_unity(NTT_LOG_MAX), wi = gf::inv(w);
        const fe W = gf::pow(w, NTT_TW_SIZE), Wi = gf::pow(wi, NTT_TW_SIZE);
        lo[0] = hi[0] = ilo[0] = ihi[0] = gf::one();
        for (int i = 1; i < NTT_TW_SIZE; ++i) {
            lo[i] = gf::mul(lo[i - 1], w);
            hi[i] = gf::mul(hi[i - 1], W);
            ilo[i] = gf::mul(ilo[i - 1], wi);
            ihi[i] = gf::mul(ihi[i - 1], Wi);
        }
        const fe r = gf::root_of_unity(NTT_RADIX_LOG_MAX), ri = gf::inv(r);
        rt[0] = irt[0] = gf::one();
        for (int i = 1; i < 1024; ++i) {
            rt[i] = gf::mul(rt[i - 1], r);
            irt[i] = gf::mul(irt[i - 1], ri);
        }
    }
    const size_t tb = NTT_TW_SIZE * sizeof(fe);
    if (hipMalloc(&ctx->tw_lo, tb) != hipSuccess || hipMalloc(&ctx->tw_hi, tb) != hipSuccess ||
        hipMalloc(&ctx->itw_lo, tb) != hipSuccess || hipMalloc(&ctx->itw_hi, tb) != hipSuccess ||
        hipMalloc(&ctx->rtw, 1024 * sizeof(fe)) != hipSuccess ||
        hipMalloc(&ctx->irtw, 1024 * sizeof(fe)) != hipSuccess)
        return fail(COSNARK_E_NOMEM);
    if (hipMemcpy(ctx->tw_lo, lo.data(), tb, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(ctx->tw_hi, hi.data(), tb, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(ctx->itw_lo, ilo.data(), tb, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(ctx->itw_hi, ihi.data(), tb, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(ctx->rtw, rt.data(), 1024 * sizeof(fe), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(ctx->irtw, irt.data(), 1024 * sizeof(fe), hipMemcpyHostToDevice) != hipSuccess)
        return fail(COSNARK_E_HIP);
    if (hipMalloc(&ctx->sc, sizeof(ScState)) != hipSuccess ||
        hipMalloc(&ctx->sc_partials, 2 * 3 * SC_MAX_BLOCKS * sizeof(fe)) != hipSuccess ||
        hipHostMalloc(&ctx->sc_host, sizeof(ScState), hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc(&ctx->co_host, CO_HOST_ELEMS * sizeof(fe), hipHostMallocDefault) != hipSuccess)
        return fail(COSNARK_E_NOMEM);
    if (hipHostGetDevicePointer((void **)&ctx->sc_host_dev, ctx->sc_host, 0) != hipSuccess)
        return fail(COSNARK_E_HIP);
    if (hipMemset(ctx->sc, 0, sizeof(ScState)) != hipSuccess) return fail(COSNARK_E_HIP);
    // this device's copy of the no-allocation entry point's scratch
    if (hipGetSymbolAddress(&ctx->noalloc_scratch, HIP_SYMBOL(scratch)) != hipSuccess || !ctx->noalloc_scratch)
        return fail(COSNARK_E_HIP);
    if (sumcheck_init(ctx) != COSNARK_OK) return fail(COSNARK_E_HIP);
    if (ntt_init(ctx) != COSNARK_OK) return fail(COSNARK_E_HIP);
    if (copcs_init(ctx) != COSNARK_OK) return fail(COSNARK_E_HIP);
    if (whir_init(ctx) != COSNARK_OK) return fail(COSNARK_E_HIP);
    if (hipDeviceSynchronize() != hipSuccess) return fail(COSNARK_E_HIP);
    *out = ctx;
    return COSNARK_OK;
}

void cosnark_destroy(cosnark_ctx *ctx) {
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    for (auto &ev : ctx->ev) if (ev) (void)hipEventDestroy(ev);
    (void)hipFree(ctx->tw_lo); (void)hipFree(ctx->tw_hi); (void)hipFree(ctx->itw_lo); (void)hipFree(ctx->itw_hi);
    (void)hipFree(ctx->rtw); (void)hipFree(ctx->irtw);
    (void)hipFree(ctx->sc); (void)hipFree(ctx->sc_partials);
    if (ctx->sc_host) (void)hipHostFree(ctx->sc_host);
    if (ctx->co_host) (void)hipHostFree(ctx->co_host);
    if (ctx->stage_host) (void)hipHostFree(ctx->stage_host);
    if (ctx->scratch) (void)hipFree(ctx->scratch);
    if (ctx->ntt_tmp) (void)hipFree(ctx->ntt_tmp);
    if (ctx->stream2) (void)hipStreamDestroy(ctx->stream2);
    if (ctx->stream3) (void)hipStreamDestroy(ctx->stream3);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

const char *cosnark_last_error(const cosnark_ctx *ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int cosnark_set_overlap(cosnark_ctx *ctx, int on) {
    if (!ctx) return COSNARK_E_ARG;
    ctx->overlap = on != 0;
    return COSNARK_OK;
}

int cosnark_sync(cosnark_ctx *ctx) {
    HIPCK(hipStreamSynchronize(ctx->stream));
    return COSNARK_OK;
}

void *cosnark_stream(cosnark_ctx *ctx) { return ctx ? (void *)ctx->stream : nullptr; }

int cosnark_dev_alloc(cosnark_ctx *ctx, uint64_t bytes, void **out) {
    HIPCK(hipMalloc(out, bytes ? bytes : 16));
    return COSNARK_OK;
}

int cosnark_dev_free(cosnark_ctx *ctx, void *p) {
    HIPCK(hipStreamSynchronize(ctx->stream));
    HIPCK(hipFree(p));
    return COSNARK_OK;
}

int cosnark_h2d(cosnark_ctx *ctx, void *d, const void *h, uint64_t bytes) {
    HIPCK(hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, ctx->stream));
    HIPCK(hipStreamSynchronize(ctx->stream));
    return COSNARK_OK;
}

int cosnark_d2h(cosnark_ctx *ctx, void *h, const void *d, uint64_t bytes) {
    HIPCK(hipMemcpyAsync(h, d, bytes, hipMemcpyDeviceToHost, ctx->stream));
    HIPCK(hipStreamSynchronize(ctx->stream));
    return COSNARK_OK;
}

int cosnark_d2d(cosnark_ctx *ctx, void *d, const void *s, uint64_t bytes) {
    HIPCK(hipMemcpyAsync(d, s, bytes, hipMemcpyDeviceToDevice, ctx->stream));
    HIPCK(hipStreamSynchronize(ctx->stream));
    return COSNARK_OK;
}

// ---------------------------------------------------------------------
// synthetic inputs
// ---------------------------------------------------------------------

static uint64_t splitmix64(uint64_t *s) {
    uint64_t z = (*s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

static inline uint64_t rotl64(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

void cosnark_synth_fill(uint64_t seed, uint64_t stream, uint64_t *out, uint64_t n) {
    uint64_t sm = seed ^ (stream * 0xD1B54A32D192ED03ull), s[4];
    for (int i = 0; i < 4; ++i) s[i] = splitmix64(&sm);
    auto next = [&]() {
        const uint64_t result = rotl64(s[1] * 5, 7) * 9;
        const uint64_t t = s[1] << 17;
        s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3];
        s[2] ^= t;
        s[3] = rotl64(s[3], 45);
        return result;
    };
    for (uint64_t i = 0; i < n; ++i) {
        uint64_t x;
        do { x = next() & gf::P; } while (x == gf::P);
        out[2 * i] = x;
        do { x = next() & gf::P; } while (x == gf::P);
        out[2 * i + 1] = x;
    }
}

// ---------------------------------------------------------------------
// NTT / RS
// ---------------------------------------------------------------------

int cosnark_ntt(cosnark_ctx *ctx, uint64_t *d_arr, int logn, int inverse) {
    if (!ctx || !d_arr) return set_err(ctx, COSNARK_E_ARG, "ntt: null argument");
    if (logn == 0) return COSNARK_OK;
    int rc = ntt_launch(ctx, (const fe *)d_arr, 1ull << logn, (fe *)d_arr, logn, inverse != 0, ctx->stream);
    if (rc) return rc;
    return cosnark_sync(ctx);
}

int cosnark_rs_encode(cosnark_ctx *ctx, const uint64_t *d_coeffs, uint64_t n, int rate, uint64_t *d_cw) {
    if (!ctx || !d_coeffs || !d_cw) return set_err(ctx, COSNARK_E_ARG, "rs_encode: null argument");
    if (!is_pow2(n) || rate < 1 || !is_pow2((uint64_t)rate))
        return set_err(ctx, COSNARK_E_ARG, "rs_encode: n and rate must be powers of two");
    const int logm = ilog2(n * (uint64_t)rate);
    if (logm == 0) return cosnark_d2d(ctx, d_cw, d_coeffs, 16);
    int rc = ntt_launch(ctx, (const fe *)d_coeffs, n, (fe *)d_cw, logm, false, ctx->stream);
    if (rc) return rc;
    return cosnark_sync(ctx);
}

// ---------------------------------------------------------------------
// Merkle
// ---------------------------------------------------------------------

uint64_t cosnark_merkle_node_count(uint64_t n_elems) {
    uint64_t total = 0;
    for (uint64_t s = n_elems / 4; s >= 1; s /= 2) total += s;
    return total;
}

int cosnark_merkle_commit(cosnark_ctx *ctx, const uint64_t *d_elems, uint64_t n_elems, uint8_t *d_nodes,
                          int dup_left) {
    if (!ctx || !d_elems || !d_nodes) return set_err(ctx, COSNARK_E_ARG, "merkle_commit: null argument");
    int rc = merkle_launch(ctx, (const fe *)d_elems, n_elems, d_nodes, dup_left != 0, ctx->stream);
    if (rc) return rc;
    return cosnark_sync(ctx);
}

int cosnark_merkle_tree_from_leaves(cosnark_ctx *ctx, uint64_t n_leaves, uint8_t *d_nodes, int dup_left) {
    if (!ctx || !d_nodes) return set_err(ctx, COSNARK_E_ARG, "merkle_tree_from_leaves: null argument");
    int rc = merkle_from_leaves_launch(ctx, n_leaves, d_nodes, dup_left != 0, ctx->stream);
    if (rc) return rc;
    return cosnark_sync(ctx);
}

int cosnark_merkle_open(cosnark_ctx *ctx, const uint8_t *d_nodes, uint64_t n_elems, uint64_t pos,
                        uint8_t *h_path, int *depth) {
    if (!ctx || !d_nodes || !h_path) return set_err(ctx, COSNARK_E_ARG, "merkle_open: null argument");
    if (!is_pow2(n_elems) || n_elems < 8) return set_err(ctx, COSNARK_E_ARG, "merkle_open: bad size");
    uint64_t s = n_elems / 4, off = 0;
    if (pos >= s) return set_err(ctx, COSNARK_E_ARG, "merkle_open: position out of range");
    int lvl = 0;
    while (s > 1) {
        const uint64_t sib = 2 * (pos / 2) + (1 - (pos % 2));
        HIPCK(hipMemcpyAsync(h_path + 32 * lvl, d_nodes + 32 * (off + sib), 32, hipMemcpyDeviceToHost,
                             ctx->stream));
        off += s;
        s /= 2;
        pos /= 2;
        ++lvl;
    }
    HIPCK(hipStreamSynchronize(ctx->stream));
    if (depth) *depth = lvl;
    return COSNARK_OK;
}

// ---------------------------------------------------------------------
// transcript (host)
// ---------------------------------------------------------------------

int cosnark_hash_to_field(cosnark_transcript *t, const uint64_t *h_elems, uint64_t n, uint64_t out[2]) {
    if (!t || (n && !h_elems)) return COSNARK_E_ARG;
    fe prev = gf::mk(t->prev[0], t->prev[1]);
    for (uint64_t i = 0; i < n; ++i) prev = ts::hash_F(prev, gf::mk(h_elems[2 * i], h_elems[2 * i + 1]));
    t->prev[0] = prev.re;
    t->prev[1] = prev.im;
    if (out) { out[0] = prev.re; out[1] = prev.im; }
    return COSNARK_OK;
}

int cosnark_get_indexes(cosnark_transcript *t, int l, uint32_t N, uint32_t M, uint32_t *idx) {
    if (!t || !idx || l < 0 || N == 0 || M == 0) return COSNARK_E_ARG;
    fe prev = gf::mk(t->prev[0], t->prev[1]);
    for (int i = 0; i < l; ++i) {
        prev = ts::hash_F(prev, prev);
        idx[2 * i] = (uint32_t)(prev.re % N);
        idx[2 * i + 1] = (uint32_t)(prev.im % M);
    }
    t->prev[0] = prev.re;
    t->prev[1] = prev.im;
    return COSNARK_OK;
}

// ---------------------------------------------------------------------
// product sumcheck
// ---------------------------------------------------------------------

// stage the sumcheck state: transcript, claim and libc-drawn tail challenges
static int sc_stage_state(cosnark_ctx *ctx, const fe &prev, const fe &y, int derive_y, int m, int fs_rounds,
                          hipStream_t s) {
    if (fs_rounds < m) {
        // four draws per round, the last is kept (src/sparse_eval.cpp:164-168)
        for (int i = fs_rounds; i < m; ++i) {
            fe r = libc_random_element();
            r = libc_random_element();
            r = libc_random_element();
            r = libc_random_element();
            ctx->sc_host->ext_r[i] = r;
        }
        // the pinned mirror is not touched again before this call's final sync
        HIPCK(hipMemcpyAsync(ctx->sc->ext_r + fs_rounds, ctx->sc_host->ext_r + fs_rounds,
                             sizeof(fe) * (m - fs_rounds), hipMemcpyHostToDevice, s));
    }
    return sumcheck_state_init(ctx, prev, y, derive_y, s);
}

// the device state written into the pinned mirror by one kernel (with a
// root, if given), then results out
static int sc_collect(cosnark_ctx *ctx, int m, cosnark_transcript *t, uint64_t *h_polys, uint64_t *h_r,
                      uint64_t *h_final, hipStream_t s, const uint8_t *root_src = nullptr) {
    int rc = sumcheck_export(ctx, root_src, s);
    if (rc) return rc;
    HIPCK(hipStreamSynchronize(s));
    const ScState &h = *ctx->sc_host;
    if (t) { t->prev[0] = h.tr_prev[m].re; t->prev[1] = h.tr_prev[m].im; }
    if (h_polys) memcpy(h_polys, h.polys, sizeof(fe) * 3 * m);
    if (h_r) memcpy(h_r, h.r, sizeof(fe) * m);
    if (h_final) memcpy(h_final, h.final_v, sizeof(fe) * 2);
    if (h.err) return set_err(ctx, COSNARK_E_CHECK, "sumcheck: H(0)+H(1) != y in round " + std::to_string(h.err - 1));
    return COSNARK_OK;
}

int cosnark_product_sumcheck(cosnark_ctx *ctx, const uint64_t *d_v1, const uint64_t *d_v2, uint64_t n,
                             const uint64_t y[2], cosnark_transcript *t, int offset, uint64_t *h_polys,
                             uint64_t *h_r, uint64_t h_final[4]) {
    if (!ctx || !d_v1 || !d_v2 || !y || !t) return set_err(ctx, COSNARK_E_ARG, "sumcheck: null argument");
    if (!is_pow2(n) || n < 2 || n > (1ull << 30)) return set_err(ctx, COSNARK_E_ARG, "sumcheck: n must be a power of two in [2, 2^30]");
    const int m = ilog2(n);
    if (offset < 0) offset = 0;
    const int fs_rounds = offset > m ? 0 : m - offset;
    void *base;
    int rc = scratch_reserve(ctx, sizeof(fe) * (n + n / 2), &base);
    if (rc) return rc;
    rc = sc_stage_state(ctx, gf::mk(t->prev[0], t->prev[1]), gf::mk(y[0], y[1]), 0, m, fs_rounds, ctx->stream);
    if (rc) return rc;
    rc = sumcheck_launch(ctx, (const fe *)d_v1, (const fe *)d_v2, n, fs_rounds, (fe *)base, ctx->stream, true);
    if (rc) return rc;
    return sc_collect(ctx, m, t, h_polys, h_r, h_final, ctx->stream);
}

// ---------------------------------------------------------------------
// plain prover
// ---------------------------------------------------------------------

uint64_t cosnark_plain_proof_bytes(uint64_t n) { return 32 + 48ull * (uint64_t)ilog2(n) + 32; }

int cosnark_prove_plain(cosnark_ctx *ctx, const uint64_t *d_f, const uint64_t *d_g, uint64_t n, int rate,
                        cosnark_transcript *t, uint8_t *h_proof, double *times) {
    return cosnark_prove_plain_ex(ctx, d_f, d_g, n, rate, t, h_proof, nullptr, nullptr, times);
}

static int prove_plain_args(cosnark_ctx *ctx, const uint64_t *d_f, const uint64_t *d_g, uint64_t n, int rate,
                            cosnark_transcript *t, uint8_t *h_proof) {
    if (!ctx || !d_f || !d_g || !t || !h_proof) return set_err(ctx, COSNARK_E_ARG, "prove_plain: null argument");
    if (!is_pow2(n) || n < 4 || rate < 1 || !is_pow2((uint64_t)rate) || n * (uint64_t)rate > (1ull << 28))
        return set_err(ctx, COSNARK_E_ARG, "prove_plain: n, rate must be powers of two, 4 <= n, n*rate <= 2^28");
    return COSNARK_OK;
}

// The plain prover on given buffers: codeword cw (rate n elements), tree nd,
// sumcheck buffers scb (n + n/2 elements), NTT temporary tmp (rate n
// elements, or null: the context's own).
static int prove_plain_core(cosnark_ctx *ctx, const uint64_t *d_f, const uint64_t *d_g, uint64_t n, int rate,
                            cosnark_transcript *t, uint8_t *h_proof, fe *cw, uint8_t *nd, fe *scb, fe *tmp,
                            double *times);

int cosnark_prove_plain_ex(cosnark_ctx *ctx, const uint64_t *d_f, const uint64_t *d_g, uint64_t n, int rate,
                           cosnark_transcript *t, uint8_t *h_proof, uint64_t *d_codeword, uint8_t *d_nodes,
                           double *times) {
    int rc = prove_plain_args(ctx, d_f, d_g, n, rate, t, h_proof);
    if (rc) return rc;
    const uint64_t cw_n = n * (uint64_t)rate;
    const uint64_t nodes = cosnark_merkle_node_count(cw_n);
    // the codeword and the tree go to the caller's buffers when given (the
    // commitment then stays openable), else to the context's scratch
    const size_t off_cw = 0;
    const size_t off_nodes = align_up(off_cw + (d_codeword ? 0 : cw_n * sizeof(fe)), 256);
    const size_t off_sc = align_up(off_nodes + (d_nodes ? 0 : nodes * 32), 256);
    const size_t total = off_sc + sizeof(fe) * (n + n / 2);
    void *base;
    rc = scratch_reserve(ctx, total, &base);
    if (rc) return rc;
    fe *cw = d_codeword ? (fe *)d_codeword : (fe *)((char *)base + off_cw);
    uint8_t *nd = d_nodes ? d_nodes : (uint8_t *)base + off_nodes;
    fe *scb = (fe *)((char *)base + off_sc);
    return prove_plain_core(ctx, d_f, d_g, n, rate, t, h_proof, cw, nd, scb, nullptr, times);
}

uint64_t cosnark_prove_plain_noalloc_max_bytes(void) { return NOALLOC_SCRATCH_BYTES; }

int cosnark_prove_plain_noalloc(cosnark_ctx *ctx, const uint64_t *d_f, const uint64_t *d_g, uint64_t n, int rate,
                                cosnark_transcript *t, uint8_t *h_proof, uint64_t *d_codeword, uint8_t *d_nodes) {
    int rc = prove_plain_args(ctx, d_f, d_g, n, rate, t, h_proof);
    if (rc) return rc;
    if (!d_codeword || !d_nodes) return set_err(ctx, COSNARK_E_ARG, "prove_plain_noalloc: null output buffer");
    if (!ctx->noalloc_scratch) return set_err(ctx, COSNARK_E_HIP, "prove_plain_noalloc: no scratch on this device");
    // scratch: the NTT temporary (rate n elements), then the sumcheck's
    // ping-pong buffers (n + n/2 elements); every byte is written before it
    // is read, so nothing is cleared
    const uint64_t cw_n = n * (uint64_t)rate;
    const size_t off_sc = align_up(cw_n * sizeof(fe), 256);
    const size_t need = off_sc + sizeof(fe) * (n + n / 2);
    if (need > NOALLOC_SCRATCH_BYTES)
        return set_err(ctx, COSNARK_E_RANGE, "prove_plain_noalloc: rate*n + 1.5 n elements exceed the " +
                                                 std::to_string(NOALLOC_SCRATCH_BYTES >> 20) + " MiB scratch");
    uint8_t *base = (uint8_t *)ctx->noalloc_scratch;
    return prove_plain_core(ctx, d_f, d_g, n, rate, t, h_proof, (fe *)d_codeword, d_nodes, (fe *)(base + off_sc),
                            (fe *)base, nullptr);
}

static int prove_plain_core(cosnark_ctx *ctx, const uint64_t *d_f, const uint64_t *d_g, uint64_t n, int rate,
                            cosnark_transcript *t, uint8_t *h_proof, fe *cw, uint8_t *nd, fe *scb, fe *tmp,
                            double *times) {
    const int m = ilog2(n);
    const uint64_t cw_n = n * (uint64_t)rate;
    const uint64_t nodes = cosnark_merkle_node_count(cw_n);
    int rc = sc_stage_state(ctx, gf::mk(t->prev[0], t->prev[1]), gf::zero(), 1, m, m, ctx->stream);
    if (rc) return rc;
    // commit (encode + Merkle) on s1, the sumcheck beside it on s2: the
    // transcript never absorbs the root, so the two are independent
    // stage events only when the caller asks for stage times
    hipStream_t s1 = ctx->stream, s2 = ctx->overlap ? ctx->stream2 : ctx->stream;
    HIPCK(hipEventRecord(ctx->ev[7], ctx->stream));
    if (s1 != ctx->stream) HIPCK(hipStreamWaitEvent(s1, ctx->ev[7], 0));
    if (times) HIPCK(hipEventRecord(ctx->ev[0], s1));
    int levels_done = 0;
    // levels 1 and 2 at most from the encode: the whole tree in its last
    // pass lengthens it more than it saves beside the sumcheck
    if ((rc = ntt_launch_leaves(ctx, (const fe *)d_f, n, cw, ilog2(cw_n), nd, &levels_done, s1, tmp, 3))) return rc;
    if (times) HIPCK(hipEventRecord(ctx->ev[1], s1));
    if (s2 != ctx->stream) HIPCK(hipStreamWaitEvent(s2, ctx->ev[7], 0));
    if (times) HIPCK(hipEventRecord(ctx->ev[4], s2));
    if ((rc = sumcheck_launch(ctx, (const fe *)d_f, (const fe *)d_g, n, m, scb, s2, true))) return rc;
    HIPCK(hipEventRecord(ctx->ev[3], s2));
    if (times) HIPCK(hipEventRecord(ctx->ev[6], s1));
    // the lowest levels came from the encode's last pass where the plan allowed it
    if ((rc = levels_done ? merkle_from_leaves_launch(ctx, cw_n / 4, nd, true, s1, levels_done)
                          : merkle_launch(ctx, cw, cw_n, nd, true, s1)))
        return rc;
    HIPCK(hipEventRecord(ctx->ev[2], s1));
    if (s1 != ctx->stream) HIPCK(hipStreamWaitEvent(ctx->stream, ctx->ev[2], 0));
    if (s2 != ctx->stream) HIPCK(hipStreamWaitEvent(ctx->stream, ctx->ev[3], 0));
    if (times) HIPCK(hipEventRecord(ctx->ev[5], ctx->stream));
    uint64_t polys[6 * 64], fin[4];
    rc = sc_collect(ctx, m, t, polys, nullptr, fin, ctx->stream, nd + 32 * (nodes - 1));
    memcpy(h_proof, ctx->sc_host->root, 32);
    memcpy(h_proof + 32, polys, 48ull * m);
    memcpy(h_proof + 32 + 48ull * m, fin, 32);
    if (times) {
        float a = 0, b = 0, c = 0, tot = 0;
        (void)hipEventElapsedTime(&a, ctx->ev[0], ctx->ev[1]);
        (void)hipEventElapsedTime(&b, ctx->ev[6], ctx->ev[2]);
        (void)hipEventElapsedTime(&c, ctx->ev[4], ctx->ev[3]);
        (void)hipEventElapsedTime(&tot, ctx->ev[0], ctx->ev[5]);
        times[0] = a; times[1] = b; times[2] = c; times[3] = tot;
    }
    return rc;
}

}  // extern "C"
