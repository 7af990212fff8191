// tune.hip -- the NTT radix plan per transform size and the Merkle schedule,
// picked by timing.
//
// A 2^L-point transform runs as ceil(L / b) LDS passes of radix at most 2^b,
// each workgroup staging a tile of T elements.  Every (b, T) computes the
// same transform (the result is exact field arithmetic), so the choice only
// moves time: fewer passes mean fewer HBM sweeps but larger in-register
// DFTs and fewer columns per tile.  The tuner times each candidate on the
// device, with `batch` transforms per launch as the batched prover issues
// them, and keeps the fastest per size in the context.
#include <vector>
#include "blake3.h"
#include "ctx.h"

using gf::fe;

namespace {

constexpr int FR_CHAINS = 8;

// Eight independent chains x <- x * b per lane: a dependency-free stream of
// GF(p^2) multiplies, the ceiling the transform kernels are measured against.
__global__ __launch_bounds__(256) void fmul_rate_kernel(fe b, int iters, fe *out) {
    const uint64_t gid = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    fe a[FR_CHAINS];
#pragma unroll
    for (int k = 0; k < FR_CHAINS; ++k) a[k] = gf::mk(gid + 1 + k, (gid * 7 + k) & 0xFFFFFFFFFFFull);
    for (int i = 0; i < iters; ++i) {
#pragma unroll
        for (int k = 0; k < FR_CHAINS; ++k) a[k] = gf::mul(a[k], b);
    }
    fe s = a[0];
#pragma unroll
    for (int k = 1; k < FR_CHAINS; ++k) s = gf::add(s, a[k]);
    out[gid] = s;
}

// Independent chains of single-block BLAKE3 compressions per lane, each
// digest fed back as the next message's first half: the hashing ceiling the
// Merkle kernels are measured against.
constexpr int HR_CHAINS = 2;
__global__ __launch_bounds__(256) void blake3_rate_kernel(int iters, uint32_t *out) {
    const uint32_t gid = blockIdx.x * 256 + threadIdx.x;
    uint32_t m[HR_CHAINS][16];
#pragma unroll
    for (int k = 0; k < HR_CHAINS; ++k)
#pragma unroll
        for (int i = 0; i < 16; ++i) m[k][i] = gid * 16 + i + 977u * k;
    for (int it = 0; it < iters; ++it) {
#pragma unroll
        for (int k = 0; k < HR_CHAINS; ++k) {
            uint32_t h[8];
            b3::compress64(m[k], h);
#pragma unroll
            for (int i = 0; i < 8; ++i) m[k][i] = h[i];
        }
    }
    uint32_t s = 0;
#pragma unroll
    for (int k = 0; k < HR_CHAINS; ++k) s ^= m[k][0];
    out[gid] = s;
}

template <class Launch>
int time_best(cosnark_ctx *ctx, Launch launch, float *best_ms) {
    hipEvent_t e0, e1;
    HIPCK(hipEventCreate(&e0));
    HIPCK(hipEventCreate(&e1));
    float best = 1e30f;
    for (int rep = 0; rep < 3; ++rep) {
        HIPCK(hipEventRecord(e0, ctx->stream));
        launch();
        LAUNCHCK();
        HIPCK(hipEventRecord(e1, ctx->stream));
        HIPCK(hipEventSynchronize(e1));
        float ms = 0;
        (void)hipEventElapsedTime(&ms, e0, e1);
        if (rep > 0 && ms < best) best = ms;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *best_ms = best;
    return COSNARK_OK;
}

}  // namespace

extern "C" {

int cosnark_blake3_rate(cosnark_ctx *ctx, double *per_s) {
    if (!ctx || !per_s) return COSNARK_E_ARG;
    HIPCK(hipSetDevice(ctx->device));
    constexpr unsigned blocks = 8192;
    constexpr int iters = 32;
    void *out;
    int rc = scratch_reserve(ctx, sizeof(uint32_t) * 256 * (size_t)blocks, &out);
    if (rc) return rc;
    float best = 0;
    rc = time_best(
        ctx,
        [&]() {
            hipLaunchKernelGGL(blake3_rate_kernel, dim3(blocks), dim3(256), 0, ctx->stream, iters, (uint32_t *)out);
        },
        &best);
    if (rc) return rc;
    *per_s = (double)blocks * 256 * iters * HR_CHAINS / (best / 1e3);
    return COSNARK_OK;
}

int cosnark_fmul_rate(cosnark_ctx *ctx, double *per_s) {
    if (!ctx || !per_s) return COSNARK_E_ARG;
    HIPCK(hipSetDevice(ctx->device));
    constexpr unsigned blocks = 8192;
    constexpr int iters = 256;
    void *out;
    int rc = scratch_reserve(ctx, sizeof(fe) * 256 * (size_t)blocks, &out);
    if (rc) return rc;
    hipEvent_t e0, e1;
    HIPCK(hipEventCreate(&e0));
    HIPCK(hipEventCreate(&e1));
    const fe b = gf::root_of_unity(20);
    float best = 1e30f;
    for (int rep = 0; rep < 3; ++rep) {
        HIPCK(hipEventRecord(e0, ctx->stream));
        hipLaunchKernelGGL(fmul_rate_kernel, dim3(blocks), dim3(256), 0, ctx->stream, b, iters, (fe *)out);
        LAUNCHCK();
        HIPCK(hipEventRecord(e1, ctx->stream));
        HIPCK(hipEventSynchronize(e1));
        float ms = 0;
        (void)hipEventElapsedTime(&ms, e0, e1);
        if (rep > 0 && ms < best) best = ms;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *per_s = (double)blocks * 256 * iters * FR_CHAINS / (best / 1e3);
    return COSNARK_OK;
}

int cosnark_autotune(cosnark_ctx *ctx, int logn_lo, int logn_hi, int batch, double *best_ms) {
    if (!ctx) return COSNARK_E_ARG;
    if (logn_lo < 1 || logn_hi > NTT_LOG_MAX || logn_lo > logn_hi || batch < 1)
        return set_err(ctx, COSNARK_E_ARG, "autotune: need 1 <= logn_lo <= logn_hi <= 28, batch >= 1");
    HIPCK(hipSetDevice(ctx->device));
    const uint64_t maxn = (uint64_t)batch << logn_hi;
    fe *in = nullptr, *out = nullptr, *tmp = nullptr;
    fe **itab = nullptr, **otab = nullptr;
    auto cleanup = [&]() {
        for (void *p : {(void *)in, (void *)out, (void *)tmp, (void *)itab, (void *)otab})
            if (p) (void)hipFree(p);
    };
    if (hipMalloc(&in, sizeof(fe) * maxn) != hipSuccess || hipMalloc(&out, sizeof(fe) * maxn) != hipSuccess ||
        hipMalloc(&tmp, sizeof(fe) * maxn) != hipSuccess || hipMalloc(&itab, sizeof(fe *) * batch) != hipSuccess ||
        hipMalloc(&otab, sizeof(fe *) * batch) != hipSuccess) {
        cleanup();
        return set_err(ctx, COSNARK_E_NOMEM, "autotune: allocation failed");
    }
    // inputs: any values do (the plan does not change what is computed)
    if (hipMemsetAsync(in, 0, sizeof(fe) * maxn, ctx->stream) != hipSuccess) {
        cleanup();
        return set_err(ctx, COSNARK_E_HIP, "autotune: memset failed");
    }
    hipEvent_t e0, e1;
    HIPCK(hipEventCreate(&e0));
    HIPCK(hipEventCreate(&e1));
    const int bits_c[] = {5, 6, 7, 8};
    const int tile_c[] = {1024, 2048, 4096};
    int rc = COSNARK_OK;
    for (int logn = logn_lo; logn <= logn_hi && rc == COSNARK_OK; ++logn) {
        const uint64_t N = 1ull << logn;
        std::vector<const fe *> ih(batch);
        std::vector<fe *> oh(batch);
        for (int y = 0; y < batch; ++y) { ih[y] = in + y * N; oh[y] = out + y * N; }
        if (hipMemcpy(itab, ih.data(), sizeof(fe *) * batch, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(otab, oh.data(), sizeof(fe *) * batch, hipMemcpyHostToDevice) != hipSuccess) {
            rc = set_err(ctx, COSNARK_E_HIP, "autotune: table copy failed");
            break;
        }
        float best = 1e30f;
        int bb = 0, bt = 0;
        for (int bits : bits_c) {
            for (int tile : tile_c) {
                if ((1 << bits) > tile) continue;
                ctx->ntt_plan[logn].passbits = bits;
                ctx->ntt_plan[logn].tile = tile;
                float t = 1e30f;
                for (int rep = 0; rep < 4 && rc == COSNARK_OK; ++rep) {
                    HIPCK(hipEventRecord(e0, ctx->stream));
                    rc = ntt_launch_batch(ctx, nullptr, (const fe *const *)itab, N, nullptr, otab, batch, logn, false,
                                          ctx->stream, tmp);
                    HIPCK(hipEventRecord(e1, ctx->stream));
                    HIPCK(hipEventSynchronize(e1));
                    float ms = 0;
                    (void)hipEventElapsedTime(&ms, e0, e1);
                    if (rep > 0 && ms < t) t = ms;      // the first run warms up
                }
                if (t < best) { best = t; bb = bits; bt = tile; }
            }
        }
        ctx->ntt_plan[logn].passbits = bb;
        ctx->ntt_plan[logn].tile = bt;
        if (best_ms) best_ms[logn - logn_lo] = best;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    cleanup();
    return rc;
}

// The Merkle schedule: how many first-level nodes one workgroup takes (2^b),
// which sets how many levels it climbs in LDS before the next launch (b) and
// how many launches a tree needs.  Fan-in stays 2: the reference's binary,
// dup-left tree (src/merkle_tree.cpp:267-299).  Every b builds the same
// tree; the tuner times a batch of trees of the given leaf counts, built in
// shared launches as the batched prover builds them, under each b.
int cosnark_autotune_merkle(cosnark_ctx *ctx, const uint64_t *leaf_counts, int count, int *best_block_log,
                            double *best_ms) {
    if (!ctx || !leaf_counts || count < 1 || count > 4096) return set_err(ctx, COSNARK_E_ARG, "autotune_merkle: arguments");
    HIPCK(hipSetDevice(ctx->device));
    std::vector<uint64_t> n_elems(count);
    uint64_t total = 0, nodes_total = 0;
    for (int i = 0; i < count; ++i) {
        if (!is_pow2(leaf_counts[i]) || leaf_counts[i] < 2 || leaf_counts[i] > (1ull << 26))
            return set_err(ctx, COSNARK_E_ARG, "autotune_merkle: leaf counts powers of two in [2, 2^26]");
        n_elems[i] = 4 * leaf_counts[i];
        total += n_elems[i];
        nodes_total += cosnark_merkle_node_count(n_elems[i]);
    }
    fe *el = nullptr;
    uint8_t *nd = nullptr, *tables = nullptr;
    auto cleanup = [&]() {
        for (void *p : {(void *)el, (void *)nd, (void *)tables})
            if (p) (void)hipFree(p);
    };
    if (hipMalloc(&el, sizeof(fe) * total) != hipSuccess || hipMalloc(&nd, 32 * nodes_total) != hipSuccess) {
        cleanup();
        return set_err(ctx, COSNARK_E_NOMEM, "autotune_merkle: allocation failed");
    }
    if (hipMemsetAsync(el, 0, sizeof(fe) * total, ctx->stream) != hipSuccess) {
        cleanup();
        return set_err(ctx, COSNARK_E_HIP, "autotune_merkle: memset failed");
    }
    std::vector<const fe *> ep(count);
    std::vector<uint8_t *> np(count);
    uint64_t eo = 0, no = 0;
    for (int i = 0; i < count; ++i) {
        ep[i] = el + eo;
        np[i] = nd + 32 * no;
        eo += n_elems[i];
        no += cosnark_merkle_node_count(n_elems[i]);
    }
    const int saved = ctx->mk_block_log;
    float best = 1e30f;
    int bb = saved, rc = COSNARK_OK;
    for (int b = 6; b <= 10 && rc == COSNARK_OK; ++b) {
        ctx->mk_block_log = b;
        TableBuf tb;
        MerkleBatchPlan plan;
        if ((rc = merkle_batch_plan(ctx, count, ep.data(), n_elems.data(), np.data(), true, tb, plan))) break;
        if (tables) (void)hipFree(tables);
        tables = nullptr;
        if (hipMalloc(&tables, tb.host.size()) != hipSuccess ||
            hipMemcpy(tables, tb.host.data(), tb.host.size(), hipMemcpyHostToDevice) != hipSuccess) {
            rc = set_err(ctx, COSNARK_E_NOMEM, "autotune_merkle: tables");
            break;
        }
        float t = 0;
        rc = time_best(ctx, [&]() { (void)merkle_batch_run(ctx, plan, tables, ctx->stream); }, &t);
        if (rc == COSNARK_OK && t < best) { best = t; bb = b; }
    }
    ctx->mk_block_log = rc == COSNARK_OK ? bb : saved;
    if (best_block_log) *best_block_log = ctx->mk_block_log;
    if (best_ms) *best_ms = best;
    (void)hipStreamSynchronize(ctx->stream);
    cleanup();
    return rc;
}

int cosnark_merkle_schedule(cosnark_ctx *ctx, int *block_log) {
    if (!ctx || !block_log) return COSNARK_E_ARG;
    *block_log = ctx->mk_block_log;
    return COSNARK_OK;
}

int cosnark_set_merkle_schedule(cosnark_ctx *ctx, int block_log) {
    if (!ctx || block_log < 6 || block_log > 10) return set_err(ctx, COSNARK_E_ARG, "set_merkle_schedule: 6 <= block_log <= 10");
    ctx->mk_block_log = block_log;
    return COSNARK_OK;
}

int cosnark_ntt_plan(cosnark_ctx *ctx, int logn, int *passbits, int *tile) {
    if (!ctx || logn < 1 || logn > NTT_LOG_MAX) return COSNARK_E_ARG;
    if (passbits) *passbits = ntt_plan_bits(ctx, logn);
    if (tile) *tile = ntt_plan_tile(ctx, logn);
    return COSNARK_OK;
}

int cosnark_set_ntt_plan(cosnark_ctx *ctx, int logn, int passbits, int tile) {
    if (!ctx || logn < 1 || logn > NTT_LOG_MAX) return COSNARK_E_ARG;
    if (passbits == 0 && tile == 0) {      // back to the context defaults
        ctx->ntt_plan[logn].passbits = 0;
        ctx->ntt_plan[logn].tile = 0;
        return COSNARK_OK;
    }
    if (passbits < 2 || passbits > NTT_PASS_LOG_MAX || tile < 256 || tile > NTT_LDS_ELEMS || !is_pow2((uint64_t)tile))
        return set_err(ctx, COSNARK_E_ARG, "set_ntt_plan: passbits in [2, 8], tile a power of two in [256, 4096]");
    ctx->ntt_plan[logn].passbits = passbits;
    ctx->ntt_plan[logn].tile = tile;
    return COSNARK_OK;
}

}  // extern "C"
