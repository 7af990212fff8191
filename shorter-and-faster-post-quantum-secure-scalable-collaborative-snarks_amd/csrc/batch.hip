// batch.hip -- many plain proofs of uneven size in shared launches, replayed
// from a captured hipGraph.
//
// Each proof is the plain single-rank prover's (cosnark_prove_plain): RS
// encode of f at the batch's blowup (Whir_Commit_initial form, src/Whir.cpp:
// 158-163), the dup-left BLAKE3 Merkle commit of the codeword
// (src/merkle_tree.cpp:195-299) and the all-Fiat-Shamir product sumcheck of
// f*g (src/Distributed_Sumcheck.cpp:204-348), so its proof bytes equal the
// plain prover's for the same inputs and transcript.
//
// A batch is planned once for a list of sizes: device slots for every
// proof's f and g, codeword, tree, sumcheck buffers and state, and the launch
// tables.  One proving pass is
//   streams 1 and 3: RS encode, one launch per NTT pass for all proofs of
//             one size; Merkle, one launch per stage for all of the lane's
//             trees;
//   stream 2: the sumchecks, two rounds per launch for every proof whose
//             vectors are still long, then one tail workgroup per proof;
//   join; one kernel packs every proof's bytes into one buffer; one copy out.
// The commits go in two lanes on two streams (the largest codewords on one,
// the rest on the other, balanced by elements encoded), each lane hashing
// its own trees.  The commit lanes and the sumchecks are captured into one
// hipGraph each at the first call and replayed after that, on their three
// streams.  Every buffer the graph touches belongs to the batch, so later work
// on the context (which may regrow its own scratch) cannot invalidate it.
#include <algorithm>
#include <cstring>
#include <map>
#include <vector>
#include "ctx.h"

using gf::fe;

namespace {

constexpr int PACK_THREADS = 64;

struct PackJob {
    const uint8_t *root;      // last digest of the tree
    const ScState *st;
    uint8_t *out;             // proof bytes, then prev (16 B), err (16 B)
    int m;
};

// proof i: root | m x (a, b, c) | v1[0], v2[0] | transcript state | err
__global__ __launch_bounds__(PACK_THREADS) void pack_kernel(const PackJob *__restrict__ jobs) {
    const PackJob J = jobs[blockIdx.x];
    const int m = J.m, nq = 2 + 3 * m + 2 + 2;
    uint4 *o = reinterpret_cast<uint4 *>(J.out);
    for (int q = threadIdx.x; q < nq; q += PACK_THREADS) {
        uint4 v;
        if (q < 2) {
            v = reinterpret_cast<const uint4 *>(J.root)[q];
        } else {
            const fe *src;
            if (q < 2 + 3 * m) src = &J.st->polys[q - 2];
            else if (q < 4 + 3 * m) src = &J.st->final_v[q - 2 - 3 * m];
            else if (q == 4 + 3 * m) src = &J.st->tr_prev[m];
            else src = nullptr;
            if (src) v = make_uint4((uint32_t)src->re, (uint32_t)(src->re >> 32), (uint32_t)src->im,
                                    (uint32_t)(src->im >> 32));
            else v = make_uint4((uint32_t)J.st->err, 0, 0, 0);
        }
        o[q] = v;
    }
}

}  // namespace

size_t align256(size_t x) { return (x + 255) / 256 * 256; }

size_t batch_work_bytes(uint64_t n, int rate, bool with_commit) {
    const uint64_t cwn = n * (uint64_t)rate;
    const size_t commit = align256(sizeof(fe) * cwn) + align256(32 * cosnark_merkle_node_count(cwn));
    return (with_commit ? commit : 0) + align256(sizeof(fe) * (n + n / 2)) +
           align256(sizeof(fe) * 6 * SC_MAX_BLOCKS) + align256(sizeof(ScState));
}

size_t batch_record_bytes(uint64_t n) { return cosnark_plain_proof_bytes(n) + 32; }

int batch_core_build(cosnark_ctx *ctx, int count, const BatchProof *pr, int rate, fe *prev_dev, uint8_t *out_dev,
                     BatchCore &C, bool two_lanes) {
    C.count = count;
    C.rate = rate;
    C.n.resize(count);
    C.out_off.resize(count);
    C.groups.clear();
    C.tb.host.clear();
    C.prev_dev = prev_dev;
    C.out_dev = out_dev;
    std::vector<const fe *> fin(count), gin(count);
    std::vector<fe *> cws(count), bufs(count), parts(count);
    std::vector<uint8_t *> nds(count);
    std::vector<ScState *> sts(count);
    std::vector<uint64_t> cwn(count);
    size_t out = 0;
    for (int i = 0; i < count; ++i) {
        const uint64_t n = pr[i].n;
        C.n[i] = n;
        cwn[i] = n * (uint64_t)rate;
        uint8_t *w = (uint8_t *)pr[i].work;
        if (pr[i].cw && pr[i].nodes) {
            cws[i] = pr[i].cw;
            nds[i] = pr[i].nodes;
        } else {
            cws[i] = (fe *)w;
            w += align256(sizeof(fe) * cwn[i]);
            nds[i] = w;
            w += align256(32 * cosnark_merkle_node_count(cwn[i]));
        }
        bufs[i] = (fe *)w;
        w += align256(sizeof(fe) * (n + n / 2));
        parts[i] = (fe *)w;
        w += align256(sizeof(fe) * 6 * SC_MAX_BLOCKS);
        sts[i] = (ScState *)w;
        fin[i] = pr[i].f;
        gin[i] = pr[i].g;
        C.out_off[i] = out;
        out += batch_record_bytes(n);
    }
    C.out_bytes = out;
    C.cw = cws;
    C.nodes = nds;
    // RS encodes: one group per codeword size, launched together
    std::map<int, std::vector<int>> by_log;
    for (int i = 0; i < count; ++i) by_log[ilog2(cwn[i])].push_back(i);
    // Where a group's plan allows it, the last NTT pass of each encode also
    // writes its tree's leaf digests and levels 1 and 2 (the rule of
    // ntt_launch_batch, asked in a dry run), and that tree starts above them.
    std::vector<int> have(count, 0);
    for (auto &kv : by_log) {
        BatchCore::Group G{kv.first, (int)kv.second.size(), C.n[kv.second[0]], 0, 0, 0, 0};
        G.in_tab = C.tb.reserve(sizeof(fe *) * G.B);
        G.out_tab = C.tb.reserve(sizeof(fe *) * G.B);
        G.lv_tab = C.tb.reserve(sizeof(uint8_t *) * G.B);
        for (int j = 0; j < G.B; ++j) {
            C.tb.at<const fe *>(G.in_tab)[j] = fin[kv.second[j]];
            C.tb.at<fe *>(G.out_tab)[j] = cws[kv.second[j]];
            C.tb.at<uint8_t *>(G.lv_tab)[j] = nds[kv.second[j]];
        }
        if (ctx->batch_leaf_levels > 0) {
            // a dry run launches and reads nothing: only the plan and the
            // presence of a node table count
            static uint8_t *const any_tab[1] = {nullptr};
            int rc = ntt_launch_batch(ctx, nullptr, nullptr, G.in_len, nullptr, nullptr, G.B, G.logn, false, nullptr,
                                      nullptr, nullptr, &G.levels, ctx->batch_leaf_levels, any_tab, true);
            if (rc) return rc;
        }
        for (int i : kv.second) have[i] = G.levels;
        C.groups.push_back(G);
    }
    // Two commit lanes (two streams): the groups, largest first, each to the
    // lane with less work so far (work ~ elements encoded).  The large
    // groups fill the device by themselves; the small ones, a few
    // workgroups per launch, then run beside them instead of one after
    // another.  Each lane hashes its own trees.
    C.glane.assign(C.groups.size(), 0);
    C.two_lanes = false;
    if (two_lanes && C.groups.size() >= 2) {
        std::vector<int> order(C.groups.size());
        for (size_t g = 0; g < order.size(); ++g) order[g] = (int)g;
        auto work = [&](int g) { return (uint64_t)C.groups[g].B << C.groups[g].logn; };
        std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return work(a) > work(b); });
        uint64_t load[2] = {0, 0};
        for (int g : order) {
            const int l = load[1] < load[0] ? 1 : 0;
            C.glane[g] = l;
            load[l] += work(g);
        }
        for (int l : C.glane) C.two_lanes = C.two_lanes || l == 1;
    }
    std::vector<int> lane_of(count, 0);
    for (size_t g = 0; g < C.groups.size(); ++g)
        for (int i : by_log[C.groups[g].logn]) lane_of[i] = C.glane[g];
    int rc = COSNARK_OK;
    C.mk1 = MerkleBatchPlan{};
    for (int l = 0; l < (C.two_lanes ? 2 : 1); ++l) {
        std::vector<const fe *> e;
        std::vector<uint64_t> ne;
        std::vector<uint8_t *> nd;
        std::vector<int> hv;
        for (int i = 0; i < count; ++i)
            if (lane_of[i] == l) {
                e.push_back(cws[i]);
                ne.push_back(cwn[i]);
                nd.push_back(nds[i]);
                hv.push_back(have[i]);
            }
        rc = merkle_batch_plan(ctx, (int)e.size(), e.data(), ne.data(), nd.data(), true, C.tb, l ? C.mk1 : C.mk,
                               hv.data());
        if (rc) return rc;
    }
    rc = sumcheck_batch_plan(ctx, count, fin.data(), gin.data(), C.n.data(), bufs.data(), parts.data(), sts.data(),
                             prev_dev, C.tb, C.sc);
    if (rc) return rc;
    C.pack = C.tb.reserve(sizeof(PackJob) * count);
    for (int i = 0; i < count; ++i) {
        PackJob &J = C.tb.at<PackJob>(C.pack)[i];
        J.root = nds[i] + 32 * (cosnark_merkle_node_count(cwn[i]) - 1);
        J.st = sts[i];
        J.out = out_dev + C.out_off[i];
        J.m = ilog2(C.n[i]);
    }
    return COSNARK_OK;
}

uint64_t batch_tmp_elems(const BatchCore &C, int lane) {
    uint64_t t = 0;
    for (size_t g = 0; g < C.groups.size(); ++g)
        if (C.glane.empty() ? lane == 0 : C.glane[g] == lane) t = std::max(t, (uint64_t)C.groups[g].B << C.groups[g].logn);
    return t;
}

// The three parts of a pass: the sumchecks (transcripts in first), the
// commits, and the packing of every proof's bytes with the copy out.
int batch_enqueue_sumchecks(cosnark_ctx *ctx, const BatchCore &C, const uint8_t *tables, const fe *prev_host,
                            hipStream_t s) {
    HIPCK(hipMemcpyAsync(C.prev_dev, prev_host, sizeof(fe) * C.count, hipMemcpyHostToDevice, s));
    return sumcheck_batch_run(ctx, C.sc, tables, s);
}

int batch_enqueue_commits(cosnark_ctx *ctx, const BatchCore &C, const uint8_t *tables, fe *tmp, uint64_t tmp_elems,
                          hipStream_t s1, int lane) {
    int rc;
    for (size_t gi = 0; gi < C.groups.size(); ++gi) {
        if ((C.glane.empty() ? 0 : C.glane[gi]) != lane) continue;
        const auto &G = C.groups[gi];
        const uint64_t chunk = tmp_elems >> G.logn;
        if (chunk == 0) return set_err(ctx, COSNARK_E_RANGE, "batch: NTT temporary smaller than one codeword");
        for (int c0 = 0; c0 < G.B; c0 += (int)chunk) {
            const int nb = (int)std::min<uint64_t>(chunk, (uint64_t)(G.B - c0));
            int lv = 0;
            rc = ntt_launch_batch(ctx, nullptr, (const fe *const *)(tables + G.in_tab) + c0, G.in_len, nullptr,
                                  (fe *const *)(tables + G.out_tab) + c0, nb, G.logn, false, s1, tmp, nullptr, &lv,
                                  G.levels, G.levels ? (uint8_t *const *)(tables + G.lv_tab) + c0 : nullptr);
            if (rc) return rc;
            // the trees were planned from the levels the encode writes
            if (lv != G.levels)
                return set_err(ctx, COSNARK_E_ARG, "batch: the NTT plan of 2^" + std::to_string(G.logn) +
                                                         " points changed after the batch was planned");
        }
    }
    return merkle_batch_run(ctx, lane ? C.mk1 : C.mk, tables, s1);
}

int batch_enqueue_pack(cosnark_ctx *ctx, const BatchCore &C, const uint8_t *tables, uint8_t *out_host,
                       hipStream_t s1) {
    hipLaunchKernelGGL(pack_kernel, dim3(C.count), dim3(PACK_THREADS), 0, s1, (const PackJob *)(tables + C.pack));
    LAUNCHCK();
    HIPCK(hipMemcpyAsync(out_host, C.out_dev, C.out_bytes, hipMemcpyDeviceToHost, s1));
    return COSNARK_OK;
}

int batch_core_enqueue(cosnark_ctx *ctx, const BatchCore &C, const uint8_t *tables, const fe *prev_host,
                       uint8_t *out_host, fe *tmp, uint64_t tmp_elems, hipStream_t s1, hipStream_t s2,
                       hipEvent_t fork, hipEvent_t join, uint64_t tmp_elems1, hipStream_t s3, hipEvent_t join3) {
    const bool lane1 = C.two_lanes;
    if (lane1 && (!s3 || !join3 || tmp_elems1 == 0))
        return set_err(ctx, COSNARK_E_ARG, "batch: two commit lanes need a third stream and their temporary");
    // the commits (the longer branch) are issued first, so they start at once
    HIPCK(hipEventRecord(fork, s1));
    HIPCK(hipStreamWaitEvent(s2, fork, 0));
    if (lane1) HIPCK(hipStreamWaitEvent(s3, fork, 0));
    int rc = batch_enqueue_commits(ctx, C, tables, tmp, tmp_elems, s1, 0);
    if (rc) return rc;
    if (lane1 && (rc = batch_enqueue_commits(ctx, C, tables, tmp + tmp_elems, tmp_elems1, s3, 1))) return rc;
    if ((rc = batch_enqueue_sumchecks(ctx, C, tables, prev_host, s2))) return rc;
    HIPCK(hipEventRecord(join, s2));
    HIPCK(hipStreamWaitEvent(s1, join, 0));
    if (lane1) {
        HIPCK(hipEventRecord(join3, s3));
        HIPCK(hipStreamWaitEvent(s1, join3, 0));
    }
    return batch_enqueue_pack(ctx, C, tables, out_host, s1);
}

// proof i's bytes and transcript out of the copied records; the index of
// the first proof whose sumcheck self-check failed, or -1
int batch_unpack(const BatchCore &C, const uint8_t *out_host, uint8_t *const *proofs, cosnark_transcript *ts,
                 int32_t *errs) {
    int bad = -1;
    for (int i = 0; i < C.count; ++i) {
        const uint64_t pb = cosnark_plain_proof_bytes(C.n[i]);
        const uint8_t *rec = out_host + C.out_off[i];
        if (proofs[i]) memcpy(proofs[i], rec, pb);
        if (ts) memcpy(ts[i].prev, rec + pb, 16);
        int32_t err;
        memcpy(&err, rec + pb + 16, 4);
        if (errs) errs[i] = err;
        if (err && bad < 0) bad = i;
    }
    return bad;
}

struct cosnark_batch {
    cosnark_ctx *ctx = nullptr;
    BatchCore core;
    std::vector<size_t> off_in;
    fe *f = nullptr, *g = nullptr, *tmp = nullptr, *prev_dev = nullptr, *prev_host = nullptr;
    uint8_t *work = nullptr, *tables = nullptr, *out_dev = nullptr, *out_host = nullptr;
    uint64_t tmp_elems = 0, tmp_elems1 = 0;   // NTT temporary of commit lanes 0 and 1 (lane 1's after lane 0's)
    hipEvent_t ev_fork = nullptr, ev_join = nullptr, ev_join3 = nullptr, ev_t0 = nullptr, ev_t1 = nullptr;
    bool use_graph = true;
    // one graph per stream: replaying one graph with several branches ran
    // them one after the other, so the commit lanes and the sumchecks are
    // captured apart and joined by events around the launches
    hipGraph_t graph = nullptr, graph_sc = nullptr, graph1 = nullptr;
    hipGraphExec_t exec = nullptr, exec_sc = nullptr, exec1 = nullptr;
};

namespace {

void batch_release(cosnark_batch *b) {
    if (!b) return;
    if (b->ctx) (void)hipSetDevice(b->ctx->device);
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    if (b->exec_sc) (void)hipGraphExecDestroy(b->exec_sc);
    if (b->graph_sc) (void)hipGraphDestroy(b->graph_sc);
    if (b->exec1) (void)hipGraphExecDestroy(b->exec1);
    if (b->graph1) (void)hipGraphDestroy(b->graph1);
    for (void *p : {(void *)b->f, (void *)b->g, (void *)b->tmp, (void *)b->prev_dev, (void *)b->work,
                    (void *)b->tables, (void *)b->out_dev})
        if (p) (void)hipFree(p);
    if (b->out_host) (void)hipHostFree(b->out_host);
    if (b->prev_host) (void)hipHostFree(b->prev_host);
    for (hipEvent_t e : {b->ev_fork, b->ev_join, b->ev_join3, b->ev_t0, b->ev_t1})
        if (e) (void)hipEventDestroy(e);
    delete b;
}

int batch_enqueue(cosnark_batch *b) {
    cosnark_ctx *ctx = b->ctx;
    return batch_core_enqueue(ctx, b->core, b->tables, b->prev_host, b->out_host, b->tmp, b->tmp_elems, ctx->stream,
                              ctx->stream2, b->ev_fork, b->ev_join, b->tmp_elems1, ctx->stream3, b->ev_join3);
}

template <class Enqueue>
int capture_one(cosnark_ctx *ctx, hipStream_t s, Enqueue enqueue, hipGraph_t *graph, hipGraphExec_t *exec) {
    HIPCK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    const int rc = enqueue();
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(s, &g);
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (e != hipSuccess) return hip_err(ctx, e, "batch: graph capture");
    *graph = g;
    HIPCK(hipGraphInstantiate(exec, g, nullptr, nullptr, 0));
    return COSNARK_OK;
}

int batch_capture(cosnark_batch *b) {
    cosnark_ctx *ctx = b->ctx;
    auto drop = [](hipGraph_t &g, hipGraphExec_t &e) {
        if (e) (void)hipGraphExecDestroy(e);
        if (g) (void)hipGraphDestroy(g);
        e = nullptr;
        g = nullptr;
    };
    int rc = capture_one(
        ctx, ctx->stream,
        [&]() { return batch_enqueue_commits(ctx, b->core, b->tables, b->tmp, b->tmp_elems, ctx->stream, 0); },
        &b->graph, &b->exec);
    if (rc) return rc;
    if (b->core.two_lanes)
        rc = capture_one(
            ctx, ctx->stream3,
            [&]() {
                return batch_enqueue_commits(ctx, b->core, b->tables, b->tmp + b->tmp_elems, b->tmp_elems1,
                                             ctx->stream3, 1);
            },
            &b->graph1, &b->exec1);
    if (!rc)
        rc = capture_one(
            ctx, ctx->stream2,
            [&]() { return batch_enqueue_sumchecks(ctx, b->core, b->tables, b->prev_host, ctx->stream2); },
            &b->graph_sc, &b->exec_sc);
    if (rc) {   // every graph or none
        drop(b->graph, b->exec);
        drop(b->graph1, b->exec1);
    }
    return rc;
}

}  // namespace

extern "C" {

int cosnark_batch_create(cosnark_ctx *ctx, int count, const uint64_t *sizes, int rate, cosnark_batch **out) {
    if (!ctx || !sizes || !out) return set_err(ctx, COSNARK_E_ARG, "batch_create: null argument");
    *out = nullptr;
    if (count < 1 || count > 4096) return set_err(ctx, COSNARK_E_ARG, "batch_create: count must be in [1, 4096]");
    if (rate < 1 || !is_pow2((uint64_t)rate)) return set_err(ctx, COSNARK_E_ARG, "batch_create: rate must be a power of two");
    for (int i = 0; i < count; ++i)
        if (!is_pow2(sizes[i]) || sizes[i] < 4 || sizes[i] * (uint64_t)rate < 8 ||
            sizes[i] * (uint64_t)rate > (1ull << 28))
            return set_err(ctx, COSNARK_E_ARG, "batch_create: sizes must be powers of two, n >= 4, 8 <= n*rate <= 2^28");
    HIPCK(hipSetDevice(ctx->device));
    auto *b = new cosnark_batch;
    b->ctx = ctx;
    auto fail = [&](int rc) { batch_release(b); return rc; };
    auto nomem = [&]() { batch_release(b); return set_err(ctx, COSNARK_E_NOMEM, "batch_create: allocation failed"); };
    size_t in_t = 0, work_t = 0, out_t = 0;
    std::vector<size_t> off_work;
    for (int i = 0; i < count; ++i) {
        b->off_in.push_back(in_t);
        in_t = align256(in_t + sizeof(fe) * sizes[i]);
        off_work.push_back(work_t);
        work_t += batch_work_bytes(sizes[i], rate);
        out_t += batch_record_bytes(sizes[i]);
    }
    if (hipMalloc(&b->f, in_t) != hipSuccess || hipMalloc(&b->g, in_t) != hipSuccess ||
        hipMalloc(&b->work, work_t) != hipSuccess || hipMalloc(&b->out_dev, out_t) != hipSuccess ||
        hipMalloc(&b->prev_dev, sizeof(fe) * count) != hipSuccess ||
        hipHostMalloc(&b->out_host, out_t, hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc(&b->prev_host, sizeof(fe) * count, hipHostMallocDefault) != hipSuccess)
        return nomem();
    if (hipEventCreateWithFlags(&b->ev_fork, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&b->ev_join, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&b->ev_join3, hipEventDisableTiming) != hipSuccess ||
        hipEventCreate(&b->ev_t0) != hipSuccess || hipEventCreate(&b->ev_t1) != hipSuccess)
        return fail(hip_err(ctx, hipErrorUnknown, "batch_create: events"));
    std::vector<BatchProof> pr(count);
    for (int i = 0; i < count; ++i) {
        pr[i].f = (const fe *)((const char *)b->f + b->off_in[i]);
        pr[i].g = (const fe *)((const char *)b->g + b->off_in[i]);
        pr[i].n = sizes[i];
        pr[i].work = b->work + off_work[i];
        b->prev_host[i] = gf::zero();
    }
    int rc = batch_core_build(ctx, count, pr.data(), rate, b->prev_dev, b->out_dev, b->core, true);
    if (rc) return fail(rc);
    // one NTT temporary per commit lane, each capped at 2^26 elements (a
    // lane's groups run in chunks that fit)
    auto lane_tmp = [&](int lane) {
        uint64_t maxn = 0;
        for (size_t g = 0; g < b->core.groups.size(); ++g)
            if (b->core.glane[g] == lane) maxn = std::max<uint64_t>(maxn, 1ull << b->core.groups[g].logn);
        return maxn ? std::max(maxn, std::min<uint64_t>(batch_tmp_elems(b->core, lane), 1ull << 26)) : 0;
    };
    b->tmp_elems = lane_tmp(0);
    b->tmp_elems1 = b->core.two_lanes ? lane_tmp(1) : 0;
    if (hipMalloc(&b->tmp, sizeof(fe) * (b->tmp_elems + b->tmp_elems1)) != hipSuccess ||
        hipMalloc(&b->tables, b->core.tb.host.size()) != hipSuccess)
        return nomem();
    if (hipMemcpy(b->tables, b->core.tb.host.data(), b->core.tb.host.size(), hipMemcpyHostToDevice) != hipSuccess)
        return fail(hip_err(ctx, hipErrorUnknown, "batch_create: tables"));
    *out = b;
    return COSNARK_OK;
}

void cosnark_batch_destroy(cosnark_batch *b) {
    if (!b) return;
    if (b->ctx) {
        (void)hipSetDevice(b->ctx->device);
        (void)hipStreamSynchronize(b->ctx->stream);
        (void)hipStreamSynchronize(b->ctx->stream2);
        (void)hipStreamSynchronize(b->ctx->stream3);
    }
    batch_release(b);
}

int cosnark_batch_input(cosnark_batch *b, int i, uint64_t **d_f, uint64_t **d_g) {
    if (!b || i < 0 || i >= b->core.count || !d_f || !d_g)
        return set_err(b ? b->ctx : nullptr, COSNARK_E_ARG, "batch_input: bad argument");
    *d_f = (uint64_t *)((char *)b->f + b->off_in[i]);
    *d_g = (uint64_t *)((char *)b->g + b->off_in[i]);
    return COSNARK_OK;
}

int cosnark_batch_commitment(cosnark_batch *b, int i, uint64_t **d_codeword, uint8_t **d_nodes) {
    if (!b || i < 0 || i >= b->core.count || !d_codeword || !d_nodes)
        return set_err(b ? b->ctx : nullptr, COSNARK_E_ARG, "batch_commitment: bad argument");
    *d_codeword = (uint64_t *)b->core.cw[i];
    *d_nodes = b->core.nodes[i];
    return COSNARK_OK;
}

uint64_t cosnark_batch_proof_offset(const cosnark_batch *b, int i) {
    if (!b || i < 0 || i > b->core.count) return 0;
    uint64_t o = 0;
    for (int j = 0; j < i; ++j) o += cosnark_plain_proof_bytes(b->core.n[j]);
    return o;
}

int cosnark_batch_set_graph(cosnark_batch *b, int on) {
    if (!b) return COSNARK_E_ARG;
    b->use_graph = on != 0;
    return COSNARK_OK;
}

int cosnark_batch_prove(cosnark_batch *b, cosnark_transcript *ts, uint8_t *h_proofs, double *ms) {
    if (!b || !ts || !h_proofs) return set_err(b ? b->ctx : nullptr, COSNARK_E_ARG, "batch_prove: null argument");
    cosnark_ctx *ctx = b->ctx;
    HIPCK(hipSetDevice(ctx->device));
    // the previous pass's copies are complete (each pass ends in a sync)
    for (int i = 0; i < b->core.count; ++i) b->prev_host[i] = gf::mk(ts[i].prev[0], ts[i].prev[1]);
    int rc;
    if (b->use_graph && !b->exec && (rc = batch_capture(b))) return rc;
    HIPCK(hipEventRecord(b->ev_t0, ctx->stream));
    if (b->use_graph) {
        const bool lane1 = b->exec1 != nullptr;
        HIPCK(hipEventRecord(b->ev_fork, ctx->stream));
        HIPCK(hipStreamWaitEvent(ctx->stream2, b->ev_fork, 0));
        if (lane1) HIPCK(hipStreamWaitEvent(ctx->stream3, b->ev_fork, 0));
        // the sumchecks' long chain of dependent rounds first: a graph
        // launch takes host time, and the later graphs' kernels start later
        HIPCK(hipGraphLaunch(b->exec_sc, ctx->stream2));
        if (lane1) HIPCK(hipGraphLaunch(b->exec1, ctx->stream3));
        HIPCK(hipGraphLaunch(b->exec, ctx->stream));
        HIPCK(hipEventRecord(b->ev_join, ctx->stream2));
        HIPCK(hipStreamWaitEvent(ctx->stream, b->ev_join, 0));
        if (lane1) {
            HIPCK(hipEventRecord(b->ev_join3, ctx->stream3));
            HIPCK(hipStreamWaitEvent(ctx->stream, b->ev_join3, 0));
        }
        if ((rc = batch_enqueue_pack(ctx, b->core, b->tables, b->out_host, ctx->stream))) return rc;
    } else if ((rc = batch_enqueue(b))) {
        return rc;
    }
    HIPCK(hipEventRecord(b->ev_t1, ctx->stream));
    HIPCK(hipStreamSynchronize(ctx->stream));
    if (ms) {
        float t = 0;
        (void)hipEventElapsedTime(&t, b->ev_t0, b->ev_t1);
        *ms = t;
    }
    std::vector<uint8_t *> dst(b->core.count);
    uint64_t o = 0;
    for (int i = 0; i < b->core.count; ++i) {
        dst[i] = h_proofs + o;
        o += cosnark_plain_proof_bytes(b->core.n[i]);
    }
    const int bad = batch_unpack(b->core, b->out_host, dst.data(), ts, nullptr);
    if (bad >= 0)
        return set_err(ctx, COSNARK_E_CHECK, "batch_prove: sumcheck self-check failed in proof " + std::to_string(bad));
    return COSNARK_OK;
}

}  // extern "C"
