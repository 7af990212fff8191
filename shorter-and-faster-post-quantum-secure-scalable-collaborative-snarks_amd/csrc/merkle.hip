// merkle.hip -- BLAKE3 Merkle commit (MT_commit_Blake + create_tree_blake).
//
// Reference: src/merkle_tree.cpp:195-233 (leaf i = BLAKE3 of elements
// 4i..4i+3, 64 raw bytes) and :267-299 (level l node i = BLAKE3 of
// level l-1 nodes 2i and 2i -- the left child twice, :287-292).  Every level
// is materialised because openings (open_tree_blake, :320-335) serve
// siblings from all levels.
//
// One message per lane.  A workgroup of 256 lanes takes a 1024-node block of
// its first level: each lane hashes four consecutive nodes and the two
// levels above them in registers, then the workgroup climbs the remaining
// eight levels of the block through LDS, writing every node (levels of at
// most 64 nodes with four lanes per compression); the block roots feed the
// next launch.  A 2^22-leaf tree takes three launches.
//
// Column mode serves distributed_MT (src/coPCS.cpp:152-183): first-level
// node j is the root of MT_commit_Blake over column j of a 32-row codeword
// matrix (rows >= nrows are zero).  With the left child hashed twice, that
// 8-leaf tree keeps only leaf 0 (rows 0..3 of the column) and self-hashes it
// three times, so a lane reads 64 bytes per column and spends four
// compressions on it.
#include <cstring>
#include <utility>
#include <vector>
#include "ctx.h"
#include "blake3.h"

using gf::fe;

namespace {

constexpr int MK_THREADS = 256;
constexpr int MK_BLOCK = 1024;      // first-level nodes per workgroup

struct LevelOffsets {
    uint64_t off[40];               // digest offset of each level in `nodes`
};

struct Digest {
    uint32_t w[8];
};

__device__ __forceinline__ void store_digest(uint8_t *nodes, uint64_t idx, const uint32_t h[8]) {
    uint4 *p = reinterpret_cast<uint4 *>(nodes + 32 * idx);
    p[0] = make_uint4(h[0], h[1], h[2], h[3]);
    p[1] = make_uint4(h[4], h[5], h[6], h[7]);
}

__device__ __forceinline__ void hash_pair(const uint32_t l[8], const uint32_t r[8], uint32_t h[8]) {
    uint32_t m[16];
#pragma unroll
    for (int j = 0; j < 8; ++j) { m[j] = l[j]; m[8 + j] = r[j]; }
    b3::compress64(m, h);
}

// MODE_ELEMS: the first level is the leaf level (input = elements).
// MODE_NODES: the first level is level `lvl0` computed from level lvl0-1.
// MODE_COLUMNS: the first level holds the column roots of a row-major matrix
// of `col_rows` rows of `col_len` elements (dup-left column trees only).
// A lane owns 2^RL consecutive first-level nodes (RL = J.reg_log) and
// climbs RL levels in registers; the workgroup then climbs the remaining
// block_log - RL levels through LDS.  All input loads of a lane are issued
// before its first compression.
constexpr int MODE_ELEMS = 0, MODE_NODES = 1, MODE_COLUMNS = 2;

// One launch's work on one tree.  A batch launch covers several trees: its
// workgroup b works on job blk_job[b], block b - first_block of it.
struct MkJob {
    const fe *elems;            // MODE_ELEMS, MODE_COLUMNS input
    const uint8_t *in_nodes;    // MODE_NODES: level lvl0 - 1
    uint8_t *nodes;             // all levels of the tree
    LevelOffsets lo;
    uint64_t col_len;
    int lvl0, block_log, dup_left, col_rows;
    int reg_log;                // RL: first-level nodes per lane 2^RL, 0 <= RL <= 2, block_log - RL <= 8
    uint32_t first_block;
};

template <int MODE>
__device__ __forceinline__ void merkle_block(const MkJob &J, uint64_t blk) {
    __shared__ Digest bufA[MK_BLOCK / 4];
    __shared__ Digest bufB[MK_BLOCK / 8];
    const fe *__restrict__ elems = J.elems;
    const uint8_t *in_nodes = J.in_nodes;
    uint8_t *nodes = J.nodes;
    const LevelOffsets &lo = J.lo;
    const int lvl0 = J.lvl0, block_log = J.block_log, dup_left = J.dup_left, col_rows = J.col_rows;
    const uint64_t col_len = J.col_len;
    const int tid = threadIdx.x;
    const int RL = J.reg_log;
    const int per = 1 << RL;                            // first-level nodes per lane
    const int lanes = (1 << block_log) >> RL;           // active lanes
    const uint64_t g0 = blk << block_log;

    if (tid < lanes) {
        const uint64_t gl = g0 + (uint64_t)tid * per;   // first node of this lane
        uint32_t msg[4][16];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            if (i < per) {
                const uint64_t g = gl + i;
                if (MODE == MODE_ELEMS) {
                    const uint4 *src = reinterpret_cast<const uint4 *>(elems + 4 * g);
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        const uint4 q = src[j];
                        msg[i][4 * j] = q.x; msg[i][4 * j + 1] = q.y; msg[i][4 * j + 2] = q.z; msg[i][4 * j + 3] = q.w;
                    }
                } else if (MODE == MODE_COLUMNS) {
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        uint4 q = make_uint4(0, 0, 0, 0);
                        if (j < col_rows) q = *reinterpret_cast<const uint4 *>(elems + (uint64_t)j * col_len + g);
                        msg[i][4 * j] = q.x; msg[i][4 * j + 1] = q.y; msg[i][4 * j + 2] = q.z; msg[i][4 * j + 3] = q.w;
                    }
                } else {
                    const uint4 *L = reinterpret_cast<const uint4 *>(in_nodes + 32 * (2 * g));
                    const uint4 *Rn = reinterpret_cast<const uint4 *>(in_nodes + 32 * (2 * g + (dup_left ? 0 : 1)));
                    const uint4 l0 = L[0], l1 = L[1], r0 = Rn[0], r1 = Rn[1];
                    uint32_t *m = msg[i];
                    m[0] = l0.x; m[1] = l0.y; m[2] = l0.z; m[3] = l0.w;
                    m[4] = l1.x; m[5] = l1.y; m[6] = l1.z; m[7] = l1.w;
                    m[8] = r0.x; m[9] = r0.y; m[10] = r0.z; m[11] = r0.w;
                    m[12] = r1.x; m[13] = r1.y; m[14] = r1.z; m[15] = r1.w;
                }
            }
        }
        uint32_t h[4][8];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            if (i < per) {
                b3::compress64(msg[i], h[i]);
                if (MODE == MODE_COLUMNS) {
                    // the column tree's levels 1..3 over its leaf 0
#pragma unroll
                    for (int l = 0; l < 3; ++l) {
                        uint32_t t[8];
                        hash_pair(h[i], h[i], t);
#pragma unroll
                        for (int j = 0; j < 8; ++j) h[i][j] = t[j];
                    }
                }
                store_digest(nodes, lo.off[lvl0] + gl + i, h[i]);
            }
        }
        // register levels: lvl0+1 (two nodes), lvl0+2 (one node)
        if (RL >= 1) {
            uint32_t p0[8], p1[8];
            hash_pair(h[0], dup_left ? h[0] : h[1], p0);
            store_digest(nodes, lo.off[lvl0 + 1] + (gl >> 1), p0);
            if (RL >= 2) {
                hash_pair(h[2], dup_left ? h[2] : h[3], p1);
                store_digest(nodes, lo.off[lvl0 + 1] + (gl >> 1) + 1, p1);
                uint32_t q[8];
                hash_pair(p0, dup_left ? p0 : p1, q);
                store_digest(nodes, lo.off[lvl0 + 2] + (gl >> 2), q);
#pragma unroll
                for (int j = 0; j < 8; ++j) bufA[tid].w[j] = q[j];
            } else {
#pragma unroll
                for (int j = 0; j < 8; ++j) bufA[tid].w[j] = p0[j];
            }
        } else {
#pragma unroll
            for (int j = 0; j < 8; ++j) bufA[tid].w[j] = h[0][j];
        }
    }

    // climb inside the block through LDS: level lvl0 + s has T >> s nodes
    Digest *src = bufA, *dst = bufB;
    for (int s = RL + 1; s <= block_log; ++s) {
        __syncthreads();
        const int cnt = (1 << block_log) >> s;
        const uint64_t gbase = g0 >> s;
        if (4 * cnt <= MK_THREADS) {
            // few nodes: four lanes per compression shortens each level's chain
            if (tid < 4 * cnt) {
                const int i = tid >> 2, q = tid & 3;
                const int ri = dup_left ? 2 * i : 2 * i + 1;
                uint32_t m[16], lo_w, hi_w;
#pragma unroll
                for (int j = 0; j < 8; ++j) { m[j] = src[2 * i].w[j]; m[8 + j] = src[ri].w[j]; }
                b3::compress64_quad(m, lo_w, hi_w);
                uint32_t *g = reinterpret_cast<uint32_t *>(nodes + 32 * (lo.off[lvl0 + s] + gbase + i));
                g[q] = lo_w;
                g[4 + q] = hi_w;
                dst[i].w[q] = lo_w;
                dst[i].w[4 + q] = hi_w;
            }
            Digest *t = src; src = dst; dst = t;
            continue;
        }
        for (int i = tid; i < cnt; i += MK_THREADS) {
            uint32_t m[16], hh[8];
#pragma unroll
            for (int j = 0; j < 8; ++j) m[j] = src[2 * i].w[j];
            const int ri = dup_left ? 2 * i : 2 * i + 1;
#pragma unroll
            for (int j = 0; j < 8; ++j) m[8 + j] = src[ri].w[j];
            b3::compress64(m, hh);
            store_digest(nodes, lo.off[lvl0 + s] + gbase + i, hh);
#pragma unroll
            for (int j = 0; j < 8; ++j) dst[i].w[j] = hh[j];
        }
        Digest *t = src; src = dst; dst = t;
    }
}

template <int MODE>
__global__ __launch_bounds__(MK_THREADS) void merkle_block_kernel(MkJob J) {
    merkle_block<MODE>(J, blockIdx.x);
}

template <int MODE>
__global__ __launch_bounds__(MK_THREADS) void merkle_batch_kernel(const MkJob *__restrict__ jobs,
                                                                  const uint16_t *__restrict__ blk_job) {
    const MkJob &J = jobs[blk_job[blockIdx.x]];
    merkle_block<MODE>(J, blockIdx.x - J.first_block);
}

}  // namespace

namespace {

// levels a lane climbs in registers: `want` (the context's schedule) within
// 0..min(2, block_log), and at least block_log - 8 (at most 256 lanes)
int reg_levels(int block_log, int want) {
    int rl = want < 2 ? want : 2;
    if (rl > block_log) rl = block_log;
    if (rl < block_log - 8) rl = block_log - 8;
    return rl < 0 ? 0 : rl;
}

LevelOffsets level_offsets(uint64_t leaves) {
    LevelOffsets lo{};
    uint64_t off = 0;
    for (int l = 0; l <= ilog2(leaves); ++l) {
        lo.off[l] = off;
        off += leaves >> l;
    }
    return lo;
}

// The launches that build one tree: the first over `leaves` first-level
// nodes, then one per up to ten further levels until the root.  Returns the
// (job, workgroups) of each launch.
std::vector<std::pair<MkJob, uint64_t>> tree_stages(int mode, const fe *elems, uint64_t leaves, uint8_t *nodes,
                                                    bool dup_left, uint64_t col_len, int col_rows, uint64_t cap,
                                                    int want_rl) {
    std::vector<std::pair<MkJob, uint64_t>> st;
    MkJob J{};
    J.elems = elems;
    J.nodes = nodes;
    J.lo = level_offsets(leaves);
    J.col_len = col_len;
    J.col_rows = col_rows;
    J.dup_left = (int)dup_left;
    J.lvl0 = 0;
    J.block_log = ilog2(leaves < cap ? leaves : cap);
    J.reg_log = reg_levels(J.block_log, want_rl);
    uint64_t blocks = leaves >> J.block_log;
    st.push_back({J, blocks});
    int top = J.block_log;                            // highest level written
    uint64_t count = blocks;                          // nodes at level `top`
    while (count > 1) {
        const uint64_t next = count / 2;              // nodes at level top+1
        MkJob U = J;
        U.elems = nullptr;
        U.in_nodes = nodes + 32 * J.lo.off[top];
        U.lvl0 = top + 1;
        U.block_log = ilog2(next < cap ? next : cap);
        const uint64_t nb = next >> U.block_log;
        // a one-workgroup stage is a latency chain: one node per lane, the
        // most lanes at its first level
        U.reg_log = reg_levels(U.block_log, nb == 1 ? 0 : want_rl);
        st.push_back({U, nb});
        top += 1 + U.block_log;
        count = nb;
    }
    return st;
}

// Levels have.. above levels 0 .. have-1 already in nodes (level 0: `leaves`
// digests).
std::vector<std::pair<MkJob, uint64_t>> upper_stages(uint64_t leaves, uint8_t *nodes, bool dup_left, int have,
                                                     uint64_t cap, int want_rl) {
    std::vector<std::pair<MkJob, uint64_t>> st;
    MkJob J{};
    J.nodes = nodes;
    J.lo = level_offsets(leaves);
    J.dup_left = (int)dup_left;
    int top = have - 1;
    uint64_t count = leaves >> top;
    while (count > 1) {
        const uint64_t next = count / 2;
        MkJob U = J;
        U.in_nodes = nodes + 32 * J.lo.off[top];
        U.lvl0 = top + 1;
        U.block_log = ilog2(next < cap ? next : cap);
        const uint64_t nb = next >> U.block_log;
        // a one-workgroup stage is a latency chain: one node per lane, the
        // most lanes at its first level
        U.reg_log = reg_levels(U.block_log, nb == 1 ? 0 : want_rl);
        st.push_back({U, nb});
        top += 1 + U.block_log;
        count = nb;
    }
    return st;
}

// first-level nodes per workgroup: the context's Merkle schedule
uint64_t block_cap(const cosnark_ctx *ctx) { return 1ull << ctx->mk_block_log; }

int build_tree(cosnark_ctx *ctx, int mode, const fe *elems, uint64_t leaves, uint8_t *nodes, bool dup_left,
               uint64_t col_len, int col_rows, hipStream_t stream) {
    const auto st = tree_stages(mode, elems, leaves, nodes, dup_left, col_len, col_rows, block_cap(ctx), ctx->mk_reg_log);
    for (size_t i = 0; i < st.size(); ++i) {
        auto k = i > 0 ? merkle_block_kernel<MODE_NODES>
                       : mode == MODE_ELEMS ? merkle_block_kernel<MODE_ELEMS> : merkle_block_kernel<MODE_COLUMNS>;
        hipLaunchKernelGGL(k, dim3((unsigned)st[i].second), dim3(MK_THREADS), 0, stream, st[i].first);
        LAUNCHCK();
    }
    return COSNARK_OK;
}

}  // namespace

int merkle_launch(cosnark_ctx *ctx, const fe *elems, uint64_t n_elems, uint8_t *nodes, bool dup_left,
                  hipStream_t stream) {
    if (!is_pow2(n_elems) || n_elems < 8 || n_elems > (1ull << 30))
        return set_err(ctx, COSNARK_E_ARG, "merkle: n_elems must be a power of two in [8, 2^30]");
    return build_tree(ctx, MODE_ELEMS, elems, n_elems / 4, nodes, dup_left, 0, 0, stream);
}

int merkle_from_leaves_launch(cosnark_ctx *ctx, uint64_t leaves, uint8_t *nodes, bool dup_left, hipStream_t stream,
                              int have) {
    if (!is_pow2(leaves) || leaves > (1ull << 30))
        return set_err(ctx, COSNARK_E_ARG, "merkle: leaf count must be a power of two <= 2^30");
    if (have < 1 || (leaves >> (have - 1)) < 1)
        return set_err(ctx, COSNARK_E_ARG, "merkle: levels already built out of range");
    for (const auto &st : upper_stages(leaves, nodes, dup_left, have, block_cap(ctx), ctx->mk_reg_log)) {
        hipLaunchKernelGGL(merkle_block_kernel<MODE_NODES>, dim3((unsigned)st.second), dim3(MK_THREADS), 0, stream,
                           st.first);
        LAUNCHCK();
    }
    return COSNARK_OK;
}

int merkle_columns_launch(cosnark_ctx *ctx, const fe *cw, int nrows, uint64_t Lc, uint8_t *nodes,
                          hipStream_t stream) {
    if (!is_pow2(Lc) || Lc > (1ull << 28))
        return set_err(ctx, COSNARK_E_ARG, "distributed_mt: Lc must be a power of two <= 2^28");
    if (nrows < 1 || nrows > 32)
        return set_err(ctx, COSNARK_E_ARG, "distributed_mt: nrows must be in [1, 32]");
    return build_tree(ctx, MODE_COLUMNS, cw, Lc, nodes, true, Lc, nrows, stream);
}

// ---- batches of trees -----------------------------------------------------

int merkle_batch_plan(cosnark_ctx *ctx, int count, const fe *const *elems, const uint64_t *n_elems,
                      uint8_t *const *nodes, bool dup_left, TableBuf &tb, MerkleBatchPlan &plan, const int *have) {
    std::vector<std::vector<std::pair<MkJob, uint64_t>>> per(count);
    std::vector<int> mode0(count);
    size_t nst = 0;
    for (int i = 0; i < count; ++i) {
        if (!is_pow2(n_elems[i]) || n_elems[i] < 8 || n_elems[i] > (1ull << 30))
            return set_err(ctx, COSNARK_E_ARG, "merkle batch: n_elems must be powers of two in [8, 2^30]");
        const uint64_t leaves = n_elems[i] / 4;
        const int h = have ? have[i] : 0;
        if (h < 0 || (h > 0 && (leaves >> (h - 1)) < 1))
            return set_err(ctx, COSNARK_E_ARG, "merkle batch: levels already built out of range");
        // a tree whose lowest levels are in place starts above them
        per[i] = h ? upper_stages(leaves, nodes[i], dup_left, h, block_cap(ctx), ctx->mk_reg_log)
                   : tree_stages(MODE_ELEMS, elems[i], leaves, nodes[i], dup_left, 0, 0, block_cap(ctx),
                                 ctx->mk_reg_log);
        mode0[i] = h ? MODE_NODES : MODE_ELEMS;
        nst = std::max(nst, per[i].size());
    }
    plan.launches.clear();
    // stage s of every tree: one launch per kernel mode present
    for (size_t s = 0; s < nst; ++s) {
        for (int mode : {MODE_ELEMS, MODE_NODES}) {
            std::vector<MkJob> jobs;
            std::vector<uint16_t> map;
            for (int i = 0; i < count; ++i) {
                if (s >= per[i].size() || (s == 0 ? mode0[i] : MODE_NODES) != mode) continue;
                MkJob J = per[i][s].first;
                J.first_block = (uint32_t)map.size();
                for (uint64_t b = 0; b < per[i][s].second; ++b) map.push_back((uint16_t)jobs.size());
                jobs.push_back(J);
            }
            if (jobs.empty()) continue;
            if (jobs.size() > 0xFFFF || map.size() > 0x7FFFFFFFu)
                return set_err(ctx, COSNARK_E_RANGE, "merkle batch: too many trees or blocks in one launch");
            MerkleBatchPlan::Launch L{mode, (uint32_t)map.size(), tb.reserve(sizeof(MkJob) * jobs.size()), 0};
            L.blk_job = tb.reserve(sizeof(uint16_t) * map.size());
            memcpy(tb.at<MkJob>(L.jobs), jobs.data(), sizeof(MkJob) * jobs.size());
            memcpy(tb.at<uint16_t>(L.blk_job), map.data(), sizeof(uint16_t) * map.size());
            plan.launches.push_back(L);
        }
    }
    return COSNARK_OK;
}

int merkle_batch_run(cosnark_ctx *ctx, const MerkleBatchPlan &plan, const uint8_t *tables, hipStream_t stream) {
    for (const auto &L : plan.launches) {
        auto k = L.mode == MODE_ELEMS ? merkle_batch_kernel<MODE_ELEMS> : merkle_batch_kernel<MODE_NODES>;
        hipLaunchKernelGGL(k, dim3(L.blocks), dim3(MK_THREADS), 0, stream, (const MkJob *)(tables + L.jobs),
                           (const uint16_t *)(tables + L.blk_job));
        LAUNCHCK();
    }
    return COSNARK_OK;
}
