// sumcheck.hip -- product sumcheck rounds with an on-device transcript.
//
// Reference: _quadratic_sumcheck (src/Distributed_Sumcheck.cpp:204-348) at
// one rank.  Round i over pairs j < L: H(x) += (d1 x + e1)(d2 x + e2) with
// e = v[2j], d = v[2j+1] - v[2j] (:224-233, polynomial.cpp:133-135), so
// a = sum d1 d2, b = sum (d1 e2 + e1 d2), c = sum e1 e2.  Check
// H(0) + H(1) = y (:243-247), challenge r = hash_to_field({a, b, c}) (:248),
// y = H(r), fold v[j] = r (v[2j+1] - v[2j]) + v[2j] (:257-261).
//
// Here b is formed as sum o1 o2 - a - c (o = v[2j+1]), three products per
// pair.  While the vectors are long, rounds go two per launch: a round's
// three sums and the next round's sums as quadratics in the challenge not
// yet drawn (nine values) are formed in one pass, so a launch finalizes two
// rounds and folds by both challenges (sc_pair01_kernel, sc_fold2_kernel).
// Each workgroup writes its partial sums; the next launch starts by reducing
// the previous launch's partials in wave 0 of every workgroup (a few KiB from
// L2, summed with DPP exchanges) and running the transcript hashes
// redundantly, on four lanes, while its other waves' loads are in flight, so
// no atomics, no last-workgroup election and no host round trip sit between
// launches.  Transcript values are kept per round (tr_prev[k], tr_y[k] =
// state before round k) so a kernel never reads a slot it writes.  Once a
// vector is small (sc_tail_elems) one workgroup runs all remaining rounds:
// its wave 0 runs the transcript while waves 1..7 fold by the challenge just
// drawn and form the next round's sums as quadratics in the challenge still
// being hashed, so a round costs the hash chain and little else.
#include <algorithm>
#include <cstddef>
#include <cstring>
#include <vector>
#include "ctx.h"
#include "transcript.h"

using gf::fe;

namespace {

constexpr int SC_THREADS = 256;
constexpr int SC_TAIL_THREADS = 512;    // wave 0: transcript; waves 1..7: folds and sums
constexpr int SC_TAIL_CT = SC_TAIL_THREADS - 64;
constexpr int SC_TAIL_WAVES = SC_TAIL_CT / 64;
constexpr int SC_ITEMS = 1;             // pairs per thread per sweep in the multi-workgroup rounds

__device__ __forceinline__ fe fold1(const fe &e, const fe &o, const fe &r) {
    return gf::add(gf::mul(r, gf::sub(o, e)), e);
}

// One 32-bit DPP move per half of each coordinate: lane i reads the lane
// that CTRL names (quad permutes and row mirrors, all rows and banks on).
template <int CTRL>
__device__ __forceinline__ uint64_t dpp64(uint64_t x) {
    const uint32_t lo = (uint32_t)__builtin_amdgcn_mov_dpp((int)(uint32_t)x, CTRL, 0xF, 0xF, false);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_mov_dpp((int)(uint32_t)(x >> 32), CTRL, 0xF, 0xF, false);
    return ((uint64_t)hi << 32) | lo;
}

template <int CTRL>
__device__ __forceinline__ fe dpp_fe(const fe &v) {
    return gf::mk(dpp64<CTRL>(v.re), dpp64<CTRL>(v.im));
}

__device__ __forceinline__ uint64_t readlane64(uint64_t x, int lane) {
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)x, lane);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(x >> 32), lane);
    return ((uint64_t)hi << 32) | lo;
}

__device__ __forceinline__ fe readlane_fe(const fe &v, int lane) {
    return gf::mk(readlane64(v.re, lane), readlane64(v.im, lane));
}

// x < 2^64 -> x mod p up to one p (<= p + 7)
__device__ __forceinline__ uint64_t fold61(uint64_t x) { return (x & gf::P) + (x >> 61); }

// v[t] summed over the 64 lanes of the wave, on every lane.  All lanes must
// be active and v canonical.  Within each row of 16: DPP exchanges with lane
// i ^ 1, i ^ 2 (quad permutes), then 7 - i and 15 - i (half-row and row
// mirrors, which pair each lane with one of the other half once the halves
// are uniform); the four row sums are then read out of lanes 0, 16, 32, 48.
// The adds stay unreduced while they cannot wrap (8 canonical values sum
// below 2^64) with one Mersenne fold before the last row step and one before
// the cross-row sum.  No LDS round trips, unlike ds_bpermute shuffles.
template <int K>
__device__ __forceinline__ void wave_allsum(fe *v) {
    uint64_t re[K], im[K];
#pragma unroll
    for (int t = 0; t < K; ++t) { re[t] = v[t].re; im[t] = v[t].im; }
#pragma unroll
    for (int t = 0; t < K; ++t) { re[t] += dpp64<0xB1>(re[t]); im[t] += dpp64<0xB1>(im[t]); }    // quad_perm(1, 0, 3, 2)
#pragma unroll
    for (int t = 0; t < K; ++t) { re[t] += dpp64<0x4E>(re[t]); im[t] += dpp64<0x4E>(im[t]); }    // quad_perm(2, 3, 0, 1)
#pragma unroll
    for (int t = 0; t < K; ++t) { re[t] += dpp64<0x141>(re[t]); im[t] += dpp64<0x141>(im[t]); }  // row_half_mirror
#pragma unroll
    for (int t = 0; t < K; ++t) {
        re[t] = fold61(re[t]); im[t] = fold61(im[t]);
        re[t] += dpp64<0x140>(re[t]); im[t] += dpp64<0x140>(im[t]);                            // row_mirror
        re[t] = fold61(re[t]); im[t] = fold61(im[t]);
    }
#pragma unroll
    for (int t = 0; t < K; ++t) {
        const uint64_t sr = readlane64(re[t], 0) + readlane64(re[t], 16) + readlane64(re[t], 32) + readlane64(re[t], 48);
        const uint64_t si = readlane64(im[t], 0) + readlane64(im[t], 16) + readlane64(im[t], 32) + readlane64(im[t], 48);
        v[t] = gf::mk(gf::canon(fold61(sr)), gf::canon(fold61(si)));
    }
}

__device__ __forceinline__ void wave_reduce3(fe &a, fe &b, fe &c) {
    fe v[3] = {a, b, c};
    wave_allsum<3>(v);
    a = v[0]; b = v[1]; c = v[2];
}

// Block-wide sum of (A, C, O); the result is valid in wave 0.  Starts and
// ends with a barrier.
template <int NT>
__device__ __forceinline__ void block_reduce3(fe &A, fe &C, fe &O, fe (*red)[NT / 64]) {
    wave_reduce3(A, C, O);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) { red[0][wv] = A; red[1][wv] = C; red[2][wv] = O; }
    __syncthreads();
    if (wv == 0) {
        A = red[0][0]; C = red[1][0]; O = red[2][0];
        for (int w = 1; w < NT / 64; ++w) {
            A = gf::add(A, red[0][w]); C = gf::add(C, red[1][w]); O = gf::add(O, red[2][w]);
        }
    }
}

// Block-wide sums of v[0..K) written to out[0..K): one DPP sum per wave,
// one barrier, then threads 0..K-1 add the waves' sums and store.  All
// threads must call it.
template <int K, int NT>
__device__ __forceinline__ void block_partials(fe *v, fe *out, fe (*red)[K]) {
    wave_allsum<K>(v);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (lane == 0)
#pragma unroll
        for (int t = 0; t < K; ++t) red[wv][t] = v[t];
    __syncthreads();
    if (threadIdx.x < K) {
        fe a = red[0][threadIdx.x];
        for (int w = 1; w < NT / 64; ++w) a = gf::add(a, red[w][threadIdx.x]);
        out[threadIdx.x] = a;
    }
}

// Wave 0: v[0..K) = sum over b < nb_in of part[K b + t], on every lane.
// A lane loads U sets before adding any, so a sweep costs one memory latency.
template <int K>
__device__ __forceinline__ void wave_partials(const fe *part, int nb_in, fe *v) {
    constexpr int U = K > 3 ? 2 : 4;
#pragma unroll
    for (int t = 0; t < K; ++t) v[t] = gf::zero();
    for (int b0 = threadIdx.x & 63; b0 < nb_in; b0 += 64 * U) {
        fe x[U][K];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int b = b0 + 64 * u;
#pragma unroll
            for (int t = 0; t < K; ++t) x[u][t] = b < nb_in ? part[K * b + t] : gf::zero();
        }
#pragma unroll
        for (int u = 0; u < U; ++u)
#pragma unroll
            for (int t = 0; t < K; ++t) v[t] = gf::add(v[t], x[u][t]);
    }
    wave_allsum<K>(v);
}

// Round `round`'s polynomial from (A = sum d1 d2, C = sum e1 e2,
// O = sum o1 o2): check H(0)+H(1) = y, derive the challenge (transcript or
// libc stream), update (prev, y), return r.  Called by lanes 0..3 of a wave
// with the same inputs (they share the four-lane hash); when `writer`,
// lane 0 records the round in st.
__device__ fe finalize_round(ScState *st, int round, int fs_rounds, int claim_mode, fe &prev, fe &y, fe A, fe C,
                             fe O, bool writer) {
    const fe a = A, c = C;
    const fe b = gf::sub(gf::sub(O, A), C);
    const fe h01 = gf::add(gf::add(a, b), gf::add(c, c));   // H(0) + H(1)
    bool bad = false;
    if (round == 0 && claim_mode == SC_CLAIM_DERIVE) y = h01;
    else if (claim_mode != SC_CLAIM_UNCHECKED) bad = !gf::eq(h01, y);
    fe r;
    if (round < fs_rounds) {
        prev = ts::hash_F_quad(prev, a);
        prev = ts::hash_F_quad(prev, b);
        prev = ts::hash_F_quad(prev, c);
        r = prev;
    } else {
        r = st->ext_r[round];
    }
    y = gf::add(gf::mul(gf::add(gf::mul(a, r), b), r), c);
    if (writer && (threadIdx.x & 3) == 0) {
        st->polys[3 * round] = a;
        st->polys[3 * round + 1] = b;
        st->polys[3 * round + 2] = c;
        st->r[round] = r;
        st->tr_prev[round + 1] = prev;
        st->tr_y[round + 1] = y;
        if (bad && st->err == 0) st->err = round + 1;
    }
    return r;
}

// Reduce round `round`'s workgroup partials and finalize it in this
// workgroup: wave 0 sums the partials (its transcript state loads issue
// first), lanes 0..3 run the round; (*sh_r, *sh_prev, *sh_y) receive the
// challenge and the state after the round, visible to all threads on return.
__device__ void finish_round(ScState *st, int round, int fs_rounds, const fe *part_in, int nb_in, fe *sh_r,
                             fe *sh_prev, fe *sh_y, bool writer) {
    if (threadIdx.x < 64) {
        fe prev = st->tr_prev[round], y = st->tr_y[round];
        const int claim = st->derive_y;
        fe v[3];
        wave_partials<3>(part_in, nb_in, v);
        if (threadIdx.x < 4) {
            // the transcript chain is serial: let it issue ahead of co-resident waves
            __builtin_amdgcn_s_setprio(3);
            const fe r = finalize_round(st, round, fs_rounds, claim, prev, y, v[0], v[1], v[2], writer);
            if (threadIdx.x == 0) { *sh_r = r; *sh_prev = prev; *sh_y = y; }
            __builtin_amdgcn_s_setprio(0);
        }
    }
    __syncthreads();
}

// One launch's work on one proof.  A batch launch covers several proofs:
// its workgroup b works on job blk_job[b], block b - first_block of it.
struct ScJob {
    const fe *s1, *s2;        // inputs
    fe *d1, *d2;              // folded outputs (FOLD rounds; the tail's first)
    fe *e1, *e2;              // the tail's second output buffers (it alternates d, e)
    uint64_t npairs;          // pairs this round (tail: elements in s1/s2)
    ScState *st;
    const fe *part_in;        // previous round's workgroup partials
    fe *part_out;
    int nb_in, nb;            // partials in part_in; this job's workgroups
    int round, m, fs_rounds;
    uint32_t first_block;
    int kind;                 // 0: one round; 1: sc_pair01_kernel; 2, 3: sc_fold2_kernel<false>, <true>
};

// One multi-workgroup round `round`.  FOLD (round > 0): inputs hold
// 4*npairs elements, folded by r[round-1] (finalized here from part_in) into
// outputs of 2*npairs; otherwise inputs hold 2*npairs elements.  Writes this
// round's per-workgroup partials to part_out.
template <bool FOLD>
__device__ __forceinline__ void sc_round(const ScJob &J, uint32_t blk) {
    __shared__ fe red[SC_THREADS / 64][3];
    __shared__ fe sh_r, sh_prev, sh_y;
    const fe *__restrict__ s1 = J.s1;
    const fe *__restrict__ s2 = J.s2;
    fe *__restrict__ d1 = J.d1;
    fe *__restrict__ d2 = J.d2;
    const uint64_t npairs = J.npairs;
    const int round = J.round, fs_rounds = J.fs_rounds, nb_in = J.nb_in;
    ScState *st = J.st;
    const fe *part_in = J.part_in;
    fe *part_out = J.part_out;
    constexpr int NV = FOLD ? 4 : 2;
    const uint64_t stride = (uint64_t)J.nb * SC_THREADS * SC_ITEMS;
    uint64_t j0 = (uint64_t)blk * SC_THREADS * SC_ITEMS + threadIdx.x;
    fe x1[SC_ITEMS][NV], x2[SC_ITEMS][NV];
    auto load = [&](uint64_t base) {
#pragma unroll
        for (int t = 0; t < SC_ITEMS; ++t) {
            const uint64_t j = base + (uint64_t)t * SC_THREADS;
            if (j < npairs) {
#pragma unroll
                for (int q = 0; q < NV; ++q) { x1[t][q] = s1[NV * j + q]; x2[t][q] = s2[NV * j + q]; }
            }
        }
    };
    // the first sweep's loads do not depend on the challenge: issue them first
    load(j0);
    fe r = gf::zero();
    if (FOLD) {
        finish_round(st, round - 1, fs_rounds, part_in, nb_in, &sh_r, &sh_prev, &sh_y, blk == 0);
        r = sh_r;
    }
    fe A = gf::zero(), C = gf::zero(), O = gf::zero();
    while (j0 < npairs) {
#pragma unroll
        for (int t = 0; t < SC_ITEMS; ++t) {
            const uint64_t j = j0 + (uint64_t)t * SC_THREADS;
            if (j >= npairs) break;
            fe e1, o1, e2, o2;
            if (FOLD) {
                e1 = fold1(x1[t][0], x1[t][1], r);
                o1 = fold1(x1[t][2], x1[t][3], r);
                e2 = fold1(x2[t][0], x2[t][1], r);
                o2 = fold1(x2[t][2], x2[t][3], r);
                d1[2 * j] = e1; d1[2 * j + 1] = o1;
                d2[2 * j] = e2; d2[2 * j + 1] = o2;
            } else {
                e1 = x1[t][0]; o1 = x1[t][1];
                e2 = x2[t][0]; o2 = x2[t][1];
            }
            A = gf::add(A, gf::mul(gf::sub(o1, e1), gf::sub(o2, e2)));
            C = gf::add(C, gf::mul(e1, e2));
            O = gf::add(O, gf::mul(o1, o2));
        }
        j0 += stride;
        if (j0 < npairs) load(j0);
    }
    fe v[3] = {A, C, O};
    block_partials<3, SC_THREADS>(v, part_out + 3 * blk, red);
}

template <bool FOLD>
__global__ __launch_bounds__(SC_THREADS) void sc_round_kernel(ScJob J) {
    sc_round<FOLD>(J, blockIdx.x);
}

// A round's sums as quadratics in the challenge r not yet drawn: for each of
// A = sum d1 d2, C = sum e1 e2, O = sum o1 o2 the value at r = 0, the value at
// r = 1 and the r^2 coefficient (q[0..2] A, q[3..5] C, q[6..8] O).  A
// next-round pair whose folded values are e = x0 + r (x1 - x0) and
// o = x2 + r (x3 - x2) (x = the four unfolded elements, y likewise) adds
// x0 y0, x1 y1 and (x1 - x0)(y1 - y0) to C; the same for O with x2, x3; and
// for A with d(0) = x2 - x0, d(1) = x3 - x1, slope (x3 - x2) - (x1 - x0).
__device__ __forceinline__ void quad_acc(fe *q, const fe *x, const fe *y) {
    const fe qe1 = gf::sub(x[1], x[0]), qo1 = gf::sub(x[3], x[2]);
    const fe qe2 = gf::sub(y[1], y[0]), qo2 = gf::sub(y[3], y[2]);
    q[0] = gf::add(q[0], gf::mul(gf::sub(x[2], x[0]), gf::sub(y[2], y[0])));
    q[1] = gf::add(q[1], gf::mul(gf::sub(x[3], x[1]), gf::sub(y[3], y[1])));
    q[2] = gf::add(q[2], gf::mul(gf::sub(qo1, qe1), gf::sub(qo2, qe2)));
    q[3] = gf::add(q[3], gf::mul(x[0], y[0]));
    q[4] = gf::add(q[4], gf::mul(x[1], y[1]));
    q[5] = gf::add(q[5], gf::mul(qe1, qe2));
    q[6] = gf::add(q[6], gf::mul(x[2], y[2]));
    q[7] = gf::add(q[7], gf::mul(x[3], y[3]));
    q[8] = gf::add(q[8], gf::mul(qo1, qo2));
}

// f(r) = f0 + r (f1 - f0 - f2) + r^2 f2
__device__ __forceinline__ fe quad_eval(const fe &f0, const fe &f1, const fe &f2, const fe &r) {
    return gf::add(f0, gf::mul(r, gf::add(gf::sub(gf::sub(f1, f0), f2), gf::mul(r, f2))));
}

// Rounds 0 and 1 from one read of the inputs: quad_acc over each quadruple
// (x0..x3) of each vector gives round 1's three terms as quadratics in the
// challenge r0 not yet drawn, and round 0's sums are sums of those
// coefficients: A0 = q5 + q8 (the pair differences' products), C0 = q3 + q6
// (values at r0 = 0), O0 = q4 + q7 (values at r0 = 1).  J.npairs =
// quadruples; nine partials per workgroup go to part_out.
__device__ __forceinline__ void sc_pair01(const ScJob &J, uint32_t blk) {
    __shared__ fe red[SC_THREADS / 64][9];
    // per wave, the 64 quadruples of each vector it takes in one step, padded
    // one element in five so the quadruple reads spread over the banks
    __shared__ fe stage[SC_THREADS / 64][2][320];
    const fe *__restrict__ s1 = J.s1;
    const fe *__restrict__ s2 = J.s2;
    const uint64_t nq = J.npairs;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    fe *const xs = stage[wave][0], *const ys = stage[wave][1];
    fe q[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) q[i] = gf::zero();
    const uint64_t stride = (uint64_t)J.nb * SC_THREADS;
    // A wave takes 64 consecutive quadruples (256 elements of each vector) a
    // step, loaded with consecutive lanes on consecutive elements (each load
    // instruction one contiguous KiB) and regrouped through its LDS slice, so
    // lane l gets quadruple l; the next step's loads are issued first.
    uint64_t wb = (uint64_t)blk * SC_THREADS + (uint64_t)wave * 64;   // the wave's first quadruple
    fe xc[4], yc[4];
    auto load = [&](uint64_t base) {
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const uint64_t i = 4 * base + lane + 64 * e;
            xc[e] = i < 4 * nq ? s1[i] : gf::zero();
            yc[e] = i < 4 * nq ? s2[i] : gf::zero();
        }
    };
    if (wb < nq) load(wb);
    while (wb < nq) {
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int i = lane + 64 * e;
            xs[i + (i >> 2)] = xc[e];
            ys[i + (i >> 2)] = yc[e];
        }
        const uint64_t wn = wb + stride;
        if (wn < nq) load(wn);
        fe x[4], y[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            x[k] = xs[5 * lane + k];
            y[k] = ys[5 * lane + k];
        }
        if (wb + lane < nq) quad_acc(q, x, y);
        wb = wn;
    }
    block_partials<9, SC_THREADS>(q, J.part_out + 9 * blk, red);
}

__global__ __launch_bounds__(SC_THREADS) void sc_pair01_kernel(ScJob J) { sc_pair01(J, blockIdx.x); }

// a batch launch: workgroup b works on job blk_job[b], block b - first_block of it
__global__ __launch_bounds__(SC_THREADS) void sc_pair01_batch_kernel(const ScJob *__restrict__ jobs,
                                                                    const uint16_t *__restrict__ blk_job) {
    const ScJob &J = jobs[blk_job[blockIdx.x]];
    sc_pair01(J, blockIdx.x - J.first_block);
}

// Two rounds per launch.  The input v(i) (J.round = i) comes with nine
// partials per workgroup of the previous launch: round i + 1's sums as
// quadratics in r_i (quad_acc's layout), from which round i's sums follow as
// in sc_pair01_kernel.  In every workgroup wave 0 reduces them and finalizes
// rounds i and i + 1 (r_i, then r_{i+1}; workgroup 0 records them) while
// the other waves' first loads are in flight; then groups of v(i) are folded
// by r_i and r_{i+1} into v(i + 2), which is written, and on v(i + 2) the
// kernel forms
//   QUAD = false: round i + 2's sums (three partials per workgroup): a lane
//     folds an 8-tuple of each vector into a pair; J.npairs = pairs of
//     v(i + 2);
//   QUAD = true: round i + 3's sums as quadratics in r_{i+2} (nine
//     partials, the layout this kernel reads): four lanes fold a 16-tuple
//     (4 elements each) into a quadruple of v(i + 2), exchange it by DPP, and
//     lane p of the four forms the three products at r = p (p = 0, 1, and
//     p = 2 for the r^2 coefficient); J.npairs = quadruples of v(i + 2).
// A chain pair01, QUAD, ..., QUAD, direct thus covers two rounds per launch
// and ends like any multi-workgroup round: v(k) in memory, round k's sums in
// three partials per workgroup.
template <bool QUAD>
__device__ __forceinline__ void sc_fold2(const ScJob &J, uint32_t blk) {
    constexpr int K = QUAD ? 9 : 3;          // partials per workgroup written
    constexpr int NL = QUAD ? 4 : 8;         // elements of v(i) per lane and vector
    __shared__ fe red[SC_THREADS / 64][K];
    __shared__ fe sh_r[2];
    const fe *__restrict__ s1 = J.s1;
    const fe *__restrict__ s2 = J.s2;
    fe *__restrict__ d1 = J.d1;
    fe *__restrict__ d2 = J.d2;
    const uint64_t lanes = QUAD ? 4 * J.npairs : J.npairs;   // lane-items
    const uint64_t stride = (uint64_t)J.nb * SC_THREADS;     // a multiple of 4: a lane keeps its place in its quad
    const int round = J.round;
    uint64_t t = (uint64_t)blk * SC_THREADS + threadIdx.x;
    fe a[NL], b[NL];
    auto load = [&](uint64_t u) {
#pragma unroll
        for (int e = 0; e < NL; ++e) { a[e] = s1[NL * u + e]; b[e] = s2[NL * u + e]; }
    };
    if (t < lanes) load(t);
    if (threadIdx.x < 64) {
        ScState *st = J.st;
        fe prev = st->tr_prev[round], yv = st->tr_y[round];
        const int claim = st->derive_y;
        fe q[9];
        wave_partials<9>(J.part_in, J.nb_in, q);
        if (threadIdx.x < 4) {
            __builtin_amdgcn_s_setprio(3);
            const bool writer = blk == 0;
            const fe A0 = gf::add(q[5], q[8]), C0 = gf::add(q[3], q[6]), O0 = gf::add(q[4], q[7]);
            const fe r0 = finalize_round(st, round, J.fs_rounds, claim, prev, yv, A0, C0, O0, writer);
            const fe A1 = quad_eval(q[0], q[1], q[2], r0), C1 = quad_eval(q[3], q[4], q[5], r0),
                     O1 = quad_eval(q[6], q[7], q[8], r0);
            const fe r1 = finalize_round(st, round + 1, J.fs_rounds, claim, prev, yv, A1, C1, O1, writer);
            if (threadIdx.x == 0) { sh_r[0] = r0; sh_r[1] = r1; }
            __builtin_amdgcn_s_setprio(0);
        }
    }
    __syncthreads();
    const fe r0 = sh_r[0], r1 = sh_r[1];
    fe v[K];
#pragma unroll
    for (int k = 0; k < K; ++k) v[k] = gf::zero();
    if (QUAD) {
        const uint32_t p = threadIdx.x & 3;
        fe acc[3] = {gf::zero(), gf::zero(), gf::zero()};   // A, C, O at r = p
        while (t < lanes) {
            // this lane's element of the quadruple: v(i + 2)[t]
            const fe x = fold1(fold1(a[0], a[1], r0), fold1(a[2], a[3], r0), r1);
            const fe z = fold1(fold1(b[0], b[1], r0), fold1(b[2], b[3], r0), r1);
            d1[t] = x;
            d2[t] = z;
            t += stride;
            if (t < lanes) load(t);
            // e(p), o(p) of the next-round pair (x0, x1), (x2, x3): p = 0: x0, x2;
            // p = 1: x1, x3; p = 2: x1 - x0, x3 - x2 (lane 3's values are unused)
            const fe ex = dpp_fe<0xD4>(x), oxh = dpp_fe<0xFE>(x);      // quad_perm(0,1,1,3), (2,3,3,3)
            const fe ez = dpp_fe<0xD4>(z), ozh = dpp_fe<0xFE>(z);
            const fe x0 = dpp_fe<0x00>(x), z0 = dpp_fe<0x00>(z);       // quad_perm(0,0,0,0)
            const bool inf = p == 2;
            const fe e1 = inf ? gf::sub(ex, x0) : ex, o1 = inf ? gf::sub(oxh, x) : oxh;
            const fe e2 = inf ? gf::sub(ez, z0) : ez, o2 = inf ? gf::sub(ozh, z) : ozh;
            acc[0] = gf::add(acc[0], gf::mul(gf::sub(o1, e1), gf::sub(o2, e2)));
            acc[1] = gf::add(acc[1], gf::mul(e1, e2));
            acc[2] = gf::add(acc[2], gf::mul(o1, o2));
        }
        // A(p) -> v[p], C(p) -> v[3 + p], O(p) -> v[6 + p]
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            v[k] = p == (uint32_t)k ? acc[0] : gf::zero();
            v[3 + k] = p == (uint32_t)k ? acc[1] : gf::zero();
            v[6 + k] = p == (uint32_t)k ? acc[2] : gf::zero();
        }
    } else {
        while (t < lanes) {
            fe f[4], g[4];
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                f[k] = fold1(a[2 * k], a[2 * k + 1], r0);
                g[k] = fold1(b[2 * k], b[2 * k + 1], r0);
            }
            const fe e1 = fold1(f[0], f[1], r1), o1 = fold1(f[2], f[3], r1);
            const fe e2 = fold1(g[0], g[1], r1), o2 = fold1(g[2], g[3], r1);
            d1[2 * t] = e1; d1[2 * t + 1] = o1;
            d2[2 * t] = e2; d2[2 * t + 1] = o2;
            t += stride;
            if (t < lanes) load(t);
            v[0] = gf::add(v[0], gf::mul(gf::sub(o1, e1), gf::sub(o2, e2)));
            v[1] = gf::add(v[1], gf::mul(e1, e2));
            v[2] = gf::add(v[2], gf::mul(o1, o2));
        }
    }
    block_partials<K, SC_THREADS>(v, J.part_out + K * blk, red);
}

template <bool QUAD>
__global__ __launch_bounds__(SC_THREADS) void sc_fold2_kernel(ScJob J) {
    sc_fold2<QUAD>(J, blockIdx.x);
}

template <bool QUAD>
__global__ __launch_bounds__(SC_THREADS) void sc_fold2_batch_kernel(const ScJob *__restrict__ jobs,
                                                                   const uint16_t *__restrict__ blk_job) {
    const ScJob &J = jobs[blk_job[blockIdx.x]];
    sc_fold2<QUAD>(J, blockIdx.x - J.first_block);
}

template <bool FOLD>
__global__ __launch_bounds__(SC_THREADS) void sc_round_batch_kernel(const ScJob *__restrict__ jobs,
                                                                     const uint16_t *__restrict__ blk_job) {
    const ScJob &J = jobs[blk_job[blockIdx.x]];
    sc_round<FOLD>(J, blockIdx.x - J.first_block);
}

// q summed over the wave (lanes without work hold zeros; P, the power of two
// covering the lanes with work, is not needed by the all-lane sum), every
// lane ends with the sum
__device__ __forceinline__ void wave_sum9(fe *q, int P) {
    (void)P;
    wave_allsum<9>(q);
}

// The nine products of a next-round pair (quad_acc's terms) as
// A_p(x) * A_p(z), A_p(v) = (v[a] - v[b]) - (v[c] - v[d]) over the pair's four
// folded values; row holds x at 0..3, z at 4..7 and zero at 8, and off
// selects x (0) or z (4).  Index 4 in the table stands for zero.
__device__ __forceinline__ fe quad_operand(const fe *row, int off, int p) {
    constexpr int8_t T[9][4] = {{2, 0, 4, 4}, {3, 1, 4, 4}, {3, 2, 1, 0}, {0, 4, 4, 4}, {1, 4, 4, 4},
                                {1, 0, 4, 4}, {2, 4, 4, 4}, {3, 4, 4, 4}, {3, 2, 4, 4}};
    int ix[4] = {4, 4, 4, 4};
#pragma unroll
    for (int k = 0; k < 9; ++k)
        if (p == k) { ix[0] = T[k][0]; ix[1] = T[k][1]; ix[2] = T[k][2]; ix[3] = T[k][3]; }
    fe v[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) v[j] = row[ix[j] == 4 ? 8 : off + ix[j]];
    return gf::sub(gf::sub(v[0], v[1]), gf::sub(v[2], v[3]));
}

__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

constexpr int SC_SPLIT_MAX = 7;    // groups whose 8 folds and 9 products fit one wave's lanes
constexpr int SC_SPLIT_NG = 128;   // phases with at most this many groups spread them over lanes
constexpr int SC_TAIL_LDS = 512;   // folded vectors up to this many elements stay in LDS

// All rounds from `round0` to m-1 in one workgroup of four waves.  The input
// (len elements) is v(round0) when round0 == 0, else v(round0 - 1), still to
// be folded by r[round0 - 1], which is finalized here from part_in.  Phase k
// finalizes round i = round0 + k on wave 0 (lanes 0..3: sums, claim check,
// transcript hashes) while waves 1..7 fold the current vector by r[i - 1]
// into the next buffer (d, e, d, ...: never the input) and accumulate round
// i + 1's sums as quadratics in r[i]; one barrier ends each phase.  A folded
// vector of at most SC_TAIL_LDS elements stays in LDS (32 KiB for both
// slots), longer ones go through global memory (L2).  Lanes 4g..4g+3 of
// wave 0 evaluate quadratic g at the last challenge in parallel before lanes
// 0..3 hash.
//
// A lane issues 64-bit multiplies slowly when it is alone on its SIMD, so
// the compute waves keep per-lane chains short: with many groups (8
// elements folded into one next-round pair) each lane takes whole groups and
// only the lanes and waves that hold groups join the reductions; with at
// most SC_SPLIT_NG groups each wave spreads its share over its lanes, one
// fold and then one product per lane, exchanged through LDS.
__device__ __forceinline__ void sc_tail(const ScJob &J) {
    __shared__ fe qpart[2][SC_TAIL_WAVES][9];
    __shared__ int sh_nw[2];                      // compute waves that wrote qpart[s]
    __shared__ fe sh_r[2];
    __shared__ fe lfold[SC_TAIL_WAVES][SC_SPLIT_MAX][9];   // per wave: folded x (0..3), z (4..7), zero (8)
    __shared__ fe lprod[SC_TAIL_WAVES][SC_SPLIT_MAX][9];
    __shared__ fe lvec[2][2][SC_TAIL_LDS];                 // [slot][vector]: short folded vectors
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const bool hash_wave = wave == 0;
    const int ct = tid - 64;
    const int round0 = J.round, m = J.m, fs_rounds = J.fs_rounds, nb_in = J.nb_in;
    ScState *st = J.st;
    const bool pending = round0 > 0;
    const fe *c1 = J.s1, *c2 = J.s2;
    fe *const o1[2] = {J.d1, J.e1};
    fe *const o2[2] = {J.d2, J.e2};
    int ob = 0;
    uint32_t cur = (uint32_t)J.npairs;      // elements in c1 / c2
    // the fold of the current vector (cur / 2 elements) goes to slot ob: LDS
    // when it fits (the next phase reads it back at LDS latency), else the
    // global ping-pong buffers
    auto out1 = [&]() -> fe * { return cur / 2 <= (uint32_t)SC_TAIL_LDS ? lvec[ob][0] : o1[ob]; };
    auto out2 = [&]() -> fe * { return cur / 2 <= (uint32_t)SC_TAIL_LDS ? lvec[ob][1] : o2[ob]; };
    fe prev = gf::zero(), y = gf::zero(), rr = gf::zero();
    int claim_mode = 0;
    // per-thread accumulation, then the sum over the lanes and waves that hold
    // work items (ng of them, item g on compute thread g mod SC_TAIL_CT)
    auto publish = [&](fe *q, uint32_t ng, int slot) {
        const int active = (int)min<uint32_t>(64u, (uint32_t)max<int>(0, (int)ng - 64 * (wave - 1)));
        if (active > 0) {
            int P = 1;
            while (P < active) P <<= 1;
            wave_sum9(q, P);
            if (lane == 0)
#pragma unroll
                for (int t = 0; t < 9; ++t) qpart[slot][wave - 1][t] = q[t];
        }
        if (tid == 64) sh_nw[slot] = (int)min<uint32_t>((uint32_t)SC_TAIL_WAVES, (ng + 63) / 64);
    };

    // ---- before phase 0 ----
    if (hash_wave) {
        claim_mode = st->derive_y;
        const int r0 = pending ? round0 - 1 : 0;
        prev = st->tr_prev[r0];
        y = st->tr_y[r0];
        if (pending) {
            fe A = gf::zero(), C = gf::zero(), O = gf::zero();
            for (int b = lane; b < nb_in; b += 64) {
                A = gf::add(A, J.part_in[3 * b]);
                C = gf::add(C, J.part_in[3 * b + 1]);
                O = gf::add(O, J.part_in[3 * b + 2]);
            }
            wave_reduce3(A, C, O);
            if (lane < 4) {
                __builtin_amdgcn_s_setprio(3);
                rr = finalize_round(st, round0 - 1, fs_rounds, claim_mode, prev, y, A, C, O, true);
                __builtin_amdgcn_s_setprio(0);
            }
            rr = readlane_fe(rr, 0);      // the wave's lanes 0..11 evaluate with it
        }
        if (lane == 0) sh_r[0] = rr;
    } else {
        fe q[9];
#pragma unroll
        for (int t = 0; t < 9; ++t) q[t] = gf::zero();
        uint32_t ng;
        if (pending) {
            // round round0's sums in r[round0 - 1], from groups of 4 of the input
            ng = cur / 4;
            for (uint32_t g = ct; g < ng; g += SC_TAIL_CT) {
                fe x[4], z[4];
#pragma unroll
                for (int e = 0; e < 4; ++e) { x[e] = c1[4 * g + e]; z[e] = c2[4 * g + e]; }
                quad_acc(q, x, z);
            }
        } else {
            // round 0's sums directly (constant terms)
            ng = cur / 2;
            for (uint32_t j = ct; j < ng; j += SC_TAIL_CT) {
                const fe e1 = c1[2 * j], od1 = c1[2 * j + 1], e2 = c2[2 * j], od2 = c2[2 * j + 1];
                q[0] = gf::add(q[0], gf::mul(gf::sub(od1, e1), gf::sub(od2, e2)));
                q[3] = gf::add(q[3], gf::mul(e1, e2));
                q[6] = gf::add(q[6], gf::mul(od1, od2));
            }
        }
        publish(q, ng, 0);
    }
    __syncthreads();

    // ---- phases ----
    for (int k = 0; round0 + k < m; ++k) {
        const int i = round0 + k, s = k & 1;
        const bool fold = pending || k > 0;    // the current vector is v(i - 1)
        if (hash_wave) {
            // lanes 4g..4g+3 sum quadratic g (A, C, O) over the compute waves
            // and evaluate it at the last challenge, in parallel; lanes 0..3
            // then run the round
            __builtin_amdgcn_s_setprio(3);
            fe E = gf::zero();
            if (lane < 12) {
                const int g = lane >> 2, nw = sh_nw[s];
                fe f0 = qpart[s][0][3 * g], f1 = qpart[s][0][3 * g + 1], f2 = qpart[s][0][3 * g + 2];
                for (int w = 1; w < nw; ++w) {
                    f0 = gf::add(f0, qpart[s][w][3 * g]);
                    f1 = gf::add(f1, qpart[s][w][3 * g + 1]);
                    f2 = gf::add(f2, qpart[s][w][3 * g + 2]);
                }
                E = fold ? quad_eval(f0, f1, f2, rr) : f0;
            }
            const fe A = readlane_fe(E, 0), C = readlane_fe(E, 4), O = readlane_fe(E, 8);
            if (lane < 4) {
                rr = finalize_round(st, i, fs_rounds, claim_mode, prev, y, A, C, O, true);
                if (lane == 0) sh_r[s ^ 1] = rr;
            }
            rr = readlane_fe(rr, 0);
            __builtin_amdgcn_s_setprio(0);
        } else {
            const bool next = i + 1 < m;       // round i + 1 exists: accumulate its sums
            fe *w1 = out1(), *w2 = out2();
            if (fold && next && cur / 8 <= (uint32_t)SC_SPLIT_NG) {
                // few groups: compute wave cw takes groups cw, cw + 7, ..., seven at a
                // time: one fold per lane, then one of the nine products per lane,
                // through the wave's own LDS rows; lanes 0..8 keep the sums
                const uint32_t ng = cur / 8;
                const int cw = wave - 1;
                const uint32_t mine =
                    ng > (uint32_t)cw ? (ng - cw + SC_TAIL_WAVES - 1) / SC_TAIL_WAVES : 0;
                const fe r = sh_r[s];
                fe acc = gf::zero();
                for (uint32_t j0 = 0; j0 < mine; j0 += SC_SPLIT_MAX) {
                    const int nc = (int)min<uint32_t>((uint32_t)SC_SPLIT_MAX, mine - j0);
                    if (lane < 8 * nc) {
                        const int gi = lane >> 3, f = lane & 7, e = f & 3;
                        const uint32_t g = cw + SC_TAIL_WAVES * (j0 + gi);
                        const fe *src = f < 4 ? c1 : c2;
                        const fe x = fold1(src[8 * g + 2 * e], src[8 * g + 2 * e + 1], r);
                        (f < 4 ? w1 : w2)[4 * g + e] = x;
                        lfold[cw][gi][f] = x;
                    }
                    if (lane < nc) lfold[cw][lane][8] = gf::zero();
                    wave_lds_sync();
                    if (lane < 9 * nc) {
                        const int gi = lane / 9, p = lane % 9;
                        lprod[cw][gi][p] = gf::mul(quad_operand(lfold[cw][gi], 0, p), quad_operand(lfold[cw][gi], 4, p));
                    }
                    wave_lds_sync();
                    if (lane < 9)
                        for (int gi = 0; gi < nc; ++gi) acc = gf::add(acc, lprod[cw][gi][lane]);
                    wave_lds_sync();     // the rows are rewritten by the next chunk
                }
                if (mine > 0 && lane < 9) qpart[s ^ 1][cw][lane] = acc;
                if (tid == 64) sh_nw[s ^ 1] = (int)min<uint32_t>((uint32_t)SC_TAIL_WAVES, ng);
            } else {
                fe q[9];
#pragma unroll
                for (int t = 0; t < 9; ++t) q[t] = gf::zero();
                uint32_t ng = 0;
                if (fold) {
                    // v(i) = fold(v(i - 1), r[i - 1]); groups of 8 -> 4 (the last phase: 4 -> 2)
                    const fe r = sh_r[s];
                    const uint32_t gs = cur >= 8 ? 8 : cur;
                    ng = cur / gs;
                    for (uint32_t g = ct; g < ng; g += SC_TAIL_CT) {
                        fe x[4], z[4];
#pragma unroll
                        for (int e = 0; e < 4; ++e) {
                            if (2 * e < (int)gs) {
                                x[e] = fold1(c1[gs * g + 2 * e], c1[gs * g + 2 * e + 1], r);
                                z[e] = fold1(c2[gs * g + 2 * e], c2[gs * g + 2 * e + 1], r);
                                w1[(gs / 2) * g + e] = x[e];
                                w2[(gs / 2) * g + e] = z[e];
                            }
                        }
                        if (next) quad_acc(q, x, z);
                    }
                } else if (next) {
                    // round 0 was direct: round 1's sums from groups of 4 of v(0)
                    ng = cur / 4;
                    for (uint32_t g = ct; g < ng; g += SC_TAIL_CT) {
                        fe x[4], z[4];
#pragma unroll
                        for (int e = 0; e < 4; ++e) { x[e] = c1[4 * g + e]; z[e] = c2[4 * g + e]; }
                        quad_acc(q, x, z);
                    }
                }
                if (next) publish(q, ng, s ^ 1);
            }
        }
        if (fold) {
            c1 = out1();
            c2 = out2();
            ob ^= 1;
            cur /= 2;
        }
        __syncthreads();
    }
    // v(m - 1) (two elements) is in c1 / c2; r[m - 1] in sh_r
    if (tid == 0) {
        const fe r = sh_r[(m - round0) & 1];
        st->final_v[0] = fold1(c1[0], c1[1], r);
        st->final_v[1] = fold1(c2[0], c2[1], r);
    }
}

__global__ __launch_bounds__(SC_TAIL_THREADS) void sc_tail_kernel(ScJob J) { sc_tail(J); }

// one workgroup per proof
__global__ __launch_bounds__(SC_TAIL_THREADS) void sc_tail_batch_kernel(const ScJob *__restrict__ jobs) {
    sc_tail(jobs[blockIdx.x]);
}

// Finalize round `round` from part_in, then fold both vectors by its
// challenge: d[j] = fold(s[2j], s[2j+1]) for j < nout (no evaluation).
__global__ __launch_bounds__(SC_THREADS) void sc_fold_kernel(
    const fe *__restrict__ s1, const fe *__restrict__ s2, fe *__restrict__ d1, fe *__restrict__ d2, uint64_t nout,
    int round, int fs_rounds, ScState *st, const fe *part_in, int nb_in) {
    __shared__ fe sh_r, sh_prev, sh_y;
    finish_round(st, round, fs_rounds, part_in, nb_in, &sh_r, &sh_prev, &sh_y, blockIdx.x == 0);
    const fe r = sh_r;
    for (uint64_t j = (uint64_t)blockIdx.x * SC_THREADS + threadIdx.x; j < nout; j += (uint64_t)gridDim.x * SC_THREADS) {
        const fe a0 = s1[2 * j], a1 = s1[2 * j + 1], b0 = s2[2 * j], b1 = s2[2 * j + 1];
        d1[j] = fold1(a0, a1, r);
        d2[j] = fold1(b0, b1, r);
    }
}

// Two rounds whose challenges are known in advance (round >= fs_rounds: r =
// ext_r, drawn by the host), fused with both folds: per quadruple j of each
// vector, round `round`'s terms on its pairs (x0, x1) and (x2, x3), the fold
// by r0 to (e, o), round + 1's terms on (e, o), and d[j] = fold(e, o, r1).
// One read of 4 nq elements and one write of nq per vector instead of two
// reads, a temporary and a fold.  Partials of the two rounds go to part0 and
// part1; sc_finish2_kernel finalizes both.
__global__ __launch_bounds__(SC_THREADS) void sc_round2_kernel(const fe *__restrict__ s1, const fe *__restrict__ s2,
                                                               fe *__restrict__ d1, fe *__restrict__ d2, uint64_t nq,
                                                               const ScState *__restrict__ st, int round,
                                                               fe *__restrict__ part0, fe *__restrict__ part1) {
    __shared__ fe red[3][SC_THREADS / 64];
    const fe r0 = st->ext_r[round], r1 = st->ext_r[round + 1];
    fe A0 = gf::zero(), C0 = gf::zero(), O0 = gf::zero(), A1 = gf::zero(), C1 = gf::zero(), O1 = gf::zero();
    const uint64_t stride = (uint64_t)gridDim.x * SC_THREADS;
    uint64_t j = (uint64_t)blockIdx.x * SC_THREADS + threadIdx.x;
    fe x[4], y[4];
    auto load = [&](uint64_t q) {
#pragma unroll
        for (int i = 0; i < 4; ++i) { x[i] = s1[4 * q + i]; y[i] = s2[4 * q + i]; }
    };
    if (j < nq) load(j);
    while (j < nq) {
        const fe x0 = x[0], x1 = x[1], x2 = x[2], x3 = x[3], y0 = y[0], y1 = y[1], y2 = y[2], y3 = y[3];
        const uint64_t jn = j + stride;
        if (jn < nq) load(jn);                  // the next quadruple's loads ahead of this one's arithmetic
        A0 = gf::add(A0, gf::add(gf::mul(gf::sub(x1, x0), gf::sub(y1, y0)), gf::mul(gf::sub(x3, x2), gf::sub(y3, y2))));
        C0 = gf::add(C0, gf::add(gf::mul(x0, y0), gf::mul(x2, y2)));
        O0 = gf::add(O0, gf::add(gf::mul(x1, y1), gf::mul(x3, y3)));
        const fe e1 = fold1(x0, x1, r0), o1 = fold1(x2, x3, r0);
        const fe e2 = fold1(y0, y1, r0), o2 = fold1(y2, y3, r0);
        A1 = gf::add(A1, gf::mul(gf::sub(o1, e1), gf::sub(o2, e2)));
        C1 = gf::add(C1, gf::mul(e1, e2));
        O1 = gf::add(O1, gf::mul(o1, o2));
        d1[j] = fold1(e1, o1, r1);
        d2[j] = fold1(e2, o2, r1);
        j = jn;
    }
    block_reduce3<SC_THREADS>(A0, C0, O0, red);
    if (threadIdx.x == 0) {
        part0[3 * blockIdx.x] = A0;
        part0[3 * blockIdx.x + 1] = C0;
        part0[3 * blockIdx.x + 2] = O0;
    }
    block_reduce3<SC_THREADS>(A1, C1, O1, red);
    if (threadIdx.x == 0) {
        part1[3 * blockIdx.x] = A1;
        part1[3 * blockIdx.x + 1] = C1;
        part1[3 * blockIdx.x + 2] = O1;
    }
}

// Finalize rounds `round` and round + 1 of sc_round2_kernel: both rounds'
// partials are loaded and reduced together, then the two rounds are
// finalized in order on the same lanes (the transcript state passes from one
// to the next in registers).
__global__ __launch_bounds__(SC_THREADS) void sc_finish2_kernel(ScState *st, int round, int fs_rounds,
                                                                const fe *part0, const fe *part1, int nb) {
    __shared__ fe red[3][SC_THREADS / 64];
    fe A0 = gf::zero(), C0 = gf::zero(), O0 = gf::zero(), A1 = gf::zero(), C1 = gf::zero(), O1 = gf::zero();
    for (int b = threadIdx.x; b < nb; b += SC_THREADS) {
        A0 = gf::add(A0, part0[3 * b]);
        C0 = gf::add(C0, part0[3 * b + 1]);
        O0 = gf::add(O0, part0[3 * b + 2]);
        A1 = gf::add(A1, part1[3 * b]);
        C1 = gf::add(C1, part1[3 * b + 1]);
        O1 = gf::add(O1, part1[3 * b + 2]);
    }
    block_reduce3<SC_THREADS>(A0, C0, O0, red);
    block_reduce3<SC_THREADS>(A1, C1, O1, red);
    if (threadIdx.x < 4) {
        fe prev = st->tr_prev[round], y = st->tr_y[round];
        finalize_round(st, round, fs_rounds, st->derive_y, prev, y, A0, C0, O0, true);
        finalize_round(st, round + 1, fs_rounds, st->derive_y, prev, y, A1, C1, O1, true);
    }
}

__global__ void sc_init_kernel(ScState *st, fe prev, fe y, int derive_y) {
    st->tr_prev[0] = prev;
    st->tr_y[0] = y;
    st->derive_y = derive_y;
    st->err = 0;
}

// proof i's state (tail[i].st): transcript from prev[i], claim derived
// from round 0
__global__ void sc_init_batch_kernel(const ScJob *tail, const fe *prev, int count) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    ScState *st = tail[i].st;
    st->tr_prev[0] = prev[i];
    st->tr_y[0] = gf::zero();
    st->derive_y = SC_CLAIM_DERIVE;
    st->err = 0;
}

}  // namespace

int sumcheck_init(cosnark_ctx *ctx) {
    (void)ctx;
    return COSNARK_OK;
}

// The state (and, when root_src is set, a 32-byte root into st->root) copied
// into host memory by vector stores: one launch in place of a device-to-
// device and a device-to-host copy at the end of a proof.
__global__ __launch_bounds__(256) void sc_export_kernel(const ScState *__restrict__ st,
                                                        const uint8_t *__restrict__ root_src, ScState *host) {
    constexpr int W = (int)(sizeof(ScState) / 16);
    const uint4 *src = reinterpret_cast<const uint4 *>(st);
    uint4 *dst = reinterpret_cast<uint4 *>(host);
    const int root_w = (int)(offsetof(ScState, root) / 16);
    for (int i = threadIdx.x; i < W; i += 256) {
        if (root_src && (i == root_w || i == root_w + 1))
            dst[i] = reinterpret_cast<const uint4 *>(root_src)[i - root_w];
        else
            dst[i] = src[i];
    }
}
static_assert(sizeof(ScState) % 16 == 0 && offsetof(ScState, root) % 16 == 0, "ScState: 16-byte words");

int sumcheck_export(cosnark_ctx *ctx, const uint8_t *root_src, hipStream_t stream) {
    hipLaunchKernelGGL(sc_export_kernel, dim3(1), dim3(256), 0, stream, (const ScState *)ctx->sc, root_src,
                       ctx->sc_host_dev);
    LAUNCHCK();
    return COSNARK_OK;
}

int sumcheck_state_init(cosnark_ctx *ctx, const fe &prev, const fe &y, int claim_mode, hipStream_t stream) {
    hipLaunchKernelGGL(sc_init_kernel, dim3(1), dim3(1), 0, stream, ctx->sc, prev, y, claim_mode);
    LAUNCHCK();
    return COSNARK_OK;
}

int sumcheck_round_launch(cosnark_ctx *ctx, const fe *s1, const fe *s2, fe *d1, fe *d2, uint64_t npairs, int round,
                          bool fold, int fs_rounds, int nb_prev, hipStream_t stream, int *nb_out) {
    fe *part[2] = {ctx->sc_partials, ctx->sc_partials + 3 * SC_MAX_BLOCKS};
    uint64_t blocks = (npairs + SC_THREADS * SC_ITEMS - 1) / (SC_THREADS * SC_ITEMS);
    if (blocks > (uint64_t)ctx->sc_blocks) blocks = ctx->sc_blocks;
    if (blocks == 0) blocks = 1;
    ScJob J{};
    J.s1 = s1; J.s2 = s2;
    J.d1 = fold ? d1 : nullptr; J.d2 = fold ? d2 : nullptr;
    J.npairs = npairs;
    J.st = ctx->sc;
    J.part_in = part[(round + 1) & 1];
    J.part_out = part[round & 1];
    J.nb_in = fold ? nb_prev : 0;
    J.nb = (int)blocks;
    J.round = round;
    J.fs_rounds = fs_rounds;
    if (fold)
        hipLaunchKernelGGL(sc_round_kernel<true>, dim3((unsigned)blocks), dim3(SC_THREADS), 0, stream, J);
    else
        hipLaunchKernelGGL(sc_round_kernel<false>, dim3((unsigned)blocks), dim3(SC_THREADS), 0, stream, J);
    LAUNCHCK();
    *nb_out = (int)blocks;
    return COSNARK_OK;
}

int sumcheck_round2_launch(cosnark_ctx *ctx, const fe *s1, const fe *s2, fe *d1, fe *d2, uint64_t nq, int round,
                           hipStream_t stream) {
    fe *part[2] = {ctx->sc_partials, ctx->sc_partials + 3 * SC_MAX_BLOCKS};
    uint64_t blocks = (nq + SC_THREADS - 1) / SC_THREADS;
    if (blocks > 1024) blocks = 1024;
    if (blocks == 0) blocks = 1;
    hipLaunchKernelGGL(sc_round2_kernel, dim3((unsigned)blocks), dim3(SC_THREADS), 0, stream, s1, s2, d1, d2, nq,
                       (const ScState *)ctx->sc, round, part[0], part[1]);
    LAUNCHCK();
    hipLaunchKernelGGL(sc_finish2_kernel, dim3(1), dim3(SC_THREADS), 0, stream, ctx->sc, round, 0, (const fe *)part[0],
                       (const fe *)part[1], (int)blocks);
    LAUNCHCK();
    return COSNARK_OK;
}

int sumcheck_fold_launch(cosnark_ctx *ctx, const fe *s1, const fe *s2, fe *d1, fe *d2, uint64_t nout, int round,
                         int fs_rounds, int nb_prev, hipStream_t stream) {
    fe *part[2] = {ctx->sc_partials, ctx->sc_partials + 3 * SC_MAX_BLOCKS};
    uint64_t blocks = (nout + SC_THREADS - 1) / SC_THREADS;
    if (blocks > (uint64_t)ctx->sc_blocks) blocks = ctx->sc_blocks;
    hipLaunchKernelGGL(sc_fold_kernel, dim3((unsigned)blocks), dim3(SC_THREADS), 0, stream, s1, s2, d1, d2, nout,
                       round, fs_rounds, ctx->sc, (const fe *)part[round & 1], nb_prev);
    LAUNCHCK();
    return COSNARK_OK;
}

namespace {

// The launches of one proof's sumcheck as sumcheck_launch issues them: the
// multi-workgroup rounds while a vector exceeds tail_elems, then the tail.
// buf: n + n/2 elements (ping n/2 per vector, pong n/4 per vector); part:
// two slots of 3 * SC_MAX_BLOCKS partials.  The tail folds into the ping-pong
// slot not holding its input first (a folded vector of m elements lands in a
// slot of at least m), then alternates; it never writes v1 / v2.
std::vector<ScJob> sumcheck_jobs(const fe *v1, const fe *v2, uint64_t n, int fs_rounds, fe *buf, fe *part0,
                                 ScState *st, int max_blocks, uint64_t tail_elems, int wide_blocks,
                                 bool fuse01 = false) {
    const int m = ilog2(n);
    fe *a1 = buf, *a2 = buf + n / 2, *b1 = buf + n, *b2 = buf + n + n / 4;
    fe *part[2] = {part0, part0 + 3 * SC_MAX_BLOCKS};
    fe *outs[2][2] = {{a1, a2}, {b1, b2}};
    std::vector<ScJob> jobs;
    const fe *s1 = v1, *s2 = v2;
    uint64_t len = n;        // elements in s1/s2
    int round = 0, which = 0, nb_prev = 0;
    bool pending = false;    // s1/s2 still need the fold with r[round-1]
    const fe *pin = part[1];  // partials the next launch reads
    // rounds 0, 1, ... two per launch while the vectors are long: sc_pair01
    // (rounds 0 and 1 as nine partials), then sc_fold2 launches, each
    // finalizing two rounds and quartering the vectors; the last one (its
    // output at most tail_elems) leaves round k's sums in three partials
    if (fuse01 && n / 2 > tail_elems) {
        // these launches stream the long vectors: as many workgroups as
        // keep every SIMD's loads in flight (wide_blocks), not the later
        // rounds' count, whose partials every workgroup re-reads
        auto width = [&](uint64_t items) {
            uint64_t b = (items + SC_THREADS - 1) / SC_THREADS;
            if (b > (uint64_t)wide_blocks) b = wide_blocks;
            return (int)(b ? b : 1);
        };
        ScJob A{};
        A.kind = 1;
        A.s1 = v1; A.s2 = v2;
        A.npairs = n / 4;
        A.st = st;
        A.part_out = part[1];
        // nine partials per workgroup: a slot of 3 * SC_MAX_BLOCKS holds
        // at most SC_MAX_BLOCKS / 3 workgroups' worth
        A.nb = std::min(width(n / 4), SC_MAX_BLOCKS / 3);
        A.round = 0; A.m = m; A.fs_rounds = fs_rounds;
        jobs.push_back(A);
        int nb_in = A.nb, w = 1, i = 0;
        uint64_t L = n;     // elements of v(i)
        for (;;) {
            const bool quad = L / 4 > tail_elems;
            ScJob B{};
            B.kind = quad ? 3 : 2;
            B.s1 = s1; B.s2 = s2;
            B.d1 = outs[w][0]; B.d2 = outs[w][1];
            B.npairs = quad ? L / 16 : L / 8;
            B.st = st;
            B.part_in = pin;
            B.nb_in = nb_in;
            B.part_out = pin == part[0] ? part[1] : part[0];
            // lanes: four per quadruple of v(i + 2) (QUAD), one per pair (direct)
            B.nb = quad ? std::min(width(L / 4), SC_MAX_BLOCKS / 3) : width(L / 8);
            B.round = i; B.m = m; B.fs_rounds = fs_rounds;
            jobs.push_back(B);
            s1 = B.d1; s2 = B.d2;
            pin = B.part_out;
            nb_in = B.nb;
            w ^= 1;
            L /= 4;
            i += 2;
            if (!quad) break;
        }
        // v(i) is in s1 / s2, round i's sums in pin: the next launch
        // finalizes round i and folds by r_i
        len = L;
        round = i + 1;
        which = w;
        nb_prev = nb_in;
        pending = true;
    }
    while (len > tail_elems) {
        const uint64_t after = pending ? len / 2 : len;
        const uint64_t npairs = after / 2;
        uint64_t blocks = (npairs + SC_THREADS * SC_ITEMS - 1) / (SC_THREADS * SC_ITEMS);
        if (blocks > (uint64_t)max_blocks) blocks = max_blocks;
        ScJob J{};
        J.s1 = s1; J.s2 = s2;
        J.npairs = npairs;
        J.st = st;
        J.part_in = pin;
        J.part_out = pin == part[0] ? part[1] : part[0];
        pin = J.part_out;
        J.nb_in = pending ? nb_prev : 0;
        J.nb = (int)blocks;
        J.round = round;
        J.m = m;
        J.fs_rounds = fs_rounds;
        if (pending) {
            J.d1 = outs[which][0]; J.d2 = outs[which][1];
            s1 = J.d1; s2 = J.d2;
            which ^= 1;
        }
        jobs.push_back(J);
        nb_prev = (int)blocks;
        len = after;
        pending = true;
        ++round;
    }
    ScJob T{};
    T.s1 = s1; T.s2 = s2;
    T.d1 = outs[which][0]; T.d2 = outs[which][1];
    T.e1 = outs[which ^ 1][0]; T.e2 = outs[which ^ 1][1];
    T.npairs = len;
    T.st = st;
    T.part_in = pin;
    T.nb_in = nb_prev;
    T.round = round;
    T.m = m;
    T.fs_rounds = fs_rounds;
    jobs.push_back(T);
    return jobs;
}

}  // namespace

int sumcheck_launch(cosnark_ctx *ctx, const fe *v1, const fe *v2, uint64_t n, int fs_rounds, fe *buf,
                    hipStream_t stream, bool fuse01) {
    // beside the commit (stream2) the long-vector launches take fewer workgroups
    const int wide = stream == ctx->stream2 ? ctx->sc_wide_shared : ctx->sc_wide_blocks;
    const auto jobs = sumcheck_jobs(v1, v2, n, fs_rounds, buf, ctx->sc_partials, ctx->sc, ctx->sc_blocks,
                                    ctx->sc_tail_elems, wide, fuse01);
    for (size_t i = 0; i + 1 < jobs.size(); ++i) {
        const ScJob &J = jobs[i];
        if (J.kind == 1)
            hipLaunchKernelGGL(sc_pair01_kernel, dim3((unsigned)J.nb), dim3(SC_THREADS), 0, stream, J);
        else if (J.kind == 2)
            hipLaunchKernelGGL(sc_fold2_kernel<false>, dim3((unsigned)J.nb), dim3(SC_THREADS), 0, stream, J);
        else if (J.kind == 3)
            hipLaunchKernelGGL(sc_fold2_kernel<true>, dim3((unsigned)J.nb), dim3(SC_THREADS), 0, stream, J);
        else if (J.round > 0)
            hipLaunchKernelGGL(sc_round_kernel<true>, dim3((unsigned)J.nb), dim3(SC_THREADS), 0, stream, J);
        else
            hipLaunchKernelGGL(sc_round_kernel<false>, dim3((unsigned)J.nb), dim3(SC_THREADS), 0, stream, J);
        LAUNCHCK();
    }
    hipLaunchKernelGGL(sc_tail_kernel, dim3(1), dim3(SC_TAIL_THREADS), 0, stream, jobs.back());
    LAUNCHCK();
    return COSNARK_OK;
}

// ---- batches of proofs ------------------------------------------------------

int sumcheck_batch_plan(cosnark_ctx *ctx, int count, const fe *const *v1, const fe *const *v2, const uint64_t *n,
                        fe *const *buf, fe *const *parts, ScState *const *st, const fe *prev, TableBuf &tb,
                        ScBatchPlan &plan) {
    std::vector<std::vector<ScJob>> per(count);
    size_t nr = 0;
    for (int i = 0; i < count; ++i) {
        const int m = ilog2(n[i]);
        per[i] = sumcheck_jobs(v1[i], v2[i], n[i], m, buf[i], parts[i], st[i], ctx->sc_batch_blocks,
                               ctx->sc_tail_elems, ctx->sc_batch_blocks, true);
        nr = std::max(nr, per[i].size() - 1);
    }
    plan.rounds.clear();
    plan.count = count;
    plan.prev = prev;
    // step r holds every proof's r-th launch; proofs at one step may need
    // different kernels (a first-rounds launch, a two-round launch of either
    // form, a single round), so a step is one launch per kernel present
    for (size_t r = 0; r < nr; ++r) {
        for (int kind = 0; kind < 4; ++kind)
            for (int fold = 0; fold < (kind == 0 ? 2 : 1); ++fold) {
                std::vector<ScJob> jobs;
                std::vector<uint16_t> map;
                for (int i = 0; i < count; ++i) {
                    if (r + 1 >= per[i].size()) continue;
                    ScJob J = per[i][r];
                    if (J.kind != kind || (kind == 0 && (J.round > 0) != (fold == 1))) continue;
                    J.first_block = (uint32_t)map.size();
                    for (int b = 0; b < J.nb; ++b) map.push_back((uint16_t)jobs.size());
                    jobs.push_back(J);
                }
                if (jobs.empty()) continue;
                if (jobs.size() > 0xFFFF) return set_err(ctx, COSNARK_E_RANGE, "sumcheck batch: too many proofs");
                ScBatchPlan::Launch L{kind, fold == 1, (uint32_t)map.size(), tb.reserve(sizeof(ScJob) * jobs.size()), 0};
                L.blk_job = tb.reserve(sizeof(uint16_t) * map.size());
                memcpy(tb.at<ScJob>(L.jobs), jobs.data(), sizeof(ScJob) * jobs.size());
                memcpy(tb.at<uint16_t>(L.blk_job), map.data(), sizeof(uint16_t) * map.size());
                plan.rounds.push_back(L);
            }
    }
    plan.tail = tb.reserve(sizeof(ScJob) * count);
    for (int i = 0; i < count; ++i) tb.at<ScJob>(plan.tail)[i] = per[i].back();
    return COSNARK_OK;
}

int sumcheck_batch_run(cosnark_ctx *ctx, const ScBatchPlan &plan, const uint8_t *tables, hipStream_t stream) {
    const ScJob *tail = (const ScJob *)(tables + plan.tail);
    hipLaunchKernelGGL(sc_init_batch_kernel, dim3((plan.count + 255) / 256), dim3(256), 0, stream, tail, plan.prev,
                       plan.count);
    LAUNCHCK();
    for (const auto &L : plan.rounds) {
        const ScJob *jobs = (const ScJob *)(tables + L.jobs);
        const uint16_t *map = (const uint16_t *)(tables + L.blk_job);
        switch (L.kind) {
            case 1: hipLaunchKernelGGL(sc_pair01_batch_kernel, dim3(L.blocks), dim3(SC_THREADS), 0, stream, jobs, map); break;
            case 2: hipLaunchKernelGGL(sc_fold2_batch_kernel<false>, dim3(L.blocks), dim3(SC_THREADS), 0, stream, jobs, map); break;
            case 3: hipLaunchKernelGGL(sc_fold2_batch_kernel<true>, dim3(L.blocks), dim3(SC_THREADS), 0, stream, jobs, map); break;
            default:
                if (L.fold)
                    hipLaunchKernelGGL(sc_round_batch_kernel<true>, dim3(L.blocks), dim3(SC_THREADS), 0, stream, jobs, map);
                else
                    hipLaunchKernelGGL(sc_round_batch_kernel<false>, dim3(L.blocks), dim3(SC_THREADS), 0, stream, jobs, map);
        }
        LAUNCHCK();
    }
    hipLaunchKernelGGL(sc_tail_batch_kernel, dim3(plan.count), dim3(SC_TAIL_THREADS), 0, stream, tail);
    LAUNCHCK();
    return COSNARK_OK;
}
