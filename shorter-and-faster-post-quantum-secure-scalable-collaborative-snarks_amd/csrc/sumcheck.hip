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
// pair.  Round i's evaluation is fused with the fold by r[i-1]: one pass
// reads four elements per vector, writes the two folded ones and sums the
// round's three terms.  Each workgroup writes its partial sums; the next
// round's kernel starts by reducing the previous round's partials in every
// workgroup (a few KiB from L2) and running the transcript hashes
// redundantly, on four lanes, so no atomics, no last-workgroup election and
// no host round trip sit between rounds.  Transcript values are kept per
// round (tr_prev[k], tr_y[k] = state before round k) so a kernel never reads
// a slot it writes.  Once a vector is small (sc_tail_elems) one workgroup
// runs all remaining rounds: its wave 0 runs the transcript while waves 1..3
// fold by the challenge just drawn and form the next round's sums as
// quadratics in the challenge still being hashed, so a round costs the hash
// chain and little else.
#include <algorithm>
#include <cstddef>
#include <cstring>
#include <vector>
#include "ctx.h"
#include "transcript.h"

using gf::fe;

namespace {

constexpr int SC_THREADS = 256;
constexpr int SC_TAIL_THREADS = 256;    // wave 0: transcript; waves 1..3: folds and sums
constexpr int SC_TAIL_CT = SC_TAIL_THREADS - 64;
constexpr int SC_TAIL_WAVES = SC_TAIL_CT / 64;
constexpr int SC_ITEMS = 1;             // pairs per thread per sweep in the multi-workgroup rounds

__device__ __forceinline__ fe fold1(const fe &e, const fe &o, const fe &r) {
    return gf::add(gf::mul(r, gf::sub(o, e)), e);
}

__device__ __forceinline__ fe shfl_xor_fe(const fe &v, int m) {
    return gf::mk(__shfl_xor(v.re, m), __shfl_xor(v.im, m));
}

__device__ __forceinline__ void wave_reduce3(fe &a, fe &b, fe &c) {
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) {
        a = gf::add(a, shfl_xor_fe(a, m));
        b = gf::add(b, shfl_xor_fe(b, m));
        c = gf::add(c, shfl_xor_fe(c, m));
    }
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
// workgroup: (*sh_r, *sh_prev, *sh_y) receive the challenge and the state
// after the round, visible to all threads on return.
template <int NT>
__device__ void finish_round(ScState *st, int round, int fs_rounds, const fe *part_in, int nb_in, fe (*red)[NT / 64],
                             fe *sh_r, fe *sh_prev, fe *sh_y, bool writer) {
    fe A = gf::zero(), C = gf::zero(), O = gf::zero();
    for (int b = threadIdx.x; b < nb_in; b += NT) {
        A = gf::add(A, part_in[3 * b]);
        C = gf::add(C, part_in[3 * b + 1]);
        O = gf::add(O, part_in[3 * b + 2]);
    }
    block_reduce3<NT>(A, C, O, red);
    if (threadIdx.x < 4) {
        // the transcript chain is serial: let it issue ahead of co-resident waves
        __builtin_amdgcn_s_setprio(3);
        fe prev = st->tr_prev[round], y = st->tr_y[round];
        const fe r = finalize_round(st, round, fs_rounds, st->derive_y, prev, y, A, C, O, writer);
        if (threadIdx.x == 0) { *sh_r = r; *sh_prev = prev; *sh_y = y; }
        __builtin_amdgcn_s_setprio(0);
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
    int kind;                 // 0: one round; 1, 2: rounds 0-2 fused (sc_pair01_kernel, sc_fold01_kernel)
};

// One multi-workgroup round `round`.  FOLD (round > 0): inputs hold
// 4*npairs elements, folded by r[round-1] (finalized here from part_in) into
// outputs of 2*npairs; otherwise inputs hold 2*npairs elements.  Writes this
// round's per-workgroup partials to part_out.
template <bool FOLD>
__device__ __forceinline__ void sc_round(const ScJob &J, uint32_t blk) {
    __shared__ fe red[3][SC_THREADS / 64];
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
        finish_round<SC_THREADS>(st, round - 1, fs_rounds, part_in, nb_in, red, &sh_r, &sh_prev, &sh_y, blk == 0);
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
    block_reduce3<SC_THREADS>(A, C, O, red);
    if (threadIdx.x == 0) {
        part_out[3 * blk] = A;
        part_out[3 * blk + 1] = C;
        part_out[3 * blk + 2] = O;
    }
}

template <bool FOLD>
__global__ __launch_bounds__(SC_THREADS) void sc_round_kernel(ScJob J) {
    sc_round<FOLD>(J, blockIdx.x);
}

// Rounds 0 and 1 from one read of the inputs.  For a quadruple (x0..x3) of
// each vector, round 1 works on e = x0 + r0 (x1 - x0), o = x2 + r0 (x3 - x2)
// with r0 still unknown, so each of its three terms is a quadratic in r0,
// kept as its values at r0 = 0 and r0 = 1 and its r0^2 coefficient (nine
// products).  Round 0's sums are sums of those: A0 = sum of the C and O r0^2
// coefficients (the pair differences' products), C0 = sum of their values
// at 0, O0 = sum of their values at 1.  J.npairs = quadruples; nine
// partials per workgroup go to part_out.
__global__ __launch_bounds__(SC_THREADS) void sc_pair01_kernel(ScJob J) {
    __shared__ fe red[3][SC_THREADS / 64];
    const fe *__restrict__ s1 = J.s1;
    const fe *__restrict__ s2 = J.s2;
    const uint64_t nq = J.npairs;
    fe q[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) q[i] = gf::zero();
    const uint64_t stride = (uint64_t)J.nb * SC_THREADS;
    uint64_t j = (uint64_t)blockIdx.x * SC_THREADS + threadIdx.x;
    while (j < nq) {
        const fe x0 = s1[4 * j], x1 = s1[4 * j + 1], x2 = s1[4 * j + 2], x3 = s1[4 * j + 3];
        const fe y0 = s2[4 * j], y1 = s2[4 * j + 1], y2 = s2[4 * j + 2], y3 = s2[4 * j + 3];
        const uint64_t jn = j + stride;
        const fe ax = gf::sub(x1, x0), bx = gf::sub(x3, x2), ay = gf::sub(y1, y0), by = gf::sub(y3, y2);
        // A: (o - e)(o' - e'), o - e = (x2 - x0) + r0 (bx - ax)
        q[0] = gf::add(q[0], gf::mul(gf::sub(x2, x0), gf::sub(y2, y0)));
        q[1] = gf::add(q[1], gf::mul(gf::sub(x3, x1), gf::sub(y3, y1)));
        q[2] = gf::add(q[2], gf::mul(gf::sub(bx, ax), gf::sub(by, ay)));
        // C: e e' = (x0 + r0 ax)(y0 + r0 ay);  O: o o' = (x2 + r0 bx)(y2 + r0 by)
        q[3] = gf::add(q[3], gf::mul(x0, y0));
        q[4] = gf::add(q[4], gf::mul(x1, y1));
        q[5] = gf::add(q[5], gf::mul(ax, ay));
        q[6] = gf::add(q[6], gf::mul(x2, y2));
        q[7] = gf::add(q[7], gf::mul(x3, y3));
        q[8] = gf::add(q[8], gf::mul(bx, by));
        j = jn;
    }
#pragma unroll
    for (int g = 0; g < 3; ++g) {
        block_reduce3<SC_THREADS>(q[3 * g], q[3 * g + 1], q[3 * g + 2], red);
        if (threadIdx.x == 0) {
            J.part_out[9 * blockIdx.x + 3 * g] = q[3 * g];
            J.part_out[9 * blockIdx.x + 3 * g + 1] = q[3 * g + 1];
            J.part_out[9 * blockIdx.x + 3 * g + 2] = q[3 * g + 2];
        }
    }
}

// Finalizes rounds 0 and 1 from sc_pair01_kernel's partials (in every
// workgroup; workgroup 0 records them), then folds each 8-tuple of the
// inputs by r0 and r1 into a pair (e, o) of the quarter-size outputs and
// forms round 2's sums on it.  J.npairs = round 2's pairs (n / 8).
__global__ __launch_bounds__(SC_THREADS) void sc_fold01_kernel(ScJob J) {
    __shared__ fe red[3][SC_THREADS / 64];
    __shared__ fe sh_r[2];
    const fe *__restrict__ s1 = J.s1;
    const fe *__restrict__ s2 = J.s2;
    fe *__restrict__ d1 = J.d1;
    fe *__restrict__ d2 = J.d2;
    const uint64_t npairs = J.npairs;
    const uint64_t stride = (uint64_t)J.nb * SC_THREADS;
    uint64_t j = (uint64_t)blockIdx.x * SC_THREADS + threadIdx.x;
    {
        fe q[9];
#pragma unroll
        for (int i = 0; i < 9; ++i) q[i] = gf::zero();
        for (int b = threadIdx.x; b < J.nb_in; b += SC_THREADS)
#pragma unroll
            for (int i = 0; i < 9; ++i) q[i] = gf::add(q[i], J.part_in[9 * b + i]);
#pragma unroll
        for (int g = 0; g < 3; ++g) block_reduce3<SC_THREADS>(q[3 * g], q[3 * g + 1], q[3 * g + 2], red);
        if (threadIdx.x < 4) {
            __builtin_amdgcn_s_setprio(3);
            ScState *st = J.st;
            const bool writer = blockIdx.x == 0;
            fe prev = st->tr_prev[0], yv = st->tr_y[0];
            // round 0 from round 1's quadratics
            const fe A0 = gf::add(q[5], q[8]), C0 = gf::add(q[3], q[6]), O0 = gf::add(q[4], q[7]);
            const fe r0 = finalize_round(st, 0, J.fs_rounds, st->derive_y, prev, yv, A0, C0, O0, writer);
            // a quadratic from its values at 0 and 1 and its square coefficient, at r0
            auto at = [&](const fe &v0, const fe &v1, const fe &c2) {
                const fe c1 = gf::sub(gf::sub(v1, v0), c2);
                return gf::add(v0, gf::mul(r0, gf::add(c1, gf::mul(r0, c2))));
            };
            const fe A1 = at(q[0], q[1], q[2]), C1 = at(q[3], q[4], q[5]), O1 = at(q[6], q[7], q[8]);
            const fe r1 = finalize_round(st, 1, J.fs_rounds, st->derive_y, prev, yv, A1, C1, O1, writer);
            if (threadIdx.x == 0) { sh_r[0] = r0; sh_r[1] = r1; }
            __builtin_amdgcn_s_setprio(0);
        }
        __syncthreads();
    }
    const fe r0 = sh_r[0], r1 = sh_r[1];
    fe A = gf::zero(), C = gf::zero(), O = gf::zero();
    // one vector's 8-tuple at a time (the scheduling barrier keeps the
    // second vector's loads after the first's folds): fewer live registers
    auto fold8 = [&](const fe *v, uint64_t t, fe &e, fe &o) {
        fe f[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) f[k] = fold1(v[8 * t + 2 * k], v[8 * t + 2 * k + 1], r0);
        e = fold1(f[0], f[1], r1);
        o = fold1(f[2], f[3], r1);
    };
    while (j < npairs) {
        fe e1, o1, e2, o2;
        fold8(s1, j, e1, o1);
        __builtin_amdgcn_sched_barrier(0);
        fold8(s2, j, e2, o2);
        d1[2 * j] = e1; d1[2 * j + 1] = o1;
        d2[2 * j] = e2; d2[2 * j + 1] = o2;
        A = gf::add(A, gf::mul(gf::sub(o1, e1), gf::sub(o2, e2)));
        C = gf::add(C, gf::mul(e1, e2));
        O = gf::add(O, gf::mul(o1, o2));
        j += stride;
    }
    block_reduce3<SC_THREADS>(A, C, O, red);
    if (threadIdx.x == 0) {
        J.part_out[3 * blockIdx.x] = A;
        J.part_out[3 * blockIdx.x + 1] = C;
        J.part_out[3 * blockIdx.x + 2] = O;
    }
}

template <bool FOLD>
__global__ __launch_bounds__(SC_THREADS) void sc_round_batch_kernel(const ScJob *__restrict__ jobs,
                                                                     const uint16_t *__restrict__ blk_job) {
    const ScJob &J = jobs[blk_job[blockIdx.x]];
    sc_round<FOLD>(J, blockIdx.x - J.first_block);
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

// q summed over lanes [0, P) of the wave (P a power of two <= 64; lanes
// without work hold zeros), every lane of the block ends with the sum
__device__ __forceinline__ void wave_sum9(fe *q, int P) {
    for (int msk = 1; msk < P; msk <<= 1)
#pragma unroll
        for (int t = 0; t < 9; ++t) q[t] = gf::add(q[t], shfl_xor_fe(q[t], msk));
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
constexpr int SC_SPLIT_NG = 64;    // phases with at most this many groups spread them over lanes

// All rounds from `round0` to m-1 in one workgroup of four waves.  The input
// (len elements) is v(round0) when round0 == 0, else v(round0 - 1), still to
// be folded by r[round0 - 1], which is finalized here from part_in.  Phase k
// finalizes round i = round0 + k on wave 0 (lanes 0..3: sums, claim check,
// transcript hashes) while waves 1..3 fold the current vector by r[i - 1]
// into the next buffer (d, e, d, ...: never the input) and accumulate round
// i + 1's sums as quadratics in r[i]; one barrier ends each phase.  Vectors
// move through global memory (L2), so the workgroup needs no LDS beyond the
// exchange slots and fits on a CU beside other kernels' workgroups.
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
            if (lane < 4) {
                __builtin_amdgcn_s_setprio(3);
                const int nw = sh_nw[s];
                fe f[9];
#pragma unroll
                for (int t = 0; t < 9; ++t) {
                    f[t] = qpart[s][0][t];
                    for (int w = 1; w < nw; ++w) f[t] = gf::add(f[t], qpart[s][w][t]);
                }
                fe A = f[0], C = f[3], O = f[6];
                if (fold) {
                    A = quad_eval(f[0], f[1], f[2], rr);
                    C = quad_eval(f[3], f[4], f[5], rr);
                    O = quad_eval(f[6], f[7], f[8], rr);
                }
                rr = finalize_round(st, i, fs_rounds, claim_mode, prev, y, A, C, O, true);
                if (lane == 0) sh_r[s ^ 1] = rr;
                __builtin_amdgcn_s_setprio(0);
            }
        } else {
            const bool next = i + 1 < m;       // round i + 1 exists: accumulate its sums
            fe *w1 = o1[ob], *w2 = o2[ob];
            if (fold && next && cur / 8 <= (uint32_t)SC_SPLIT_NG) {
                // few groups: compute wave cw takes groups cw, cw + 3, ..., seven at a
                // time: one fold per lane, then one of the nine products per lane,
                // through the wave's own LDS rows; lanes 0..8 keep the sums
                const uint32_t ng = cur / 8;
                const int cw = wave - 1;
                const uint32_t mine = ng > (uint32_t)cw ? (ng - cw + 2) / 3 : 0;
                const fe r = sh_r[s];
                fe acc = gf::zero();
                for (uint32_t j0 = 0; j0 < mine; j0 += SC_SPLIT_MAX) {
                    const int nc = (int)min<uint32_t>((uint32_t)SC_SPLIT_MAX, mine - j0);
                    if (lane < 8 * nc) {
                        const int gi = lane >> 3, f = lane & 7, e = f & 3;
                        const uint32_t g = cw + 3 * (j0 + gi);
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
            c1 = o1[ob];
            c2 = o2[ob];
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
    __shared__ fe red[3][SC_THREADS / 64];
    __shared__ fe sh_r, sh_prev, sh_y;
    finish_round<SC_THREADS>(st, round, fs_rounds, part_in, nb_in, red, &sh_r, &sh_prev, &sh_y, blockIdx.x == 0);
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
                                 ScState *st, int max_blocks, uint64_t tail_elems, bool fuse01 = false) {
    const int m = ilog2(n);
    fe *a1 = buf, *a2 = buf + n / 2, *b1 = buf + n, *b2 = buf + n + n / 4;
    fe *part[2] = {part0, part0 + 3 * SC_MAX_BLOCKS};
    fe *outs[2][2] = {{a1, a2}, {b1, b2}};
    std::vector<ScJob> jobs;
    const fe *s1 = v1, *s2 = v2;
    uint64_t len = n;        // elements in s1/s2
    int round = 0, which = 0, nb_prev = 0;
    bool pending = false;    // s1/s2 still need the fold with r[round-1]
    // rounds 0-2 as two launches where all three are multi-workgroup rounds:
    // round 2's folded pairs (n / 4 elements) land in outs[1]
    if (fuse01 && n / 2 > tail_elems) {
        auto width = [&](uint64_t items) {
            uint64_t b = (items + SC_THREADS - 1) / SC_THREADS;
            if (b > (uint64_t)max_blocks) b = max_blocks;
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
        ScJob B{};
        B.kind = 2;
        B.s1 = v1; B.s2 = v2;
        B.d1 = outs[1][0]; B.d2 = outs[1][1];
        B.npairs = n / 8;
        B.st = st;
        B.part_in = part[1];
        B.nb_in = A.nb;
        B.part_out = part[0];
        B.nb = width(n / 8);
        B.round = 2; B.m = m; B.fs_rounds = fs_rounds;
        jobs.push_back(B);
        s1 = outs[1][0]; s2 = outs[1][1];
        len = n / 4;
        round = 3;
        which = 0;
        nb_prev = B.nb;
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
        J.part_in = part[(round + 1) & 1];
        J.part_out = part[round & 1];
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
    T.part_in = part[(round + 1) & 1];
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
    const auto jobs = sumcheck_jobs(v1, v2, n, fs_rounds, buf, ctx->sc_partials, ctx->sc, ctx->sc_blocks,
                                    ctx->sc_tail_elems, fuse01);
    for (size_t i = 0; i + 1 < jobs.size(); ++i) {
        const ScJob &J = jobs[i];
        if (J.kind == 1)
            hipLaunchKernelGGL(sc_pair01_kernel, dim3((unsigned)J.nb), dim3(SC_THREADS), 0, stream, J);
        else if (J.kind == 2)
            hipLaunchKernelGGL(sc_fold01_kernel, dim3((unsigned)J.nb), dim3(SC_THREADS), 0, stream, J);
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
        per[i] = sumcheck_jobs(v1[i], v2[i], n[i], m, buf[i], parts[i], st[i], ctx->sc_blocks, ctx->sc_tail_elems);
        nr = std::max(nr, per[i].size() - 1);
    }
    plan.rounds.clear();
    plan.count = count;
    plan.prev = prev;
    for (size_t r = 0; r < nr; ++r) {
        std::vector<ScJob> jobs;
        std::vector<uint16_t> map;
        for (int i = 0; i < count; ++i) {
            if (r + 1 >= per[i].size()) continue;
            ScJob J = per[i][r];
            J.first_block = (uint32_t)map.size();
            for (int b = 0; b < J.nb; ++b) map.push_back((uint16_t)jobs.size());
            jobs.push_back(J);
        }
        if (jobs.size() > 0xFFFF) return set_err(ctx, COSNARK_E_RANGE, "sumcheck batch: too many proofs");
        ScBatchPlan::Launch L{r > 0, (uint32_t)map.size(), tb.reserve(sizeof(ScJob) * jobs.size()), 0};
        L.blk_job = tb.reserve(sizeof(uint16_t) * map.size());
        memcpy(tb.at<ScJob>(L.jobs), jobs.data(), sizeof(ScJob) * jobs.size());
        memcpy(tb.at<uint16_t>(L.blk_job), map.data(), sizeof(uint16_t) * map.size());
        plan.rounds.push_back(L);
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
        if (L.fold)
            hipLaunchKernelGGL(sc_round_batch_kernel<true>, dim3(L.blocks), dim3(SC_THREADS), 0, stream,
                               (const ScJob *)(tables + L.jobs), (const uint16_t *)(tables + L.blk_job));
        else
            hipLaunchKernelGGL(sc_round_batch_kernel<false>, dim3(L.blocks), dim3(SC_THREADS), 0, stream,
                               (const ScJob *)(tables + L.jobs), (const uint16_t *)(tables + L.blk_job));
        LAUNCHCK();
    }
    hipLaunchKernelGGL(sc_tail_batch_kernel, dim3(plan.count), dim3(SC_TAIL_THREADS), 0, stream, tail);
    LAUNCHCK();
    return COSNARK_OK;
}
