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
// pair.  Round i+1's evaluation is fused with round i's fold: one pass reads
// four elements per vector, writes the two folded ones and accumulates the
// next round's sums.  Workgroup partials are combined by the last workgroup
// to finish (atomic counter), which also runs the transcript hashes, so the
// large rounds need one launch each and no host round trip.  Once a vector
// fits in LDS (<= 4096 elements) one workgroup runs all remaining rounds.
#include "ctx.h"
#include "transcript.h"

using gf::fe;

namespace {

constexpr int SC_THREADS = 256;
constexpr int SC_TAIL_THREADS = 1024;
constexpr int SC_TAIL_ELEMS = 4096;

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

// Scalar transcript state carried through the rounds by one thread.
struct ScRound {
    fe prev, y;
    int err, derive_y;
};

// Combine (A = sum d1 d2, C = sum e1 e2, O = sum o1 o2) into round `round`'s
// polynomial, check it, derive the challenge, update y.  One thread.
// st->polys / st->r are written only; st->ext_r is read only.
__device__ fe finalize_round(ScRound &s, ScState *st, int round, int fs_rounds, fe A, fe C, fe O) {
    const fe a = A, c = C;
    const fe b = gf::sub(gf::sub(O, A), C);
    const fe h01 = gf::add(gf::add(a, b), gf::add(c, c));   // H(0) + H(1)
    if (round == 0 && s.derive_y) {
        s.y = h01;
    } else if (!gf::eq(h01, s.y) && s.err == 0) {
        s.err = round + 1;
    }
    fe r;
    if (round < fs_rounds) {
        s.prev = ts::hash_F(s.prev, a);
        s.prev = ts::hash_F(s.prev, b);
        s.prev = ts::hash_F(s.prev, c);
        r = s.prev;
    } else {
        r = st->ext_r[round];
    }
    s.y = gf::add(gf::mul(gf::add(gf::mul(a, r), b), r), c);
    st->polys[3 * round] = a;
    st->polys[3 * round + 1] = b;
    st->polys[3 * round + 2] = c;
    st->r[round] = r;
    return r;
}

__device__ __forceinline__ ScRound load_round_state(const ScState *st) {
    ScRound s;
    s.prev = st->prev; s.y = st->y; s.err = st->err; s.derive_y = st->derive_y;
    return s;
}

__device__ __forceinline__ void store_round_state(ScState *st, const ScRound &s) {
    st->prev = s.prev; st->y = s.y; st->err = s.err;
}

// One multi-workgroup round.  FOLD: inputs have 4*npairs elements and are
// folded with r[round-1] into outputs of 2*npairs; otherwise inputs have
// 2*npairs elements and are used as they are.
template <bool FOLD>
__global__ __launch_bounds__(SC_THREADS) void sc_round_kernel(
    const fe *__restrict__ s1, const fe *__restrict__ s2, fe *__restrict__ d1, fe *__restrict__ d2,
    uint64_t npairs, int round, int fs_rounds, ScState *st, fe *partials) {
    __shared__ fe red[3][SC_THREADS / 64];
    __shared__ int is_last;
    fe r = gf::zero();
    if (FOLD) r = st->r[round - 1];
    fe A = gf::zero(), C = gf::zero(), O = gf::zero();
    const uint64_t stride = (uint64_t)gridDim.x * SC_THREADS;
    for (uint64_t j = (uint64_t)blockIdx.x * SC_THREADS + threadIdx.x; j < npairs; j += stride) {
        fe e1, o1, e2, o2;
        if (FOLD) {
            e1 = fold1(s1[4 * j], s1[4 * j + 1], r);
            o1 = fold1(s1[4 * j + 2], s1[4 * j + 3], r);
            e2 = fold1(s2[4 * j], s2[4 * j + 1], r);
            o2 = fold1(s2[4 * j + 2], s2[4 * j + 3], r);
            d1[2 * j] = e1; d1[2 * j + 1] = o1;
            d2[2 * j] = e2; d2[2 * j + 1] = o2;
        } else {
            e1 = s1[2 * j]; o1 = s1[2 * j + 1];
            e2 = s2[2 * j]; o2 = s2[2 * j + 1];
        }
        A = gf::add(A, gf::mul(gf::sub(o1, e1), gf::sub(o2, e2)));
        C = gf::add(C, gf::mul(e1, e2));
        O = gf::add(O, gf::mul(o1, o2));
    }
    wave_reduce3(A, C, O);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (lane == 0) { red[0][wv] = A; red[1][wv] = C; red[2][wv] = O; }
    __syncthreads();
    if (threadIdx.x == 0) {
        fe a = red[0][0], c = red[1][0], o = red[2][0];
        for (int w = 1; w < SC_THREADS / 64; ++w) {
            a = gf::add(a, red[0][w]); c = gf::add(c, red[1][w]); o = gf::add(o, red[2][w]);
        }
        partials[3 * blockIdx.x] = a;
        partials[3 * blockIdx.x + 1] = c;
        partials[3 * blockIdx.x + 2] = o;
        __threadfence();
        const uint32_t prev = atomicAdd(&st->counter, 1u);
        is_last = (prev == gridDim.x - 1);
    }
    __syncthreads();
    if (!is_last) return;
    __threadfence();
    // last workgroup: reduce all partials
    A = gf::zero(); C = gf::zero(); O = gf::zero();
    for (unsigned b = threadIdx.x; b < gridDim.x; b += SC_THREADS) {
        const volatile uint64_t *p = reinterpret_cast<const volatile uint64_t *>(partials + 3 * b);
        A = gf::add(A, gf::mk(p[0], p[1]));
        C = gf::add(C, gf::mk(p[2], p[3]));
        O = gf::add(O, gf::mk(p[4], p[5]));
    }
    wave_reduce3(A, C, O);
    __syncthreads();
    if (lane == 0) { red[0][wv] = A; red[1][wv] = C; red[2][wv] = O; }
    __syncthreads();
    if (threadIdx.x == 0) {
        fe a = red[0][0], c = red[1][0], o = red[2][0];
        for (int w = 1; w < SC_THREADS / 64; ++w) {
            a = gf::add(a, red[0][w]); c = gf::add(c, red[1][w]); o = gf::add(o, red[2][w]);
        }
        ScRound rs = load_round_state(st);
        finalize_round(rs, st, round, fs_rounds, a, c, o);
        store_round_state(st, rs);
        st->counter = 0;
        __threadfence();
    }
}

// All rounds from `round0` to m-1 in one workgroup, vectors held in LDS.
// If round0 > 0 the input (len elements) still needs the fold with
// r[round0-1].  Writes the final folded values to st->final_v.
__global__ __launch_bounds__(SC_TAIL_THREADS) void sc_tail_kernel(
    const fe *__restrict__ s1, const fe *__restrict__ s2, uint32_t len, int round0, int m,
    int fs_rounds, ScState *st) {
    extern __shared__ fe lds[];
    fe *x1 = lds, *x2 = lds + SC_TAIL_ELEMS;
    __shared__ fe red[3][SC_TAIL_THREADS / 64];
    __shared__ ScRound srs;
    __shared__ fe sh_r;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    for (uint32_t i = tid; i < len; i += SC_TAIL_THREADS) { x1[i] = s1[i]; x2[i] = s2[i]; }
    if (tid == 0) {
        srs = load_round_state(st);
        sh_r = round0 > 0 ? st->r[round0 - 1] : gf::zero();
    }
    __syncthreads();
    uint32_t cur = len;      // elements currently held (before pending fold)
    bool pending = round0 > 0;
    for (int round = round0; round <= m; ++round) {
        // apply pending fold with r[round-1], then evaluate round `round`
        const fe r = sh_r;
        const uint32_t after = pending ? cur / 2 : cur;     // elements after fold
        if (round == m) {
            // final fold only
            if (tid == 0) {
                st->final_v[0] = pending ? fold1(x1[0], x1[1], r) : x1[0];
                st->final_v[1] = pending ? fold1(x2[0], x2[1], r) : x2[0];
                store_round_state(st, srs);
            }
            break;
        }
        const uint32_t npairs = after / 2;
        fe A = gf::zero(), C = gf::zero(), O = gf::zero();
        fe w[2][4];
        int nmine = 0;
        for (uint32_t j = tid; j < npairs; j += SC_TAIL_THREADS) {
            fe e1, o1, e2, o2;
            if (pending) {
                e1 = fold1(x1[4 * j], x1[4 * j + 1], r);
                o1 = fold1(x1[4 * j + 2], x1[4 * j + 3], r);
                e2 = fold1(x2[4 * j], x2[4 * j + 1], r);
                o2 = fold1(x2[4 * j + 2], x2[4 * j + 3], r);
            } else {
                e1 = x1[2 * j]; o1 = x1[2 * j + 1];
                e2 = x2[2 * j]; o2 = x2[2 * j + 1];
            }
            A = gf::add(A, gf::mul(gf::sub(o1, e1), gf::sub(o2, e2)));
            C = gf::add(C, gf::mul(e1, e2));
            O = gf::add(O, gf::mul(o1, o2));
            w[nmine][0] = e1; w[nmine][1] = o1; w[nmine][2] = e2; w[nmine][3] = o2;
            ++nmine;
        }
        wave_reduce3(A, C, O);
        if (lane == 0) { red[0][wv] = A; red[1][wv] = C; red[2][wv] = O; }
        __syncthreads();   // all reads of x done; partials visible
        if (pending) {
            int q = 0;
            for (uint32_t j = tid; j < npairs; j += SC_TAIL_THREADS, ++q) {
                x1[2 * j] = w[q][0]; x1[2 * j + 1] = w[q][1];
                x2[2 * j] = w[q][2]; x2[2 * j + 1] = w[q][3];
            }
        }
        if (tid == 0) {
            fe a = red[0][0], c = red[1][0], o = red[2][0];
            for (int k = 1; k < SC_TAIL_THREADS / 64; ++k) {
                a = gf::add(a, red[0][k]); c = gf::add(c, red[1][k]); o = gf::add(o, red[2][k]);
            }
            sh_r = finalize_round(srs, st, round, fs_rounds, a, c, o);
        }
        __syncthreads();
        cur = after;
        pending = true;
    }
}

}  // namespace

int sumcheck_init(cosnark_ctx *ctx) {
    // the tail kernel keeps two 4096-element vectors (128 KiB) in LDS
    HIPCK(hipFuncSetAttribute((const void *)sc_tail_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)(2 * SC_TAIL_ELEMS * sizeof(fe))));
    return COSNARK_OK;
}

int sumcheck_launch(cosnark_ctx *ctx, const fe *v1, const fe *v2, uint64_t n, int fs_rounds, fe *buf) {
    const int m = ilog2(n);
    // buffers: ping (n/2 per vector) and pong (n/4 per vector)
    fe *a1 = buf, *a2 = buf + n / 2, *b1 = buf + n, *b2 = buf + n + n / 4;
    const fe *s1 = v1, *s2 = v2;
    uint64_t len = n;        // elements in s1/s2
    int round = 0;
    bool pending = false;    // s1/s2 still need the fold with r[round-1]
    fe *outs[2][2] = {{a1, a2}, {b1, b2}};
    int which = 0;
    while (len > (uint64_t)SC_TAIL_ELEMS) {
        const uint64_t after = pending ? len / 2 : len;
        const uint64_t npairs = after / 2;
        uint64_t blocks = (npairs + SC_THREADS - 1) / SC_THREADS;
        if (blocks > (uint64_t)SC_MAX_BLOCKS) blocks = SC_MAX_BLOCKS;
        if (pending) {
            fe *o1 = outs[which][0], *o2 = outs[which][1];
            hipLaunchKernelGGL(sc_round_kernel<true>, dim3((unsigned)blocks), dim3(SC_THREADS), 0, ctx->stream,
                               s1, s2, o1, o2, npairs, round, fs_rounds, ctx->sc, ctx->sc_partials);
            s1 = o1; s2 = o2;
            which ^= 1;
        } else {
            hipLaunchKernelGGL(sc_round_kernel<false>, dim3((unsigned)blocks), dim3(SC_THREADS), 0, ctx->stream,
                               s1, s2, (fe *)nullptr, (fe *)nullptr, npairs, round, fs_rounds, ctx->sc,
                               ctx->sc_partials);
        }
        LAUNCHCK();
        len = after;
        pending = true;
        ++round;
    }
    const size_t lds = 2 * SC_TAIL_ELEMS * sizeof(fe);
    hipLaunchKernelGGL(sc_tail_kernel, dim3(1), dim3(SC_TAIL_THREADS), lds, ctx->stream,
                       s1, s2, (uint32_t)len, round, m, fs_rounds, ctx->sc);
    LAUNCHCK();
    return COSNARK_OK;
}
