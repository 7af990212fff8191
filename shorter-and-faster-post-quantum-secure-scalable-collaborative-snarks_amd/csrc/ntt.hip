// ntt.hip -- number-theoretic transform over GF((2^61-1)^2) for gfx950.
//
// Computes the reference's fft(arr, logn, flag) (src/utils.cpp:904-964):
// X[k] = sum_j x[j] w^(jk), w = getRootOfUnity(logn), natural order in and
// out; the inverse uses w^-1 and scales by 2^-logn.  pfft
// (src/utils.cpp:1583-1673) is the same forward transform.
//
// Design: a 2^L-point transform is split into passes of radix 2^r (r <= 8).
// Pass t transforms one index digit of stride S_t = 2^(L - r_0 - .. - r_t)
// for every (prefix, suffix) pair: a workgroup stages R x C elements
// (R = 2^r points for C adjacent columns, C*R = 4096) in LDS, runs radix-2
// decimation-in-time stages there, multiplies by the inter-pass twiddle
// w_{N_t}^(lo*k) and writes back in place.  The last pass (stride 1) writes
// in natural order; its workgroup takes C rows whose leading output digit
// is consecutive, so each store is a C-element contiguous run.  Inside a
// pass values are weak (gf.h: unreduced sums, one fold per radix-4 layer);
// the twiddle products and the last pass's stores are canonical.  Zero padding
// (Reed-Solomon encoding of n coefficients into rate*n points) is handled by
// the first pass, which reads only the n real inputs.  Where padding by
// E = 2^e would cost a pass, the first pass fans out instead: its digit is
// e bits wider than its transform, and since only the first R of its E*R
// inputs are non-zero, X[k0 + E k1] = DFT_R(x[d] w_(ER)^(d k0))[k1]; the
// workgroup runs the R-point transform once per k0 < E on its twisted tile.
#include "blake3.h"
#include "ctx.h"
#include "dft.h"

using gf::fe;

namespace {

using dft::brev_c;
using dft::dif_w;
using dft::static_for;
using dft::tw28;

struct NttPass {
    // batch item y = blockIdx.y reads in_tab[y] (or in + y * in_stride) and
    // writes out_tab[y] (or out + y * out_stride)
    const fe *const *in_tab;
    fe *const *out_tab;
    uint64_t in_stride, out_stride;
    uint64_t in_len;    // elements of `in` that are non-zero inputs
    int logN, logR, logS, logC;
    int last;           // 1: stride-1 pass writing natural order
    int fan;            // first pass of a transform padded by 2^fan: fan out (see above)
    int npass;
    int rbits[8];       // radix bits per pass, first pass first
    int pre[8];         // pre[t] = rbits[0] + .. + rbits[t-1]
    uint8_t *leaves;    // last pass, batch item 0: Merkle leaf digests of the output (or null)
};

constexpr int Y_PAD = 1;   // one padding element per LDS row keeps 16-byte lanes off shared banks

// One pass of radix R = R1 * R2 over C columns held in LDS as a 2-D
// (four-step) transform: step 1 does R2*C in-register R1-point DFTs over
// stride-R2 inputs and the w_R^(d2 k1) twiddle, step 2 does R1*C R2-point
// DFTs, applies the inter-pass twiddle w_{N_t}^(lo k) and stores.
//
// SPLIT: every thread holds at most one step-1 and one step-2 item, and the
// LDS exchange moves the real words first and the imaginary words second
// through one buffer of 8-byte words.  That halves the tile's LDS (about
// 42 KiB for the 4096-element tile instead of 76), so three workgroups fit
// on a CU instead of two, for two more barriers per pass.
template <int R1, int R2, bool LAST, bool INV, bool SPLIT>
__global__ __launch_bounds__(NTT_THREADS) void ntt_pass_kernel(
    const fe *in0, fe *out0, NttPass p, const fe *__restrict__ twlo, const fe *__restrict__ twhi,
    const fe *__restrict__ rtab, fe scale) {
    constexpr int R = R1 * R2;
    constexpr int LOGR = (R >= 256) ? 8 : (R >= 128) ? 7 : (R >= 64) ? 6 : (R >= 32) ? 5 : (R >= 16) ? 4 : (R >= 8) ? 3 : (R >= 4) ? 2 : (R >= 2) ? 1 : 0;
    constexpr int LOGR2 = (R2 >= 16) ? 4 : (R2 >= 8) ? 3 : (R2 >= 4) ? 2 : (R2 >= 2) ? 1 : 0;
    extern __shared__ fe ntt_lds[];
    fe *wr = ntt_lds;          // w_R^e, e < R
    fe *Y = ntt_lds + R;       // (R * C + R * Y_PAD) elements (SPLIT: as many 8-byte words)
    uint64_t *Yw = reinterpret_cast<uint64_t *>(Y);
    const int tid = threadIdx.x;
    const int logC = p.logC, C = 1 << logC;
    const int CP = C + Y_PAD;   // padded LDS row length
    const uint32_t by = blockIdx.y;
    const fe *in = p.in_tab ? p.in_tab[by] : in0 + by * p.in_stride;
    fe *out = p.out_tab ? p.out_tab[by] : out0 + by * p.out_stride;
    for (int j = tid; j < R; j += NTT_THREADS) wr[j] = rtab[j << (NTT_RADIX_LOG_MAX - LOGR)];

    const uint64_t bid = blockIdx.x;
    // block -> positions
    uint64_t hi = 0, lo0 = 0;              // non-last: pos = hi*(R*S) + d*S + lo0 + c
    uint64_t rest = 0, c0 = 0, W0 = 0;      // last: row(c) = rest + (c0 + c)*W0, pos = row*R + d
    const int logS = p.logS;
    if (!LAST) {
        const int lo_blocks_log = logS - logC;
        hi = bid >> lo_blocks_log;
        lo0 = (bid & ((1ull << lo_blocks_log) - 1)) << logC;
    } else if (p.npass >= 2) {
        const int cblocks_log = p.rbits[0] - logC;
        rest = bid >> cblocks_log;
        c0 = (bid & ((1ull << cblocks_log) - 1)) << logC;
        W0 = 1ull << (p.logN - p.rbits[0] - p.rbits[p.npass - 1]);
    }
    __syncthreads();

    // fan-out (first pass of a padded forward transform): the pass digit has
    // LOGR + fan bits, and the steps below run once per k0 < 2^fan
    // (the host never picks the split exchange for a fan-out pass: the k0
    // loop around it costs the split kernel half its occupancy in registers)
    const int fan = (LAST || INV || SPLIT) ? 0 : p.fan;
    const int logNt = LOGR + fan + logS;
    const int items1 = R2 << logC;
    const int items2 = R1 << logC;
    uint64_t nat_rest = 0;
    int kshift = 0;
    if (LAST) {
        uint64_t x = rest;
        for (int t = p.npass - 2; t >= 1; --t) {
            const uint64_t dgt = x & ((1ull << p.rbits[t]) - 1);
            x >>= p.rbits[t];
            nat_rest += dgt << p.pre[t];
        }
        kshift = p.logN - LOGR;
    }
    // the fused leaf hashing below needs the last item's outputs after the
    // loop (it is enabled only when each thread has at most one item)
    const bool leaves_on = LAST && p.leaves && by == 0 && items2 <= NTT_THREADS;
    fe v[R2];
    int my_c = 0, my_k1 = 0;
    // fan-out: w_(E R)^(d k0) for the current k0, d < R (R * CP is even)
    fe *tz = SPLIT ? Y + (size_t)R * CP / 2 : Y + (size_t)R * CP;
    for (int k0 = 0; k0 < (1 << fan); ++k0) {
    if (k0 > 0) __syncthreads();   // the previous k0's step 2 is done with Y

    // ---- step 1: R1-point DFTs over d1 for each (d2, c) ----
    // item -> x[i] = the DFT output k1 = brev(i) after the w_R^(d2 k1)
    // twiddle, bound for exchange row k1 * R2 + d2, column c (*off = d2 * CP + c)
    auto step1 = [&](int it, fe *x, int *off) {
        int c, d2;
        uint64_t base;
        if (!LAST) {
            c = it & (C - 1);
            d2 = it >> logC;
            base = (hi << logNt) + lo0 + c;
        } else {
            d2 = it & (R2 - 1);
            c = it >> LOGR2;
            base = (rest + (c0 + c) * W0) << LOGR;
        }
#pragma unroll
        for (int d1 = 0; d1 < R1; ++d1) {
            const uint64_t d = (uint64_t)(R2 * d1 + d2);
            const uint64_t pos = base + (LAST ? d : (d << logS));
            x[d1] = pos < p.in_len ? in[pos] : gf::zero();
            if (k0 > 0) x[d1] = gf::mul_w(tz[d], x[d1]);   // twist by w_(E R)^(d k0); inputs are canonical
        }
        dif_w<R1, INV>(x);
        static_for<0, R1>([&](auto ii) {
            constexpr int i = decltype(ii)::value;
            constexpr int k1 = brev_c<R1>(i);
            if (k1 != 0 && d2 != 0) x[i] = gf::mul_w(wr[d2 * k1], x[i]);
        });
        *off = d2 * CP + c;
    };
    // the next k0's twists, while step 2 runs (called once step 1 is done
    // with tz; the barrier at the top of the next k0 orders these writes
    // before its reads)
    auto next_twists = [&]() {
        if (k0 + 1 < (1 << fan))
            for (int d = tid; d < R; d += NTT_THREADS)
                tz[d] = tw28(twlo, twhi, (uint32_t)(d * (k0 + 1)) << (NTT_LOG_MAX - LOGR - fan));
    };
    // ---- step 2: R2-point DFTs over d2 for each (k1, c), twiddle, store ----
    // (v holds the item's exchange column: v[d2] = row k1 * R2 + d2, column c)
    auto step2 = [&](int it) {
        const int c = it & (C - 1), k1 = it >> logC;
        my_c = c;
        my_k1 = k1;
        dif_w<R2, INV>(v);
        if (!LAST) {
            // inter-pass twiddle w_{N_t}^(lo k), k = k0 + 2^fan (k1 + R1 k2): one table
            // product for w^(lo (k0 + 2^fan k1)) and one for w^(lo 2^fan R1), then a
            // running product over k2
            const uint64_t lo = lo0 + c;
            const int sh = NTT_LOG_MAX - logNt;
            const uint32_t eb = (uint32_t)((lo * (uint64_t)(k0 + (k1 << fan))) << sh);
            const uint32_t es = (uint32_t)((lo * (uint64_t)(R1 << fan)) << sh);
            fe t = tw28(twlo, twhi, eb);
            const fe step = tw28(twlo, twhi, es);
            static_for<0, R2>([&](auto kk) {
                constexpr int k2 = decltype(kk)::value;
                constexpr int i = brev_c<R2>(k2);
                const uint64_t k = (uint64_t)k0 + ((uint64_t)(k1 + R1 * k2) << fan);
                out[(hi << logNt) + (k << logS) + lo] = gf::mul_w(t, v[i]);   // weak: read by the next pass
                if (k2 + 1 < R2) t = gf::mul(t, step);
            });
        } else {
            static_for<0, R2>([&](auto ii) {
                constexpr int i = decltype(ii)::value;
                constexpr int k2 = brev_c<R2>(i);
                const uint64_t k = (uint64_t)(k1 + R1 * k2);
                v[i] = INV ? gf::mul(scale, v[i]) : gf::canon_w(v[i]);
                out[nat_rest + c0 + c + (k << kshift)] = v[i];
            });
        }
    };

    if constexpr (!SPLIT) {
        for (int it = tid; it < items1; it += NTT_THREADS) {
            fe x[R1];
            int off;
            step1(it, x, &off);
            static_for<0, R1>([&](auto ii) {
                constexpr int i = decltype(ii)::value;
                Y[brev_c<R1>(i) * R2 * CP + off] = x[i];
            });
        }
        __syncthreads();
        next_twists();
        for (int it = tid; it < items2; it += NTT_THREADS) {
            const int c = it & (C - 1), k1 = it >> logC;
#pragma unroll
            for (int d2 = 0; d2 < R2; ++d2) v[d2] = Y[(k1 * R2 + d2) * CP + c];
            step2(it);
        }
    } else {
        // at most one item of each step per thread (items1, items2 <= NTT_THREADS)
        const bool act1 = tid < items1, act2 = tid < items2;
        fe x[R1];
        int off = 0;
        if (act1) step1(tid, x, &off);
        const int c2 = tid & (C - 1), k12 = tid >> logC;
        if (act1)
            static_for<0, R1>([&](auto ii) {
                constexpr int i = decltype(ii)::value;
                Yw[brev_c<R1>(i) * R2 * CP + off] = x[i].re;
            });
        __syncthreads();
        next_twists();
        if (act2)
#pragma unroll
            for (int d2 = 0; d2 < R2; ++d2) v[d2].re = Yw[(k12 * R2 + d2) * CP + c2];
        __syncthreads();
        if (act1)
            static_for<0, R1>([&](auto ii) {
                constexpr int i = decltype(ii)::value;
                Yw[brev_c<R1>(i) * R2 * CP + off] = x[i].im;
            });
        __syncthreads();
        if (act2) {
#pragma unroll
            for (int d2 = 0; d2 < R2; ++d2) v[d2].im = Yw[(k12 * R2 + d2) * CP + c2];
            step2(tid);
        }
    }
    }   // k0
    if (leaves_on) {
        // Merkle leaves of the output (MT_commit_Blake, src/merkle_tree.cpp:
        // 211-220: leaf = BLAKE3 of 4 consecutive elements).  A leaf is 4
        // consecutive c at one k.  Each thread's outputs are still in v:
        // restage them as Z[k][c] in the exchange buffer (after every
        // thread's step-2 reads of it), then each thread hashes whole leaves.
        // The condition is uniform over the workgroup, so are the barriers.
        // SPLIT halves the buffer, so it restages and hashes rows k < R/2
        // (k2 < R2/2) first and the rest second.
        constexpr int NH = SPLIT ? 2 : 1, RH = R / NH;
        const int qpr = C / 4, nleaf = (RH * C) / 4;
        static_for<0, NH>([&](auto hh) {
            constexpr int h = decltype(hh)::value;
            __syncthreads();   // the exchange buffer is free (step-2 reads, or the previous half's hashes)
            if (tid < items2) {
                static_for<0, R2>([&](auto ii) {
                    constexpr int i = decltype(ii)::value;
                    constexpr int k2 = brev_c<R2>(i);
                    const int hc = decltype(hh)::value;
                    if (k2 / (R2 / NH) == hc) Y[(my_k1 + R1 * k2 - hc * RH) * CP + my_c] = v[i];
                });
            }
            __syncthreads();
            for (int l = tid; l < nleaf; l += NTT_THREADS) {
                const int k = l / qpr, q = l % qpr;
                uint32_t m[16], hv[8];
                const fe *src = Y + k * CP + 4 * q;
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    m[4 * e] = (uint32_t)src[e].re;
                    m[4 * e + 1] = (uint32_t)(src[e].re >> 32);
                    m[4 * e + 2] = (uint32_t)src[e].im;
                    m[4 * e + 3] = (uint32_t)(src[e].im >> 32);
                }
                b3::compress64(m, hv);
                const uint64_t pos = nat_rest + c0 + 4 * q + ((uint64_t)(k + h * RH) << kshift);
                uint4 *dst = reinterpret_cast<uint4 *>(p.leaves + 32 * (pos >> 2));
                dst[0] = make_uint4(hv[0], hv[1], hv[2], hv[3]);
                dst[1] = make_uint4(hv[4], hv[5], hv[6], hv[7]);
            }
        });
    }
}

using PassKernel = void (*)(const fe *, fe *, NttPass, const fe *, const fe *, const fe *, fe);

template <int R1, int R2, bool SPLIT>
PassKernel pass_kernel_s(bool last, bool inverse) {
    return last ? (inverse ? ntt_pass_kernel<R1, R2, true, true, SPLIT> : ntt_pass_kernel<R1, R2, true, false, SPLIT>)
                : (inverse ? ntt_pass_kernel<R1, R2, false, true, SPLIT>
                           : ntt_pass_kernel<R1, R2, false, false, SPLIT>);
}

template <int R1, int R2>
PassKernel pass_kernel(bool last, bool inverse, bool split) {
    return split ? pass_kernel_s<R1, R2, true>(last, inverse) : pass_kernel_s<R1, R2, false>(last, inverse);
}

PassKernel pass_kernel_for(int r, bool last, bool inverse, bool split) {
    switch (r) {
        case 1: return pass_kernel<2, 1>(last, inverse, split);
        case 2: return pass_kernel<2, 2>(last, inverse, split);
        case 3: return pass_kernel<4, 2>(last, inverse, split);
        case 4: return pass_kernel<4, 4>(last, inverse, split);
        case 5: return pass_kernel<8, 4>(last, inverse, split);
        case 6: return pass_kernel<8, 8>(last, inverse, split);
        case 7: return pass_kernel<16, 8>(last, inverse, split);
        case 8: return pass_kernel<16, 16>(last, inverse, split);
        default: return nullptr;
    }
}

template <int R1, int R2>
int launch_pass(cosnark_ctx *ctx, hipStream_t stream, uint64_t blocks, int batch, const fe *in, fe *out,
                const NttPass &q, bool last, bool inverse, const fe *twlo, const fe *twhi,
                const fe *rtab, fe scale) {
    // the split exchange needs at most one item of each step per thread
    const bool split = ctx->ntt_split && !q.fan && ((R1 << q.logC) <= NTT_THREADS) &&
                       ((R2 << q.logC) <= NTT_THREADS);
    PassKernel k = pass_kernel<R1, R2>(last, inverse, split);
    // radix table, padded tile (split: 8-byte words), and for a fan-out pass the R twists
    const size_t R = (size_t)R1 * R2, CP = (1u << q.logC) + Y_PAD;
    const size_t lds = sizeof(fe) * (R + (q.fan ? R : 0)) + (split ? 8 : 16) * R * CP;
    hipLaunchKernelGGL(k, dim3((unsigned)blocks, (unsigned)batch), dim3(NTT_THREADS), lds, stream, in, out, q, twlo,
                       twhi, rtab, scale);
    LAUNCHCK();
    return COSNARK_OK;
}

int dispatch_pass(cosnark_ctx *ctx, hipStream_t stream, int r, uint64_t blocks, int batch, const fe *in, fe *out,
                  const NttPass &q, bool last, bool inverse, const fe *twlo, const fe *twhi, const fe *rtab,
                  fe scale) {
#define PASS(R1, R2) launch_pass<R1, R2>(ctx, stream, blocks, batch, in, out, q, last, inverse, twlo, twhi, rtab, scale)
    switch (r) {
        case 1: return PASS(2, 1);
        case 2: return PASS(2, 2);
        case 3: return PASS(4, 2);
        case 4: return PASS(4, 4);
        case 5: return PASS(8, 4);
        case 6: return PASS(8, 8);
        case 7: return PASS(16, 8);
        case 8: return PASS(16, 16);
        default: return set_err(ctx, COSNARK_E_RANGE, "ntt: unsupported radix");
    }
#undef PASS
}

}  // namespace

int ntt_init(cosnark_ctx *ctx) {
    // the largest tile: 4096 elements + padding + the radix table + a fan-out
    // pass's twists (~76 KiB)
    const int bytes = (int)(sizeof(fe) * (NTT_LDS_ELEMS + 256 * (Y_PAD + 2)));
    for (int r = 1; r <= NTT_PASS_LOG_MAX; ++r)
        for (int last = 0; last < 2; ++last)
            for (int inv = 0; inv < 2; ++inv)
                for (int split = 0; split < 2; ++split)
                    HIPCK(hipFuncSetAttribute((const void *)pass_kernel_for(r, last, inv, split),
                                              hipFuncAttributeMaxDynamicSharedMemorySize, bytes));
    return COSNARK_OK;
}

int ntt_plan_bits(const cosnark_ctx *ctx, int logn) {
    const int b = ctx->ntt_plan[logn].passbits;
    return b ? b : ctx->ntt_passbits;
}

int ntt_plan_tile(const cosnark_ctx *ctx, int logn) {
    const int t = ctx->ntt_plan[logn].tile;
    return t ? t : ctx->ntt_tile;
}

int ntt_launch_batch(cosnark_ctx *ctx, const fe *in, const fe *const *in_tab, uint64_t in_len, fe *out,
                     fe *const *out_tab, int batch, int logn, bool inverse, hipStream_t stream, fe *tmp_own,
                     uint8_t *leaves, int *levels_done) {
    if (levels_done) *levels_done = 0;
    if (logn < 1 || logn > NTT_LOG_MAX) return set_err(ctx, COSNARK_E_RANGE, "ntt: logn out of range");
    if (batch < 1 || batch > 65535) return set_err(ctx, COSNARK_E_RANGE, "ntt: batch out of range");
    // radix plan: fewest passes of at most `passbits` bits, spread evenly, larger first
    const int pmax = ntt_plan_bits(ctx, logn);
    const int tile = ntt_plan_tile(ctx, logn);
    const uint64_t N = 1ull << logn;
    int m = (logn + pmax - 1) / pmax;
    // a forward transform padded by E = 2^fan (in_len = N / E, E <= 8) fans
    // out in its first pass where that saves a pass and the first pass keeps
    // enough workgroups (in_len / tile of them, each running E transforms)
    int fan = 0;
    if (!inverse && in_len > 0 && in_len < N && (in_len & (in_len - 1)) == 0 &&
        (in_len / (uint64_t)tile) >= (uint64_t)ctx->ntt_fan_min_blocks) {
        const int e = logn - ilog2(in_len);
        const int mf = (logn - e + pmax - 1) / pmax;
        if (e <= NTT_FAN_LOG_MAX && mf >= 2 && mf < m) {
            fan = e;
            m = mf;
        }
    }
    NttPass p{};
    p.in_len = in_len;
    p.logN = logn;
    p.npass = m;
    const int lb = logn - fan;
    for (int t = 0; t < m; ++t) p.rbits[t] = lb / m + (t < lb % m ? 1 : 0);
    p.rbits[0] += fan;
    p.pre[0] = 0;
    for (int t = 1; t < m; ++t) p.pre[t] = p.pre[t - 1] + p.rbits[t - 1];
    const fe *twlo = inverse ? ctx->itw_lo : ctx->tw_lo;
    const fe *twhi = inverse ? ctx->itw_hi : ctx->tw_hi;
    const fe *rtab = inverse ? ctx->irtw : ctx->rtw;
    const fe scale = gf::inv(gf::from_u64(N));
    // The last pass scatters rows into natural order, so it cannot run in
    // place: passes 0..m-2 work in a context-owned temporary (one N-element
    // slice per batch item) and the last one writes `out`.  (A single pass
    // is one workgroup per item: reads finish before its writes start.)
    fe *tmp = tmp_own;
    if (m >= 2 && !tmp) {
        int rc = ntt_tmp_reserve(ctx, (sizeof(fe) << logn) * (size_t)batch, (void **)&tmp);
        if (rc) return rc;
    }
    int done_bits = 0;
    for (int t = 0; t < m; ++t) {
        const int r = p.rbits[t];                 // the pass digit's bits
        const int rd = r - (t == 0 ? fan : 0);    // its transform's
        NttPass q = p;
        q.logR = r;
        q.fan = t == 0 ? fan : 0;
        const bool last = (t == m - 1);
        q.last = last;
        q.logS = logn - done_bits - r;
        // step 2's first radix, as pass_kernel_for splits the digit
        const int r1 = r >= 7 ? 16 : r >= 5 ? 8 : r >= 3 ? 4 : 2;
        int logC = ilog2((uint64_t)tile) - rd;
        if (logC < 0) logC = 0;
        if (!last) {
            if (logC > q.logS) logC = q.logS;
        } else {
            if (m == 1) logC = 0;
            else if (logC > p.rbits[0]) logC = p.rbits[0];
            // with leaf digests wanted, narrow the tile until each thread
            // holds at most one step-2 item (R1 * C <= NTT_THREADS)
            if (leaves && m >= 2)
                while (logC > 2 && (r1 << logC) > NTT_THREADS) --logC;
        }
        q.logC = logC;
        if (t > 0) q.in_len = N;
        const uint64_t blocks = N >> (r + logC);
        const fe *src;
        fe *dst;
        if (t == 0) {
            src = in;
            q.in_tab = in_tab;
            q.in_stride = in_len;
        } else {
            src = tmp;
            q.in_tab = nullptr;
            q.in_stride = N;
        }
        if (last) {
            dst = out;
            q.out_tab = out_tab;
            q.out_stride = N;
            // fused leaf digests: whole leaves per tile (C >= 4) and at most
            // one step-2 item per thread (R1 * C <= NTT_THREADS)
            if (leaves && m >= 2 && logC >= 2 && (r1 << logC) <= NTT_THREADS) {
                q.leaves = leaves;
                if (levels_done) *levels_done = 1;
            }
        } else {
            dst = tmp;
            q.out_tab = nullptr;
            q.out_stride = N;
        }
        int rc = dispatch_pass(ctx, stream, rd, blocks, batch, src, dst, q, last, inverse, twlo, twhi, rtab, scale);
        if (rc) return rc;
        done_bits += r;
    }
    return COSNARK_OK;
}

int ntt_launch_leaves(cosnark_ctx *ctx, const fe *in, uint64_t in_len, fe *out, int logn, uint8_t *nodes,
                      int *levels_done, hipStream_t stream, fe *tmp) {
    return ntt_launch_batch(ctx, in, nullptr, in_len, out, nullptr, 1, logn, false, stream, tmp, nodes,
                            levels_done);
}

int ntt_launch(cosnark_ctx *ctx, const fe *in, uint64_t in_len, fe *out, int logn, bool inverse,
               hipStream_t stream) {
    return ntt_launch_batch(ctx, in, nullptr, in_len, out, nullptr, 1, logn, inverse, stream, nullptr);
}
