// ntt.hip -- number-theoretic transform over GF((2^61-1)^2) for gfx950.
//
// Computes the reference's fft(arr, logn, flag) (src/utils.cpp:904-964):
// X[k] = sum_j x[j] w^(jk), w = getRootOfUnity(logn), natural order in and
// out; the inverse uses w^-1 and scales by 2^-logn.  pfft
// (src/utils.cpp:1583-1673) is the same forward transform.
//
// Design: a 2^L-point transform is split into passes of radix 2^r (r <= 8).
// Pass t transforms one index digit of stride S_t = 2^(L - r_0 - .. - r_t)
// for every (prefix, suffix) pair: a 256-thread workgroup stages R x C
// elements (R = 2^r points for C adjacent columns, C*R <= 4096) and runs the
// R-point DFT as a four-step R1 x R2 in registers (step 1, twiddle, LDS
// exchange, step 2).  The inter-pass twiddle w_{N_t}^(lo*k) is either
// applied at the store (two-pass plans) or left to the passes that read the
// output (deferred twiddles, see ntt_pass_body).  The last pass (stride 1)
// writes in natural order; its workgroup takes C rows whose leading output
// digit is consecutive, so each store is a C-element contiguous run, and it
// can hash the Merkle leaves of its tile.  Inside a pass values are weak
// (gf.h: unreduced sums, one fold per radix-4 layer); the last pass's stores
// are canonical.  Zero padding (Reed-Solomon encoding of n coefficients into
// rate*n points) is handled by the first pass, which reads only the n real
// inputs.  Where padding by E = 2^e would cost a pass, the first pass fans
// out instead: its digit is e bits wider than its transform, and since only
// the first R of its E*R inputs are non-zero, X[k0 + E k1] =
// DFT_R(x[d] w_(ER)^(d k0))[k1]; the R-point transform runs once per
// k0 < E, in one workgroup per k0 (block form) or in a loop.
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
    int fan_blk;        // fan-out with k0 from the block index (one k0 per workgroup, see below)
    int npass;
    int pass;           // this pass's index
    int defer;          // inter-pass twiddles applied by the consuming pass (see below)
    int ut_elems;       // elements of the per-tile input-twiddle table (0 if none)
    int ut_lo;          // 1: that table also holds w_N^(k_0 lo) per column (R1 x C values)
    int rbits[8];       // radix bits per pass, first pass first
    int pre[8];         // pre[t] = rbits[0] + .. + rbits[t-1]
    uint8_t *leaves;    // last pass, batch item 0: Merkle leaf digests of the output (or null)
    uint8_t *const *leaves_tab;  // last pass: item y's node array leaves_tab[y] (null: `leaves` for item 0)
    int leaf_lv;        // tree levels the last pass writes at `leaves`: 1, 3 (levels 1 and 2 too), or
                        // more than 3: every level up to the root
};

// loads and stores through global-address-space pointers (global_load /
// global_store rather than flat accesses, which also count against the LDS
// wait counter)
using gfe = __attribute__((address_space(1))) fe;
__device__ __forceinline__ fe gload(const gfe *p) {
    fe r;
    r.re = p->re;
    r.im = p->im;
    return r;
}
__device__ __forceinline__ void gstore(gfe *p, const fe &v) {
    p->re = v.re;
    p->im = v.im;
}
// A wave-uniform base plus a 32-bit per-lane element offset (< 2^28, so its
// byte offset fits 32 bits): one offset register per access instead of a
// 64-bit address per access, and the base stays in scalar registers.
using gbyte = __attribute__((address_space(1))) uint8_t;
__device__ __forceinline__ fe gload_at(const gfe *base, uint32_t off) {
    return gload((const gfe *)((const gbyte *)base + 16u * off));
}
__device__ __forceinline__ void gstore_at(gfe *base, uint32_t off, const fe &v) {
    gstore((gfe *)((gbyte *)base + 16u * off), v);
}

constexpr int Y_PAD = 1;   // one padding element per LDS row keeps 16-byte lanes off shared banks

// One pass of radix R = R1 * R2 over C columns held in LDS as a 2-D
// (four-step) transform: step 1 does R2*C in-register R1-point DFTs over
// stride-R2 inputs and the w_R^(d2 k1) twiddle, step 2 does R1*C R2-point
// DFTs, applies the inter-pass twiddle w_{N_t}^(lo k) and stores.
//
// Deferred twiddles (plans of three or more passes).  Pass t's inter-pass
// twiddle on the element at (previous frequencies, k_t, lo) is
// w_N^(k_t 2^pre[t] lo).  Instead of multiplying it in at pass t's store, a
// later pass applies it: in pass j's tile every earlier frequency k_t is
// fixed, so the pending factor on input digit d and column lo is
// w_N^(E (d S + lo)) with E = sum_t k_t 2^pre[t] (the digit-reversed hi).
// Its d-dependent part is w_N^(E S R2 d1) on step 1's inputs (a per-tile
// table of R1 values) times w_N^(E S d2), which is constant along step 1's
// DFT and so joins the w_R^(d2 k1) twiddle (a per-tile table of R values
// replacing w_R's); its lo part is left for the next pass.  The last pass's
// columns are values of k_0, so k_0's share of the last digit's factor,
// w_N^(k_0 lo), is applied one pass earlier, folded into that pass's input
// table (R1 x C values).  Per element that costs about one multiply and a
// table read where the running product cost two.
//
// SPLIT: every thread holds at most one step-1 and one step-2 item, and the
// LDS exchange moves the real words first and the imaginary words second
// through one buffer of 8-byte words.  That halves the tile's LDS (about
// 42 KiB for the 4096-element tile instead of 76), so three workgroups fit
// on a CU instead of two, for two more barriers per pass.
// DEFER (non-last passes; p.defer set): the pass stores without the
// inter-pass twiddle and, past the first pass, applies the pending ones.  The
// last pass takes p.defer at run time.
template <int R1, int R2, bool LAST, bool INV, bool SPLIT, bool DEFER>
__device__ __forceinline__ void ntt_pass_body(
    const fe *in0, fe *out0, const NttPass &p, const fe *__restrict__ twlo, const fe *__restrict__ twhi,
    const fe *__restrict__ rtab, fe scale) {
    constexpr int R = R1 * R2;
    constexpr int LOGR = (R >= 256) ? 8 : (R >= 128) ? 7 : (R >= 64) ? 6 : (R >= 32) ? 5 : (R >= 16) ? 4 : (R >= 8) ? 3 : (R >= 4) ? 2 : (R >= 2) ? 1 : 0;
    constexpr int LOGR2 = (R2 >= 16) ? 4 : (R2 >= 8) ? 3 : (R2 >= 4) ? 2 : (R2 >= 2) ? 1 : 0;
    extern __shared__ fe ntt_lds[];
    fe *wr = ntt_lds;          // step 1's twiddle for (d2, k1) at d2 * R1 + k1
    fe *ut = ntt_lds + R;      // deferred twiddles: step 1's input factors (p.ut_elems)
    fe *Y = ut + p.ut_elems;   // (R * C + R * Y_PAD) elements (SPLIT: as many 8-byte words)
    uint64_t *Yw = reinterpret_cast<uint64_t *>(Y);
    const int tid = threadIdx.x;
    const int logC = p.logC, C = 1 << logC;
    const int CP = C + Y_PAD;   // padded LDS row length
    const uint32_t by = blockIdx.y;
    const gfe *in = (const gfe *)(p.in_tab ? p.in_tab[by] : in0 + by * p.in_stride);
    gfe *out = (gfe *)(p.out_tab ? p.out_tab[by] : out0 + by * p.out_stride);

    // A fan-out pass in block form gives each of the 2^fan values of k0 a
    // workgroup of its own (block = tile * 2^fan + k0): the twists
    // w_(E R)^(d k0) are then one more per-tile factor on step 1's inputs,
    // applied like the deferred twiddles, and the split exchange needs no k0
    // loop around it.  The tile's inputs are read by each of its 2^fan
    // workgroups, which run side by side.
    const bool fan_blk = !LAST && !INV && p.fan_blk;
    const uint32_t k0_blk = fan_blk ? (uint32_t)(blockIdx.x & ((1u << p.fan) - 1)) : 0u;
    uint64_t bid = fan_blk ? (uint64_t)(blockIdx.x >> p.fan) : (uint64_t)blockIdx.x;
    // A last pass that climbs the whole tree takes its tiles in bit-reversed
    // order: the tiles whose level-2 nodes carry the longest chains (column
    // bases with the most trailing zeros) then start first instead of last.
    if (LAST && p.leaf_lv > 3) {
        const int tl = p.logN - LOGR - logC;
        bid = tl > 0 ? (uint64_t)(__brev((uint32_t)bid) >> (32 - tl)) : 0;
    }
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
    // E of the deferred twiddles: the earlier passes' frequencies held in
    // hi (non-last) or rest (last; k_0 is the column there and its share was
    // applied one pass earlier), digit-reversed
    const bool consumer = ((LAST || DEFER) && p.defer && p.pass >= 1) || (DEFER && fan_blk);
    uint64_t E = 0, k_0 = 0;
    if (consumer) {
        uint64_t x = LAST ? rest : hi;
        for (int t = LAST ? p.npass - 2 : p.pass - 1; t >= (LAST ? 1 : 0); --t) {
            const uint64_t dgt = x & ((1ull << p.rbits[t]) - 1);
            x >>= p.rbits[t];
            E += dgt << p.pre[t];
            k_0 = dgt;   // the last digit read (t = 0) is the top one
        }
    }
    if (!consumer) {
        for (int j = tid; j < R; j += NTT_THREADS) {
            const int d2 = j / R1, k1 = j % R1;
            wr[j] = rtab[((d2 * k1) & (R - 1)) << (NTT_RADIX_LOG_MAX - LOGR)];
        }
    } else {
        // w_R^(d2 k1) w_N^(E S d2), and w_N^(E S R2 d1) (times w_N^(k_0 lo) for
        // the table of R1 x C values), as powers of the order-2^28 root
        const uint32_t m28 = (1u << NTT_LOG_MAX) - 1;
        const int sN = NTT_LOG_MAX - p.logN;
        // w_N^(E S) as a power of w_28; for a fan-out block, w_(E R)^(k0)
        const uint32_t es = fan_blk ? (k0_blk << (NTT_LOG_MAX - LOGR - p.fan)) : (uint32_t)((E << logS) << sN);
        // Each entry is a product of factors that many lanes of a wave share
        // (w_N^(E S d2): R1 lanes; w_N^(E S R2 d1): C lanes) or that come from
        // the small radix table, so the table loads touch few cache lines.
        for (int j = tid; j < R; j += NTT_THREADS) {
            const uint32_t d2 = j / R1, k1 = j % R1;
            const fe v = tw28(twlo, twhi, (d2 * es) & m28);
            wr[j] = gf::mul(v, rtab[((d2 * k1) & (R - 1)) << (NTT_RADIX_LOG_MAX - LOGR)]);
        }
        for (int j = tid; j < p.ut_elems; j += NTT_THREADS) {
            const uint32_t d1 = p.ut_lo ? j >> logC : j;
            const fe u = tw28(twlo, twhi, (d1 * R2 * es) & m28);
            ut[j] = p.ut_lo ? gf::mul(u, tw28(twlo, twhi, (uint32_t)((k_0 * (lo0 + (j & (C - 1)))) << sN) & m28))
                            : u;
        }
    }
    __syncthreads();

    // fan-out (first pass of a padded forward transform): the pass digit has
    // LOGR + fan bits, and the steps below run once per k0 < 2^fan, or once
    // for this block's k0 in block form (the only fan-out the split exchange
    // takes: the k0 loop around it costs the split kernel half its occupancy
    // in registers)
    const int fan = (LAST || INV || (SPLIT && !fan_blk)) ? 0 : p.fan;
    const int nk0 = fan_blk ? 1 : (1 << fan);
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
    // The split form gives each thread up to 16 elements in each step:
    // 16 / R1 step-1 items and 16 / R2 step-2 items (one each for 16 x 16,
    // one and two for 16 x 8, two and two for 8 x 8, ...), so a 4096-element
    // tile keeps every thread busy in both steps.
    constexpr int MI1 = (SPLIT && R1 < 16) ? 16 / R1 : 1;
    constexpr int MI2 = (SPLIT && R2 < 16) ? 16 / R2 : 1;
    uint8_t *const lvp = !LAST ? nullptr : p.leaves_tab ? p.leaves_tab[by] : by == 0 ? p.leaves : nullptr;
    const bool leaves_on = LAST && lvp && items2 <= MI2 * NTT_THREADS;
    fe vv[MI2][R2];
    int my_c[MI2] = {}, my_k1[MI2] = {};
    // zero padding (only the first pass reads fewer than 2^logN inputs)
    const bool padded = p.in_len < (1ull << p.logN);
    // fan-out: w_(E R)^(d k0) for the current k0, d < R (R * CP is even)
    fe *tz = SPLIT ? Y + (size_t)R * CP / 2 : Y + (size_t)R * CP;
    for (int kk = 0; kk < nk0; ++kk) {
    const int k0 = fan_blk ? (int)k0_blk : kk;
    if (kk > 0) __syncthreads();   // the previous k0's step 2 is done with Y

    // ---- step 1: R1-point DFTs over d1 for each (d2, c) ----
    // item -> x[i] = the DFT output k1 = brev(i) after the w_R^(d2 k1)
    // twiddle, bound for exchange row k1 * R2 + d2, column c (*off = d2 * CP + c)
    auto step1 = [&](int it, fe *x, int *off) {
        int c, d2;
        uint64_t base;
        uint64_t ubase;      // the workgroup's part of base (wave-uniform)
        uint32_t loff;       // the lane's part of the input offset at d1 = 0
        if (!LAST) {
            c = it & (C - 1);
            d2 = it >> logC;
            ubase = (hi << logNt) + lo0;
            base = ubase + c;
            loff = (uint32_t)c + ((uint32_t)d2 << logS);
        } else {
            d2 = it & (R2 - 1);
            c = it >> LOGR2;
            ubase = (rest + c0 * W0) << LOGR;
            base = (rest + (c0 + c) * W0) << LOGR;
            loff = (uint32_t)((c * W0) << LOGR) + (uint32_t)d2;
        }
        if (padded) {
#pragma unroll
            for (int d1 = 0; d1 < R1; ++d1) {
                const uint64_t d = (uint64_t)(R2 * d1 + d2);
                const uint64_t pos = base + (LAST ? d : (d << logS));
                x[d1] = pos < p.in_len ? gload(in + pos) : gf::zero();
                // twist by w_(E R)^(d k0) (block form: in the deferred tables); inputs are canonical
                if (!fan_blk && k0 > 0) x[d1] = gf::mul_w(tz[d], x[d1]);
            }
        } else {
#pragma unroll
            for (int d1 = 0; d1 < R1; ++d1) {
                const uint64_t ud = (uint64_t)(R2 * d1);
                x[d1] = gload_at(in + ubase + (LAST ? ud : (ud << logS)), loff);
            }
        }
        // deferred twiddles (inputs weak <= p + 7 from the previous pass's
        // store), in a block of their own after every load is issued
        if (consumer) {
            if (p.ut_lo) {
#pragma unroll
                for (int d1 = 0; d1 < R1; ++d1) x[d1] = gf::mul_w(ut[(d1 << logC) + c], x[d1]);
            } else {
#pragma unroll
                for (int d1 = 1; d1 < R1; ++d1) x[d1] = gf::mul_w(ut[d1], x[d1]);
            }
        }
        dif_w<R1, INV>(x);
        static_for<0, R1>([&](auto ii) {
            constexpr int i = decltype(ii)::value;
            constexpr int k1 = brev_c<R1>(i);
            // (where a pass may consume deferred twiddles, column k1 = 0 is
            // multiplied too: by 1 unless it does)
            if ((k1 != 0 || LAST || DEFER) && d2 != 0) x[i] = gf::mul_w(wr[d2 * R1 + k1], x[i]);
        });
        *off = d2 * CP + c;
    };
    // the next k0's twists, while step 2 runs (called once step 1 is done
    // with tz; the barrier at the top of the next k0 orders these writes
    // before its reads)
    auto next_twists = [&]() {
        if (kk + 1 < nk0)
            for (int d = tid; d < R; d += NTT_THREADS)
                tz[d] = tw28(twlo, twhi, (uint32_t)(d * (k0 + 1)) << (NTT_LOG_MAX - LOGR - fan));
    };
    // ---- step 2: R2-point DFTs over d2 for each (k1, c), twiddle, store ----
    // (v holds the item's exchange column: v[d2] = row k1 * R2 + d2, column c;
    // q: the thread's item slot)
    auto step2 = [&](int it, fe *v, int q) {
        const int c = it & (C - 1), k1 = it >> logC;
        my_c[q] = c;
        my_k1[q] = k1;
        dif_w<R2, INV>(v);
        if constexpr (!LAST && DEFER) {
            // the inter-pass twiddle is left to the next passes; weak stores
            static_for<0, R2>([&](auto kk) {
                constexpr int k2 = decltype(kk)::value;
                constexpr int i = brev_c<R2>(k2);
                const uint64_t uk = (uint64_t)k0 + ((uint64_t)(R1 * k2) << fan);
                gstore_at(out + (hi << logNt) + (uk << logS) + lo0, (uint32_t)c + ((uint32_t)k1 << (fan + logS)), v[i]);
            });
        } else if constexpr (!LAST) {
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
                const uint64_t uk = (uint64_t)k0 + ((uint64_t)(R1 * k2) << fan);
                // weak: read by the next pass
                gstore_at(out + (hi << logNt) + (uk << logS) + lo0, (uint32_t)c + ((uint32_t)k1 << (fan + logS)),
                          gf::mul_w(t, v[i]));
                if (k2 + 1 < R2) t = gf::mul(t, step);
            });
        } else {
            static_for<0, R2>([&](auto ii) {
                constexpr int i = decltype(ii)::value;
                constexpr int k2 = brev_c<R2>(i);
                const uint64_t uk = (uint64_t)(R1 * k2);
                v[i] = INV ? gf::mul(scale, v[i]) : gf::canon_w(v[i]);
                gstore_at(out + nat_rest + c0 + (uk << kshift), (uint32_t)c + ((uint32_t)k1 << kshift), v[i]);
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
            for (int d2 = 0; d2 < R2; ++d2) vv[0][d2] = Y[(k1 * R2 + d2) * CP + c];
            step2(it, vv[0], 0);
        }
    } else {
        // at most MI1 step-1 and MI2 step-2 items per thread
        // (items1 <= MI1 * NTT_THREADS, items2 <= MI2 * NTT_THREADS)
        fe x[MI1][R1];
        int off[MI1] = {};
        static_for<0, MI1>([&](auto qq) {
            constexpr int q = decltype(qq)::value;
            if (tid + q * NTT_THREADS < items1) step1(tid + q * NTT_THREADS, x[q], &off[q]);
        });
        static_for<0, MI1>([&](auto qq) {
            constexpr int q = decltype(qq)::value;
            if (tid + q * NTT_THREADS < items1)
                static_for<0, R1>([&](auto ii) {
                    constexpr int i = decltype(ii)::value;
                    Yw[brev_c<R1>(i) * R2 * CP + off[q]] = x[q][i].re;
                });
        });
        __syncthreads();
        next_twists();
        static_for<0, MI2>([&](auto qq) {
            constexpr int q = decltype(qq)::value;
            const int it = tid + q * NTT_THREADS;
            if (it < items2) {
                const int c2 = it & (C - 1), k12 = it >> logC;
#pragma unroll
                for (int d2 = 0; d2 < R2; ++d2) vv[q][d2].re = Yw[(k12 * R2 + d2) * CP + c2];
            }
        });
        __syncthreads();
        static_for<0, MI1>([&](auto qq) {
            constexpr int q = decltype(qq)::value;
            if (tid + q * NTT_THREADS < items1)
                static_for<0, R1>([&](auto ii) {
                    constexpr int i = decltype(ii)::value;
                    Yw[brev_c<R1>(i) * R2 * CP + off[q]] = x[q][i].im;
                });
        });
        __syncthreads();
        static_for<0, MI2>([&](auto qq) {
            constexpr int q = decltype(qq)::value;
            const int it = tid + q * NTT_THREADS;
            if (it < items2) {
                const int c2 = it & (C - 1), k12 = it >> logC;
#pragma unroll
                for (int d2 = 0; d2 < R2; ++d2) vv[q][d2].im = Yw[(k12 * R2 + d2) * CP + c2];
                step2(it, vv[q], q);
            }
        });
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
        //
        // With p.leaf_lv >= 3 the pass also writes tree levels 1 and 2.  The
        // tree hashes a node's left child twice (src/merkle_tree.cpp:287-292),
        // so level-1 node j is BLAKE3(leaf 2j || leaf 2j) and level-2 node i
        // depends on leaf 4i alone.  The host sets it only where a half holds
        // one unit of two consecutive leaves per thread (R C = 16 x 256), a
        // row holds whole groups of four leaves (C a multiple of 16) and rows
        // start at multiples of four leaves: a thread hashes its two leaves
        // and the level-1 node of the first, and after the second half the
        // level-2 nodes go one per thread, the odd units' lanes taking the
        // second half's from their even neighbour by DPP.
        constexpr int NH = (SPLIT && R2 >= 2) ? 2 : 1, RH = R / NH;
        const int qpr = C / 4, nleaf = (RH * C) / 4;
        const bool lv3 = NH == 2 && p.leaf_lv >= 3;
        const uint64_t n_lf = 1ull << (p.logN - 2);    // leaves: level 1 starts there, level 2 at 3/2 of it
        uint32_t carry[8];                              // first half: the unit's level-1 digest
        static_for<0, NH>([&](auto hh) {
            constexpr int h = decltype(hh)::value;
            __syncthreads();   // the exchange buffer is free (step-2 reads, or the previous half's hashes)
            static_for<0, MI2>([&](auto qq) {
                constexpr int q = decltype(qq)::value;
                if (tid + q * NTT_THREADS < items2) {
                    static_for<0, R2>([&](auto ii) {
                        constexpr int i = decltype(ii)::value;
                        constexpr int k2 = brev_c<R2>(i);
                        const int hc = decltype(hh)::value;
                        if (k2 / (R2 / NH) == hc) Y[(my_k1[q] + R1 * k2 - hc * RH) * CP + my_c[q]] = vv[q][i];
                    });
                }
            });
            __syncthreads();
            auto leaf_msg = [&](const fe *src, uint32_t *m) {
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    m[4 * e] = (uint32_t)src[e].re;
                    m[4 * e + 1] = (uint32_t)(src[e].re >> 32);
                    m[4 * e + 2] = (uint32_t)src[e].im;
                    m[4 * e + 3] = (uint32_t)(src[e].im >> 32);
                }
            };
            auto self_pair = [](const uint32_t *d, uint32_t *out) {
                uint32_t m[16];
#pragma unroll
                for (int j = 0; j < 8; ++j) { m[j] = d[j]; m[8 + j] = d[j]; }
                b3::compress64(m, out);
            };
            if (lv3) {
                const int upr = qpr / 2, k = tid / upr, up = tid % upr;
                const fe *src = Y + k * CP + 8 * up;
                uint32_t m[16], h0[8], h1[8], l1[8];
                leaf_msg(src, m);
                b3::compress64(m, h0);
                leaf_msg(src + 4, m);
                b3::compress64(m, h1);
                const uint64_t li = (nat_rest + c0 + 8 * up + ((uint64_t)(k + h * RH) << kshift)) >> 2;
                uint4 *dst = reinterpret_cast<uint4 *>(lvp + 32 * li);
                dst[0] = make_uint4(h0[0], h0[1], h0[2], h0[3]);
                dst[1] = make_uint4(h0[4], h0[5], h0[6], h0[7]);
                dst[2] = make_uint4(h1[0], h1[1], h1[2], h1[3]);
                dst[3] = make_uint4(h1[4], h1[5], h1[6], h1[7]);
                self_pair(h0, l1);
                uint4 *d1 = reinterpret_cast<uint4 *>(lvp + 32 * (n_lf + (li >> 1)));
                d1[0] = make_uint4(l1[0], l1[1], l1[2], l1[3]);
                d1[1] = make_uint4(l1[4], l1[5], l1[6], l1[7]);
                if (h == 0) {
#pragma unroll
                    for (int j = 0; j < 8; ++j) carry[j] = l1[j];
                } else {
                    // lane 2i + 1 reads lane 2i's level-1 digest (quad_perm 0, 0, 2, 2)
                    const bool odd = (up & 1) != 0;
                    uint32_t in[8], l2[8];
#pragma unroll
                    for (int j = 0; j < 8; ++j) {
                        const uint32_t nb = (uint32_t)__builtin_amdgcn_mov_dpp((int)l1[j], 0xA0, 0xF, 0xF, false);
                        in[j] = odd ? nb : carry[j];
                    }
                    self_pair(in, l2);
                    const int row = odd ? k + RH : k;
                    const uint64_t i2 = (nat_rest + c0 + 8 * (up & ~1) + ((uint64_t)row << kshift)) >> 4;
                    uint4 *d2 = reinterpret_cast<uint4 *>(lvp + 32 * (n_lf + (n_lf >> 1) + i2));
                    d2[0] = make_uint4(l2[0], l2[1], l2[2], l2[3]);
                    d2[1] = make_uint4(l2[4], l2[5], l2[6], l2[7]);
                    if (p.leaf_lv > 3) {
                        // the rest of the tree: node j of level 2 + s is the
                        // chain from level-2 node j 2^s, so the lane holding
                        // level-2 node i2 climbs while i2 has s trailing
                        // zeros, up to the root (lanes leave the loop at
                        // their own level; nothing below waits on them)
                        uint64_t off = n_lf + (n_lf >> 1), cnt = n_lf >> 2, j = i2;
                        while (cnt > 1 && (j & 1) == 0) {
                            off += cnt;
                            cnt >>= 1;
                            j >>= 1;
                            uint32_t nx[8];
                            self_pair(l2, nx);
#pragma unroll
                            for (int w = 0; w < 8; ++w) l2[w] = nx[w];
                            uint4 *dn = reinterpret_cast<uint4 *>(lvp + 32 * (off + j));
                            dn[0] = make_uint4(l2[0], l2[1], l2[2], l2[3]);
                            dn[1] = make_uint4(l2[4], l2[5], l2[6], l2[7]);
                        }
                    }
                }
            } else {
                for (int l = tid; l < nleaf; l += NTT_THREADS) {
                    const int k = l / qpr, q = l % qpr;
                    uint32_t m[16], hv[8];
                    leaf_msg(Y + k * CP + 4 * q, m);
                    b3::compress64(m, hv);
                    const uint64_t pos = nat_rest + c0 + 4 * q + ((uint64_t)(k + h * RH) << kshift);
                    uint4 *dst = reinterpret_cast<uint4 *>(lvp + 32 * (pos >> 2));
                    dst[0] = make_uint4(hv[0], hv[1], hv[2], hv[3]);
                    dst[1] = make_uint4(hv[4], hv[5], hv[6], hv[7]);
                }
            }
        });
    }
}

// The split form's LDS lets three workgroups share a CU, so its kernels are
// held to the registers of three waves per SIMD (168 VGPRs); the whole-element
// form's LDS allows two.
template <int R1, int R2, bool LAST, bool INV, bool DEFER>
__global__ __launch_bounds__(NTT_THREADS, 3) void ntt_pass_split_kernel(
    const fe *in0, fe *out0, NttPass p, const fe *__restrict__ twlo, const fe *__restrict__ twhi,
    const fe *__restrict__ rtab, fe scale) {
    ntt_pass_body<R1, R2, LAST, INV, true, DEFER>(in0, out0, p, twlo, twhi, rtab, scale);
}

template <int R1, int R2, bool LAST, bool INV, bool DEFER>
__global__ __launch_bounds__(NTT_THREADS, 2) void ntt_pass_kernel(
    const fe *in0, fe *out0, NttPass p, const fe *__restrict__ twlo, const fe *__restrict__ twhi,
    const fe *__restrict__ rtab, fe scale) {
    ntt_pass_body<R1, R2, LAST, INV, false, DEFER>(in0, out0, p, twlo, twhi, rtab, scale);
}

using PassKernel = void (*)(const fe *, fe *, NttPass, const fe *, const fe *, const fe *, fe);

template <int R1, int R2, bool LAST, bool INV, bool SPLIT, bool DEFER>
constexpr PassKernel pass_fn() {
    if constexpr (SPLIT) return ntt_pass_split_kernel<R1, R2, LAST, INV, DEFER>;
    else return ntt_pass_kernel<R1, R2, LAST, INV, DEFER>;
}

template <int R1, int R2, bool SPLIT>
PassKernel pass_kernel_s(bool last, bool inverse, bool defer) {
    if (last) return inverse ? pass_fn<R1, R2, true, true, SPLIT, false>() : pass_fn<R1, R2, true, false, SPLIT, false>();
    if (defer) return inverse ? pass_fn<R1, R2, false, true, SPLIT, true>() : pass_fn<R1, R2, false, false, SPLIT, true>();
    return inverse ? pass_fn<R1, R2, false, true, SPLIT, false>() : pass_fn<R1, R2, false, false, SPLIT, false>();
}

template <int R1, int R2>
PassKernel pass_kernel(bool last, bool inverse, bool split, bool defer) {
    return split ? pass_kernel_s<R1, R2, true>(last, inverse, defer)
                 : pass_kernel_s<R1, R2, false>(last, inverse, defer);
}

PassKernel pass_kernel_for(int r, bool last, bool inverse, bool split, bool defer) {
    switch (r) {
        case 1: return pass_kernel<2, 1>(last, inverse, split, defer);
        case 2: return pass_kernel<2, 2>(last, inverse, split, defer);
        case 3: return pass_kernel<4, 2>(last, inverse, split, defer);
        case 4: return pass_kernel<4, 4>(last, inverse, split, defer);
        case 5: return pass_kernel<8, 4>(last, inverse, split, defer);
        case 6: return pass_kernel<8, 8>(last, inverse, split, defer);
        case 7: return pass_kernel<16, 8>(last, inverse, split, defer);
        case 8: return pass_kernel<16, 16>(last, inverse, split, defer);
        default: return nullptr;
    }
}

// the split exchange needs at most one item of each step per thread (up to
// 16 elements per thread in each step: 16 / R1 step-1 items, 16 / R2 step-2
// items)
bool pass_split(const cosnark_ctx *ctx, int R1, int R2, const NttPass &q) {
    return ctx->ntt_split && (!q.fan || q.fan_blk) && R2 >= 2 && ((R2 << q.logC) <= (16 / R1) * NTT_THREADS) &&
           ((R1 << q.logC) <= (16 / R2) * NTT_THREADS);
}

template <int R1, int R2>
int launch_pass(cosnark_ctx *ctx, hipStream_t stream, uint64_t blocks, int batch, const fe *in, fe *out,
                const NttPass &q, bool last, bool inverse, const fe *twlo, const fe *twhi,
                const fe *rtab, fe scale) {
    const bool split = pass_split(ctx, R1, R2, q);
    PassKernel k = pass_kernel<R1, R2>(last, inverse, split, q.defer != 0);
    // radix table, padded tile (split: 8-byte words), and for a fan-out pass the R twists
    const size_t R = (size_t)R1 * R2, CP = (1u << q.logC) + Y_PAD;
    const size_t lds = sizeof(fe) * (R + (size_t)q.ut_elems + (q.fan && !q.fan_blk ? R : 0)) + (split ? 8 : 16) * R * CP;
    hipLaunchKernelGGL(k, dim3((unsigned)blocks, (unsigned)batch), dim3(NTT_THREADS), lds, stream, in, out, q, twlo,
                       twhi, rtab, scale);
    LAUNCHCK();
    return COSNARK_OK;
}

// step-2 items one thread may hold in a pass whose transform has rd bits:
// 16 / R2 where the split exchange applies (rd >= 2), else one
int items2_per_thread(const cosnark_ctx *ctx, int rd) {
    if (!ctx->ntt_split || rd < 2) return 1;
    const int r1 = rd >= 7 ? 16 : rd >= 5 ? 8 : rd >= 3 ? 4 : 2;
    return 16 / (1 << (rd - ilog2((uint64_t)r1)));
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
    // pass's twists or a deferred-twiddle table of up to 1024 values (~92 KiB)
    const int bytes = (int)(sizeof(fe) * (NTT_LDS_ELEMS + 256 * (Y_PAD + 2) + 1024));
    for (int r = 1; r <= NTT_PASS_LOG_MAX; ++r)
        for (int last = 0; last < 2; ++last)
            for (int inv = 0; inv < 2; ++inv)
                for (int split = 0; split < 2; ++split)
                    for (int defer = 0; defer < 2; ++defer)
                        HIPCK(hipFuncSetAttribute((const void *)pass_kernel_for(r, last, inv, split, defer),
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
                     uint8_t *leaves, int *levels_done, int max_levels, uint8_t *const *leaves_tab, bool dry) {
    if (levels_done) *levels_done = 0;
    const bool want_leaves = leaves || leaves_tab;
    const int lv_cap = max_levels < ctx->ntt_leaf_levels ? max_levels : ctx->ntt_leaf_levels;
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
    // deferred inter-pass twiddles: plans of three or more passes whose
    // transforms all have at least 4 bits (step 2 radix >= 4, so the table of
    // the pass before the last stays within 1024 values)
    bool defer = ctx->ntt_defer && m >= 3;
    for (int t = 0; t < m; ++t)
        if (p.rbits[t] - (t == 0 ? fan : 0) < 4) defer = false;
    p.defer = defer ? 1 : 0;
    // The last pass scatters rows into natural order, so it cannot run in
    // place: passes 0..m-2 work in a context-owned temporary (one N-element
    // slice per batch item) and the last one writes `out`.  (A single pass
    // is one workgroup per item: reads finish before its writes start.)
    fe *tmp = tmp_own;
    if (m >= 2 && !tmp && !dry) {
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
            // holds at most one step-2 item (R1 * C <= NTT_THREADS; two in
            // the split form where R1 = 2 R2)
            if (want_leaves && m >= 2)
                while (logC > 2 && (r1 << logC) > items2_per_thread(ctx, rd) * NTT_THREADS) --logC;
        }
        q.logC = logC;
        q.pass = t;
        q.ut_lo = (defer && t >= 1 && t == m - 2) ? 1 : 0;
        q.ut_elems = (defer && t >= 1) ? (q.ut_lo ? (r1 << logC) : r1) : 0;
        // a fan-out first pass takes the block form (one k0 per workgroup,
        // the twists in the deferred tables) where the split exchange fits
        // (the step radices of the transform's rd bits, as pass_kernel_for splits them)
        const int t1 = rd >= 7 ? 16 : rd >= 5 ? 8 : rd >= 3 ? 4 : 2, t2 = 1 << (rd - ilog2((uint64_t)t1));
        q.fan_blk = (t == 0 && fan && defer && ctx->ntt_split && t2 >= 2 &&
                     (t1 << logC) <= items2_per_thread(ctx, rd) * NTT_THREADS &&
                     (t2 << logC) <= (16 / t1) * NTT_THREADS) ? 1 : 0;
        if (q.fan_blk) q.ut_elems = t1;
        if (t > 0) q.in_len = N;
        const uint64_t blocks = (N >> (r + logC)) << (q.fan_blk ? fan : 0);
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
            // one step-2 item per thread (two in the split form where R1 =
            // 2 R2; the pass then takes that form, see launch_pass)
            if (want_leaves && m >= 2 && logC >= 2 && (r1 << logC) <= items2_per_thread(ctx, rd) * NTT_THREADS) {
                q.leaves = leaves;
                q.leaves_tab = leaves_tab;
                q.leaf_lv = 1;
                // levels 1 and 2 as well (dup-left trees): the split form with
                // one two-leaf unit per thread in each half (R C = 4096), whole
                // four-leaf groups per row (C >= 16) and rows at multiples of
                // four leaves (logn - r >= 4)
                const int r2 = 1 << (r - ilog2((uint64_t)r1));
                if (lv_cap >= 3 && pass_split(ctx, r1, r2, q) && ((r1 * r2) << logC) == 16 * NTT_THREADS &&
                    logC >= 4 && logn - r >= 4)
                    q.leaf_lv = lv_cap > 3 ? logn - 1 : 3;
                if (levels_done) *levels_done = q.leaf_lv;
            }
        } else {
            dst = tmp;
            q.out_tab = nullptr;
            q.out_stride = N;
        }
        if (dry) {
            done_bits += r;
            continue;
        }
        int rc = dispatch_pass(ctx, stream, rd, blocks, batch, src, dst, q, last, inverse, twlo, twhi, rtab, scale);
        if (rc) return rc;
        done_bits += r;
    }
    return COSNARK_OK;
}

int ntt_launch_leaves(cosnark_ctx *ctx, const fe *in, uint64_t in_len, fe *out, int logn, uint8_t *nodes,
                      int *levels_done, hipStream_t stream, fe *tmp, int max_levels) {
    return ntt_launch_batch(ctx, in, nullptr, in_len, out, nullptr, 1, logn, false, stream, tmp, nodes,
                            levels_done, max_levels);
}

int ntt_launch(cosnark_ctx *ctx, const fe *in, uint64_t in_len, fe *out, int logn, bool inverse,
               hipStream_t stream) {
    return ntt_launch_batch(ctx, in, nullptr, in_len, out, nullptr, 1, logn, inverse, stream, nullptr);
}
