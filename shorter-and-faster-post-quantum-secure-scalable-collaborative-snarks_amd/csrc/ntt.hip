// ntt.hip -- number-theoretic transform over GF((2^61-1)^2) for gfx950.
//
// Computes the reference's fft(arr, logn, flag) (src/utils.cpp:904-964):
// X[k] = sum_j x[j] w^(jk), w = getRootOfUnity(logn), natural order in and
// out; the inverse uses w^-1 and scales by 2^-logn.  pfft
// (src/utils.cpp:1583-1673) is the same forward transform.
//
// Design: a 2^L-point transform is split into passes of radix 2^r (r <= 10).
// Pass t transforms one index digit of stride S_t = 2^(L - r_0 - .. - r_t)
// for every (prefix, suffix) pair: a workgroup stages R x C elements
// (R = 2^r points for C adjacent columns, C*R = 4096) in LDS, runs radix-2
// decimation-in-time stages there, multiplies by the inter-pass twiddle
// w_{N_t}^(lo*k) and writes back in place.  The last pass (stride 1) writes
// in natural order; its workgroup takes C rows whose leading output digit
// is consecutive, so each store is a C-element contiguous run.  Zero padding
// (Reed-Solomon encoding of n coefficients into rate*n points) is handled by
// the first pass, which reads only the n real inputs.
#include "ctx.h"

using gf::fe;

namespace {

struct NttPass {
    uint64_t in_len;    // elements of `in` that are non-zero inputs
    int logN, logR, logS, logC;
    int last;           // 1: stride-1 pass writing natural order
    int npass;
    int rbits[8];       // radix bits per pass, first pass first
    int pre[8];         // pre[t] = rbits[0] + .. + rbits[t-1]
};

__device__ __forceinline__ uint32_t brev(uint32_t x, int bits) {
    return bits == 0 ? 0u : (__brev(x) >> (32 - bits));
}

__device__ __forceinline__ fe tw28(const fe *lo, const fe *hi, uint32_t e) {
    return gf::mul(lo[e & (NTT_TW_SIZE - 1)], hi[e >> NTT_TW_BITS]);
}

__global__ __launch_bounds__(NTT_THREADS) void ntt_pass_kernel(
    const fe *in, fe *out, NttPass p, const fe *__restrict__ twlo, const fe *__restrict__ twhi,
    const fe *__restrict__ rtw, fe scale, int apply_scale) {
    __shared__ fe tile[NTT_LDS_ELEMS];
    __shared__ fe rt[512];
    const int tid = threadIdx.x;
    const int logR = p.logR, logC = p.logC;
    const int R = 1 << logR, C = 1 << logC;
    const int nelem = R << logC;

    for (int j = tid; j < R / 2; j += NTT_THREADS) rt[j] = rtw[j << (NTT_RADIX_LOG_MAX - logR)];

    const uint64_t bid = blockIdx.x;
    if (!p.last) {
        // positions: hi * (R*S) + d * S + lo, this block: fixed hi, lo in [lo0, lo0 + C)
        const int logS = p.logS;
        const uint64_t lo_blocks_log = (uint64_t)(logS - logC);
        const uint64_t hi = bid >> lo_blocks_log;
        const uint64_t lo0 = (bid & ((1ull << lo_blocks_log) - 1)) << logC;
        const uint64_t base = (hi << (logR + logS)) + lo0;
        for (int idx = tid; idx < nelem; idx += NTT_THREADS) {
            const int d = idx >> logC, c = idx & (C - 1);
            const uint64_t pos = base + ((uint64_t)d << logS) + c;
            fe v = pos < p.in_len ? in[pos] : gf::zero();
            tile[(brev(d, logR) << logC) + c] = v;
        }
        __syncthreads();
        for (int s = 0; s < logR; ++s) {
            for (int b = tid; b < nelem / 2; b += NTT_THREADS) {
                const int c = b & (C - 1), bb = b >> logC;
                const int k = bb & ((1 << s) - 1), grp = bb >> s;
                const int i0 = (grp << (s + 1)) + k, i1 = i0 + (1 << s);
                const fe w = rt[k << (logR - 1 - s)];
                const fe u = tile[(i0 << logC) + c];
                const fe v = gf::mul(tile[(i1 << logC) + c], w);
                tile[(i0 << logC) + c] = gf::add(u, v);
                tile[(i1 << logC) + c] = gf::sub(u, v);
            }
            __syncthreads();
        }
        const int logNt = logR + logS;
        for (int idx = tid; idx < nelem; idx += NTT_THREADS) {
            const int k = idx >> logC, c = idx & (C - 1);
            const uint64_t lo = lo0 + c;
            const uint32_t e = (uint32_t)((lo * (uint64_t)k) << (NTT_LOG_MAX - logNt));
            fe v = tile[idx];
            if (e) v = gf::mul(v, tw28(twlo, twhi, e));
            out[(hi << logNt) + ((uint64_t)k << logS) + lo] = v;
        }
    } else {
        // rows hi_c = rest + (c0 + c) * W0 of R contiguous elements
        const int L = p.logN;
        const int m = p.npass;
        uint64_t rest = 0, c0 = 0, W0 = 0;
        if (m >= 2) {
            const int logW0 = L - p.rbits[0] - p.rbits[m - 1];
            const int cblocks_log = p.rbits[0] - logC;
            rest = bid >> cblocks_log;
            c0 = (bid & ((1ull << cblocks_log) - 1)) << logC;
            W0 = 1ull << logW0;
        }
        for (int idx = tid; idx < nelem; idx += NTT_THREADS) {
            const int c = idx >> logR, d = idx & (R - 1);
            const uint64_t hic = rest + (c0 + c) * W0;
            const uint64_t pos = (hic << logR) + d;
            fe v = pos < p.in_len ? in[pos] : gf::zero();
            tile[(brev(d, logR) << logC) + c] = v;
        }
        __syncthreads();
        for (int s = 0; s < logR; ++s) {
            for (int b = tid; b < nelem / 2; b += NTT_THREADS) {
                const int c = b & (C - 1), bb = b >> logC;
                const int k = bb & ((1 << s) - 1), grp = bb >> s;
                const int i0 = (grp << (s + 1)) + k, i1 = i0 + (1 << s);
                const fe w = rt[k << (logR - 1 - s)];
                const fe u = tile[(i0 << logC) + c];
                const fe v = gf::mul(tile[(i1 << logC) + c], w);
                tile[(i0 << logC) + c] = gf::add(u, v);
                tile[(i1 << logC) + c] = gf::sub(u, v);
            }
            __syncthreads();
        }
        // digit reversal of `rest` (digits of passes 1..m-2)
        uint64_t nat_rest = 0;
        {
            uint64_t x = rest;
            for (int t = m - 2; t >= 1; --t) {
                const uint64_t dgt = x & ((1ull << p.rbits[t]) - 1);
                x >>= p.rbits[t];
                nat_rest += dgt << p.pre[t];
            }
        }
        const int kshift = L - p.rbits[m - 1];
        for (int idx = tid; idx < nelem; idx += NTT_THREADS) {
            const int k = idx >> logC, c = idx & (C - 1);
            fe v = tile[idx];
            if (apply_scale) v = gf::mul(v, scale);
            out[nat_rest + c0 + c + ((uint64_t)k << kshift)] = v;
        }
    }
}

}  // namespace

int ntt_launch(cosnark_ctx *ctx, const fe *in, uint64_t in_len, fe *out, int logn, bool inverse) {
    if (logn < 1 || logn > NTT_LOG_MAX) return set_err(ctx, COSNARK_E_RANGE, "ntt: logn out of range");
    // radix plan: fewest passes with r <= 10, bits spread evenly, larger first
    int m = (logn + NTT_RADIX_LOG_MAX - 1) / NTT_RADIX_LOG_MAX;
    if (logn > 16 && m < 3) m = 3;          // keep r <= 8 (C = 16) for large transforms
    if (m > logn) m = logn;
    NttPass p{};
    p.in_len = in_len;
    p.logN = logn;
    p.npass = m;
    for (int t = 0; t < m; ++t) p.rbits[t] = logn / m + (t < logn % m ? 1 : 0);
    p.pre[0] = 0;
    for (int t = 1; t < m; ++t) p.pre[t] = p.pre[t - 1] + p.rbits[t - 1];
    const fe *twlo = inverse ? ctx->itw_lo : ctx->tw_lo;
    const fe *twhi = inverse ? ctx->itw_hi : ctx->tw_hi;
    const fe *rtw = inverse ? ctx->irtw : ctx->rtw;
    fe scale = gf::inv(gf::from_u64(1ull << logn));
    int done_bits = 0;
    for (int t = 0; t < m; ++t) {
        const int r = p.rbits[t];
        NttPass q = p;
        q.logR = r;
        q.last = (t == m - 1);
        q.logS = logn - done_bits - r;
        int logC = ilog2(NTT_LDS_ELEMS) - r;
        if (!q.last) {
            if (logC > q.logS) logC = q.logS;
        } else {
            if (m == 1) logC = 0;
            else if (logC > p.rbits[0]) logC = p.rbits[0];
        }
        q.logC = logC;
        if (t > 0) q.in_len = 1ull << logn;
        const uint64_t blocks = (1ull << logn) >> (r + logC);
        hipLaunchKernelGGL(ntt_pass_kernel, dim3((unsigned)blocks), dim3(NTT_THREADS), 0, ctx->stream,
                           t == 0 ? in : out, out, q, twlo, twhi, rtw, scale,
                           (int)(inverse && q.last));
        LAUNCHCK();
        done_bits += r;
    }
    return COSNARK_OK;
}
