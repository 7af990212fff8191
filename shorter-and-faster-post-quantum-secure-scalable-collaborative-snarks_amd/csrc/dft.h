// dft.h -- in-register small DFTs and compile-time helpers for the
// transform kernels (NTT passes, coPCS block transforms).
#pragma once
#include <type_traits>
#include "ctx.h"

namespace dft {

using gf::fe;

template <int N>
constexpr int brev_c(int x) {
    int r = 0;
    for (int b = 1, q = N >> 1; b < N; b <<= 1, q >>= 1)
        if (x & b) r |= q;
    return r;
}

__device__ __forceinline__ fe tw28(const fe *lo, const fe *hi, uint32_t e) {
    return gf::mul(lo[e & (NTT_TW_SIZE - 1)], hi[e >> NTT_TW_BITS]);
}

template <int B, int E, class F>
__host__ __device__ __forceinline__ void static_for(F &&f) {
    if constexpr (B < E) {
        f(std::integral_constant<int, B>{});
        static_for<B + 1, E>(f);
    }
}

// In-register DFT of N points, decimation in frequency: natural order in,
// bit-reversed order out (v[i] = X[brev(i)]).  All twiddles are compile-time
// constants; powers of w_8 and +-I cost no general multiply.
template <int N, bool INV>
__device__ __forceinline__ void dif(fe *v) {
    if constexpr (N > 1) {
        static_for<0, N / 2>([&](auto kk) {
            constexpr int k = decltype(kk)::value;
            const fe a = v[k], b = v[k + N / 2];
            v[k] = gf::add(a, b);
            v[k + N / 2] = gf::mul_root<N, INV ? -k : k>(gf::sub(a, b));
        });
        dif<N / 2, INV>(v);
        dif<N / 2, INV>(v + N / 2);
    }
}

// The same DFT on weak inputs (gf.h), weak outputs: radix-4 layers whose
// butterflies add and subtract without reducing and fold each output once,
// with the w_N^(k n1) twiddles between layers.  For N = 4M, the layer takes
// x[k + M j] (j < 4) to Y_k[n1] = sum_j x[k + M j] w_4^(j n1) and places
// w_N^(k n1) Y_k[n1] at k + M brev4(n1); the M-point DFT of block b then
// yields X[brev4(b) + 4 brev_M(i)], i.e. v[i] = X[brev_N(i)] overall.
// Bounds (W = 2^61 + 2^32): a layer's sums stay below 4W, its differences
// a + (2p - b) and a + (4p - b) below 2^64.
template <bool INV>
__host__ __device__ __forceinline__ void bfly4_w(fe &x0, fe &x1, fe &x2, fe &x3) {
    using gf::P2;
    using gf::P4;
    using gf::fold;
    const uint64_t ar = x0.re + x2.re, ai = x0.im + x2.im;            // x0 + x2
    const uint64_t cr = x0.re + (P2 - x2.re), ci = x0.im + (P2 - x2.im);  // x0 - x2
    const uint64_t br = x1.re + x3.re, bi = x1.im + x3.im;            // x1 + x3
    // d = (x1 - x3) * w_4, w_4 = -I (inverse: I)
    const uint64_t dr = INV ? x3.im + (P2 - x1.im) : x1.im + (P2 - x3.im);
    const uint64_t di = INV ? x1.re + (P2 - x3.re) : x3.re + (P2 - x1.re);
    x0 = gf::mk(fold(ar + br), fold(ai + bi));                        // Y0
    x2 = gf::mk(fold(ar + (P4 - br)), fold(ai + (P4 - bi)));          // Y2
    x1 = gf::mk(fold(cr + dr), fold(ci + di));                        // Y1
    x3 = gf::mk(fold(cr + (P4 - dr)), fold(ci + (P4 - di)));          // Y3
}

template <int N, bool INV>
__host__ __device__ __forceinline__ void dif_w(fe *v) {
    if constexpr (N == 2) {
        const fe a = v[0], b = v[1];
        v[0] = gf::mk(gf::fold(a.re + b.re), gf::fold(a.im + b.im));
        v[1] = gf::mk(gf::fold(a.re + (gf::P2 - b.re)), gf::fold(a.im + (gf::P2 - b.im)));
    } else if constexpr (N >= 4) {
        constexpr int M = N / 4;
        static_for<0, M>([&](auto kk) {
            constexpr int k = decltype(kk)::value;
            fe y0 = v[k], y1 = v[k + M], y2 = v[k + 2 * M], y3 = v[k + 3 * M];
            bfly4_w<INV>(y0, y1, y2, y3);    // now Y0, Y1, Y2, Y3
            v[k] = y0;
            v[k + 2 * M] = gf::mul_root_w<N, INV ? -k : k>(y1);
            v[k + M] = gf::mul_root_w<N, INV ? -2 * k : 2 * k>(y2);
            v[k + 3 * M] = gf::mul_root_w<N, INV ? -3 * k : 3 * k>(y3);
        });
        if constexpr (M > 1) {
            dif_w<M, INV>(v);
            dif_w<M, INV>(v + M);
            dif_w<M, INV>(v + 2 * M);
            dif_w<M, INV>(v + 3 * M);
        }
    }
}

}  // namespace dft
