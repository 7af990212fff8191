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
__device__ __forceinline__ void static_for(F &&f) {
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

}  // namespace dft
