// Host check of the transform kernels' weak arithmetic (gf.h, dft.h):
// dif_w<N> on weak inputs against the DFT written out as a sum,
// X[n] = sum_j x[j] w_N^(jn), on random and adversarial inputs (components
// at the top of the weak range, just above p, both forms of zero); every
// output must stay weak.  Also mul_root_w against a general multiply, mul
// and mul_w with a weak second operand, canon_w at the range edges, and
// mul_pow2_w (multiply by 2^K, e.g. 1/d = 2^(61 - log d)) against mul.
#include <cstdio>
#include <random>
#include "dft.h"

using gf::fe;
static std::mt19937_64 rng(12345);
static const uint64_t W = (1ull << 61) + (1ull << 32);

static uint64_t component(int mode) {
    switch (mode) {
        case 0: return rng() % gf::P;                 // canonical
        case 1: return W - (rng() % 16);              // top of the weak range
        case 2: return gf::P + (rng() % 16);          // just above p
        case 3: return (rng() & 1) ? 0 : gf::P;       // both forms of zero
        default: return rng() % (W + 1);              // anywhere weak
    }
}

template <int N, bool INV>
static int check_dft(int trials) {
    int bad = 0;
    const fe root = gf::root_of_unity(__builtin_ctz(N));
    const fe w = INV ? gf::inv(root) : root;
    for (int t = 0; t < trials; ++t) {
        fe x[N], v[N];
        // every component draws its own kind, so small and large inputs meet in
        // one butterfly (a difference a + (kp - b) then needs its full offset)
        for (int i = 0; i < N; ++i) {
            x[i] = gf::mk(component((int)(rng() % 5)), component((int)(rng() % 5)));
            v[i] = x[i];
        }
        dft::dif_w<N, INV>(v);
        for (int i = 0; i < N; ++i) {
            if (v[i].re > W || v[i].im > W) { ++bad; continue; }
            const int n = dft::brev_c<N>(i);
            const fe wn = gf::pow(w, (uint64_t)n);
            fe X = gf::zero(), wp = gf::one();
            for (int j = 0; j < N; ++j) {
                X = gf::add(X, gf::mul(gf::canon_w(x[j]), wp));
                wp = gf::mul(wp, wn);
            }
            if (!gf::eq(gf::canon_w(v[i]), X)) ++bad;
        }
    }
    if (bad) printf("dif_w<%d,%d>: %d bad\n", N, (int)INV, bad);
    return bad;
}

template <int M, int E>
static int check_root(int trials) {
    int bad = 0;
    const fe w = gf::pow(gf::root_of_unity(__builtin_ctz(M)), (uint64_t)(((E % M) + M) % M));
    for (int t = 0; t < trials; ++t) {
        const fe x = gf::mk(component((int)(rng() % 5)), component((int)(rng() % 5)));
        const fe y = gf::mul_root_w<M, E>(x);
        if (y.re > W || y.im > W || !gf::eq(gf::canon_w(y), gf::mul(w, gf::canon_w(x)))) ++bad;
    }
    if (bad) printf("mul_root_w<%d,%d>: %d bad\n", M, E, bad);
    return bad;
}

template <int K>
static int check_pow2(int trials) {
    int bad = 0;
    const fe w = gf::mk((uint64_t)1 << K, 0);
    for (int t = 0; t < trials; ++t) {
        const fe x = gf::mk(component((int)(rng() % 5)), component((int)(rng() % 5)));
        const fe y = gf::mul_pow2_w<K>(x);
        if (y.re > W || y.im > W || !gf::eq(gf::canon_w(y), gf::mul(w, gf::canon_w(x)))) ++bad;
    }
    if (bad) printf("mul_pow2_w<%d>: %d bad\n", K, bad);
    return bad;
}

int main() {
    int bad = 0;
    bad += check_dft<2, false>(20000) + check_dft<2, true>(20000);
    bad += check_dft<4, false>(20000) + check_dft<4, true>(20000);
    bad += check_dft<8, false>(10000) + check_dft<8, true>(10000);
    bad += check_dft<16, false>(5000) + check_dft<16, true>(5000);
    bad += check_root<16, 1>(50000) + check_root<16, 2>(50000) + check_root<16, 3>(50000) + check_root<16, 4>(50000);
    bad += check_root<16, 6>(50000) + check_root<16, 8>(50000) + check_root<16, 10>(50000);
    bad += check_root<16, 12>(50000) + check_root<16, 14>(50000) + check_root<16, -9>(50000);
    bad += check_root<8, 1>(50000) + check_root<8, 3>(50000) + check_root<8, 5>(50000) + check_root<8, 7>(50000);
    bad += check_pow2<1>(50000) + check_pow2<30>(50000) + check_pow2<31>(50000) + check_pow2<32>(50000);
    bad += check_pow2<57>(50000) + check_pow2<58>(50000) + check_pow2<59>(50000) + check_pow2<60>(50000);
    int mbad = 0;
    for (int t = 0; t < 200000; ++t) {
        const fe a = gf::mk(rng() % gf::P, rng() % gf::P);
        const fe x = gf::mk(component(t % 5), component((t / 5) % 5));
        const fe y = gf::mul(a, x);
        if (y.re >= gf::P || y.im >= gf::P || !gf::eq(y, gf::mul(a, gf::canon_w(x)))) ++mbad;
    }
    if (mbad) printf("mul(canonical, weak): %d bad\n", mbad);
    bad += mbad;
    // mul_w: the same product folded once (<= p + 7); its first operand may be p
    mbad = 0;
    for (int t = 0; t < 200000; ++t) {
        fe a = gf::mk(rng() % gf::P, rng() % gf::P);
        if (t % 7 == 0) a.re = gf::P;
        if (t % 11 == 0) a.im = gf::P;
        const fe x = gf::mk(component((int)(rng() % 5)), component((int)(rng() % 5)));
        const fe y = gf::mul_w(a, x);
        if (y.re > gf::P + 7 || y.im > gf::P + 7 || !gf::eq(gf::canon_w(y), gf::mul(gf::canon_w(a), gf::canon_w(x))))
            ++mbad;
    }
    if (mbad) printf("mul_w(canonical or p, weak): %d bad\n", mbad);
    bad += mbad;
    const uint64_t edges[] = {0, gf::P - 1, gf::P, gf::P + 1, (uint64_t)1 << 61, W - 1, W};
    for (uint64_t x : edges)
        if (gf::canon_w(x) != x % gf::P) {
            printf("canon_w(%llx)\n", (unsigned long long)x);
            ++bad;
        }
    printf("weak forms: %d bad\n", bad);
    return bad != 0;
}
