// copcs.hip -- rank-local kernels of the coPCS commitment and opening:
// the coset encode of the committed rows, the block sumcheck's two phases
// and RS_multifold5.  (The column tree of distributed_MT is the column mode
// of the Merkle kernel.)
//
// Encode (src/coPCS.cpp:131-147): per row, iNTT_L, times w^-j with
// w = root(1 + log2 rate_offset + log2 L), zero-pad to rate*L, NTT.  The
// twist is one elementwise kernel between two library transforms.
//
// Block sumcheck.  pcs_sumcheck_phase1 (src/coPCS.cpp:547-569) and the
// unmasked block sum of quadratic_coSumcheck_phase1 (src/coSumcheck.cpp:
// 150-180) compute, over blocks j of d elements of p1 and p2,
//     H[k] = sum_j P_j(w_2d^k) Q_j(w_2d^k),   k < 2d,
// with P_j, Q_j the interpolants of the blocks on the d-th roots (iNTT_d,
// pad, NTT_2d).  At even k = 2i an interpolant takes its block's own value
// i, so H[2i] = sum_j p1[jd+i] p2[jd+i] needs no transform.  The odd points
// w_2d^(2i+1) come from one d-point transform of the twisted coefficients:
//     Y = NTT_d(t . iNTT_d(x)),  t_s = w_2d^s,  Y_i = d P(w_2d^(2i+1))
// (both transforms unscaled; the 1/d^2 is applied once to the sums).  The
// reference skips blocks whose p1 part is all zero (src/coSumcheck.cpp:
// 161-167); their products are zero, so the sums agree without that test.
//   d <= 32: one lane per (block, vector), the transforms in registers with
//   compile-time twiddles; the two lanes of a block meet in LDS.
//   64 <= d <= 2048: one workgroup per block, radix-2 stages in LDS.
// Both write per-workgroup column sums; a small kernel adds them up.
//
// Phase 2 (src/coPCS.cpp:571-587, src/coSumcheck.cpp:201-213):
// p[j] = sum_k L_k p[jd+k] with L = compute_lagrange_coeff(r, d), the
// inverse transform of (1, r, ..., r^(d-1)) (src/utils.cpp:169-173).
//
// RS_multifold5 (src/coPCS_utils.cpp:386-437).  Its partial codewords after
// log2(fd) butterfly levels are, for output i < n/fd and x_s = cw[i + s n/fd],
//     part[i][k] = w_n^(-ik) X_k,   X = inverse DFT_fd(x) (unscaled),
// and out[i] = (1/fd) sum_k a_k part[i][k]: one fd-point transform per
// output and a Horner pass in w_n^-i.
#include <algorithm>
#include <cstring>
#include <vector>
#include "ctx.h"
#include "transcript.h"
#include "dft.h"

using gf::fe;

namespace {

constexpr int BSC_TB = 64;              // blocks per tile, small d
constexpr int BSC_SMALL_MAX = 32;
constexpr int BSC_D_MAX = 2048;
constexpr int BSC3_TILE_LOG = 9;        // cubic form: elements of each vector per tile, d <= 512
constexpr int BSC3_D_MAX = 1024;
constexpr int BSC3_SMALL_MAX = 16;      // cubic form: register transforms up to this d
constexpr int BSC_MAX_WG = 1024;
constexpr int CO_THREADS = 256;
constexpr int MF_SMALL_MAX = 32;
constexpr int MF_FD_MAX = 4096;

// ---- Encode: x[j] *= w^-j, w = root(k), e = j << (28 - k) ----------------
__global__ __launch_bounds__(CO_THREADS) void coset_twist_kernel(fe *x, uint64_t L, int shift,
                                                                 const fe *__restrict__ itlo,
                                                                 const fe *__restrict__ ithi) {
    const uint64_t j = (uint64_t)blockIdx.x * CO_THREADS + threadIdx.x;
    if (j < L) x[j] = gf::mul(x[j], dft::tw28(itlo, ithi, (uint32_t)(j << shift)));
}

// ---- block sumcheck, phase 1 ----------------------------------------------

// x (natural order) -> Y_i = d P(w_2d^(2i+1)), natural order, in place
template <int D>
__device__ __forceinline__ void odd_points(fe *x) {
    dft::dif<D, true>(x);                                   // x[i] = X[brev(i)]
    fe y[D];
    dft::static_for<0, D>([&](auto ii) {
        constexpr int i = decltype(ii)::value;
        constexpr int s = dft::brev_c<D>(i);
        y[s] = gf::mul_root<2 * D, s>(x[i]);                // X[s] w_2d^s
    });
    dft::dif<D, false>(y);                                  // y[i] = Y[brev(i)]
    dft::static_for<0, D>([&](auto ii) {
        constexpr int i = decltype(ii)::value;
        x[dft::brev_c<D>(i)] = y[i];
    });
}

// 128 lanes: lanes 0..63 take p1 of the tile's 64 blocks, lanes 64..127 the
// same blocks of p2.  LDS row b (2D + 1 elements, padded) collects block b's
// products: columns 0..D-1 even points, D..2D-1 odd points.
template <int D>
__global__ __launch_bounds__(2 * BSC_TB) void bsc_small_kernel(const fe *__restrict__ p1,
                                                               const fe *__restrict__ p2,
                                                               uint64_t nblocks, fe *__restrict__ part) {
    extern __shared__ fe S[];
    constexpr int ROW = 2 * D + 1, COLS = 2 * D, G = 2 * BSC_TB / COLS;
    const int t = threadIdx.x, lane = t & (BSC_TB - 1), vec = t / BSC_TB;
    const int col = t % COLS, grp = t / COLS;
    const fe *src = vec ? p2 : p1;
    fe *row = S + lane * ROW;
    fe acc = gf::zero();
    const uint64_t ntiles = (nblocks + BSC_TB - 1) / BSC_TB;
    for (uint64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        const uint64_t blk = tile * BSC_TB + lane;
        fe x[D];
        if (blk < nblocks) {
            const fe *b = src + blk * D;
#pragma unroll
            for (int k = 0; k < D; ++k) x[k] = b[k];
        } else {
#pragma unroll
            for (int k = 0; k < D; ++k) x[k] = gf::zero();
        }
        if (vec) {
#pragma unroll
            for (int k = 0; k < D; ++k) row[k] = x[k];
        }
        __syncthreads();
        if (!vec) {
#pragma unroll
            for (int k = 0; k < D; ++k) row[k] = gf::mul(x[k], row[k]);
        }
        odd_points<D>(x);
        if (vec) {
#pragma unroll
            for (int k = 0; k < D; ++k) row[D + k] = x[k];
        }
        __syncthreads();
        if (!vec) {
#pragma unroll
            for (int k = 0; k < D; ++k) row[D + k] = gf::mul(x[k], row[D + k]);
        }
        __syncthreads();
        for (int r = grp; r < BSC_TB; r += G) acc = gf::add(acc, S[r * ROW + col]);
        __syncthreads();
    }
    S[grp * COLS + col] = acc;
    __syncthreads();
    if (t < COLS) {
        fe s = gf::zero();
        for (int g = 0; g < G; ++g) s = gf::add(s, S[g * COLS + t]);
        part[(uint64_t)blockIdx.x * COLS + t] = s;
    }
}

__device__ __forceinline__ uint32_t brev_bits(uint32_t x, int bits) {
    return bits ? __brev(x) >> (32 - bits) : 0;
}

// Radix-2 decimation-in-frequency stages over NA arrays of 2^logd elements
// in LDS (natural order in, bit-reversed out), unscaled; the twiddle tables
// passed in (root of order 2^28 or its inverse) set the direction.  With
// logn > logd each array holds 2^(logn - logd) consecutive independent
// blocks (a stage's pairs never cross a block).
template <int NA>
__device__ void lds_dif(fe *const *A, int logd, const fe *twlo, const fe *twhi, int logn = -1) {
    const int n = 1 << (logn < 0 ? logd : logn);
    for (int lh = logd - 1; lh >= 0; --lh) {
        const int h = 1 << lh;
        for (int p = threadIdx.x; p < n / 2; p += CO_THREADS) {
            const int off = p & (h - 1), i = ((p >> lh) << (lh + 1)) + off;
            const fe w = dft::tw28(twlo, twhi, (uint32_t)off << (NTT_LOG_MAX - lh - 1));
#pragma unroll
            for (int a = 0; a < NA; ++a) {
                const fe u = A[a][i], v = A[a][i + h];
                A[a][i] = gf::add(u, v);
                A[a][i + h] = gf::mul(gf::sub(u, v), w);
            }
        }
        __syncthreads();
    }
}

// Radix-2 decimation-in-time stages: bit-reversed order in, natural out.
template <int NA>
__device__ void lds_dit(fe *const *A, int logd, const fe *twlo, const fe *twhi, int logn = -1) {
    const int n = 1 << (logn < 0 ? logd : logn);
    for (int lh = 0; lh < logd; ++lh) {
        const int h = 1 << lh;
        for (int p = threadIdx.x; p < n / 2; p += CO_THREADS) {
            const int off = p & (h - 1), i = ((p >> lh) << (lh + 1)) + off;
            const fe w = dft::tw28(twlo, twhi, (uint32_t)off << (NTT_LOG_MAX - lh - 1));
#pragma unroll
            for (int a = 0; a < NA; ++a) {
                const fe u = A[a][i], v = gf::mul(A[a][i + h], w);
                A[a][i] = gf::add(u, v);
                A[a][i + h] = gf::sub(u, v);
            }
        }
        __syncthreads();
    }
}

// one workgroup per block (grid-stride); thread t owns columns t + 256 q
__global__ __launch_bounds__(CO_THREADS) void bsc_large_kernel(const fe *__restrict__ p1,
                                                               const fe *__restrict__ p2, uint64_t nblocks,
                                                               int logd, const fe *twlo, const fe *twhi,
                                                               const fe *itlo, const fe *ithi,
                                                               fe *__restrict__ part) {
    extern __shared__ fe S[];
    constexpr int QMAX = 2 * BSC_D_MAX / CO_THREADS;
    const int d = 1 << logd, t = threadIdx.x, cols = 2 * d;
    fe *A = S, *B = S + d;
    fe *const AB[2] = {A, B};
    fe acc[QMAX];
#pragma unroll
    for (int q = 0; q < QMAX; ++q) acc[q] = gf::zero();
    for (uint64_t j = blockIdx.x; j < nblocks; j += gridDim.x) {
        for (int k = t; k < d; k += CO_THREADS) {
            A[k] = p1[j * d + k];
            B[k] = p2[j * d + k];
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < QMAX; ++q) {
            const int c = t + CO_THREADS * q;
            if (c < d) acc[q] = gf::add(acc[q], gf::mul(A[c], B[c]));
        }
        __syncthreads();
        lds_dif<2>(AB, logd, itlo, ithi);             // A[i] = X[brev(i)]
        for (int i = t; i < d; i += CO_THREADS) {
            const uint32_t s = brev_bits((uint32_t)i, logd);
            const fe w = dft::tw28(twlo, twhi, s << (NTT_LOG_MAX - logd - 1));
            A[i] = gf::mul(A[i], w);
            B[i] = gf::mul(B[i], w);
        }
        __syncthreads();
        lds_dit<2>(AB, logd, twlo, twhi);                   // A[k] = Y[k]
#pragma unroll
        for (int q = 0; q < QMAX; ++q) {
            const int c = t + CO_THREADS * q;
            if (c >= d && c < cols) acc[q] = gf::add(acc[q], gf::mul(A[c - d], B[c - d]));
        }
        __syncthreads();
    }
#pragma unroll
    for (int q = 0; q < QMAX; ++q) {
        const int c = t + CO_THREADS * q;
        if (c < cols) part[(uint64_t)blockIdx.x * cols + c] = acc[q];
    }
}

// d = 256 with the transforms in registers: the four-step 16 x 16 (natural
// order in and out, as the NTT passes) on weak values (gf.h, dft.h).  A
// workgroup takes 8 blocks of both vectors per tile; thread (a, b, col)
// holds one 16-element column of vector a's block b.  Per block: the
// inverse transform (unscaled), the twist of coefficient k by w_512^k, the
// forward transform, so M[k] = d P(w_512 w_256^k) as bsc_large_kernel
// forms it; thread t then adds point t's products over the 8 blocks (even
// points from the block values, odd points from the transforms).  Eight
// barriers per tile of 8 blocks instead of 16 per block, every thread busy.
constexpr int L256_B = 8, L256_RS = 17;        // blocks per tile, padded row (elements)
constexpr int L256_LDS = 2 * L256_B * 16 * L256_RS + 2 * 256;   // elements: matrices, w_256 and w_512 tables

template <bool INV>
__device__ __forceinline__ void four_step256(fe *Mb, int col, const fe *wd, const fe *wt) {
    fe v[16];
    // step 1: column n2 = col, 16-point DFTs over n1, twiddle w_256^(n2 k1)
#pragma unroll
    for (int n1 = 0; n1 < 16; ++n1) v[n1] = Mb[n1 * L256_RS + col];
    __syncthreads();
    dft::dif_w<16, INV>(v);
    dft::static_for<0, 16>([&](auto ii) {
        constexpr int i = decltype(ii)::value;
        constexpr int k1 = dft::brev_c<16>(i);
        fe x = v[i];
        if (k1 != 0 && col != 0) {
            const int e = col * k1;
            x = gf::mul_w(wd[INV ? (256 - e) & 255 : e], x);
        }
        Mb[k1 * L256_RS + col] = x;             // Y[n2][k1] at row k1, column n2
    });
    __syncthreads();
    // step 2: row k1 = col, 16-point DFTs over n2; X[k1 + 16 k2] to natural order
#pragma unroll
    for (int n2 = 0; n2 < 16; ++n2) v[n2] = Mb[col * L256_RS + n2];
    __syncthreads();
    dft::dif_w<16, INV>(v);
    dft::static_for<0, 16>([&](auto ii) {
        constexpr int i = decltype(ii)::value;
        constexpr int k2 = dft::brev_c<16>(i);
        fe x = v[i];
        if (INV) x = gf::mul_w(wt[col + 16 * k2], x);     // twist coefficient k by w_512^k
        Mb[k2 * L256_RS + col] = x;             // natural k = col + 16 k2 at row k2, column col
    });
    __syncthreads();
}

__global__ __launch_bounds__(CO_THREADS) void bsc_large256_kernel(const fe *__restrict__ p1,
                                                                  const fe *__restrict__ p2, uint64_t nblocks,
                                                                  const fe *twlo, const fe *twhi,
                                                                  fe *__restrict__ part) {
    extern __shared__ fe S[];
    fe *wd = S + 2 * L256_B * 16 * L256_RS, *wt = wd + 256;
    const int t = threadIdx.x;
    wd[t] = dft::tw28(twlo, twhi, (uint32_t)t << (NTT_LOG_MAX - 8));     // w_256^t
    wt[t] = dft::tw28(twlo, twhi, (uint32_t)t << (NTT_LOG_MAX - 9));     // w_512^t
    const int a = t >> 7, bl = (t >> 4) & (L256_B - 1), col = t & 15;
    fe *Mb = S + (a * L256_B + bl) * 16 * L256_RS;
    const fe *src = a ? p2 : p1;
    const int pr = (t >> 4) * L256_RS + (t & 15);   // point t of a block matrix
    fe acc0 = gf::zero(), acc1 = gf::zero();
    const uint64_t ntiles = (nblocks + L256_B - 1) / L256_B;
    for (uint64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        const uint64_t blk = tile * L256_B + bl;
#pragma unroll
        for (int n1 = 0; n1 < 16; ++n1)
            Mb[n1 * L256_RS + col] = blk < nblocks ? src[blk * 256 + n1 * 16 + col] : gf::zero();
        __syncthreads();
        // even points: the block values themselves
#pragma unroll
        for (int b = 0; b < L256_B; ++b)
            acc0 = gf::add(acc0, gf::mul(S[b * 16 * L256_RS + pr], S[(L256_B + b) * 16 * L256_RS + pr]));
        four_step256<true>(Mb, col, wd, wt);
        four_step256<false>(Mb, col, wd, wt);
        // odd points: the coset values (weak; the first factor made canonical)
#pragma unroll
        for (int b = 0; b < L256_B; ++b)
            acc1 = gf::add(acc1, gf::mul(gf::canon_w(S[b * 16 * L256_RS + pr]), S[(L256_B + b) * 16 * L256_RS + pr]));
        __syncthreads();
    }
    part[(uint64_t)blockIdx.x * 512 + t] = acc0;
    part[(uint64_t)blockIdx.x * 512 + 256 + t] = acc1;
}

// column sums over the workgroups; 16 columns per workgroup, 16 lanes each.
// Column c = co * d + i is the point w^(ncos i + co) (w of order ncos * d);
// columns with co > 0 are multiplied by `scale`.
constexpr int FIN_COLS = 16, FIN_LANES = CO_THREADS / FIN_COLS;
__global__ __launch_bounds__(CO_THREADS) void bsc_final_kernel(const fe *__restrict__ part, int nwg, int d, int ncos,
                                                               fe scale, fe *__restrict__ out) {
    __shared__ fe S[FIN_LANES][FIN_COLS];
    const int cols = ncos * d, cl = threadIdx.x % FIN_COLS, g = threadIdx.x / FIN_COLS;
    const int c = blockIdx.x * FIN_COLS + cl;
    fe s = gf::zero();
    if (c < cols)
        for (int w = g; w < nwg; w += FIN_LANES) s = gf::add(s, part[(uint64_t)w * cols + c]);
    S[g][cl] = s;
    __syncthreads();
    if (g == 0 && c < cols) {
        for (int h = 1; h < FIN_LANES; ++h) s = gf::add(s, S[h][cl]);
        const int co = c / d, i = c - co * d;
        out[ncos * i + co] = co ? gf::mul(s, scale) : s;
    }
}

// Cubic block sum (cubic_coSumcheck_phase1, src/coSumcheck.cpp:11-45):
//     H[k] = sum_j (P1 P2 - P4)(w_4d^k) P3(w_4d^k),   k < 4d,
// with P1, P2, P3, P4 the interpolants of block j of p1, p2, b1, p3.  For
// k = 4i + s, s = 0 is the blocks' own values; s = 1, 2, 3 are the
// d-point transforms NTT_d(c_m w_4d^(s m) / d) of the unscaled inverse
// transform c of each block, so each block takes one inverse and three
// forward transforms per vector.  A workgroup takes tiles of T = max(d, 512)
// elements of each vector (T/d blocks); each thread accumulates the four
// cosets of the elements it owns, and the tile's blocks are added up per
// column at the end.
__global__ __launch_bounds__(CO_THREADS) void bsc3_kernel(const fe *__restrict__ p1, const fe *__restrict__ p2,
                                                          const fe *__restrict__ p3, const fe *__restrict__ b1,
                                                          uint64_t M, int logd, int logT, const fe *twlo,
                                                          const fe *twhi, const fe *itlo, const fe *ithi, fe inv_d,
                                                          fe *__restrict__ part) {
    extern __shared__ fe S[];
    constexpr int QMAX = BSC3_D_MAX / CO_THREADS;
    const int d = 1 << logd, T = 1 << logT, t = threadIdx.x, QT = T / CO_THREADS;
    fe *const A[4] = {S, S + T, S + 2 * T, S + 3 * T};      // p1, p2, p3, b1
    fe *const B[1] = {S + 4 * T};
    const fe *const src[4] = {p1, p2, p3, b1};
    fe acc[4][QMAX], tmp[QMAX], f[QMAX];
#pragma unroll
    for (int c = 0; c < 4; ++c)
#pragma unroll
        for (int q = 0; q < QMAX; ++q) acc[c][q] = gf::zero();
    const uint64_t ntiles = (M + T - 1) >> logT;
    const int G = T >> logd;
    int *live = (int *)(S + 5 * T);      // per block of the tile: its p1 part has a non-zero element
    for (uint64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        const uint64_t base = tile << logT;
        for (int g = t; g < G; g += CO_THREADS) live[g] = 0;
        __syncthreads();
#pragma unroll
        for (int q = 0; q < QMAX; ++q) {
            if (q >= QT) break;
            const int e = t + CO_THREADS * q;
            const uint64_t pos = base + e;
#pragma unroll
            for (int v = 0; v < 4; ++v) A[v][e] = pos < M ? src[v][pos] : gf::zero();
            if (A[0][e].re | A[0][e].im) atomicOr(&live[e >> logd], 1);
        }
        __syncthreads();
        // the reference skips blocks whose p1 part is zero (src/coSumcheck.cpp:
        // 22-28); zeroing such a block's b1 part removes it from every coset
#pragma unroll
        for (int q = 0; q < QMAX; ++q) {
            if (q >= QT) break;
            const int e = t + CO_THREADS * q;
            if (!live[e >> logd]) {
                A[3][e] = gf::zero();
            } else {
                const fe x = gf::sub(gf::mul(A[0][e], A[1][e]), A[2][e]);
                acc[0][q] = gf::add(acc[0][q], gf::mul(x, A[3][e]));
            }
        }
        __syncthreads();
        lds_dif<4>(A, logd, itlo, ithi, logT);                  // A[v][e] = c[brev(e mod d)] (unscaled)
        dft::static_for<1, 4>([&](auto cc) {
            constexpr int co = decltype(cc)::value;
#pragma unroll
            for (int q = 0; q < QMAX; ++q) {
                if (q >= QT) break;
                const uint32_t m = brev_bits((uint32_t)((t + CO_THREADS * q) & (d - 1)), logd);
                f[q] = gf::mul(dft::tw28(twlo, twhi, (co * m) << (NTT_LOG_MAX - logd - 2)), inv_d);
            }
            dft::static_for<0, 4>([&](auto vv) {
                constexpr int v = decltype(vv)::value;
#pragma unroll
                for (int q = 0; q < QMAX; ++q) {
                    if (q >= QT) break;
                    const int e = t + CO_THREADS * q;
                    B[0][e] = gf::mul(A[v][e], f[q]);
                }
                __syncthreads();
                lds_dit<1>(B, logd, twlo, twhi, logT);          // B[e] = P_v(w_4d^(4 (e mod d) + co))
#pragma unroll
                for (int q = 0; q < QMAX; ++q) {
                    if (q >= QT) break;
                    const fe y = B[0][t + CO_THREADS * q];
                    if (v == 0) tmp[q] = y;
                    else if (v == 1) tmp[q] = gf::mul(tmp[q], y);
                    else if (v == 2) tmp[q] = gf::sub(tmp[q], y);
                    else acc[co][q] = gf::add(acc[co][q], gf::mul(tmp[q], y));
                }
                __syncthreads();
            });
        });
    }
    // element slots -> columns: column (co, i) adds the slots co*T + g*d + i
#pragma unroll
    for (int co = 0; co < 4; ++co)
#pragma unroll
        for (int q = 0; q < QMAX; ++q) {
            if (q >= QT) break;
            S[co * T + t + CO_THREADS * q] = acc[co][q];
        }
    __syncthreads();
    for (int c = t; c < 4 * d; c += CO_THREADS) {
        const int co = c >> logd, i = c & (d - 1);
        fe sum = gf::zero();
        for (int g = 0; g < G; ++g) sum = gf::add(sum, S[co * T + (g << logd) + i]);
        part[(uint64_t)blockIdx.x * 4 * d + c] = sum;
    }
}

// Cubic block sum for d = 512 with the transforms in registers.  A
// 512-point DFT is the four-step 8 x 64 with the 64-point part as 8 x 8
// (natural order in and out, radix-8 register DFTs on weak values, twiddles
// w_512^e from an LDS table).  A workgroup takes one block of the four
// vectors per iteration, 64 threads per vector; thread q keeps the
// coefficients c[q + 64 n] (n < 8) of its vector in registers after the
// inverse transform (its last pass leaves exactly those), so each coset
// s = 1, 2, 3 is a twist by w_2048^(s m) / d and a forward transform from
// registers.  LDS carries the exchanges between passes and the four
// vectors' values at each point for the products.  Blocks whose p1 part is
// zero are skipped, as the reference skips them (src/coSumcheck.cpp:22-28).
constexpr int C512_RS = 65;                     // padded row of the 8 x 64 exchange
constexpr int C512_VEC = 8 * C512_RS;           // elements per vector in LDS (>= 512)
constexpr int C512_LDS = 4 * C512_VEC + 512 + 3 * 512;

// One 512-point transform of the calling thread's column q (64 threads per
// vector, L = that vector's LDS area): x[n1] = x_in[q + 64 n1] on entry;
// on return thread q holds X[q + 64 j] in out[j] (j < 8).  Begins by
// writing L, so L's previous readers must have passed a barrier; every
// thread of the workgroup calls it.
template <bool INV>
__device__ __forceinline__ void t512(fe *x, fe *out, fe *L, int q, const fe *w512) {
    // pass A: 8 points over n1 for column n2 = q, twiddle w_512^(q k1)
    dft::dif_w<8, INV>(x);
    dft::static_for<0, 8>([&](auto ii) {
        constexpr int i = decltype(ii)::value;
        constexpr int k1 = dft::brev_c<8>(i);
        fe y = x[i];
        if (k1 != 0 && q != 0) {
            const int e = q * k1;
            y = gf::mul_w(w512[INV ? (512 - e) & 511 : e], y);
        }
        L[k1 * C512_RS + q] = y;
    });
    __syncthreads();
    // pass B: thread (k1, m2): 8 points over m1 of Z[8 m1 + m2], twiddle w_64^(m2 j1), in place
    {
        const int k1 = q >> 3, m2 = q & 7;
        fe v[8];
#pragma unroll
        for (int m1 = 0; m1 < 8; ++m1) v[m1] = L[k1 * C512_RS + 8 * m1 + m2];
        dft::dif_w<8, INV>(v);
        dft::static_for<0, 8>([&](auto ii) {
            constexpr int i = decltype(ii)::value;
            constexpr int j1 = dft::brev_c<8>(i);
            fe y = v[i];
            if (j1 != 0 && m2 != 0) {
                const int e = 8 * m2 * j1;
                y = gf::mul_w(w512[INV ? (512 - e) & 511 : e], y);
            }
            L[k1 * C512_RS + 8 * j1 + m2] = y;
        });
    }
    __syncthreads();
    // pass C: thread (k1, j1), q = k1 + 8 j1: 8 points over m2; X[q + 64 j2]
    {
        const int k1 = q & 7, j1 = q >> 3;
        fe v[8];
#pragma unroll
        for (int m2 = 0; m2 < 8; ++m2) v[m2] = L[k1 * C512_RS + 8 * j1 + m2];
        dft::dif_w<8, INV>(v);
        dft::static_for<0, 8>([&](auto ii) {
            constexpr int i = decltype(ii)::value;
            out[dft::brev_c<8>(i)] = v[i];
        });
    }
}

__global__ __launch_bounds__(CO_THREADS) void bsc3_512_kernel(const fe *__restrict__ p1, const fe *__restrict__ p2,
                                                              const fe *__restrict__ p3, const fe *__restrict__ b1,
                                                              uint64_t nblocks, const fe *twlo, const fe *twhi,
                                                              fe inv_d, fe *__restrict__ part) {
    extern __shared__ fe S[];
    __shared__ int live;
    fe *w512 = S + 4 * C512_VEC, *tw = w512 + 512;
    const int t = threadIdx.x, vec = t >> 6, q = t & 63;
    fe *L = S + vec * C512_VEC;
    const fe *src = vec == 0 ? p1 : vec == 1 ? p2 : vec == 2 ? p3 : b1;
    for (int e = t; e < 512; e += CO_THREADS) {
        w512[e] = dft::tw28(twlo, twhi, (uint32_t)e << (NTT_LOG_MAX - 9));
#pragma unroll
        for (int co = 1; co < 4; ++co)      // w_2048^(co e) / d
            tw[(co - 1) * 512 + e] = gf::mul(inv_d, dft::tw28(twlo, twhi, (uint32_t)(co * e) << (NTT_LOG_MAX - 11)));
    }
    fe acc[4][2];
#pragma unroll
    for (int c = 0; c < 4; ++c) acc[c][0] = acc[c][1] = gf::zero();
    // (P1 P2 - P3) B1 at points t and t + 256 from the natural-order rows
    // (every accumulator takes y or zero, so acc stays in registers)
    auto products = [&](int co) {
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            const int e = t + 256 * j;
            const fe a = gf::canon_w(S[e]), b = S[C512_VEC + e];
            const fe x = gf::sub(gf::mul(a, b), gf::canon_w(S[2 * C512_VEC + e]));
            const fe y = gf::mul(x, S[3 * C512_VEC + e]);
            dft::static_for<0, 4>([&](auto cc) {
                constexpr int c = decltype(cc)::value;
                acc[c][j] = gf::add(acc[c][j], co == c ? y : gf::zero());
            });
        }
    };
    for (uint64_t blk = blockIdx.x; blk < nblocks; blk += gridDim.x) {
        fe x[8];
        int nz = 0;
#pragma unroll
        for (int n = 0; n < 8; ++n) {
            x[n] = src[blk * 512 + q + 64 * n];
            nz |= (x[n].re | x[n].im) != 0;
            L[q + 64 * n] = x[n];
        }
        if (t == 0) live = 0;
        __syncthreads();
        if (vec == 0 && nz) live = 1;
        __syncthreads();
        if (!live) continue;                    // uniform: the block's p1 part is zero
        products(0);                            // coset 0: the block's own values
        __syncthreads();
        fe c[8];
        t512<true>(x, c, L, q, w512);           // c[n] = d * coefficient (q + 64 n)
#pragma unroll 1
        for (int co = 1; co < 4; ++co) {
            fe y[8], z[8];
#pragma unroll
            for (int n = 0; n < 8; ++n) y[n] = gf::mul_w(tw[(co - 1) * 512 + q + 64 * n], c[n]);
            __syncthreads();                    // the previous readers of L are done
            t512<false>(y, z, L, q, w512);      // z[j] = P(w_2048^(co + 4 (q + 64 j)))
            __syncthreads();
#pragma unroll
            for (int j = 0; j < 8; ++j) L[q + 64 * j] = z[j];
            __syncthreads();
            products(co);
        }
        __syncthreads();                        // before the next block's loads
    }
#pragma unroll
    for (int c = 0; c < 4; ++c)
#pragma unroll
        for (int j = 0; j < 2; ++j) part[(uint64_t)blockIdx.x * 2048 + c * 512 + t + 256 * j] = acc[c][j];
}

// Cubic block sum for d <= 16, transforms in registers.  Every wave works
// on its own: it takes 16 blocks, lane l = 16 v + b holding vector v (p1,
// p2, p3, b1) of block b.  A lane takes its block's inverse transform once,
// then for each coset s = 1, 2, 3 the d-point transform of the coefficients
// twisted by w_4d^(s m) / d (compile-time roots); coset 0 is the block
// itself.  Each coset's values meet in the wave's own LDS rows, so the waves
// synchronise only with themselves (no workgroup barrier until the final
// column sums) and one wave's transforms overlap another's products.  Lane
// l forms (P1 P2 - P4) P3 for point k = l mod d of blocks l/d + (64/d) j
// of its wave.
__device__ __forceinline__ void co_wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <int D>
__global__ __launch_bounds__(CO_THREADS) void bsc3_wave_kernel(const fe *__restrict__ p1, const fe *__restrict__ p2,
                                                               const fe *__restrict__ p3, const fe *__restrict__ b1,
                                                               uint64_t nblocks, fe *__restrict__ part) {
    extern __shared__ fe S[];
    constexpr int WB = 16, RP = D + 1, VS = WB * RP, WS = 4 * VS;   // blocks per wave, row pitch, vector and wave spans
    constexpr int NP = WB * D, J = (NP + 63) / 64;                   // (block, point) pairs per coset and wave
    __shared__ int live[CO_THREADS / 64][WB];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int vec = lane / WB, b = lane & (WB - 1);
    const fe *src = vec == 0 ? p1 : vec == 1 ? p2 : vec == 2 ? p3 : b1;
    fe *rows = S + w * WS;
    fe *row = rows + vec * VS + b * RP;
    const int k = lane & (D - 1);
    fe acc[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) acc[c] = gf::zero();
    auto products = [&](fe &a) {
#pragma unroll
        for (int j = 0; j < J; ++j) {
            const int p = lane + 64 * j;
            if (p >= NP) break;
            const int bb = p / D;
            if (live[w][bb]) {
                const fe *c = rows + bb * RP + k;
                a = gf::add(a, gf::mul(gf::sub(gf::mul(c[0], c[VS]), c[2 * VS]), c[3 * VS]));
            }
        }
    };
    const uint64_t ngroups = (nblocks + WB - 1) / WB;
    const uint64_t gw = (uint64_t)blockIdx.x * (CO_THREADS / 64) + w, nw = (uint64_t)gridDim.x * (CO_THREADS / 64);
    for (uint64_t grp = gw; grp < ngroups; grp += nw) {
        const uint64_t blk = grp * WB + b;
        fe x[D];
        int nz = 0;
#pragma unroll
        for (int i = 0; i < D; ++i) {
            x[i] = blk < nblocks ? src[blk * D + i] : gf::zero();
            nz |= (x[i].re | x[i].im) != 0;
            row[i] = x[i];
        }
        if (vec == 0) live[w][b] = nz;     // blocks whose p1 part is zero are skipped
        co_wave_sync();
        products(acc[0]);
        co_wave_sync();                     // the rows are rewritten below
        dft::dif_w<D, true>(x);             // x[i] = D c[brev(i)]
        constexpr int LOGD = D >= 16 ? 4 : D >= 8 ? 3 : D >= 4 ? 2 : 1;
#pragma unroll
        for (int i = 0; i < D; ++i) x[i] = gf::mul_pow2_w<61 - LOGD>(x[i]);   // / D: a rotation
        dft::static_for<1, 4>([&](auto cc) {
            constexpr int co = decltype(cc)::value;
            fe y[D];
            dft::static_for<0, D>([&](auto ii) {
                constexpr int i = decltype(ii)::value;
                constexpr int m = dft::brev_c<D>(i);
                y[m] = gf::mul_root_w<4 * D, co * m>(x[i]);
            });
            dft::dif_w<D, false>(y);        // y[i] = P(w_4d^(4 brev(i) + co))
            dft::static_for<0, D>([&](auto ii) {
                constexpr int i = decltype(ii)::value;
                row[dft::brev_c<D>(i)] = gf::canon_w(y[i]);
            });
            co_wave_sync();
            products(acc[co]);
            co_wave_sync();
        });
    }
    // column (co, k) adds the threads with t mod D = k
    __syncthreads();
#pragma unroll
    for (int c = 0; c < 4; ++c) S[c * CO_THREADS + t] = acc[c];
    __syncthreads();
    if (t < 4 * D) {
        const int co = t / D, kk = t & (D - 1);
        fe sum = gf::zero();
        for (int u = kk; u < CO_THREADS; u += D) sum = gf::add(sum, S[co * CO_THREADS + u]);
        part[(uint64_t)blockIdx.x * 4 * D + t] = sum;
    }
}

// ---- block sumcheck, phase 2 ----------------------------------------------

struct FoldPtrs {
    const fe *in[4];
    fe *out[4];
};

__device__ __forceinline__ fe shfl_xor_fe(const fe &x, int m) {
    return gf::mk(__shfl_xor(x.re, m), __shfl_xor(x.im, m));
}

// Phase 2 for d <= 64 over NA vectors: lane l of a d-lane group takes
// element k = l mod d of one output (loads coalesced across the wave), and
// the group adds its d products by butterfly shuffles.
template <int NA>
__global__ __launch_bounds__(CO_THREADS) void bsc_fold_seg_kernel(FoldPtrs P, uint64_t total, int logd,
                                                                  const fe *__restrict__ L) {
    const int d = 1 << logd;
    const fe l = L[threadIdx.x & (d - 1)];
    const uint64_t step = (uint64_t)gridDim.x * CO_THREADS;
    for (uint64_t e = (uint64_t)blockIdx.x * CO_THREADS + threadIdx.x; e < total; e += step) {
        fe s[NA];
#pragma unroll
        for (int a = 0; a < NA; ++a) s[a] = gf::mul(l, P.in[a][e]);
        for (int m = d >> 1; m >= 1; m >>= 1)
#pragma unroll
            for (int a = 0; a < NA; ++a)
                s[a] = gf::add(s[a], gf::mk(__shfl_xor(s[a].re, m), __shfl_xor(s[a].im, m)));
        if ((e & (d - 1)) == 0)
#pragma unroll
            for (int a = 0; a < NA; ++a) P.out[a][e >> logd] = s[a];
    }
}

// one wave per output (d >= 128): lane l takes k = l + 64 q
template <int NA>
__global__ __launch_bounds__(CO_THREADS) void bsc_fold_wave_kernel(FoldPtrs P, uint64_t nout, int d,
                                                                   const fe *__restrict__ L) {
    const int lane = threadIdx.x & 63;
    const uint64_t w0 = ((uint64_t)blockIdx.x * CO_THREADS + threadIdx.x) >> 6;
    const uint64_t nw = ((uint64_t)gridDim.x * CO_THREADS) >> 6;
    for (uint64_t j = w0; j < nout; j += nw) {
        fe s[NA];
#pragma unroll
        for (int a = 0; a < NA; ++a) s[a] = gf::zero();
        for (int k = lane; k < d; k += 64) {
            const fe l = L[k];
#pragma unroll
            for (int a = 0; a < NA; ++a) s[a] = gf::add(s[a], gf::mul(l, P.in[a][j * d + k]));
        }
#pragma unroll
        for (int m = 32; m >= 1; m >>= 1)
#pragma unroll
            for (int a = 0; a < NA; ++a) s[a] = gf::add(s[a], shfl_xor_fe(s[a], m));
        if (lane == 0)
#pragma unroll
            for (int a = 0; a < NA; ++a) P.out[a][j] = s[a];
    }
}

// ---- RS_multifold5 ----------------------------------------------------------

struct FoldWeights {
    fe a[MF_SMALL_MAX];     // a_k / fd
};

// Output q is the fold at position i = idx[q] (idx null: i = q, all nout
// outputs); seg (optional) receives the fd gathered elements of each output.
template <int FD>
__global__ __launch_bounds__(CO_THREADS) void multifold_small_kernel(const fe *__restrict__ cw, uint64_t nout,
                                                                     int shift, const fe *__restrict__ itlo,
                                                                     const fe *__restrict__ ithi,
                                                                     FoldWeights wts, fe *__restrict__ out,
                                                                     const uint32_t *__restrict__ idx, uint64_t nq,
                                                                     fe *__restrict__ seg) {
    const uint64_t q = (uint64_t)blockIdx.x * CO_THREADS + threadIdx.x;
    if (q >= nq) return;
    const uint64_t i = idx ? (uint64_t)idx[q] : q;
    fe x[FD];
#pragma unroll
    for (int s = 0; s < FD; ++s) x[s] = cw[i + (uint64_t)s * nout];
    if (seg) {
#pragma unroll
        for (int s = 0; s < FD; ++s) seg[q * FD + s] = x[s];
    }
    dft::dif<FD, true>(x);                                  // x[j] = X[brev(j)]
    const fe w = dft::tw28(itlo, ithi, (uint32_t)(i << shift));   // w_n^-i
    fe acc = gf::mul(wts.a[FD - 1], x[dft::brev_c<FD>(FD - 1)]);
    dft::static_for<1, FD>([&](auto kk) {
        constexpr int k = FD - 1 - decltype(kk)::value;
        acc = gf::add(gf::mul(acc, w), gf::mul(wts.a[k], x[dft::brev_c<FD>(k)]));
    });
    out[q] = acc;
}

// one workgroup per output (grid-stride), the fd-point transform in LDS
__global__ __launch_bounds__(CO_THREADS) void multifold_large_kernel(const fe *__restrict__ cw, uint64_t nout,
                                                                     int logfd, int logn, const fe *itlo,
                                                                     const fe *ithi, const fe *__restrict__ a,
                                                                     fe *__restrict__ out,
                                                                     const uint32_t *__restrict__ idx, uint64_t nq,
                                                                     fe *__restrict__ seg) {
    extern __shared__ fe S[];
    const int fd = 1 << logfd, t = threadIdx.x;
    fe *X = S, *R = S + fd;
    fe *const XA[1] = {X};
    for (uint64_t q = blockIdx.x; q < nq; q += gridDim.x) {
        const uint64_t i = idx ? (uint64_t)idx[q] : q;
        for (int s = t; s < fd; s += CO_THREADS) {
            X[s] = cw[i + (uint64_t)s * nout];
            if (seg) seg[q * fd + s] = X[s];
        }
        __syncthreads();
        lds_dif<1>(XA, logfd, itlo, ithi);            // X[j] = X_brev(j)
        fe acc = gf::zero();
        for (int j = t; j < fd; j += CO_THREADS) {
            const uint32_t k = brev_bits((uint32_t)j, logfd);
            const uint64_t e = (i * k) << (NTT_LOG_MAX - logn);     // i k < n
            const fe w = dft::tw28(itlo, ithi, (uint32_t)e);
            acc = gf::add(acc, gf::mul(gf::mul(a[k], w), X[j]));
        }
        R[t] = acc;
        __syncthreads();
        for (int h = CO_THREADS / 2; h >= 1; h >>= 1) {
            if (t < h) R[t] = gf::add(R[t], R[t + h]);
            __syncthreads();
        }
        if (t == 0) out[q] = R[0];
        __syncthreads();
    }
}

// ---- the opening's batching step: out[k] = sum_i c_i rows[i][k] -------------
struct RowWeights {
    fe c[32];
};

__global__ __launch_bounds__(CO_THREADS) void row_combine_kernel(const fe *__restrict__ rows, int nrows, uint64_t len,
                                                                 RowWeights w, fe *__restrict__ out) {
    for (uint64_t k = (uint64_t)blockIdx.x * CO_THREADS + threadIdx.x; k < len;
         k += (uint64_t)gridDim.x * CO_THREADS) {
        fe acc = gf::zero();
        for (int i = 0; i < nrows; ++i) acc = gf::add(acc, gf::mul(w.c[i], rows[(uint64_t)i * len + k]));
        out[k] = acc;
    }
}

// ---- encode_concatenated: segment transforms --------------------------------
// Segment j of cprime viewed as degree x S (column j: cprime[j + k S]),
// zero-padded to DP = degree * rate1 points, forward DFT, stored as row j of
// C.  A workgroup takes G consecutive segments (G * DP <= 4096): loads run
// over j (coalesced), the transform runs as radix-2 decimation-in-time
// stages over bit-reversed LDS positions, and the G rows leave as one
// contiguous run.
__global__ __launch_bounds__(CO_THREADS) void seg_ntt_kernel(const fe *__restrict__ cp, uint64_t S, int degree,
                                                             int logdp, int G, const fe *twlo, const fe *twhi,
                                                             fe *__restrict__ C) {
    extern __shared__ fe X[];
    const int DP = 1 << logdp, t = threadIdx.x;
    const uint64_t ngroups = S / (uint64_t)G;
    for (uint64_t grp = blockIdx.x; grp < ngroups; grp += gridDim.x) {
        const uint64_t j0 = grp * G;
        for (int e = t; e < G * DP; e += CO_THREADS) {
            const int g = e % G, k = e / G;
            fe v = gf::zero();
            if (k < degree) v = cp[j0 + g + (uint64_t)k * S];
            X[g * DP + (int)brev_bits((uint32_t)k, logdp)] = v;
        }
        __syncthreads();
        for (int lh = 0; lh < logdp; ++lh) {
            const int h = 1 << lh;
            for (int p = t; p < G * DP / 2; p += CO_THREADS) {
                const int g = p >> (logdp - 1), q = p & (DP / 2 - 1);
                const int off = q & (h - 1), i = g * DP + ((q >> lh) << (lh + 1)) + off;
                const fe w = dft::tw28(twlo, twhi, (uint32_t)off << (NTT_LOG_MAX - lh - 1));
                const fe u = X[i], v = gf::mul(X[i + h], w);
                X[i] = gf::add(u, v);
                X[i + h] = gf::sub(u, v);
            }
            __syncthreads();
        }
        for (int e = t; e < G * DP; e += CO_THREADS) C[j0 * DP + e] = X[e];
        __syncthreads();
    }
}

// ---- host helpers -------------------------------------------------------------

// the reference's fft (src/utils.cpp:904-964) on the host, for d-element
// weight vectors: natural order in and out, inverse scaled by 1/n
void host_fft(std::vector<fe> &a, bool inverse) {
    const int n = (int)a.size(), logn = ilog2((uint64_t)n);
    for (int i = 1, j = 0; i < n; ++i) {
        int bit = n >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) std::swap(a[i], a[j]);
    }
    // one table w_n^k (k < n/2; w_n^-k for the inverse): stage `len` uses
    // every (n/len)-th entry, since root_of_unity(lg) = root_of_unity(logn)^(n/len)
    std::vector<fe> tw(n > 1 ? n / 2 : 1);
    fe wn = gf::root_of_unity(logn);
    if (inverse) wn = gf::inv(wn);
    tw[0] = gf::one();
    for (int k = 1; k < n / 2; ++k) tw[k] = gf::mul(tw[k - 1], wn);
    for (int lg = 1; lg <= logn; ++lg) {
        const int len = 1 << lg, step = n >> lg;
        for (int s = 0; s < n; s += len) {
            for (int k = 0; k < len / 2; ++k) {
                const fe u = a[s + k], v = k ? gf::mul(a[s + k + len / 2], tw[k * step]) : a[s + k + len / 2];
                a[s + k] = gf::add(u, v);
                a[s + k + len / 2] = gf::sub(u, v);
            }
        }
    }
    if (inverse) {
        const fe ni = gf::inv(gf::from_u64((uint64_t)n));
        for (auto &x : a) x = gf::mul(x, ni);
    }
}

// compute_lagrange_coeff(., r, d), src/utils.cpp:169-173
std::vector<fe> lagrange_weights(const fe &r, int d) {
    std::vector<fe> L(d);
    L[0] = gf::one();
    for (int i = 1; i < d; ++i) L[i] = gf::mul(L[i - 1], r);
    host_fft(L, true);
    return L;
}

template <int D>
void launch_bsc_small(const fe *p1, const fe *p2, uint64_t nblocks, fe *part, int nwg, hipStream_t s) {
    const size_t lds = sizeof(fe) * BSC_TB * (2 * D + 1);
    hipLaunchKernelGGL(bsc_small_kernel<D>, dim3(nwg), dim3(2 * BSC_TB), lds, s, p1, p2, nblocks, part);
}

template <int FD>
void launch_mf_small(const fe *cw, uint64_t nout, int shift, const fe *itlo, const fe *ithi,
                     const FoldWeights &w, fe *out, const uint32_t *idx, uint64_t nq, fe *seg, hipStream_t s) {
    const unsigned g = (unsigned)((nq + CO_THREADS - 1) / CO_THREADS);
    hipLaunchKernelGGL(multifold_small_kernel<FD>, dim3(g), dim3(CO_THREADS), 0, s, cw, nout, shift, itlo,
                       ithi, w, out, idx, nq, seg);
}

}  // namespace

int copcs_init(cosnark_ctx *) {
    const int small_lds = (int)(sizeof(fe) * BSC_TB * (2 * BSC_SMALL_MAX + 1));
    const int large_lds = (int)(sizeof(fe) * 2 * BSC_D_MAX);
    const int mf_lds = (int)(sizeof(fe) * (MF_FD_MAX + CO_THREADS));
    if (hipFuncSetAttribute((const void *)seg_ntt_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)(sizeof(fe) * 4096)) != hipSuccess)
        return COSNARK_E_HIP;
    if (hipFuncSetAttribute((const void *)bsc_small_kernel<32>, hipFuncAttributeMaxDynamicSharedMemorySize,
                            small_lds) != hipSuccess ||
        hipFuncSetAttribute((const void *)bsc_large_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                            large_lds) != hipSuccess ||
        hipFuncSetAttribute((const void *)bsc_large256_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)(sizeof(fe) * L256_LDS)) != hipSuccess ||
        hipFuncSetAttribute((const void *)multifold_large_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                            mf_lds) != hipSuccess ||
        hipFuncSetAttribute((const void *)bsc3_wave_kernel<16>, hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)(sizeof(fe) * 4 * 64 * 17)) != hipSuccess ||
        hipFuncSetAttribute((const void *)bsc3_512_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)(sizeof(fe) * C512_LDS)) != hipSuccess ||
        hipFuncSetAttribute((const void *)bsc3_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)(sizeof(fe) * 5 * BSC3_D_MAX + sizeof(int) * (1 << BSC3_TILE_LOG))) != hipSuccess)
        return COSNARK_E_HIP;
    return COSNARK_OK;
}

// p[j] = sum_k L_k p[jd + k] for each of the na vectors, in place (through
// scratch: a block's outputs overwrite inputs of other blocks)
static int block_fold(cosnark_ctx *ctx, uint64_t *const *v, int na, uint64_t M, int d, const uint64_t *h_L,
                      const char *what) {
    for (int a = 0; a < na; ++a)
        if (!v[a]) return set_err(ctx, COSNARK_E_ARG, std::string(what) + ": null argument");
    if (d < 2 || !is_pow2((uint64_t)d) || !is_pow2(M) || M < (uint64_t)d)
        return set_err(ctx, COSNARK_E_ARG, std::string(what) + ": d and M must be powers of two, 2 <= d <= M");
    if (d > CO_HOST_ELEMS) return set_err(ctx, COSNARK_E_RANGE, std::string(what) + ": d too large");
    const uint64_t nout = M / d;
    void *base;
    int rc = scratch_reserve(ctx, sizeof(fe) * (na * nout + d), &base);
    if (rc) return rc;
    FoldPtrs P{};
    for (int a = 0; a < na; ++a) {
        P.in[a] = (const fe *)v[a];
        P.out[a] = (fe *)base + a * nout;
    }
    fe *Ld = (fe *)base + na * nout;
    memcpy(ctx->co_host, h_L, sizeof(fe) * d);
    hipStream_t s = ctx->stream;
    HIPCK(hipMemcpyAsync(Ld, ctx->co_host, sizeof(fe) * d, hipMemcpyHostToDevice, s));
    if (d <= 64) {
        const uint64_t g = (M + CO_THREADS - 1) / CO_THREADS;
        const dim3 grid((unsigned)(g < 16384 ? g : 16384));
        const int logd = ilog2((uint64_t)d);
        if (na == 2) hipLaunchKernelGGL(bsc_fold_seg_kernel<2>, grid, dim3(CO_THREADS), 0, s, P, M, logd, (const fe *)Ld);
        else hipLaunchKernelGGL(bsc_fold_seg_kernel<4>, grid, dim3(CO_THREADS), 0, s, P, M, logd, (const fe *)Ld);
    } else {
        const uint64_t wg = (nout + 3) / 4;
        const dim3 g((unsigned)(wg < 4096 ? wg : 4096));
        if (na == 2) hipLaunchKernelGGL(bsc_fold_wave_kernel<2>, g, dim3(CO_THREADS), 0, s, P, nout, d, (const fe *)Ld);
        else hipLaunchKernelGGL(bsc_fold_wave_kernel<4>, g, dim3(CO_THREADS), 0, s, P, nout, d, (const fe *)Ld);
    }
    LAUNCHCK();
    for (int a = 0; a < na; ++a)
        HIPCK(hipMemcpyAsync(v[a], P.out[a], sizeof(fe) * nout, hipMemcpyDeviceToDevice, s));
    return cosnark_sync(ctx);
}

// unpack_secret_share on the host (src/MPI_utils.cpp:2769-2785): the n
// shares are values of a polynomial of degree < n at w_2n^(2j+1); returns
// its values at the n-th roots of unity in place of the shares
static void unpack_values(std::vector<fe> &x) {
    const int n = (int)x.size();
    host_fft(x, true);
    const fe wi = gf::inv(gf::root_of_unity(1 + ilog2((uint64_t)n)));
    fe m = gf::one();
    for (int i = 0; i < n; ++i) {
        x[i] = gf::mul(x[i], m);
        m = gf::mul(m, wi);
    }
    host_fft(x, false);
}

extern "C" {

int cosnark_copcs_encode(cosnark_ctx *ctx, const uint64_t *d_rows, int nrows, uint64_t L, int rate,
                         int log_rate_offset, uint64_t *d_codewords) {
    if (!ctx || !d_rows || !d_codewords) return set_err(ctx, COSNARK_E_ARG, "copcs_encode: null argument");
    if (nrows < 1 || !is_pow2(L) || L < 2 || rate < 1 || !is_pow2((uint64_t)rate) || log_rate_offset < 0)
        return set_err(ctx, COSNARK_E_ARG, "copcs_encode: need nrows >= 1, L >= 2 and rate powers of two");
    const int logL = ilog2(L), logM = ilog2(L * (uint64_t)rate), k = 1 + log_rate_offset + logL;
    if (logM > NTT_LOG_MAX || k > NTT_LOG_MAX)
        return set_err(ctx, COSNARK_E_RANGE, "copcs_encode: transform or coset root beyond 2^28");
    void *base;
    int rc = scratch_reserve(ctx, sizeof(fe) * L, &base);
    if (rc) return rc;
    fe *tmp = (fe *)base;
    const uint64_t M = L * (uint64_t)rate;
    for (int i = 0; i < nrows; ++i) {
        rc = ntt_launch(ctx, (const fe *)d_rows + (uint64_t)i * L, L, tmp, logL, true, ctx->stream);
        if (rc) return rc;
        hipLaunchKernelGGL(coset_twist_kernel, dim3((unsigned)((L + CO_THREADS - 1) / CO_THREADS)),
                           dim3(CO_THREADS), 0, ctx->stream, tmp, L, NTT_LOG_MAX - k,
                           (const fe *)ctx->itw_lo, (const fe *)ctx->itw_hi);
        LAUNCHCK();
        rc = ntt_launch(ctx, tmp, L, (fe *)d_codewords + (uint64_t)i * M, logM, false, ctx->stream);
        if (rc) return rc;
    }
    return cosnark_sync(ctx);
}

int cosnark_encode_concatenated(cosnark_ctx *ctx, const uint64_t *d_data, uint64_t n, int rate0, int rate1,
                                int degree, uint64_t *d_cprime, uint64_t *d_C) {
    if (!ctx || !d_data || !d_cprime || !d_C) return set_err(ctx, COSNARK_E_ARG, "encode_concatenated: null argument");
    const uint64_t m = n * (uint64_t)rate0, DP = (uint64_t)degree * (uint64_t)rate1;
    if (!is_pow2(n) || rate0 < 1 || !is_pow2((uint64_t)rate0) || rate1 < 1 || !is_pow2((uint64_t)rate1) ||
        degree < 1 || !is_pow2((uint64_t)degree) || (uint64_t)degree > m || DP < 2)
        return set_err(ctx, COSNARK_E_ARG, "encode_concatenated: n, rates, degree powers of two, degree <= rate0*n, "
                                           "degree*rate1 >= 2");
    if (m > (1ull << NTT_LOG_MAX) || DP > 4096)
        return set_err(ctx, COSNARK_E_RANGE, "encode_concatenated: rate0*n > 2^28 or degree*rate1 > 4096");
    hipStream_t s = ctx->stream;
    if (m == 1) {
        HIPCK(hipMemcpyAsync(d_cprime, d_data, sizeof(fe), hipMemcpyDeviceToDevice, s));
    } else {
        int rc = ntt_launch(ctx, (const fe *)d_data, n, (fe *)d_cprime, ilog2(m), false, s);
        if (rc) return rc;
    }
    const uint64_t S = m / degree;
    const uint64_t G = std::min<uint64_t>(S, 4096 / DP);
    const uint64_t groups = S / G;
    hipLaunchKernelGGL(seg_ntt_kernel, dim3((unsigned)(groups < 4096 ? groups : 4096)), dim3(CO_THREADS),
                       sizeof(fe) * G * DP, s, (const fe *)d_cprime, S, degree, ilog2(DP), (int)G,
                       (const fe *)ctx->tw_lo, (const fe *)ctx->tw_hi, (fe *)d_C);
    LAUNCHCK();
    return cosnark_sync(ctx);
}

int cosnark_distributed_mt(cosnark_ctx *ctx, const uint64_t *d_cw, int nrows, uint64_t Lc, uint8_t *d_nodes) {
    if (!ctx || !d_cw || !d_nodes) return set_err(ctx, COSNARK_E_ARG, "distributed_mt: null argument");
    int rc = merkle_columns_launch(ctx, (const fe *)d_cw, nrows, Lc, d_nodes, ctx->stream);
    if (rc) return rc;
    return cosnark_sync(ctx);
}

int cosnark_cosumcheck_phase1(cosnark_ctx *ctx, const uint64_t *d_p1, const uint64_t *d_p2, uint64_t M, int d,
                              uint64_t *h_poly) {
    if (!ctx || !d_p1 || !d_p2 || !h_poly) return set_err(ctx, COSNARK_E_ARG, "cosumcheck_phase1: null argument");
    if (d < 2 || !is_pow2((uint64_t)d) || !is_pow2(M) || M < (uint64_t)d)
        return set_err(ctx, COSNARK_E_ARG, "cosumcheck_phase1: d and M must be powers of two, 2 <= d <= M");
    if (d > BSC_D_MAX) return set_err(ctx, COSNARK_E_RANGE, "cosumcheck_phase1: d > 2048");
    const uint64_t nblocks = M / d;
    const int cols = 2 * d;
    int nwg;
    if (d <= BSC_SMALL_MAX) {
        const uint64_t tiles = (nblocks + BSC_TB - 1) / BSC_TB;
        nwg = (int)(tiles < BSC_MAX_WG ? tiles : BSC_MAX_WG);
    } else if (d == 256) {
        const uint64_t tiles = (nblocks + L256_B - 1) / L256_B;
        nwg = (int)(tiles < BSC_MAX_WG ? tiles : BSC_MAX_WG);
    } else {
        nwg = (int)(nblocks < BSC_MAX_WG ? nblocks : BSC_MAX_WG);
    }
    void *base;
    int rc = scratch_reserve(ctx, sizeof(fe) * ((size_t)nwg * cols + cols), &base);
    if (rc) return rc;
    fe *part = (fe *)base, *out = part + (size_t)nwg * cols;
    const fe *a = (const fe *)d_p1, *b = (const fe *)d_p2;
    hipStream_t s = ctx->stream;
    switch (d) {
        case 2: launch_bsc_small<2>(a, b, nblocks, part, nwg, s); break;
        case 4: launch_bsc_small<4>(a, b, nblocks, part, nwg, s); break;
        case 8: launch_bsc_small<8>(a, b, nblocks, part, nwg, s); break;
        case 16: launch_bsc_small<16>(a, b, nblocks, part, nwg, s); break;
        case 32: launch_bsc_small<32>(a, b, nblocks, part, nwg, s); break;
        case 256:
            hipLaunchKernelGGL(bsc_large256_kernel, dim3(nwg), dim3(CO_THREADS), sizeof(fe) * L256_LDS, s, a, b,
                               nblocks, (const fe *)ctx->tw_lo, (const fe *)ctx->tw_hi, part);
            break;
        default:
            hipLaunchKernelGGL(bsc_large_kernel, dim3(nwg), dim3(CO_THREADS), sizeof(fe) * 2 * d, s, a, b,
                               nblocks, ilog2((uint64_t)d), (const fe *)ctx->tw_lo, (const fe *)ctx->tw_hi,
                               (const fe *)ctx->itw_lo, (const fe *)ctx->itw_hi, part);
    }
    LAUNCHCK();
    const fe di = gf::inv(gf::from_u64((uint64_t)d));
    hipLaunchKernelGGL(bsc_final_kernel, dim3((cols + FIN_COLS - 1) / FIN_COLS), dim3(CO_THREADS), 0, s, (const fe *)part, nwg, d, 2,
                       gf::mul(di, di), out);
    LAUNCHCK();
    HIPCK(hipMemcpyAsync(ctx->co_host, out, sizeof(fe) * cols, hipMemcpyDeviceToHost, s));
    rc = cosnark_sync(ctx);
    if (rc) return rc;
    memcpy(h_poly, ctx->co_host, sizeof(fe) * cols);
    return COSNARK_OK;
}

int cosnark_cosumcheck_fold(cosnark_ctx *ctx, uint64_t *d_p1, uint64_t *d_p2, uint64_t M, int d,
                            const uint64_t *h_L) {
    if (!ctx || !d_p1 || !d_p2 || !h_L) return set_err(ctx, COSNARK_E_ARG, "cosumcheck_fold: null argument");
    uint64_t *v[4] = {d_p1, d_p2, nullptr, nullptr};
    return block_fold(ctx, v, 2, M, d, h_L, "cosumcheck_fold");
}

int cosnark_cosumcheck_phase2(cosnark_ctx *ctx, uint64_t *d_p1, uint64_t *d_p2, uint64_t M, int d,
                              const uint64_t r[2]) {
    if (!ctx || !r) return set_err(ctx, COSNARK_E_ARG, "cosumcheck_phase2: null argument");
    if (d < 2 || !is_pow2((uint64_t)d) || d > CO_HOST_ELEMS)
        return set_err(ctx, COSNARK_E_ARG, "cosumcheck_phase2: d must be a power of two in [2, 8192]");
    const std::vector<fe> L = lagrange_weights(gf::mk(r[0], r[1]), d);
    return cosnark_cosumcheck_fold(ctx, d_p1, d_p2, M, d, (const uint64_t *)L.data());
}

int cosnark_cosumcheck3_phase1(cosnark_ctx *ctx, const uint64_t *d_p1, const uint64_t *d_p2, const uint64_t *d_p3,
                               const uint64_t *d_b1, uint64_t M, int d, uint64_t *h_poly) {
    if (!ctx || !d_p1 || !d_p2 || !d_p3 || !d_b1 || !h_poly)
        return set_err(ctx, COSNARK_E_ARG, "cosumcheck3_phase1: null argument");
    if (d < 2 || !is_pow2((uint64_t)d) || !is_pow2(M) || M < (uint64_t)d)
        return set_err(ctx, COSNARK_E_ARG, "cosumcheck3_phase1: d and M must be powers of two, 2 <= d <= M");
    if (d > BSC3_D_MAX) return set_err(ctx, COSNARK_E_RANGE, "cosumcheck3_phase1: d > 1024");
    const int logd = ilog2((uint64_t)d), logT = std::max(logd, BSC3_TILE_LOG);
    const uint64_t ntiles = d <= BSC3_SMALL_MAX ? (M / d + 63) / 64 : (M + (1ull << logT) - 1) >> logT;
    const int nwg = (int)(ntiles < 1024 ? ntiles : 1024), cols = 4 * d;
    void *base;
    int rc = scratch_reserve(ctx, sizeof(fe) * ((size_t)nwg * cols + cols), &base);
    if (rc) return rc;
    fe *part = (fe *)base, *out = part + (size_t)nwg * cols;
    hipStream_t s = ctx->stream;
    const fe di = gf::inv(gf::from_u64((uint64_t)d));
    const fe *a1 = (const fe *)d_p1, *a2 = (const fe *)d_p2, *a3 = (const fe *)d_p3, *a4 = (const fe *)d_b1;
    if (d <= BSC3_SMALL_MAX) {
        const uint64_t nb = M / d;
        // the tile's rows, and the four per-thread coset sums at the end
        const size_t lds = sizeof(fe) * std::max<size_t>(4 * 64 * (d + 1), 4 * CO_THREADS);
        switch (d) {
            case 2: hipLaunchKernelGGL(bsc3_wave_kernel<2>, dim3(nwg), dim3(CO_THREADS), lds, s, a1, a2, a3, a4, nb, part); break;
            case 4: hipLaunchKernelGGL(bsc3_wave_kernel<4>, dim3(nwg), dim3(CO_THREADS), lds, s, a1, a2, a3, a4, nb, part); break;
            case 8: hipLaunchKernelGGL(bsc3_wave_kernel<8>, dim3(nwg), dim3(CO_THREADS), lds, s, a1, a2, a3, a4, nb, part); break;
            default: hipLaunchKernelGGL(bsc3_wave_kernel<16>, dim3(nwg), dim3(CO_THREADS), lds, s, a1, a2, a3, a4, nb, part); break;
        }
    } else if (d == 512) {
        hipLaunchKernelGGL(bsc3_512_kernel, dim3(nwg), dim3(CO_THREADS), sizeof(fe) * C512_LDS, s, a1, a2, a3, a4,
                           M / 512, (const fe *)ctx->tw_lo, (const fe *)ctx->tw_hi, di, part);
    } else {
        const size_t lds = sizeof(fe) * 5 * ((size_t)1 << logT) + sizeof(int) * ((size_t)1 << (logT - logd));
        hipLaunchKernelGGL(bsc3_kernel, dim3(nwg), dim3(CO_THREADS), lds, s, a1, a2, a3, a4, M, logd, logT,
                           (const fe *)ctx->tw_lo, (const fe *)ctx->tw_hi, (const fe *)ctx->itw_lo,
                           (const fe *)ctx->itw_hi, di, part);
    }
    LAUNCHCK();
    hipLaunchKernelGGL(bsc_final_kernel, dim3((cols + FIN_COLS - 1) / FIN_COLS), dim3(CO_THREADS), 0, s, (const fe *)part, nwg, d, 4,
                       gf::one(), out);
    LAUNCHCK();
    HIPCK(hipMemcpyAsync(ctx->co_host, out, sizeof(fe) * cols, hipMemcpyDeviceToHost, s));
    rc = cosnark_sync(ctx);
    if (rc) return rc;
    memcpy(h_poly, ctx->co_host, sizeof(fe) * cols);
    return COSNARK_OK;
}

int cosnark_cosumcheck3_fold(cosnark_ctx *ctx, uint64_t *d_p1, uint64_t *d_p2, uint64_t *d_p3, uint64_t *d_b1,
                             uint64_t M, int d, const uint64_t *h_L) {
    if (!ctx || !h_L) return set_err(ctx, COSNARK_E_ARG, "cosumcheck3_fold: null argument");
    uint64_t *v[4] = {d_p1, d_p2, d_p3, d_b1};
    return block_fold(ctx, v, 4, M, d, h_L, "cosumcheck3_fold");
}

int cosnark_cosumcheck3_phase2(cosnark_ctx *ctx, uint64_t *d_p1, uint64_t *d_p2, uint64_t *d_p3, uint64_t *d_b1,
                               uint64_t M, int d, const uint64_t r[2]) {
    if (!ctx || !r) return set_err(ctx, COSNARK_E_ARG, "cosumcheck3_phase2: null argument");
    if (d < 2 || !is_pow2((uint64_t)d) || d > CO_HOST_ELEMS)
        return set_err(ctx, COSNARK_E_ARG, "cosumcheck3_phase2: d must be a power of two in [2, 8192]");
    const std::vector<fe> L = lagrange_weights(gf::mk(r[0], r[1]), d);
    return cosnark_cosumcheck3_fold(ctx, d_p1, d_p2, d_p3, d_b1, M, d, (const uint64_t *)L.data());
}

int cosnark_share_unpack(uint64_t *h_secret, int n, int kk, uint64_t *h_out) {
    if (!h_secret || !h_out) return COSNARK_E_ARG;
    if (n < 1 || n > CO_HOST_ELEMS || !is_pow2((uint64_t)n) || kk < 1 || kk > n) return COSNARK_E_ARG;
    std::vector<fe> x(n);
    for (int i = 0; i < n; ++i) x[i] = gf::mk(h_secret[2 * i], h_secret[2 * i + 1]);
    unpack_values(x);
    for (int i = 0; i < n; ++i) {
        h_secret[2 * i] = x[i].re;
        h_secret[2 * i + 1] = x[i].im;
    }
    for (int j = 0; j < kk; ++j) {
        const fe y = x[(uint64_t)n * j / kk];
        h_out[2 * j] = y.re;
        h_out[2 * j + 1] = y.im;
    }
    return COSNARK_OK;
}

int cosnark_sumcheck_aggr(const uint64_t *h_polys, int nranks, int len, const uint64_t *h_b, uint64_t *h_out) {
    if (!h_polys || !h_out || (nranks >= 4 && !h_b)) return COSNARK_E_ARG;
    if (nranks < 1 || nranks > CO_HOST_ELEMS || !is_pow2((uint64_t)nranks) || len < 0) return COSNARK_E_ARG;
    // out[i] = sum_{j < N/4} b_j U(P[.][i])[2j] with U the unpack map; U is
    // linear, so out[i] = sum_m w_m P[m][i], w_m = sum_j b_j U(e_m)[2j]
    const int N = nranks;
    std::vector<fe> w(N, gf::zero()), e(N);
    for (int m = 0; m < N && N >= 4; ++m) {
        std::fill(e.begin(), e.end(), gf::zero());
        e[m] = gf::one();
        unpack_values(e);
        for (int j = 0; j < N / 4; ++j)
            w[m] = gf::add(w[m], gf::mul(gf::mk(h_b[2 * j], h_b[2 * j + 1]), e[2 * j]));
    }
    for (int i = 0; i < len; ++i) {
        fe acc = gf::zero();
        for (int m = 0; m < N; ++m) {
            const uint64_t *p = h_polys + 2 * ((uint64_t)m * len + i);
            acc = gf::add(acc, gf::mul(w[m], gf::mk(p[0], p[1])));
        }
        h_out[2 * i] = acc.re;
        h_out[2 * i + 1] = acc.im;
    }
    return COSNARK_OK;
}

}  // extern "C"

// RS_multifold5 outputs at positions idx[0..nq) (idx null: all n/fd), with
// their gathered segments in seg (optional); everything on ctx->stream.
static int multifold_launch(cosnark_ctx *ctx, const fe *cw, uint64_t n, const uint64_t *h_a, int fd, fe *out,
                            const uint32_t *idx, uint64_t nq, fe *seg, fe *a_dev) {
    const int logn = ilog2(n), logfd = ilog2((uint64_t)fd);
    const uint64_t nout = n / fd;
    const fe fi = gf::inv(gf::from_u64((uint64_t)fd));
    hipStream_t s = ctx->stream;
    const fe *itlo = (const fe *)ctx->itw_lo, *ithi = (const fe *)ctx->itw_hi;
    const int shift = NTT_LOG_MAX - logn;
    if (fd <= MF_SMALL_MAX) {
        FoldWeights w{};
        for (int k = 0; k < fd; ++k) w.a[k] = gf::mul(gf::mk(h_a[2 * k], h_a[2 * k + 1]), fi);
        switch (fd) {
            case 2: launch_mf_small<2>(cw, nout, shift, itlo, ithi, w, out, idx, nq, seg, s); break;
            case 4: launch_mf_small<4>(cw, nout, shift, itlo, ithi, w, out, idx, nq, seg, s); break;
            case 8: launch_mf_small<8>(cw, nout, shift, itlo, ithi, w, out, idx, nq, seg, s); break;
            case 16: launch_mf_small<16>(cw, nout, shift, itlo, ithi, w, out, idx, nq, seg, s); break;
            default: launch_mf_small<32>(cw, nout, shift, itlo, ithi, w, out, idx, nq, seg, s); break;
        }
        LAUNCHCK();
    } else {
        for (int k = 0; k < fd; ++k) ctx->co_host[k] = gf::mul(gf::mk(h_a[2 * k], h_a[2 * k + 1]), fi);
        HIPCK(hipMemcpyAsync(a_dev, ctx->co_host, sizeof(fe) * fd, hipMemcpyHostToDevice, s));
        const unsigned g = (unsigned)(nq < 4096 ? nq : 4096);
        hipLaunchKernelGGL(multifold_large_kernel, dim3(g), dim3(CO_THREADS), sizeof(fe) * (fd + CO_THREADS), s, cw,
                           nout, logfd, logn, itlo, ithi, (const fe *)a_dev, out, idx, nq, seg);
        LAUNCHCK();
    }
    return COSNARK_OK;
}

extern "C" {

int cosnark_rs_multifold5(cosnark_ctx *ctx, const uint64_t *d_cw, uint64_t n, const uint64_t *h_a, int fd,
                          uint64_t *d_out) {
    if (!ctx || !d_cw || !h_a || !d_out) return set_err(ctx, COSNARK_E_ARG, "rs_multifold5: null argument");
    if (fd < 2 || !is_pow2((uint64_t)fd) || !is_pow2(n) || n < (uint64_t)fd)
        return set_err(ctx, COSNARK_E_ARG, "rs_multifold5: fd and n must be powers of two, 2 <= fd <= n");
    if (ilog2(n) > NTT_LOG_MAX || fd > MF_FD_MAX)
        return set_err(ctx, COSNARK_E_RANGE, "rs_multifold5: n > 2^28 or fd > 4096");
    if ((const void *)d_cw == (const void *)d_out)
        return set_err(ctx, COSNARK_E_ARG, "rs_multifold5: output must not alias the codeword");
    fe *ad = nullptr;
    if (fd > MF_SMALL_MAX) {
        void *base;
        int rc = scratch_reserve(ctx, sizeof(fe) * fd, &base);
        if (rc) return rc;
        ad = (fe *)base;
    }
    int rc = multifold_launch(ctx, (const fe *)d_cw, n, h_a, fd, (fe *)d_out, nullptr, n / fd, nullptr, ad);
    if (rc) return rc;
    return cosnark_sync(ctx);
}

int cosnark_multifold_segments(cosnark_ctx *ctx, const uint64_t *d_cw, uint64_t n, const uint64_t *h_L, int fd,
                               const uint32_t *h_idx, int l, uint64_t *h_segments, uint64_t *h_y) {
    if (!ctx || !d_cw || !h_L || !h_idx || !h_y) return set_err(ctx, COSNARK_E_ARG, "multifold_segments: null argument");
    if (fd < 2 || !is_pow2((uint64_t)fd) || !is_pow2(n) || n < (uint64_t)fd || l < 0)
        return set_err(ctx, COSNARK_E_ARG, "multifold_segments: fd and n must be powers of two, 2 <= fd <= n");
    if (ilog2(n) > NTT_LOG_MAX || fd > MF_FD_MAX || l > 65536)
        return set_err(ctx, COSNARK_E_RANGE, "multifold_segments: n > 2^28, fd > 4096 or more than 65536 queries");
    if (l == 0) return COSNARK_OK;
    for (int q = 0; q < l; ++q)
        if ((uint64_t)h_idx[q] >= n / fd) return set_err(ctx, COSNARK_E_ARG, "multifold_segments: index out of range");
    // device block: indexes, weights (fd > 32), outputs, segments
    const size_t off_a = align256(sizeof(uint32_t) * l), off_y = off_a + align256(sizeof(fe) * fd);
    const size_t off_seg = off_y + align256(sizeof(fe) * l);
    const size_t total = off_seg + (h_segments ? sizeof(fe) * (size_t)l * fd : 0);
    void *base;
    int rc = scratch_reserve(ctx, total, &base);
    if (rc) return rc;
    uint8_t *b = (uint8_t *)base;
    hipStream_t s = ctx->stream;
    HIPCK(hipMemcpyAsync(b, h_idx, sizeof(uint32_t) * l, hipMemcpyHostToDevice, s));
    HIPCK(hipStreamSynchronize(s));     // h_idx is pageable; the weights below reuse the staging buffer
    fe *seg = h_segments ? (fe *)(b + off_seg) : nullptr;
    rc = multifold_launch(ctx, (const fe *)d_cw, n, h_L, fd, (fe *)(b + off_y), (const uint32_t *)b, (uint64_t)l,
                          seg, (fe *)(b + off_a));
    if (rc) return rc;
    HIPCK(hipMemcpyAsync(h_y, b + off_y, sizeof(fe) * l, hipMemcpyDeviceToHost, s));
    if (h_segments) HIPCK(hipMemcpyAsync(h_segments, seg, sizeof(fe) * (size_t)l * fd, hipMemcpyDeviceToHost, s));
    return cosnark_sync(ctx);
}

int cosnark_row_combine(cosnark_ctx *ctx, const uint64_t *d_rows, int nrows, uint64_t len, const uint64_t *h_c,
                        uint64_t *d_out) {
    if (!ctx || !d_rows || !h_c || !d_out) return set_err(ctx, COSNARK_E_ARG, "row_combine: null argument");
    if (nrows < 1 || nrows > 32 || len == 0) return set_err(ctx, COSNARK_E_ARG, "row_combine: 1 <= nrows <= 32, len > 0");
    RowWeights w{};
    for (int i = 0; i < nrows; ++i) w.c[i] = gf::mk(h_c[2 * i], h_c[2 * i + 1]);
    const uint64_t g = (len + CO_THREADS - 1) / CO_THREADS;
    hipLaunchKernelGGL(row_combine_kernel, dim3((unsigned)(g < 8192 ? g : 8192)), dim3(CO_THREADS), 0, ctx->stream,
                       (const fe *)d_rows, nrows, len, w, (fe *)d_out);
    LAUNCHCK();
    return cosnark_sync(ctx);
}

}  // extern "C"

// ---- coSumcheck drivers (src/coSumcheck.cpp:233-448), one rank ------------
//
// The rank's block-sum rounds run on the GPU (phase 1, fold); the masks, the
// Lagrange weights, the transcript and rank 0's share arithmetic run on the
// host.  The reference's gathers to rank 0 followed by a broadcast of the
// combined result (sumcheck_aggr, src/MPI_utils.cpp:2604-2670;
// send_final_sumcheck_poly, :2672-2736) are one all-gather through the
// caller's callback: every rank then forms the same combination itself.

namespace {

struct CoDriver {
    cosnark_ctx *ctx;
    int N, rank;
    cosnark_allgather_fn gather;
    void *user;
    cosnark_transcript *ts;
    std::vector<std::vector<fe>> R;
    int stride;                          // mask stride: 2 quadratic, 4 cubic
    fe c;
    std::vector<fe> ip;                  // ip_vec, the rounds' vectors back to back
    std::vector<int> lens;
    int fails = 0;
    double ps = 0;

    fe absorb(const fe *x, size_t n) {
        fe prev = gf::mk(ts->prev[0], ts->prev[1]);
        for (size_t i = 0; i < n; ++i) prev = ts::hash_F(prev, x[i]);
        ts->prev[0] = prev.re;
        ts->prev[1] = prev.im;
        return prev;
    }
    // H[t] += c H_r[t] (coSumcheck.cpp:182-193 / :48-58)
    void mask(std::vector<fe> &H, int iter) const {
        fe h0 = R[iter][0];
        for (size_t j = iter + 1; j < R.size(); ++j)
            for (size_t k = 0; k < R[j].size() / stride; ++k) h0 = gf::add(h0, R[j][stride * k]);
        for (size_t t = 0; t < H.size(); ++t) H[t] = gf::add(H[t], gf::mul(c, t == 0 ? h0 : R[iter][t]));
    }
    // every rank's values of `mine`, rank-major
    int allgather(const std::vector<fe> &mine, std::vector<fe> &all) {
        all.assign((size_t)N * mine.size(), gf::zero());
        const int rc = gather(user, (const uint64_t *)mine.data(), mine.size(), (uint64_t *)all.data());
        return rc ? set_err(ctx, COSNARK_E_ARG, "coSumcheck: the all-gather callback failed") : COSNARK_OK;
    }
    // sumcheck_aggr(H, b, N): the gathered polynomials combined with b
    int aggr(std::vector<fe> &H, const std::vector<fe> &b) {
        std::vector<fe> all;
        int rc = allgather(H, all);
        if (rc) return rc;
        if (cosnark_sumcheck_aggr((const uint64_t *)all.data(), N, (int)H.size(), (const uint64_t *)b.data(),
                                  (uint64_t *)H.data()))
            return set_err(ctx, COSNARK_E_ARG, "coSumcheck: sumcheck_aggr sizes");
        if (rank == 0) ps += (double)H.size() * 16 / 1024.0;
        return COSNARK_OK;
    }
    // one round's bookkeeping after the aggregated polynomial H of `len`
    // entries is known: the check, the challenge, ip_vec, K, the sparse mask
    // fold (coSumcheck.cpp:215-228); returns r
    fe round(const std::vector<fe> &H, int i, fe &K, fe &final_eval_r) {
        fe eval = gf::zero();
        for (size_t j = 0; j < H.size() / stride; ++j) eval = gf::add(eval, H[stride * j]);
        if (!gf::eq(eval, K)) fails |= 1 << i;
        const fe r = absorb(H.data(), H.size());
        const std::vector<fe> L = lagrange_weights(r, (int)H.size());
        size_t off = 0;
        for (size_t j = 0; j < lens.size(); ++j) {
            for (int k = 0; k < lens[j]; ++k) ip[off + k] = gf::mul(ip[off + k], (int)j == i ? L[k] : L[0]);
            off += lens[j];
        }
        K = gf::zero();
        for (size_t j = 0; j < H.size(); ++j) K = gf::add(K, gf::mul(L[j], H[j]));
        fe s = gf::zero();
        for (size_t j = 0; j < R[i].size(); ++j) s = gf::add(s, gf::mul(L[j], R[i][j]));
        for (size_t j = i + 1; j < R.size(); ++j)
            for (auto &x : R[j]) x = gf::mul(x, L[0]);
        if ((size_t)i + 1 < R.size()) R[i + 1][0] = gf::add(R[i + 1][0], s);
        final_eval_r = s;
        return r;
    }
    // send_final_sumcheck_poly(s, N): coordinate i's N shares -> N/2 values
    int final_values(const std::vector<fe> &mine, std::vector<std::vector<fe>> &H) {
        std::vector<fe> all;
        int rc = allgather(mine, all);
        if (rc) return rc;
        H.assign(mine.size(), std::vector<fe>(N / 2));
        std::vector<fe> P(N);
        for (size_t i = 0; i < mine.size(); ++i) {
            for (int m = 0; m < N; ++m) P[m] = all[(size_t)m * mine.size() + i];
            cosnark_share_unpack((uint64_t *)P.data(), N, N / 2, (uint64_t *)H[i].data());
        }
        if (rank == 0) ps += (double)mine.size() * (N / 2) * 16 / 1024.0;
        return COSNARK_OK;
    }
    std::vector<fe> closing(const fe &r) const {
        return N / 4 == 1 ? std::vector<fe>{gf::one()} : lagrange_weights(r, N / 4);
    }
};

int codriver_args(cosnark_ctx *ctx, uint64_t M, const uint64_t *const *h_R, const uint64_t *R_sizes, int nR,
                  const int *degrees, int ndeg, int nranks, int rank, cosnark_allgather_fn gather, void *ts,
                  int rmul, int dmax, int final_len, const char *what) {
    if (!ctx || !h_R || !R_sizes || !degrees || !gather || !ts) return set_err(ctx, COSNARK_E_ARG, std::string(what) + ": null argument");
    if (nranks < 4 || nranks > CO_HOST_ELEMS || !is_pow2((uint64_t)nranks) || rank < 0 || rank >= nranks)
        return set_err(ctx, COSNARK_E_ARG, std::string(what) + ": nranks a power of two >= 4 (the reference divides by N/4)");
    if (ndeg < 1 || nR != ndeg || !is_pow2(M)) return set_err(ctx, COSNARK_E_ARG, std::string(what) + ": one mask per round, M a power of two");
    uint64_t Mc = M;
    for (int i = 0; i < ndeg; ++i) {
        if (degrees[i] < 2 || degrees[i] > dmax || !is_pow2((uint64_t)degrees[i]) || Mc < (uint64_t)degrees[i])
            return set_err(ctx, COSNARK_E_ARG, std::string(what) + ": degrees powers of two within the vector");
        if (R_sizes[i] != (uint64_t)rmul * degrees[i] || !h_R[i])
            return set_err(ctx, COSNARK_E_ARG, std::string(what) + ": mask i must hold the round's polynomial size");
        Mc /= (uint64_t)degrees[i];
    }
    if (Mc < (uint64_t)final_len) return set_err(ctx, COSNARK_E_ARG, std::string(what) + ": vectors folded below the final values");
    return COSNARK_OK;
}

}  // namespace

extern "C" {

int cosnark_quadratic_cosumcheck(cosnark_ctx *ctx, uint64_t *d_p1, uint64_t *d_p2, uint64_t M,
                                 const uint64_t *const *h_R, const uint64_t *R_sizes, int nR, const uint64_t K_in[2],
                                 const uint64_t K_r[2], const int *degrees, int ndeg, int nranks, int rank,
                                 cosnark_allgather_fn gather, void *user, cosnark_transcript *ts, uint64_t *h_claims,
                                 uint64_t *h_r, uint64_t *h_ip_vec, double *ps, int *check_failures) {
    int rc = codriver_args(ctx, M, h_R, R_sizes, nR, degrees, ndeg, nranks, rank, gather, ts, 2, BSC_D_MAX, 2,
                           "quadratic_coSumcheck");
    if (rc) return rc;
    if (!d_p1 || !d_p2 || !K_in || !K_r || !h_claims || !h_r)
        return set_err(ctx, COSNARK_E_ARG, "quadratic_coSumcheck: null argument");
    CoDriver D{ctx, nranks, rank, gather, user, ts, {}, 2, gf::zero(), {}, {}};
    for (int i = 0; i < nR; ++i) D.R.emplace_back((const fe *)h_R[i], (const fe *)h_R[i] + R_sizes[i]);
    fe K = gf::mk(K_in[0], K_in[1]);
    const fe kk[2] = {K, gf::mk(K_r[0], K_r[1])};
    D.c = D.absorb(kk, 2);                                               // :234
    K = gf::add(K, gf::mul(D.c, kk[1]));                                 // :235
    for (int i = 0; i < ndeg; ++i) D.lens.push_back(2 * degrees[i]);
    for (int l : D.lens) D.ip.insert(D.ip.end(), (size_t)l, gf::one()); // :244-247
    const std::vector<fe> ones(nranks / 4, gf::one());
    std::vector<fe> rs;
    fe final_eval_r = gf::zero();
    uint64_t Mc = M;
    for (int i = 0; i < ndeg; ++i) {
        const int d = degrees[i];
        std::vector<fe> H(2 * d);
        if ((rc = cosnark_cosumcheck_phase1(ctx, d_p1, d_p2, Mc, d, (uint64_t *)H.data()))) return rc;
        D.mask(H, i);
        if ((rc = D.aggr(H, ones))) return rc;                           // :197
        const fe r = D.round(H, i, K, final_eval_r);                     // :252-284
        rs.push_back(r);
        const uint64_t rr[2] = {r.re, r.im};
        if ((rc = cosnark_cosumcheck_phase2(ctx, d_p1, d_p2, Mc, d, rr))) return rc;   // :286-289
        Mc /= (uint64_t)d;
    }
    fe v[4];
    HIPCK(hipMemcpy(v, d_p1, 2 * sizeof(fe), hipMemcpyDeviceToHost));
    HIPCK(hipMemcpy(v + 2, d_p2, 2 * sizeof(fe), hipMemcpyDeviceToHost));
    std::vector<std::vector<fe>> H;
    if ((rc = D.final_values({v[0], v[2], v[1], v[3], final_eval_r}, H))) return rc;   // :293
    for (int i = 0; i < nranks / 4; ++i)                                 // :298-300
        K = gf::sub(K, gf::add(gf::add(gf::mul(H[0][i], H[1][i]), gf::mul(H[2][i], H[3][i])), gf::mul(D.c, H[4][i])));
    if (!gf::eq(K, gf::zero())) D.fails |= 1 << ndeg;
    fe r = D.absorb(H[0].data(), H[0].size());                           // :306-307
    for (size_t i = 1; i < H.size(); ++i) r = D.absorb(H[i].data(), H[i].size());
    const std::vector<fe> L = D.closing(r);                              // :311-316
    for (size_t i = 0; i < H.size(); ++i) {                              // :318-325
        fe e = gf::zero();
        for (int j = 0; j < nranks / 4; ++j) e = gf::add(e, gf::mul(L[j], H[i][j]));
        h_claims[2 * i] = e.re;
        h_claims[2 * i + 1] = e.im;
    }
    rs.insert(rs.begin(), r);
    memcpy(h_r, rs.data(), sizeof(fe) * rs.size());
    if (h_ip_vec) memcpy(h_ip_vec, D.ip.data(), sizeof(fe) * D.ip.size());
    if (ps) *ps += D.ps;
    if (check_failures) *check_failures = D.fails;
    return COSNARK_OK;
}

int cosnark_cubic_cosumcheck(cosnark_ctx *ctx, uint64_t *d_p1, uint64_t *d_p2, uint64_t *d_p3, uint64_t M,
                             const uint64_t *h_r1, int nr1, const uint64_t *h_r2, int nr2, const uint64_t *const *h_R,
                             const uint64_t *R_sizes, int nR, const uint64_t K_in[2], const uint64_t K_r[2],
                             const int *degrees, int ndeg, int nranks, int rank, cosnark_allgather_fn gather,
                             void *user, cosnark_transcript *ts, uint64_t *h_claims, uint64_t *h_r, uint64_t *h_ip_vec,
                             double *ps, int *check_failures) {
    int rc = codriver_args(ctx, M, h_R, R_sizes, nR, degrees, ndeg, nranks, rank, gather, ts, 4, BSC3_D_MAX, 1,
                           "cubic_coSumcheck");
    if (rc) return rc;
    if (!d_p1 || !d_p2 || !d_p3 || !h_r1 || !h_r2 || !K_in || !K_r || !h_claims || !h_r)
        return set_err(ctx, COSNARK_E_ARG, "cubic_coSumcheck: null argument");
    if (nr1 < 1 || nr1 > NTT_LOG_MAX || (1ull << nr1) != M || nr2 < 0 || nr2 > 13 || (1 << nr2) < nranks / 4)
        return set_err(ctx, COSNARK_E_ARG, "cubic_coSumcheck: |p| = 2^|r1| and 2^|r2| >= N/4");
    int logsum = 0;
    for (int i = 0; i < ndeg; ++i) logsum += ilog2((uint64_t)degrees[i]);
    if (logsum > nr1) return set_err(ctx, COSNARK_E_ARG, "cubic_coSumcheck: degrees beyond |r1|");
    CoDriver D{ctx, nranks, rank, gather, user, ts, {}, 4, gf::zero(), {}, {}};
    for (int i = 0; i < nR; ++i) D.R.emplace_back((const fe *)h_R[i], (const fe *)h_R[i] + R_sizes[i]);
    fe K = gf::mk(K_in[0], K_in[1]);
    const fe kk[2] = {K, gf::mk(K_r[0], K_r[1])};
    D.c = D.absorb(kk, 2);                                               // :336
    K = gf::add(K, gf::mul(D.c, kk[1]));                                 // :337
    // b1 = precompute_beta(r1) on the device (M elements); b2 on the host
    uint64_t *d_b1 = nullptr;
    HIPCK(hipMalloc(&d_b1, sizeof(fe) * M));
    struct Free { uint64_t *p; ~Free() { if (p) (void)hipFree(p); } } fb{d_b1};
    if ((rc = cosnark_precompute_beta(ctx, h_r1, nr1, d_b1))) return rc;  // :342
    std::vector<fe> b2(1u << nr2);                                       // :343
    {
        b2[0] = gf::one();
        const fe *r2 = (const fe *)h_r2;
        for (int i = 0; i < nr2; ++i) {                                  // precompute_beta, r2[0] on bit 0
            const fe ri = r2[nr2 - 1 - i];
            for (size_t j = (size_t)1 << i; j-- > 0;) {
                const fe t = gf::mul(ri, b2[j]);
                b2[2 * j] = gf::sub(b2[j], t);
                b2[2 * j + 1] = t;
            }
        }
    }
    for (int i = 0; i < ndeg; ++i) D.lens.push_back(4 * degrees[i]);
    for (int l : D.lens) D.ip.insert(D.ip.end(), (size_t)l, gf::one());
    std::vector<fe> rs;
    std::vector<std::vector<fe>> LV;
    fe final_eval_r = gf::zero();
    uint64_t Mc = M;
    for (int i = 0; i < ndeg; ++i) {
        const int d = degrees[i];
        std::vector<fe> H(4 * d);
        if ((rc = cosnark_cosumcheck3_phase1(ctx, d_p1, d_p2, d_p3, d_b1, Mc, d, (uint64_t *)H.data()))) return rc;
        D.mask(H, i);
        if ((rc = D.aggr(H, b2))) return rc;                             // :61
        const fe r = D.round(H, i, K, final_eval_r);                     // :356-388
        rs.push_back(r);
        LV.push_back(lagrange_weights(r, d));                            // :390-392
        if ((rc = cosnark_cosumcheck3_fold(ctx, d_p1, d_p2, d_p3, d_b1, Mc, d, (const uint64_t *)LV.back().data())))
            return rc;                                                   // :393
        Mc /= (uint64_t)d;
    }
    fe v[4];
    HIPCK(hipMemcpy(v, d_p1, sizeof(fe), hipMemcpyDeviceToHost));
    HIPCK(hipMemcpy(v + 1, d_p2, sizeof(fe), hipMemcpyDeviceToHost));
    HIPCK(hipMemcpy(v + 2, d_p3, sizeof(fe), hipMemcpyDeviceToHost));
    HIPCK(hipMemcpy(v + 3, d_b1, sizeof(fe), hipMemcpyDeviceToHost));
    const fe b10 = v[3];
    std::vector<std::vector<fe>> H;
    if ((rc = D.final_values({v[0], v[1], v[2], final_eval_r}, H))) return rc;   // :398
    for (int i = 0; i < nranks / 4; ++i)                                 // :403-405
        K = gf::sub(K, gf::mul(b2[i], gf::add(gf::mul(b10, gf::sub(gf::mul(H[0][i], H[1][i]), H[2][i])),
                                              gf::mul(D.c, H[3][i]))));
    if (!gf::eq(K, gf::zero())) D.fails |= 1 << ndeg;
    fe r = D.absorb(H[0].data(), H[0].size());                           // :410: H[3] is not absorbed
    r = D.absorb(H[1].data(), H[1].size());
    r = D.absorb(H[2].data(), H[2].size());
    const std::vector<fe> L = D.closing(r);
    for (size_t i = 0; i < H.size(); ++i) {                              // :420-427
        fe e = gf::zero();
        for (int j = 0; j < nranks / 4; ++j) e = gf::add(e, gf::mul(L[j], H[i][j]));
        h_claims[2 * i] = e.re;
        h_claims[2 * i + 1] = e.im;
    }
    // verification 2 (:429-443): b1[0] = prod_i <beta(r1 slice i), L_i>
    {
        fe y = gf::one();
        int ctr = 0;
        const fe *r1 = (const fe *)h_r1;
        for (int i = 0; i < ndeg; ++i) {
            const int lg = ilog2((uint64_t)degrees[i]);
            std::vector<fe> tb((size_t)1 << lg);
            tb[0] = gf::one();
            for (int a = 0; a < lg; ++a) {
                const fe ri = r1[ctr + lg - 1 - a];
                for (size_t j = (size_t)1 << a; j-- > 0;) {
                    const fe t = gf::mul(ri, tb[j]);
                    tb[2 * j] = gf::sub(tb[j], t);
                    tb[2 * j + 1] = t;
                }
            }
            ctr += lg;
            fe s = gf::zero();
            for (int j = 0; j < degrees[i]; ++j) s = gf::add(s, gf::mul(tb[j], LV[i][j]));
            y = gf::mul(y, s);
        }
        if (!gf::eq(y, b10)) D.fails |= 1 << (ndeg + 1);
    }
    rs.insert(rs.begin(), r);
    memcpy(h_r, rs.data(), sizeof(fe) * rs.size());
    if (h_ip_vec) memcpy(h_ip_vec, D.ip.data(), sizeof(fe) * D.ip.size());
    if (ps) *ps += D.ps;
    if (check_failures) *check_failures = D.fails;
    return COSNARK_OK;
}

}  // extern "C"

// ---- PCS_commit at one rank (src/coPCS.cpp:236-250) ------------------------

namespace {

constexpr int PCS_BATCH_FACTOR = 32;     // BATCH_FACTOR, src/New_PCS.cpp:9

// prepare_data's layout (src/coPCS.cpp:76-128) as copies: rows of B =
// n / (BATCH_FACTOR / 2) elements; each holds B - q0 shares then q0 random
// values until the shares run out (the row where they do is zero-filled
// after its last share), and the next row takes random values only
struct Seg {
    int row;
    uint64_t col, src, len;
    bool random;
};

int pcs_prepare_plan(uint64_t n, uint64_t nr, uint64_t q0, std::vector<Seg> &segs, int *nrows, uint64_t *B_out) {
    const uint64_t B = n / (PCS_BATCH_FACTOR / 2);
    if (B <= q0) return COSNARK_E_ARG;
    uint64_t data = 0, rnd = 0;
    int last = -1;
    for (int i = 0; i < PCS_BATCH_FACTOR; ++i) {
        const uint64_t take = std::min<uint64_t>(B - q0, n - data);
        segs.push_back({i, 0, data, take, false});
        data += take;
        segs.push_back({i, B - q0, rnd, q0, true});
        rnd += q0;
        if (data == n) { last = i; break; }
    }
    if (last < 0) return COSNARK_E_RANGE;                   // more batches needed (the reference exits, :106-109)
    ++last;                                                 // :110
    if (last >= PCS_BATCH_FACTOR) return COSNARK_E_RANGE;   // no row left for the random row
    // :113-126; the reference compares Shares.size() - rand_ctr as an
    // unsigned value, so more random values used than shares is case 1
    uint64_t cnt;
    if (rnd > n || n - rnd > B) {
        if (nr <= rnd) return COSNARK_E_ARG;                // it would read past Random
        cnt = std::min<uint64_t>(B, nr - rnd);              // stops when Random runs out
    } else {
        cnt = B - (n - rnd);
    }
    segs.push_back({last, 0, rnd, cnt, true});
    if (rnd + cnt > nr) return COSNARK_E_ARG;               // the reference reads past Random there
    *nrows = last + 1;
    *B_out = B;
    return COSNARK_OK;
}

}  // namespace

extern "C" {

int cosnark_pcs_prepare_rows(uint64_t n_shares, uint64_t n_random, int q0, int *nrows, uint64_t *batch_size) {
    if (!nrows || !batch_size || q0 < 0 || !is_pow2(n_shares) || n_shares < PCS_BATCH_FACTOR / 2) return COSNARK_E_ARG;
    std::vector<Seg> segs;
    return pcs_prepare_plan(n_shares, n_random, (uint64_t)q0, segs, nrows, batch_size);
}

int cosnark_pcs_commit(cosnark_ctx *ctx, const uint64_t *d_shares, uint64_t n_shares, const uint64_t *d_random,
                       uint64_t n_random, int q0, int rate, int log_rate_offset, uint64_t *d_rows,
                       uint64_t *d_codewords, uint8_t *d_nodes, int *nrows) {
    if (!ctx || !d_shares || !d_random || !d_rows || !d_codewords || !d_nodes || !nrows)
        return set_err(ctx, COSNARK_E_ARG, "pcs_commit: null argument");
    if (!is_pow2(n_shares) || n_shares < PCS_BATCH_FACTOR / 2 || q0 < 0 || rate < 1 || !is_pow2((uint64_t)rate))
        return set_err(ctx, COSNARK_E_ARG, "pcs_commit: shares a power of two >= 16, rate a power of two");
    std::vector<Seg> segs;
    uint64_t B = 0;
    int rows = 0;
    int rc = pcs_prepare_plan(n_shares, n_random, (uint64_t)q0, segs, &rows, &B);
    if (rc == COSNARK_E_RANGE) return set_err(ctx, rc, "pcs_commit: the shares do not fit BATCH_FACTOR rows");
    if (rc) return set_err(ctx, rc, "pcs_commit: q0 >= the row size, or too few random values");
    if (B * (uint64_t)rate > (1ull << NTT_LOG_MAX)) return set_err(ctx, COSNARK_E_RANGE, "pcs_commit: codeword too long");
    hipStream_t s = ctx->stream;
    // prepare_data: rows zero-filled, then the share and random segments
    HIPCK(hipMemsetAsync(d_rows, 0, sizeof(fe) * B * rows, s));
    for (const Seg &g : segs) {
        if (!g.len) continue;
        const uint64_t *src = (g.random ? d_random : d_shares) + 2 * g.src;
        HIPCK(hipMemcpyAsync(d_rows + 2 * (g.row * B + g.col), src, sizeof(fe) * g.len, hipMemcpyDeviceToDevice, s));
    }
    // Encode (:131-147), then distributed_MT at one rank (:152-183)
    if ((rc = cosnark_copcs_encode(ctx, d_rows, rows, B, rate, log_rate_offset, d_codewords))) return rc;
    if ((rc = cosnark_distributed_mt(ctx, d_codewords, rows, B * (uint64_t)rate, d_nodes))) return rc;
    *nrows = rows;
    return COSNARK_OK;
}

}  // extern "C"
