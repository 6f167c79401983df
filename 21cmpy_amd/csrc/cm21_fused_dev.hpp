// Device bodies of the fused, pruned inverse transforms, shared by the
// single-box radius loop (cm21_fused.hip) and the batched loop
// (cm21_batched.hip) so that both run the same arithmetic.
//
// Layout conventions: S is a c2r-ready half spectrum S[a][b][c] (c fastest,
// H = D/2+1 entries).  The columns and rows passes write T[p][q][c] with a row
// stride `ho` >= cmax+1 (H for the single-box loop, cmax+1 for the compact
// per-radius storage of the batched loop).  The lines pass reads the prefix
// c < cl of each line and produces the D real values of the line.
#pragma once

#include <cmath>

#include "cm21_dft.hpp"
#include "cm21_fused_api.hpp"
#include "cm21_internal.hpp"
#include "cm21_window.hpp"

// The transforms may contract a*b+c into FMAs; the parity-critical float32
// steps use explicit __f*_rn intrinsics, which are never contracted.
#pragma clang fp contract(fast)

namespace cm21 {

enum LineMode { kLineSum = 0, kLineThreshold = 1, kLineFinal = 2 };

// Window of one mode from its integer |n|^2.  Inside the shell m2 >= m2lo the
// reference's float32 cut test decides (find_bubbles.cu:22-35); the Gaussian
// factor is exp2(m2 * gcoef); `scale` folds 1/N (and the growth factor of
// nonlinear input).
__device__ inline float2 fused_window(float2 x, int a, int b, int c, int m2, int D, const FusedParams& fp,
                                      float scale) {
    if (fp.exact_cut && m2 >= fp.m2lo) {
        const float kR = __fmul_rn(kmag_of(a, b, c, D, fp.dk, 0), fp.R);
        if ((float)((double)kR * 0.413566994) > 1.0f) return make_float2(0.0f, 0.0f);
    }
    if (fp.gcoef != 0.0f) {
        const float w2 = exp2f((float)m2 * fp.gcoef);
        x = make_float2(__fmul_rn(x.x, w2), __fmul_rn(x.y, w2));
    }
    return make_float2(__fmul_rn(x.x, scale), __fmul_rn(x.y, scale));
}

// Second half of the four-step DFT of W columns of length D = D1*D2: thread
// (j, n1) holds the D2-point DFT outputs v[k2] of column j; they are
// twiddled by w_D^(n1 k2), transposed through LDS, and the D1-point DFTs over
// n1 run per (column, k2) item, which hands element k2 + D2*k1 of column jj to
// store(jj, k2, k1, value).  With ROUNDS = 1 the transpose goes through a
// float2 tile of W*(D+1); with ROUNDS = 2 in two halves of k2 through a tile
// half that size, which keeps the unwritten half of v live (and spills at
// eight waves per SIMD).
//
// With ROUNDS = 0 the whole transpose goes through a float tile of W*(D+1)
// floats, real parts first and imaginary parts second: half the bytes of the
// one-round tile, with no more registers live than in one round (the real
// parts read back take the place of the real parts written).  The radius
// kernel and the forward transform use it.
template <int D1, int D2, int W, class Store>
__device__ __forceinline__ void transpose_finish_reim(float2 (&v)[D2], const float2* __restrict__ tw, float* lds,
                                                      Store store) {
    constexpr int NT = W * D1, LD = D2 * D1 + 1;
    constexpr int ITEMS = (W * D2 + NT - 1) / NT;
    const int t = threadIdx.x;
    const int j = t % W, n1 = t / W;
#pragma unroll
    for (int k2 = 0; k2 < D2; ++k2)
        if (n1 * k2 != 0) v[k2] = cmul(v[k2], tw[n1 * k2]);
#pragma unroll
    for (int k2 = 0; k2 < D2; ++k2) lds[j * LD + k2 * D1 + n1] = v[k2].x;
    __syncthreads();
    float2 u[ITEMS][D1];
#pragma unroll
    for (int q = 0; q < ITEMS; ++q) {
        const int it = t + q * NT;
        const int jj = it % W, k2 = it / W;
        if (it < W * D2) {
#pragma unroll
            for (int n = 0; n < D1; ++n) u[q][n].x = lds[jj * LD + k2 * D1 + n];
        }
    }
    __syncthreads();
#pragma unroll
    for (int k2 = 0; k2 < D2; ++k2) lds[j * LD + k2 * D1 + n1] = v[k2].y;
    __syncthreads();
#pragma unroll
    for (int q = 0; q < ITEMS; ++q) {
        const int it = t + q * NT;
        const int jj = it % W, k2 = it / W;
        if (it < W * D2) {
#pragma unroll
            for (int n = 0; n < D1; ++n) u[q][n].y = lds[jj * LD + k2 * D1 + n];
            dft<D1>(u[q]);
#pragma unroll
            for (int k1 = 0; k1 < D1; ++k1) store(jj, k2, k1, u[q][k1]);
        }
    }
}

template <int D1, int D2, int W, int ROUNDS, class Store>
__device__ __forceinline__ void transpose_finish(float2 (&v)[D2], const float2* __restrict__ tw, float2* lds,
                                                 Store store) {
    if constexpr (ROUNDS == 0) {
        transpose_finish_reim<D1, D2, W>(v, tw, reinterpret_cast<float*>(lds), store);
        return;
    }
    constexpr int NT = W * D1, K2R = D2 / (ROUNDS > 0 ? ROUNDS : 1), LD = K2R * D1 + 1;
    static_assert(ROUNDS == 0 || D2 % ROUNDS == 0, "rounds must split k2 evenly");
    const int t = threadIdx.x;
    const int j = t % W, n1 = t / W;
#pragma unroll
    for (int h = 0; h < ROUNDS; ++h) {
        if (h) __syncthreads();  // the previous round's reads are done
#pragma unroll
        for (int kk = 0; kk < K2R; ++kk) {
            const int k2 = h * K2R + kk;
            lds[j * LD + kk * D1 + n1] = (n1 * k2 == 0) ? v[k2] : cmul(v[k2], tw[n1 * k2]);
        }
        __syncthreads();
        for (int it = t; it < W * K2R; it += NT) {
            const int jj = it % W, kk = it / W;
            float2 u[D1];
#pragma unroll
            for (int n = 0; n < D1; ++n) u[n] = lds[jj * LD + kk * D1 + n];
            dft<D1>(u);
#pragma unroll
            for (int k1 = 0; k1 < D1; ++k1) store(jj, h * K2R + kk, k1, u[k1]);
        }
    }
}

// Columns pass of one block: window + inverse DFT along axis 0 for the W
// columns (b, c0..c0+W-1), b given by its slot among the surviving rows
// (0..bpos positive, then the negative modes).  lds: W*(D/ROUNDS+1) float2
// (ROUNDS >= 1), or W*(D+1) floats (ROUNDS = 0).
// S_BMAJOR: S is stored b-major, S[b][a][c] with rows padded to
// padded_row(D/2+1) modes (the fused loop's own forward transform writes it
// so: a column group along a is then D aligned 128-byte pieces of one
// contiguous plane), else in the r2c layout S[a][b][c] (stride D/2+1).
template <int D1, int D2, int W, int ROUNDS = 1, bool S_BMAJOR = false>
__device__ __forceinline__ void columns_body(const float2* __restrict__ S, float2* __restrict__ T, int ho,
                                             const float2* __restrict__ tw, const FusedParams& fp, float scale,
                                             int slot, int cgroup, float2* lds) {
    constexpr int D = D1 * D2, H = D / 2 + 1;
    const int t = threadIdx.x;
    const int j = t % W, n1 = t / W;
    const int bpos = fp.bm < D / 2 ? fp.bm : D / 2;
    const int b = slot <= bpos ? slot : D - (slot - bpos);
    const int bs = mode_of(b, D);
    const int c0 = cgroup * W;
    if (bs * bs + c0 * c0 > fp.m2max) return;  // whole group outside the disc (uniform)
    const int c = c0 + j;
    const bool col_ok = c < H && bs * bs + c * c <= fp.m2max;
    float2 v[D2];
#pragma unroll
    for (int n2 = 0; n2 < D2; ++n2) {
        const int a = n1 + D1 * n2;
        const int as = mode_of(a, D);
        const int m2 = as * as + bs * bs + c * c;
        float2 x = make_float2(0.0f, 0.0f);
        const size_t si = S_BMAJOR ? ((size_t)b * D + a) * padded_row(H) + c : ((size_t)a * D + b) * H + c;
        if (col_ok && m2 <= fp.m2max) x = fused_window(S[si], a, b, c, m2, D, fp, scale);
        v[n2] = x;
    }
    dft<D2>(v);
    transpose_finish<D1, D2, W, ROUNDS>(v, tw, lds, [&](int jj, int k2, int k1, float2 u) {
        const int cc = c0 + jj;
        if (cc < H && bs * bs + cc * cc <= fp.m2max) T[((size_t)(k2 + D2 * k1) * D + b) * ho + cc] = u;
    });
}

// Rows pass of one block: inverse DFT along axis 1, in place, for the W
// columns (p, c0..c0+W-1).  lds as for the columns pass.
template <int D1, int D2, int W, int ROUNDS = 1>
__device__ __forceinline__ void rows_body(float2* __restrict__ T, int ho, const float2* __restrict__ tw,
                                          const FusedParams& fp, int p, int cgroup, float2* lds) {
    constexpr int D = D1 * D2, H = D / 2 + 1;
    const int t = threadIdx.x;
    const int j = t % W, n1 = t / W;
    const int c0 = cgroup * W;
    const int c = c0 + j;
    float2 v[D2];
#pragma unroll
    for (int n2 = 0; n2 < D2; ++n2) {
        const int b = n1 + D1 * n2;
        const int bs = mode_of(b, D);
        v[n2] = (c < H && bs * bs + c * c <= fp.m2max) ? T[((size_t)p * D + b) * ho + c] : make_float2(0.0f, 0.0f);
    }
    dft<D2>(v);
    transpose_finish<D1, D2, W, ROUNDS>(v, tw, lds, [&](int jj, int k2, int k1, float2 u) {
        const int cc = c0 + jj;
        if (cc <= fp.cmax) T[((size_t)p * D + (k2 + D2 * k1)) * ho + cc] = u;
    });
}

// Forward DFT along a column axis for the pruned r2c: the inverse four-step
// of the conjugated inputs, conjugated on the way out (forward(x) =
// conj(inverse(conj x))).  Rows: plane p, in place along b, every b loaded
// for c <= cmax, stored only where b^2 + c^2 <= m2max (the columns pass reads
// nothing else).  Columns: T (rows layout, stride ho) along a into the
// half spectrum S, stored b-major (S[b][a][c], stride padded_row(D/2+1)), only columns
// inside the disc and modes inside the sphere |n|^2 <= m2max (no radius
// reads beyond it).
template <int D1, int D2, int W, int ROUNDS = 1>
__device__ __forceinline__ void fwd_rows_body(float2* __restrict__ T, int ho, const float2* __restrict__ tw, int cmax,
                                              int m2max, int p, int cgroup, float2* lds) {
    constexpr int D = D1 * D2;
    const int t = threadIdx.x;
    const int j = t % W, n1 = t / W;
    const int c0 = cgroup * W;
    const int c = c0 + j;
    float2 v[D2];
#pragma unroll
    for (int n2 = 0; n2 < D2; ++n2) {
        const int b = n1 + D1 * n2;
        v[n2] = c <= cmax ? cconj(T[((size_t)p * D + b) * ho + c]) : make_float2(0.0f, 0.0f);
    }
    dft<D2>(v);
    transpose_finish<D1, D2, W, ROUNDS>(v, tw, lds, [&](int jj, int k2, int k1, float2 u) {
        const int cc = c0 + jj, b = k2 + D2 * k1, bs = mode_of(b, D);
        if (cc <= cmax && bs * bs + cc * cc <= m2max) T[((size_t)p * D + b) * ho + cc] = cconj(u);
    });
}

template <int D1, int D2, int W, int ROUNDS = 1>
__device__ __forceinline__ void fwd_columns_body(const float2* __restrict__ T, int ho, float2* __restrict__ S,
                                                 const float2* __restrict__ tw, int bm, int cmax, int m2max, int slot,
                                                 int cgroup, float2* lds) {
    constexpr int D = D1 * D2, H = D / 2 + 1;
    const int t = threadIdx.x;
    const int j = t % W, n1 = t / W;
    const int bpos = bm < D / 2 ? bm : D / 2;
    const int b = slot <= bpos ? slot : D - (slot - bpos);
    const int bs = mode_of(b, D);
    const int c0 = cgroup * W;
    if (bs * bs + c0 * c0 > m2max) return;  // whole group outside the disc (uniform)
    const int c = c0 + j;
    const bool col_ok = c <= cmax && bs * bs + c * c <= m2max;
    float2 v[D2];
#pragma unroll
    for (int n2 = 0; n2 < D2; ++n2) {
        const int a = n1 + D1 * n2;
        v[n2] = col_ok ? cconj(T[((size_t)a * D + b) * ho + c]) : make_float2(0.0f, 0.0f);
    }
    dft<D2>(v);
    transpose_finish<D1, D2, W, ROUNDS>(v, tw, lds, [&](int jj, int k2, int k1, float2 u) {
        const int cc = c0 + jj, a = k2 + D2 * k1, as = mode_of(a, D);
        if (cc <= cmax && as * as + bs * bs + cc * cc <= m2max) S[((size_t)b * D + a) * padded_row(H) + cc] = cconj(u);
    });
}

// c2r of L consecutive lines along the last axis.  The length-D real output
// is the length-N = D/2 complex transform of
// Z[k] = (X[k] + conj X[N-k]) + i (X[k] - conj X[N-k]) w_D^k, whose outputs
// hold x[2m] (real part) and x[2m+1] (imaginary part).  One LDS buffer
// (L*HP float2) holds, in turn, the half lines, the transposed partial
// transforms and the real outputs; rows are padded to an odd number of
// float2 so that lanes walking different lines hit different banks.
template <int N1, int N2, int L>
struct LinesCore {
    static constexpr int N = N1 * N2, D = 2 * N, H = N + 1, NT = L * N1;
    static constexpr int HP = H | 1;
    static constexpr int HO = padded_row(H);  // line stride of the intermediates
    static constexpr int LOADS = (L * H + NT - 1) / NT;
    static constexpr int ITEMS = (L * N2 + NT - 1) / NT;  // step-3 transforms per thread
    static constexpr int WORDS = L * D / 64;              // mask words of the block's cells
    static_assert(NT % 64 == 0 && (L * D) % 64 == 0, "waves must cover whole mask words");
    static_assert(L * D / 64 <= NT, "one thread per mask word");

    // Global loads of the prefixes c < cl of L lines starting at U (line stride hs).
    __device__ static __forceinline__ void load(const float2* __restrict__ U, int hs, int cl, float2 (&tmp)[LOADS]) {
        const int t = threadIdx.x;
#pragma unroll
        for (int q = 0; q < LOADS; ++q) {
            const int f = t + q * NT;
            const int i = f / H, c = f - i * H;
            tmp[q] = (f < L * H && c < cl) ? U[(size_t)i * hs + c] : make_float2(0.0f, 0.0f);
        }
    }

    __device__ static __forceinline__ void store(const float2 (&tmp)[LOADS], float2* buf) {
        const int t = threadIdx.x;
#pragma unroll
        for (int q = 0; q < LOADS; ++q) {
            const int f = t + q * NT;
            const int i = f / H, c = f - i * H;
            if (f < L * H) buf[i * HP + c] = tmp[q];
        }
    }

    // In: buf holds the half lines (after a barrier).  Out: buf holds the real
    // lines, line i at ((float*)buf)[i * 2 * HP + r], followed by a barrier.
    // cl: the modes c >= cl of every line are zero (buf holds zeros there).
    // With cl <= N1 only the first and the last of a thread's N2 inputs Z[k]
    // can be non-zero (k = n1 has X[k]; k = n1 + N1 (N2-1) has conj X[N-k]),
    // and the N2-point DFT reduces to V[k2] = Z0 + Z_last w_N2^((N2-1) k2).
    __device__ static __forceinline__ void transform(float2* buf, const float2* stw, int cl = H) {
        const int t = threadIdx.x;
        {
            const int i = t % L, n1 = t / L;  // lanes walk lines: distinct banks
            float2 v[N2];
            if (cl <= N1) {
                const int k0 = n1, k1 = n1 + N1 * (N2 - 1);
                const float2 X0 = buf[i * HP + k0];
                const float2 z0 = cadd(X0, times_i(k0 == 0 ? X0 : cmul(X0, stw[k0])));
                const float2 Xm = cconj(buf[i * HP + (N - k1)]);
                const float2 z1 = csub(Xm, times_i(cmul(Xm, stw[k1])));
#pragma unroll
                for (int k2 = 0; k2 < N2; ++k2)
                    v[k2] = k2 == 0 ? cadd(z0, z1) : cadd(z0, cmul(z1, twiddle((long)(N2 - 1) * k2, N2)));
            } else {
#pragma unroll
                for (int n2 = 0; n2 < N2; ++n2) {
                    const int k = n1 + N1 * n2;
                    const float2 X = buf[i * HP + k];
                    const float2 Xm = cconj(buf[i * HP + (N - k)]);
                    const float2 s = cadd(X, Xm);
                    const float2 d = csub(X, Xm);
                    v[n2] = cadd(s, times_i(k == 0 ? d : cmul(d, stw[k])));
                }
                dft<N2>(v);
            }
            __syncthreads();
#pragma unroll
            for (int k2 = 0; k2 < N2; ++k2)
                buf[i * HP + k2 * N1 + n1] = (n1 * k2 == 0) ? v[k2] : cmul(v[k2], stw[2 * n1 * k2]);
        }
        __syncthreads();
        float2 u[ITEMS][N1];
#pragma unroll
        for (int q = 0; q < ITEMS; ++q) {
            const int it = t + q * NT;
            const int ii = it % L, k2 = it / L;  // lanes walk lines: distinct banks
            if (it < L * N2) {
#pragma unroll
                for (int n = 0; n < N1; ++n) u[q][n] = buf[ii * HP + k2 * N1 + n];
                dft<N1>(u[q]);
            }
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < ITEMS; ++q) {
            const int it = t + q * NT;
            const int ii = it % L, k2 = it / L;
            if (it < L * N2) {
#pragma unroll
                for (int k1 = 0; k1 < N1; ++k1) buf[ii * HP + k2 + N2 * k1] = u[q][k1];
            }
        }
        __syncthreads();
    }
};

// Forward transform of real lines, for the pruned r2c of the fused loop.
// buf holds L lines of D reals as N = D/2 packed pairs z[m] = (x[2m],
// x[2m+1]) at buf[i*HP + m] (after a barrier); stw holds w_D^q, q < D.  The inverse four-step DFT of
// conj(z) leaves W = conj(Z), Z the forward length-N DFT of z, at
// buf[i*HP + k] (followed by a barrier); line_spectrum() then forms the
// half-spectrum mode X[k] = (Z[k] + conj Z[N-k])/2 - i (Z[k] - conj Z[N-k])
// w_D^-k / 2 for 0 <= k <= N.
template <int N1, int N2, int L>
__device__ __forceinline__ void forward_pairs(float2* buf, const float2* stw) {
    constexpr int N = N1 * N2, HP = (N + 1) | 1, NT = L * N1;
    constexpr int ITEMS = (L * N2 + NT - 1) / NT;
    const int t = threadIdx.x;
    {
        const int i = t % L, n1 = t / L;
        float2 v[N2];
#pragma unroll
        for (int n2 = 0; n2 < N2; ++n2) v[n2] = cconj(buf[i * HP + n1 + N1 * n2]);
        dft<N2>(v);
        __syncthreads();
#pragma unroll
        for (int k2 = 0; k2 < N2; ++k2)
            buf[i * HP + k2 * N1 + n1] = (n1 * k2 == 0) ? v[k2] : cmul(v[k2], stw[2 * n1 * k2]);
    }
    __syncthreads();
    float2 u[ITEMS][N1];
#pragma unroll
    for (int q = 0; q < ITEMS; ++q) {
        const int it = t + q * NT;
        const int ii = it % L, k2 = it / L;
        if (it < L * N2) {
#pragma unroll
            for (int n = 0; n < N1; ++n) u[q][n] = buf[ii * HP + k2 * N1 + n];
            dft<N1>(u[q]);
        }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < ITEMS; ++q) {
        const int it = t + q * NT;
        const int ii = it % L, k2 = it / L;
        if (it < L * N2) {
#pragma unroll
            for (int k1 = 0; k1 < N1; ++k1) buf[ii * HP + k2 + N2 * k1] = u[q][k1];
        }
    }
    __syncthreads();
}

// X[k] of a line from W = conj(Z) (see forward_pairs); wk = w_D^k.
__device__ __forceinline__ float2 line_spectrum(float2 Wk, float2 Wmk, float2 wk) {
    const float2 zk = cconj(Wk);                // Z[k]
    const float2 s = cadd(zk, Wmk);             // Z[k] + conj Z[N-k]
    const float2 d = csub(zk, Wmk);             // Z[k] - conj Z[N-k]
    const float2 e = cmul(d, cconj(wk));        // ... w_D^-k
    return make_float2(0.5f * (s.x + e.y), 0.5f * (s.y - e.x));  // (s - i e) / 2
}

__device__ inline float fcoll_of(float smoothed, float deltac, float denom) {
    return erfcf(__fdiv_rn(__fsub_rn(deltac, smoothed), denom));
}

// erfc for the box-mean sum only (decisions use erfcf through fcoll_of).
// For a = |x|: erfc(a) = exp(-a^2) P(t) / (1 + 2a), t = (a - K)/(a + K),
// K = 2.5, the change of variable of Shepherd & Laframboise (Math. Comp. 36
// (1981) 249-253) for (1 + 2a) exp(a^2) erfc(a), with P a degree-9
// polynomial fitted here to it on a in [0, 10.5] by iteratively reweighted
// least squares on the relative error (2e-8 in exact arithmetic; 3.5e-7 in
// float32 arithmetic over x in [-9, 9]).  One reciprocal serves both
// quotients; exp(-a^2) keeps the rounding residues of a^2 and of the scaling
// by log2(e).  erfc(x) = 2 - erfc(-x) for x < 0; beyond |x| = 9 (results
// below 5e-37) erfcf is called.  The FP64 box mean moves by at most 3.5e-7
// relative, far inside the 1e-5 decision band.
__device__ __forceinline__ float erfc_sum(float x) {
    const float a = fabsf(x);
    if (a > 9.0f) return erfcf(x);
    constexpr float K = 2.5f;
    constexpr float kL2E = 1.44269502f, kL2E_lo = 1.92596299e-08f, kLN2 = 0.693147182f;
    const float one2a = fmaf(2.0f, a, 1.0f);
    const float ap = a + K;
    const float r = __builtin_amdgcn_rcpf(ap * one2a);
    const float t = ((a - K) * one2a) * r;
    const float inv12 = ap * r;
    float p = 4.49154089e-04f;
    p = fmaf(p, t, 3.18190403e-04f);
    p = fmaf(p, t, -4.28856490e-03f);
    p = fmaf(p, t, 1.03365746e-03f);
    p = fmaf(p, t, 3.27421837e-02f);
    p = fmaf(p, t, -1.02989852e-01f);
    p = fmaf(p, t, 1.57631412e-01f);
    p = fmaf(p, t, -9.90227982e-02f);
    p = fmaf(p, t, -1.22356802e-01f);
    p = fmaf(p, t, 1.26483822e+00f);
    const float a2h = a * a;
    const float a2l = fmaf(a, a, -a2h);
    const float ph = a2h * kL2E;
    float pl = fmaf(a2h, kL2E, -ph);
    pl = fmaf(a2h, kL2E_lo, pl);
    pl = fmaf(a2l, kL2E, pl);
    const float e = __builtin_amdgcn_exp2f(-ph);
    const float v = ((p * inv12) * e) * fmaf(-pl, kLN2, 1.0f);
    return x < 0.0f ? 2.0f - v : v;
}

// f_coll for the box-mean sum only: erfc_sum of the quotient (deltac -
// smoothed)/denom, formed with the reciprocal of the denominator and one
// residual correction, so that it matches the correctly rounded quotient of
// fcoll_of to within one ulp (erfc's relative condition number is about
// 2 x^2, so a cruder quotient would move the tail cells' f_coll by up to
// 1e-5 at x = 9).  Decisions use fcoll_of.
__device__ inline float fcoll_for_sum(float smoothed, float deltac, float denom, float inv_denom) {
    const float num = __fsub_rn(deltac, smoothed);
    const float q = num * inv_denom;
    return erfc_sum(fmaf(fmaf(-q, denom, num), inv_denom, q));
}

// Threshold of the centre-pixel rule in density space (DensityThreshold):
// d* = deltac - denom * erfcinv(1/(zeta*scale)), band 1e-4 (1 + |deltac| +
// denom + |d*|).
__device__ inline DensityThreshold density_threshold(float deltac, float denom, double zeta, float scale) {
    const double y = 1.0 / (zeta * (double)scale);
    DensityThreshold t;
    if (!(y > 1e-30 && y < 1.999)) {
        // degenerate targets (a zero box mean, erfc's end points): every cell
        // takes the exact float32 test
        t.dstar = 0.0f;
        t.band = INFINITY;
        return t;
    }
    const double ds = (double)deltac - (double)denom * erfcinv(y);
    t.dstar = (float)ds;
    t.band = (float)(1e-4 * (1.0 + fabs((double)deltac) + (double)denom + fabs(ds)));
    return t;
}

// The centre-pixel decision of one cell (find_bubbles.cu:53-70): zeta*f >= 1,
// tested as (double)f >= 1/zeta with f = erfcf(...) (times the box-mean
// scale for nonlinear input).
__device__ __forceinline__ bool cell_hit(float delta, const DensityThreshold& th, float deltac, float denom,
                                         bool use_scale, float scale, double inv_zeta) {
    if (fabsf(delta - th.dstar) <= th.band) {
        float fc = fcoll_of(delta, deltac, denom);
        if (use_scale) fc = __fmul_rn(fc, scale);
        return (double)fc >= inv_zeta;
    }
    return delta > th.dstar;
}

// Final-radius value of a cell not yet ionized (find_bubbles.cu:142-158):
// 1 when zeta*f >= 1, else the partial zeta*f.
__device__ __forceinline__ float final_value(float delta, float deltac, float denom, bool use_scale, float scale,
                                             double zeta, double inv_zeta) {
    float fc = fcoll_of(delta, deltac, denom);
    if (use_scale) fc = __fmul_rn(fc, scale);
    return ((double)fc >= inv_zeta) ? 1.0f : (float)((double)fc * zeta);
}

}  // namespace cm21
