// Fused, pruned inverse transforms of the find_HII_bubbles radius loop.
//
// For each radius the filtered spectrum W(kR) S(k) is transformed back to real
// space in three passes over the half spectrum S[a][b][c] (a, b full axes, c the
// halved axis, c fastest):
//   columns: window multiply + inverse DFT along a, written to T[p][b][c];
//   rows:    inverse DFT along b, in place, T[p][q][c];
//   lines:   c2r along c (as a length-D/2 complex DFT) fused with the f_coll
//            epilogue (FP64 box sum, threshold into an ionization bitmask, or
//            the final paint into the float output).
// With the k-space top-hat (HII_FILTER 0 and 1 cut modes with kR > 2.418 in
// the reference's fall-through chain, Programs/find_bubbles.cu:33-37) every
// mode with |n|^2 > m2max is zero, so the columns pass reads and writes only
// columns inside the disc b^2 + c^2 <= m2max, the rows pass only c <= cmax and
// the lines pass only c <= cmax of each line.  Large radii keep a few modes;
// the work per radius follows the number of surviving modes.
//
// Each DFT of length D = D1*D2 is a four-step transform: D2-point DFTs in
// registers (one per thread), a twiddle by w_D^(n1 k2), an LDS transpose, then
// D1-point DFTs in registers.  The window is evaluated per mode from its
// integer |n|^2: the Gaussian factor as exp2(m2 * gcoef) and the cut by an
// integer bound, except in a thin shell around the cut where the reference's
// float32 test (find_bubbles.cu:26-35) decides, so cut decisions are
// bit-identical to the reference arithmetic.
#include <algorithm>

#include "cm21_fused_dev.hpp"

#pragma clang fp contract(fast)

namespace cm21 {

// Box-sum reduction of one radius by the NT threads of one block: strided
// per-thread FP64 sums, the wavefront tree, then the waves in order (the order
// is fixed, so the total is bitwise reproducible).  With `scaled` thread 0
// also derives the box-mean scale float32(fc_mean / (sum / ncells)) and its
// density threshold, once for every block of the threshold pass; with
// `active` the block lists the line groups whose maximum is not below
// d* - band (the order of the list does not matter: each group writes only
// its own mask words).
template <int NT>
__device__ void finish_box_sum(const double* partial, int n, double* out, const SumScaleArgs& sa) {
    __shared__ double part[NT / 64];
    __shared__ DensityThreshold sth;
    __shared__ int count;
    double acc = 0.0;
    for (int i = threadIdx.x; i < n; i += NT) acc += partial[i];
    acc = wave_sum(acc);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double r = 0.0;
        for (int w = 0; w < NT / 64; ++w) r += part[w];
        *out = r;
        if (sa.scaled) {
            const float scale = (float)((double)sa.fc_mean / (r / sa.ncells));
            ScaledThreshold s;
            s.scale = scale;
            s.th = density_threshold(sa.deltac, sa.denom, sa.zeta, scale);
            s.pad = 0.0f;
            *sa.scaled = s;
            sth = s.th;
        }
        count = 0;
    }
    if (!sa.scaled || !sa.active) return;
    __syncthreads();
    const float lim = sth.dstar - sth.band;
    for (int i = threadIdx.x; i < n; i += NT)
        if (!(sa.bmax[i] < lim)) sa.active[atomicAdd(&count, 1)] = i;
    __syncthreads();
    if (threadIdx.x == 0) *sa.nactive = count;
}

// Pass 3: c2r along axis 2 for L consecutive lines, then the f_coll epilogue:
// FP64 block partial of the box sum and the block's maximum of the smoothed
// field (MODE kLineSum), threshold into the ionization bitmask
// (kLineThreshold) or the final paint (kLineFinal).  In sum mode a one-block
// kernel adds the partials in a fixed order (deterministic), and the kernel
// boundary makes them visible across the chip's L2 slices.
//
// Before any transform work a threshold block returns when it cannot change
// the mask: every one of its cells is already ionized, or the box-sum pass
// found its maximum below d* - band (no cell can reach the threshold; cells
// inside the band take the exact test, so the skip never drops a hit).  A
// final-paint block whose cells are all ionized writes 1 without transforming.
// `blk` is the index of the L-line group the block works on.
template <int N1, int N2, int L, int MODE>
__device__ __forceinline__ void lines_block(const float2* __restrict__ U, const float2* __restrict__ tw,
                                            const FusedParams& fp, const LineEpilogue& ep, int blk, float2* buf,
                                            float2* stw, unsigned long long* smask) {
    using C = LinesCore<N1, N2, L>;
    constexpr int D = C::D, NT = C::NT, WORDS = C::WORDS;
    const int t = threadIdx.x;
    const size_t line0 = (size_t)blk * L;
    const size_t word0 = line0 * D / 64;
    const size_t cell0 = line0 * D;
    const bool use_scale = ep.scaled != nullptr;
    float scale = 1.0f;
    DensityThreshold th = ep.th;
    unsigned long long mword = 0ull;
    if (MODE != kLineSum) {
        if (use_scale) {
            const ScaledThreshold s = *ep.scaled;
            scale = s.scale;
            th = s.th;
        }
        mword = t < WORDS ? ep.mask[word0 + t] : ~0ull;
        const bool all_ionized = __syncthreads_and(mword == ~0ull);
        if (all_ionized) {
            if (MODE == kLineFinal)
                for (int q = t; q < L * D; q += NT) ep.out[cell0 + q] = 1.0f;
            return;
        }
        if (MODE == kLineThreshold && ep.bmax != nullptr && ep.bmax[blk] < th.dstar - th.band) return;
    }
    // all global loads in flight before the first LDS store
    float2 tmp[C::LOADS];
    C::load(U + line0 * C::HO, C::HO, fp.cmax + 1, tmp);
    for (int q = t; q < D; q += NT) stw[q] = tw[q];
    C::store(tmp, buf);
    if (MODE != kLineSum && t < WORDS) smask[t] = mword;
    __syncthreads();
    C::transform(buf, stw, fp.cmax + 1);
    double acc = 0.0;
    float vmax = -INFINITY;
    const int lane = t & 63;
    const float* xrf = reinterpret_cast<const float*>(buf);
    const float inv_denom = 1.0f / ep.denom;
    // box sum: four float32 partials per thread (a quarter of its cells each,
    // a few roundings of values in [0, 2]), added in FP64 after the loop
    static_assert((L * D) % (4 * NT) == 0, "whole sets of four cells per thread");
    float part[4] = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll 4
    for (int q = 0; q < L * D / NT; ++q) {
        const int f = t + q * NT;
        const int i = f / D, r = f - i * D;
        const float delta = xrf[i * 2 * C::HP + r];
        if (MODE == kLineSum) {
            part[q & 3] += fcoll_for_sum(delta, ep.deltac, ep.denom, inv_denom);
            vmax = fmaxf(vmax, delta);
        } else {
            const bool hit = cell_hit(delta, th, ep.deltac, ep.denom, use_scale, scale, ep.inv_zeta);
            const unsigned long long word = smask[f >> 6];
            if (MODE == kLineThreshold) {
                const unsigned long long bits = __ballot(hit);
                if (lane == 0) smask[f >> 6] = word | bits;
            } else {
                const bool was = (word >> (f & 63)) & 1ull;
                float val = 1.0f;
                if (!was && !hit) val = final_value(delta, ep.deltac, ep.denom, use_scale, scale, ep.zeta, ep.inv_zeta);
                ep.out[cell0 + f] = val;
            }
        }
    }
    if (MODE == kLineSum) {
        acc = ((double)part[0] + (double)part[1]) + ((double)part[2] + (double)part[3]);
        if (ep.bmax != nullptr) {
            __shared__ float wmax[NT / 64];
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) vmax = fmaxf(vmax, __shfl_xor(vmax, off, 64));
            if (lane == 0) wmax[t >> 6] = vmax;
            __syncthreads();
            if (t == 0) {
                float m = wmax[0];
                for (int w = 1; w < NT / 64; ++w) m = fmaxf(m, wmax[w]);
                ep.bmax[blk] = m;
            }
        }
        acc = block_sum(acc);
        if (t == 0) ep.partial[blk] = acc;
    }
    if (MODE == kLineThreshold) {
        __syncthreads();
        if (t < WORDS && smask[t] != mword) ep.mask[word0 + t] = smask[t];
    }
}

// With an active list (threshold pass after a box sum) block b works on the
// b-th listed line group and the blocks past the end of the list return at
// once, so line groups that cannot change cost one scalar load each and the
// groups that can are spread over the first blocks of the dispatch;
// otherwise block b works on line group b.
template <int N1, int N2, int L, int MODE>
__global__ __launch_bounds__(L* N1) void k_fused_lines(const float2* __restrict__ U, const float2* __restrict__ tw,
                                                      FusedParams fp, LineEpilogue ep) {
    using C = LinesCore<N1, N2, L>;
    __shared__ float2 buf[L * C::HP];
    __shared__ float2 stw[C::D];
    __shared__ unsigned long long smask[C::WORDS];
    int blk = blockIdx.x;
    if (MODE == kLineThreshold && ep.active != nullptr) {
        if (blk >= *ep.nactive) return;
        blk = ep.active[blk];
    }
    lines_block<N1, N2, L, MODE>(U, tw, fp, ep, blk, buf, stw, smask);
}

// One launch of the pipelined radius loop.  Its blocks take one of five
// roles: the columns pass of radius r+2 and the rows pass of radius r+1 (into
// intermediates no other role of the launch touches), the one-block box-sum
// reduction of radius r-1 (scale, threshold and active list), the threshold
// pass of radius r-2 and the box-sum pass of radius r.  Every role reads only
// what earlier launches wrote and writes buffers no other role of the launch
// reads or writes, so the roles need no ordering among themselves; in-order
// launches on one stream order the radii.
// Block order: the reduction, then for nonlinear input the box-sum blocks,
// the columns/rows blocks and the threshold blocks (block i of these takes
// the i-th listed line group; those past the end of the list return at
// once), for linear input the threshold blocks and then the columns/rows
// blocks.  The VALU-bound box sums fill every CU first and the
// memory-bound transposes take the slots they free; against the columns/rows
// blocks interleaved one for one with the box sums after the threshold
// blocks, 578 -> 603 boxes/s (literal) and 450 -> 457 (partly ionized) at
// 256^3; thresholds first stay better for linear input (128^3: 2530
// against 2399 boxes/s with thresholds last).
// The kernel asks for eight waves per SIMD, the box-sum role's full
// occupancy (about 52 VGPRs, 19.5 KB of LDS at D = 256); the columns/rows
// roles fit it without spilling through the re/im-staged transpose.
template <int D1, int D2, int W, int N1, int N2, int L>
__global__ __launch_bounds__(W* D1, 8) void k_fused_radius(RadiusRoles a, const float2* __restrict__ tw) {
    using C = LinesCore<N1, N2, L>;
    constexpr int D = D1 * D2, NT = W * D1;
    static_assert(C::D == D && C::NT == NT, "columns/rows and lines tiles must share D and the block size");
    // the columns/rows transposes stage real and imaginary parts in turn
    // (a float tile) when their one-round float2 tile would need more LDS
    // than the lines pass, and so lower the occupancy of the box sums
    constexpr int kLdsLines = L * C::HP + D + C::WORDS;
    constexpr int kRounds = W * (D + 1) > kLdsLines + kLdsLines / 4 ? 0 : 1;
    constexpr int kLdsCR = kRounds == 0 ? (W * (D + 1) + 1) / 2 : W * (D + 1);
    __shared__ float2 lds[kLdsCR > kLdsLines ? kLdsCR : kLdsLines];
    float2* buf = lds;
    float2* stw = lds + L * C::HP;
    unsigned long long* smask = reinterpret_cast<unsigned long long*>(lds + L * C::HP + D);
    int b = blockIdx.x;
    if (b < a.n_red) {
        finish_box_sum<NT>(a.partial, a.npart, a.total, a.sa);
        return;
    }
    b -= a.n_red;
    // nonlinear input: box sums, transposes, thresholds; linear input (no
    // box sums): thresholds, transposes
    const int thr0 = a.n_sum > 0 ? a.n_sum + a.n_mem : 0;
    if (b >= thr0 && b < thr0 + a.n_thr) {
        b -= thr0;
        if (a.ep_thr.active != nullptr) {
            if (b >= *a.ep_thr.nactive) return;
            b = a.ep_thr.active[b];
        }
        lines_block<N1, N2, L, kLineThreshold>(a.Uthr, tw, a.fp_thr, a.ep_thr, b, buf, stw, smask);
        return;
    }
    if (b >= thr0 + a.n_thr) b -= a.n_thr;
    const bool mem = b >= a.n_sum;
    int i = mem ? b - a.n_sum : b;
    if (!mem) {
        lines_block<N1, N2, L, kLineSum>(a.Usum, tw, a.fp_sum, a.ep_sum, i, buf, stw, smask);
    } else if (i < a.n_cols) {
        columns_body<D1, D2, W, kRounds, true>(a.S, a.Tcols, C::HO, tw, a.fp_cols, a.fp_cols.scale, i / a.ncg_cols,
                                i % a.ncg_cols, lds);
    } else {
        i -= a.n_cols;
        rows_body<D1, D2, W, kRounds>(a.Trows, C::HO, tw, a.fp_rows, i / a.ncg_rows, i % a.ncg_rows, lds);
    }
}

// Pruned r2c of the fused loop (three launches): only the modes that the
// smallest radius keeps (|n|^2 <= m2max, c <= cmax) are formed.
// Lines: the r2c along c of L lines per block, from the real box into T0
// (stride HO, c <= cmax).
template <int N1, int N2, int L>
__global__ __launch_bounds__(L* N1) void k_fwd_lines(const float* __restrict__ delta, float2* __restrict__ T0,
                                                    const float2* __restrict__ tw, int cmax) {
    using C = LinesCore<N1, N2, L>;
    constexpr int N = C::N, D = C::D, NT = C::NT, HP = C::HP;
    __shared__ float2 buf[L * HP];
    __shared__ float2 stw[D];
    const int t = threadIdx.x;
    const size_t line0 = (size_t)blockIdx.x * L;
    const float2* src = reinterpret_cast<const float2*>(delta + line0 * D);
    for (int f = t; f < L * N; f += NT) {
        const int i = f / N, m = f - i * N;
        buf[i * HP + m] = src[f];
    }
    for (int q = t; q < D; q += NT) stw[q] = tw[q];
    __syncthreads();
    forward_pairs<N1, N2, L>(buf, stw);
    const int nk = cmax + 1;
    for (int f = t; f < L * nk; f += NT) {
        const int i = f / nk, k = f - i * nk;
        const float2 Wk = buf[i * HP + (k == N ? 0 : k)];
        const float2 Wmk = buf[i * HP + (k == 0 ? 0 : N - k)];
        T0[(line0 + i) * C::HO + k] = line_spectrum(Wk, Wmk, stw[k]);
    }
}

// Rows and columns of the forward transform: the re/im-staged transpose (a
// float tile of W*(D+1)) keeps the LDS of a block at half the
// one-round tile; at six waves per SIMD they need no scratch (at eight they
// spill a few registers at D = 256 and run slower).
template <int D1, int D2, int W>
__global__ __launch_bounds__(W* D1, 6) void k_fwd_rows(float2* __restrict__ T0, const float2* __restrict__ tw,
                                                      int cmax, int m2max) {
    constexpr int D = D1 * D2;
    __shared__ float lds[W * (D + 1)];
    fwd_rows_body<D1, D2, W, 0>(T0, padded_row(D / 2 + 1), tw, cmax, m2max, blockIdx.y, blockIdx.x,
                                reinterpret_cast<float2*>(lds));
}

template <int D1, int D2, int W>
__global__ __launch_bounds__(W* D1, 6) void k_fwd_columns(const float2* __restrict__ T0, float2* __restrict__ S,
                                                         const float2* __restrict__ tw, int bm, int cmax, int m2max) {
    constexpr int D = D1 * D2;
    __shared__ float lds[W * (D + 1)];
    fwd_columns_body<D1, D2, W, 0>(T0, padded_row(D / 2 + 1), S, tw, bm, cmax, m2max, blockIdx.y, blockIdx.x,
                                   reinterpret_cast<float2*>(lds));
}

__global__ void k_fill_u64(unsigned long long* __restrict__ p, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0ull;
}


// ------------------------------------------------------------------ host side

template <int N1, int N2, int L>
int launch_lines(int mode, const float2* U, const float2* tw, const FusedParams& fp, const LineEpilogue& ep,
                 hipStream_t st) {
    constexpr int D = 2 * N1 * N2;
    const unsigned blocks = (unsigned)((size_t)D * D / L);
    if (mode == kLineSum)
        k_fused_lines<N1, N2, L, kLineSum><<<blocks, L * N1, 0, st>>>(U, tw, fp, ep);
    else if (mode == kLineThreshold)
        k_fused_lines<N1, N2, L, kLineThreshold><<<blocks, L * N1, 0, st>>>(U, tw, fp, ep);
    else
        k_fused_lines<N1, N2, L, kLineFinal><<<blocks, L * N1, 0, st>>>(U, tw, fp, ep);
    CM21_LAUNCH_CHECK("k_fused_lines");
    return CM21_OK;
}

template <int D1, int D2, int W, int N1, int N2, int L>
int launch_radius(const RadiusRoles& a, const float2* tw, hipStream_t st) {
    const int blocks = a.n_red + a.n_thr + a.n_mem + a.n_sum;
    if (blocks <= 0) return CM21_OK;
    k_fused_radius<D1, D2, W, N1, N2, L><<<blocks, W * D1, 0, st>>>(a, tw);
    CM21_LAUNCH_CHECK("k_fused_radius");
    return CM21_OK;
}

template <int D1, int D2, int W, int N1, int N2, int L>
int launch_forward(const float* delta, float2* T0, float2* S, const float2* tw, const FusedParams& fp,
                   hipStream_t st) {
    constexpr int D = D1 * D2, H = D / 2 + 1;
    const int cmax = fp.cmax < H - 1 ? fp.cmax : H - 1;
    k_fwd_lines<N1, N2, L><<<D * D / L, L * N1, 0, st>>>(delta, T0, tw, cmax);
    CM21_LAUNCH_CHECK("k_fwd_lines");
    const int ncg = cmax / W + 1;
    k_fwd_rows<D1, D2, W><<<dim3(ncg, D), W * D1, 0, st>>>(T0, tw, cmax, fp.m2max);
    CM21_LAUNCH_CHECK("k_fwd_rows");
    const int bpos = fp.bm < D / 2 ? fp.bm : D / 2;
    const int bneg = fp.bm < D / 2 - 1 ? fp.bm : D / 2 - 1;
    k_fwd_columns<D1, D2, W><<<dim3(ncg, bpos + 1 + bneg), W * D1, 0, st>>>(T0, S, tw, fp.bm, cmax, fp.m2max);
    CM21_LAUNCH_CHECK("k_fwd_columns");
    return CM21_OK;
}

// D, lines (N1, N2, L), radius (columns/rows D1, D2, W; lines N1, N2, L), forward
const FusedShape kShapes[] = {
    {64, launch_lines<4, 8, 64>, launch_radius<8, 8, 32, 4, 8, 64>, launch_forward<8, 8, 32, 4, 8, 64>, 32, 64},
    {96, launch_lines<4, 12, 64>, launch_radius<8, 12, 32, 4, 12, 64>, launch_forward<8, 12, 32, 4, 12, 64>, 32, 64},
    {128, launch_lines<8, 8, 32>, launch_radius<8, 16, 32, 8, 8, 32>, launch_forward<8, 16, 32, 8, 8, 32>, 32, 32},
    {160, launch_lines<8, 10, 32>, launch_radius<16, 10, 16, 8, 10, 32>, launch_forward<16, 10, 16, 8, 10, 32>, 16, 32},
    {256, launch_lines<16, 8, 16>, launch_radius<16, 16, 16, 16, 8, 16>, launch_forward<16, 16, 16, 16, 8, 16>, 16, 16},
    {512, launch_lines<16, 16, 8>, launch_radius<16, 32, 8, 16, 16, 8>, launch_forward<16, 32, 8, 16, 16, 8>, 8, 8},
};

const FusedShape* fused_shape(int D) {
    for (const auto& s : kShapes)
        if (s.D == D) return &s;
    return nullptr;
}

size_t fused_partials(int D) {
    const FusedShape* s = fused_shape(D);
    return s ? (size_t)D * D / s->L : 0;
}

// Per-radius bounds of the pruned transforms.
FusedParams fused_params(int D, float dk, float R, int filter_type, int ref_fallthrough, float scale) {
    FusedParams fp{};
    const int h = D / 2;
    const int all = 3 * h * h;
    fp.dk = dk;
    fp.R = R;
    fp.scale = scale;
    const double ln2 = 0.69314718055994530942;
    const bool cut = filter_type == 1;  // type 0 stays on the generic path
    if (cut) {
        const double kcut = 1.0 / (0.413566994 * (double)dk * (double)R);  // in units of dk
        const double m2c = kcut * kcut;
        const double hi = std::floor(m2c * (1.0 + 1e-5)), lo = std::ceil(m2c * (1.0 - 1e-5));
        fp.m2max = hi >= (double)all ? all : (int)hi;
        fp.m2lo = lo >= (double)all ? all + 1 : (int)lo;
        fp.exact_cut = 1;
    } else {
        fp.m2max = all;
        fp.m2lo = all + 1;
        fp.exact_cut = 0;
    }
    double kappa = 0.0;
    if (ref_fallthrough)
        kappa = filter_type == 1 ? 0.413566994 * 0.643 : 0.643;
    else
        kappa = filter_type == 2 ? 0.643 : 0.0;
    const double g = kappa * (double)dk * (double)R;
    fp.gcoef = (float)(-(g * g) / 2.0 / ln2);
    int bm = (int)std::floor(std::sqrt((double)fp.m2max));
    while ((long)(bm + 1) * (bm + 1) <= fp.m2max) ++bm;
    while ((long)bm * bm > fp.m2max) --bm;
    fp.bm = bm < h ? bm : h;
    fp.cmax = fp.bm;
    return fp;
}

void column_groups(const FusedShape& sh, const FusedParams& fp, int* ncg, int* nb) {
    const int D = sh.D, H = D / 2 + 1;
    const int bpos = fp.bm < D / 2 ? fp.bm : D / 2;
    const int bneg = fp.bm < D / 2 - 1 ? fp.bm : D / 2 - 1;
    *nb = bpos + 1 + bneg;
    const int ncg_all = (H + sh.W - 1) / sh.W;
    *ncg = std::min(fp.cmax / sh.W + 1, ncg_all);
}

int fused_lines(const FusedShape& sh, int mode, const float2* U, const float2* tw, const FusedParams& fp,
                const LineEpilogue& ep, hipStream_t st, int stage) {
    ProfToken t = prof_begin(stage, st);
    CM21_TRY(sh.lines(mode, U, tw, fp, ep, st));
    prof_end(t, st);
    return CM21_OK;
}

int fused_forward(const FusedShape& sh, const float* delta, float2* T0, float2* S, const float2* tw,
                  const FusedParams& fp, hipStream_t st) {
    ProfToken t = prof_begin(CM21_STAGE_R2C, st);
    CM21_TRY(sh.forward(delta, T0, S, tw, fp, st));
    prof_end(t, st);
    return CM21_OK;
}

int fused_radius_launch(const FusedShape& sh, const RadiusRoles& a, const float2* tw, hipStream_t st) {
    ProfToken t = prof_begin(CM21_STAGE_FUSED_RADIUS, st);
    CM21_TRY(sh.radius(a, tw, st));
    prof_end(t, st);
    return CM21_OK;
}

int fused_mask_clear(unsigned long long* mask, size_t words, hipStream_t st) {
    k_fill_u64<<<grid_for(words), kThreads, 0, st>>>(mask, words);
    CM21_LAUNCH_CHECK("k_fill_u64");
    return CM21_OK;
}

// erfc^-1 on the host: bisection on std::erfc (monotone decreasing) to the
// last bit, for y in (0, 2).
static double erfcinv_host(double y) {
    double lo = -6.0, hi = 27.3;  // erfc(-6) = 2 - 2e-17, erfc(27.3) < 1e-300
    for (int i = 0; i < 200 && hi - lo > 0.0; ++i) {
        const double mid = 0.5 * (lo + hi);
        if (mid == lo || mid == hi) break;
        if (std::erfc(mid) > y)
            lo = mid;
        else
            hi = mid;
    }
    return 0.5 * (lo + hi);
}

DensityThreshold density_threshold_host(float deltac, float denom, double zeta, float scale) {
    const double y = 1.0 / (zeta * (double)scale);
    DensityThreshold t;
    if (!(y > 1e-30 && y < 1.999)) {
        t.dstar = 0.0f;
        t.band = INFINITY;
        return t;
    }
    const double ds = (double)deltac - (double)denom * erfcinv_host(y);
    t.dstar = (float)ds;
    t.band = (float)(1e-4 * (1.0 + std::fabs((double)deltac) + (double)denom + std::fabs(ds)));
    return t;
}

LineEpilogue line_epilogue(float deltac, float denom, double zeta, const double* sum, float fc_mean,
                           unsigned long long* mask, float* out, double* partial, double ncells, float* bmax) {
    LineEpilogue ep{};
    ep.bmax = bmax;
    ep.deltac = deltac;
    ep.denom = denom;
    ep.zeta = zeta;
    ep.inv_zeta = 1.0 / zeta;
    ep.sum = sum;
    ep.fc_mean = fc_mean;
    ep.mask = mask;
    ep.out = out;
    ep.partial = partial;
    ep.ncells = ncells;
    return ep;
}

}  // namespace cm21
