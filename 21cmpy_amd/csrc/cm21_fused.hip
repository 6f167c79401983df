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
#include <cmath>

#include "cm21_dft.hpp"
#include "cm21_internal.hpp"
#include "cm21_window.hpp"
#include "cm21_fused_api.hpp"

#pragma clang fp contract(fast)

namespace cm21 {

enum LineMode { kLineSum = 0, kLineThreshold = 1, kLineFinal = 2 };

__device__ inline float2 fused_window(float2 x, int a, int b, int c, int m2, int D, const FusedParams& fp) {
    if (fp.exact_cut && m2 >= fp.m2lo) {
        const float kR = __fmul_rn(kmag_of(a, b, c, D, fp.dk, 0), fp.R);
        if ((float)((double)kR * 0.413566994) > 1.0f) return make_float2(0.0f, 0.0f);
    }
    if (fp.gcoef != 0.0f) {
        const float w2 = exp2f((float)m2 * fp.gcoef);
        x = make_float2(__fmul_rn(x.x, w2), __fmul_rn(x.y, w2));
    }
    return make_float2(__fmul_rn(x.x, fp.scale), __fmul_rn(x.y, fp.scale));
}

// Pass 1: window + inverse DFT along axis 0 for W columns (b, c0..c0+W-1).
template <int D1, int D2, int W>
__global__ __launch_bounds__(W* D1) void k_fused_columns(const float2* __restrict__ S, float2* __restrict__ T,
                                                        const float2* __restrict__ tw, FusedParams fp) {
    constexpr int D = D1 * D2, H = D / 2 + 1, NT = W * D1, LD = D + 1;
    __shared__ float2 lds[W * LD];
    const int t = threadIdx.x;
    const int j = t % W, n1 = t / W;
    const int bpos = fp.bm < D / 2 ? fp.bm : D / 2;
    const int slot = blockIdx.y;
    const int b = slot <= bpos ? slot : D - (slot - bpos);
    const int bs = mode_of(b, D);
    const int c0 = blockIdx.x * W;
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
        if (col_ok && m2 <= fp.m2max) x = fused_window(S[((size_t)a * D + b) * H + c], a, b, c, m2, D, fp);
        v[n2] = x;
    }
    dft<D2>(v);
#pragma unroll
    for (int k2 = 0; k2 < D2; ++k2) lds[j * LD + k2 * D1 + n1] = (n1 * k2 == 0) ? v[k2] : cmul(v[k2], tw[n1 * k2]);
    __syncthreads();
    for (int it = t; it < W * D2; it += NT) {
        const int jj = it % W, k2 = it / W;
        const int cc = c0 + jj;
        float2 u[D1];
#pragma unroll
        for (int n = 0; n < D1; ++n) u[n] = lds[jj * LD + k2 * D1 + n];
        dft<D1>(u);
        if (cc < H && bs * bs + cc * cc <= fp.m2max) {
#pragma unroll
            for (int k1 = 0; k1 < D1; ++k1) T[((size_t)(k2 + D2 * k1) * D + b) * H + cc] = u[k1];
        }
    }
}

// Pass 2: inverse DFT along axis 1, in place, for W columns (p, c0..c0+W-1).
template <int D1, int D2, int W>
__global__ __launch_bounds__(W* D1) void k_fused_rows(float2* __restrict__ T, const float2* __restrict__ tw,
                                                     FusedParams fp) {
    constexpr int D = D1 * D2, H = D / 2 + 1, NT = W * D1, LD = D + 1;
    __shared__ float2 lds[W * LD];
    const int t = threadIdx.x;
    const int j = t % W, n1 = t / W;
    const int p = blockIdx.y;
    const int c0 = blockIdx.x * W;
    const int c = c0 + j;
    float2 v[D2];
#pragma unroll
    for (int n2 = 0; n2 < D2; ++n2) {
        const int b = n1 + D1 * n2;
        const int bs = mode_of(b, D);
        v[n2] = (c < H && bs * bs + c * c <= fp.m2max) ? T[((size_t)p * D + b) * H + c] : make_float2(0.0f, 0.0f);
    }
    dft<D2>(v);
#pragma unroll
    for (int k2 = 0; k2 < D2; ++k2) lds[j * LD + k2 * D1 + n1] = (n1 * k2 == 0) ? v[k2] : cmul(v[k2], tw[n1 * k2]);
    __syncthreads();
    for (int it = t; it < W * D2; it += NT) {
        const int jj = it % W, k2 = it / W;
        const int cc = c0 + jj;
        float2 u[D1];
#pragma unroll
        for (int n = 0; n < D1; ++n) u[n] = lds[jj * LD + k2 * D1 + n];
        dft<D1>(u);
        if (cc <= fp.cmax) {
#pragma unroll
            for (int k1 = 0; k1 < D1; ++k1) T[((size_t)p * D + (k2 + D2 * k1)) * H + cc] = u[k1];
        }
    }
}

__device__ inline float fcoll_of(float smoothed, float deltac, float denom) {
    return erfcf(__fdiv_rn(__fsub_rn(deltac, smoothed), denom));
}

// Threshold of the centre-pixel rule in density space.  f = erfcf((deltac -
// d)/denom) * scale >= 1/zeta holds exactly when d >= d*, with
// d* = deltac - denom * erfcinv(1/(zeta*scale)), up to the float32 rounding of
// the left side; cells within `band` of d* take the exact float32 test, the
// others are decided by comparing d with d*.
struct DensityThreshold {
    float dstar;  // +inf: no cell can reach the threshold
    float band;
};

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

// Pass 3: c2r along axis 2 for L consecutive lines, then the f_coll epilogue.
// The length-D real output is the length-N = D/2 complex transform of
// Z[k] = (X[k] + conj X[N-k]) + i (X[k] - conj X[N-k]) w_D^k, whose outputs
// hold x[2m] (real part) and x[2m+1] (imaginary part).  One LDS buffer holds,
// in turn, the half lines, the transposed partial transforms and the real
// outputs; rows are padded to an odd number of float2 so that lanes walking
// different lines hit different banks.  In sum mode each block writes its FP64
// partial; a one-block kernel adds them in a fixed order (deterministic), and
// the kernel boundary makes them visible across the chip's L2 slices.
template <int N1, int N2, int L, int MODE>
__global__ __launch_bounds__(L* N1) void k_fused_lines(const float2* __restrict__ U, const float2* __restrict__ tw,
                                                      FusedParams fp, LineEpilogue ep) {
    constexpr int N = N1 * N2, D = 2 * N, H = N + 1, NT = L * N1;
    constexpr int HP = H | 1;  // padded line stride (float2); also holds N+1 real pairs
    static_assert(NT % 64 == 0 && (L * D) % 64 == 0, "waves must cover whole mask words");
    static_assert(L * D / 64 <= NT, "one thread per mask word");
    constexpr int LOADS = (L * H + NT - 1) / NT;
    constexpr int ITEMS = (L * N2 + NT - 1) / NT;  // step-3 transforms per thread
    constexpr int WORDS = L * D / 64;              // mask words of the block's cells
    __shared__ float2 buf[L * HP];
    __shared__ float2 stw[D];
    __shared__ unsigned long long smask[WORDS];
    const int t = threadIdx.x;
    const size_t line0 = (size_t)blockIdx.x * L;
    const size_t word0 = line0 * D / 64;
    const int cl = fp.cmax + 1;  // loaded prefix of each line
    // all global loads in flight before the first LDS store
    float2 tmp[LOADS];
#pragma unroll
    for (int q = 0; q < LOADS; ++q) {
        const int f = t + q * NT;
        const int i = f / H, c = f - i * H;
        tmp[q] = (f < L * H && c < cl) ? U[(line0 + i) * H + c] : make_float2(0.0f, 0.0f);
    }
    unsigned long long mword = 0ull;
    if (MODE != kLineSum && t < WORDS) mword = ep.mask[word0 + t];
    for (int q = t; q < D; q += NT) stw[q] = tw[q];
#pragma unroll
    for (int q = 0; q < LOADS; ++q) {
        const int f = t + q * NT;
        const int i = f / H, c = f - i * H;
        if (f < L * H) buf[i * HP + c] = tmp[q];
    }
    if (MODE != kLineSum && t < WORDS) smask[t] = mword;
    __syncthreads();
    {
        const int i = t % L, n1 = t / L;  // lanes walk lines: distinct banks
        float2 v[N2];
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
    const size_t cell0 = line0 * D;
    float scale = 1.0f;
    if (MODE != kLineSum && ep.sum) scale = (float)((double)ep.fc_mean / (*ep.sum / ep.ncells));
    __shared__ DensityThreshold sth;
    if (MODE != kLineSum && t == 0) sth = density_threshold(ep.deltac, ep.denom, ep.zeta, scale);
    __syncthreads();
    const DensityThreshold th = sth;
    double acc = 0.0;
    const int lane = t & 63;
    const float* xrf = reinterpret_cast<const float*>(buf);
#pragma unroll 4
    for (int f = t; f < L * D; f += NT) {
        const int i = f / D, r = f - i * D;
        const float delta = xrf[i * 2 * HP + r];
        if (MODE == kLineSum) {
            acc += (double)fcoll_of(delta, ep.deltac, ep.denom);
        } else {
            bool hit;
            if (fabsf(delta - th.dstar) <= th.band) {
                float fc = fcoll_of(delta, ep.deltac, ep.denom);
                if (ep.sum) fc = __fmul_rn(fc, scale);
                hit = (double)fc >= ep.inv_zeta;
            } else {
                hit = delta > th.dstar;
            }
            const unsigned long long word = smask[f >> 6];
            if (MODE == kLineThreshold) {
                const unsigned long long bits = __ballot(hit);
                if (lane == 0) smask[f >> 6] = word | bits;
            } else {
                const bool was = (word >> (f & 63)) & 1ull;
                float val = 1.0f;
                if (!was && !hit) {
                    float fc = fcoll_of(delta, ep.deltac, ep.denom);
                    if (ep.sum) fc = __fmul_rn(fc, scale);
                    val = ((double)fc >= ep.inv_zeta) ? 1.0f : (float)((double)fc * ep.zeta);
                }
                ep.out[cell0 + f] = val;
            }
        }
    }
    if (MODE == kLineSum) {
        acc = block_sum(acc);
        if (t == 0) ep.partial[blockIdx.x] = acc;
    }
    if (MODE == kLineThreshold) {
        __syncthreads();
        if (t < WORDS && smask[t] != mword) ep.mask[word0 + t] = smask[t];
    }
}

__global__ void k_fill_u64(unsigned long long* __restrict__ p, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0ull;
}


__global__ void k_sum_partials_lines(const double* __restrict__ partial, int n, double* __restrict__ out) {
    double acc = 0.0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) acc += partial[i];
    acc = block_sum(acc);
    if (threadIdx.x == 0) *out = acc;
}

// ------------------------------------------------------------------ host side

template <int D1, int D2, int W>
int launch_columns(const float2* S, float2* T, const float2* tw, const FusedParams& fp, int ncg, int nb,
                   hipStream_t st) {
    k_fused_columns<D1, D2, W><<<dim3(ncg, nb), W * D1, 0, st>>>(S, T, tw, fp);
    CM21_LAUNCH_CHECK("k_fused_columns");
    return CM21_OK;
}

template <int D1, int D2, int W>
int launch_rows(float2* T, const float2* tw, const FusedParams& fp, int ncg, hipStream_t st) {
    k_fused_rows<D1, D2, W><<<dim3(ncg, D1 * D2), W * D1, 0, st>>>(T, tw, fp);
    CM21_LAUNCH_CHECK("k_fused_rows");
    return CM21_OK;
}

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

// D, columns/rows (D1, D2, W), lines (N1, N2, L)
const FusedShape kShapes[] = {
    {64, launch_columns<8, 8, 32>, launch_rows<8, 8, 32>, launch_lines<4, 8, 64>, 32, 64},
    {96, launch_columns<8, 12, 32>, launch_rows<8, 12, 32>, launch_lines<4, 12, 64>, 32, 64},
    {128, launch_columns<8, 16, 32>, launch_rows<8, 16, 32>, launch_lines<8, 8, 32>, 32, 32},
    {160, launch_columns<16, 10, 16>, launch_rows<16, 10, 16>, launch_lines<8, 10, 32>, 16, 32},
    {256, launch_columns<16, 16, 16>, launch_rows<16, 16, 16>, launch_lines<16, 8, 16>, 16, 16},
    {512, launch_columns<16, 32, 8>, launch_rows<16, 32, 8>, launch_lines<16, 16, 8>, 8, 8},
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

int fused_radius(const FusedShape& sh, const float2* S, float2* T, const float2* tw, const FusedParams& fp,
                 hipStream_t st) {
    const int D = sh.D, H = D / 2 + 1;
    const int bpos = fp.bm < D / 2 ? fp.bm : D / 2;
    const int bneg = fp.bm < D / 2 - 1 ? fp.bm : D / 2 - 1;
    const int nb = bpos + 1 + bneg;
    int ncg = fp.cmax / sh.W + 1;
    const int ncg_all = (H + sh.W - 1) / sh.W;
    if (ncg > ncg_all) ncg = ncg_all;
    ProfToken t = prof_begin(CM21_STAGE_FUSED_COLUMNS, st);
    CM21_TRY(sh.columns(S, T, tw, fp, ncg, nb, st));
    prof_end(t, st);
    t = prof_begin(CM21_STAGE_FUSED_ROWS, st);
    CM21_TRY(sh.rows(T, tw, fp, ncg, st));
    prof_end(t, st);
    return CM21_OK;
}

int fused_lines(const FusedShape& sh, int mode, const float2* U, const float2* tw, const FusedParams& fp,
                const LineEpilogue& ep, hipStream_t st) {
    ProfToken t = prof_begin(CM21_STAGE_FUSED_LINES, st);
    CM21_TRY(sh.lines(mode, U, tw, fp, ep, st));
    prof_end(t, st);
    return CM21_OK;
}

int fused_mask_clear(unsigned long long* mask, size_t words, hipStream_t st) {
    k_fill_u64<<<grid_for(words), kThreads, 0, st>>>(mask, words);
    CM21_LAUNCH_CHECK("k_fill_u64");
    return CM21_OK;
}

LineEpilogue line_epilogue(float deltac, float denom, double zeta, const double* sum, float fc_mean,
                           unsigned long long* mask, float* out, double* partial, double ncells) {
    LineEpilogue ep{};
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

int fused_sum(const double* partial, int n, double* out, hipStream_t st) {
    k_sum_partials_lines<<<1, kThreads, 0, st>>>(partial, n, out);
    CM21_LAUNCH_CHECK("k_sum_partials_lines");
    return CM21_OK;
}

}  // namespace cm21
