// 3D U-Net surrogate of the excursion-set ionization field: the layer
// kernels of its forward and backward passes and the AdamW update.
//
// No counterpart in the reference (Programs/ computes every box with the
// excursion-set loop); the surrogate is trained on that loop's outputs and
// checked against it.
//
// Tensors are channels-last (B, D, D, D, C) with periodic boundaries (the
// boxes are periodic), activations in bf16 (uint16 bit patterns, round to
// nearest even), weights in fp32 masters with bf16 copies for the matrix
// cores, accumulation in fp32, reductions in a fixed order.
//
//   3^3 convolution, forward:  implicit GEMM on v_mfma_f32_16x16x32_bf16.  A
//     block computes TZ x TY x 16 output voxels for all output channels: the
//     input tile with a one-voxel halo sits in LDS (or a ring of z-planes for
//     the sliding form); for each 32-wide slice of K = 27 * cin (tap-major,
//     channel-minor) every lane reads 8 consecutive K of one output channel
//     (the A operand, 16 bytes of the bf16 weights) and 8 consecutive
//     channels of one voxel (B, 16 bytes from LDS); the 16 x 16 x 32 product
//     accumulates in registers as channels x voxels, so each lane stores 4
//     consecutive channels of one voxel.
//   3^3 convolution, input gradient: the same kernel on the output gradient
//     (masked by the ReLU of the forward output) with the weights transposed
//     and spatially flipped.
//   3^3 convolution, weight gradient: v_mfma_f32_16x16x32_bf16 with K = 32
//     voxels of an x-line, the masked output gradient and the input written
//     to LDS channel-major (the input in three dx-shifted copies) so that
//     every fragment is one aligned 16-byte read; fp32 VALU for boxes of
//     side 16.  Per-block partials reduced in a fixed order.
//   Input layer (density plus two constant conditioning channels, the
//   latter folded into a per-sample bias): 3^3 taps on the fp32 matrix cores
//   (v_mfma_f32_16x16x4f32), forward and weight gradient.
//   2x2x2 stride-2 transposed convolution: bf16 matrix cores forward and
//   input gradient, fp32 matrix cores (K = 4 voxels) weight gradient.
//   2x2x2 average pooling, the 1^3 head with sigmoid + binary cross entropy
//   (fused into the last convolution for inference), and AdamW: fp32 VALU.
#include <cmath>
#include <cstring>

#include "cm21.h"
#include "cm21_internal.hpp"

namespace cm21 {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
typedef unsigned short bf16_t;

__device__ __forceinline__ float bf2f(bf16_t h) { return __uint_as_float((unsigned)h << 16); }

typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef float floatx2 __attribute__((ext_vector_type(2)));
typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
// two floats to two bf16 packed low/high, round to nearest even (one
// v_cvt_pk_bf16_f32)
__device__ __forceinline__ unsigned f2bf2(float lo, float hi) {
    return __builtin_bit_cast(unsigned, __builtin_convertvector(floatx2{lo, hi}, bf16x2));
}
__device__ __forceinline__ bf16_t f2bf(float f) { return (bf16_t)(f2bf2(f, 0.0f) & 0xffffu); }

__device__ __forceinline__ int wrap(int i, int D) { return i < 0 ? i + D : (i >= D ? i - D : i); }

// 8 bf16 of x kept where the matching mask value is > 0 (ReLU derivative of
// the forward output), else 0.
__device__ __forceinline__ u32x4 mask8(u32x4 x, u32x4 m) {
    u32x4 r;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const unsigned lo = m[i] & 0xffffu, hi = m[i] >> 16;
        const unsigned klo = (lo != 0 && !(lo & 0x8000u)) ? 0x0000ffffu : 0u;
        const unsigned khi = (hi != 0 && !(hi & 0x8000u)) ? 0xffff0000u : 0u;
        r[i] = x[i] & (klo | khi);
    }
    return r;
}

__device__ __forceinline__ bool pos_bf(bf16_t m) { return m != 0 && !(m & 0x8000u); }

// ------------------------------------------------------------------ 3^3 conv (MFMA)

struct ConvArgs {
    int B, D, c0, o0;
    const bf16_t* x0;
    const bf16_t* x1;
    const bf16_t* mask;
    const bf16_t* w;  // [cout][27][cin]
    const float* bias;
    int bias_stride;  // 0 (shared) or cout (per sample)
    int relu;
    bf16_t* y0;
    bf16_t* y1;
    const float* head;  // fused 1^3 head (cout = 16): w[16], b; or NULL
    float* prob;        // its sigmoid output (B,D,D,D) fp32
    bf16_t* pool;       // 2x2x2 mean of y0 (B,D/2,D/2,D/2,cout), written by the POOL form
};

// The MFMA takes the weights as its A operand and the activations as its B
// operand, so that lane (row, grp) ends with output channels 4 grp .. 4 grp+3
// of voxel `row` of the 16-voxel line and writes them as one 8-byte store.
//
// K = 27 * cin runs tap-major, channel-minor, padded to a multiple of 32.
// Lane group grp of slice kc covers k = 32 kc + 8 grp .. +7: for cin >= 32
// the slice lies inside one tap, for cin = 16 groups 0-1 take tap 2 kc and
// groups 2-3 tap 2 kc + 1 (tap 27 is the padding, read as zeros).  After the
// slice loop is unrolled, the tap offsets are compile-time constants and a
// lane only selects between two of them.
template <int CIN>
struct SliceLanes {
    bool hi;  // cin = 16 and lane group 2 or 3
    int ci;   // the lane's first channel inside its tap (for cin >= 32 plus the slice's channel base)
    __device__ explicit SliceLanes(int grp) : hi(CIN == 16 && grp >= 2), ci(CIN == 16 ? (grp & 1) * 8 : grp * 8) {}
    static constexpr int tap(int kc) { return kc * 32 / CIN; }
    static constexpr int cbase(int kc) { return CIN == 16 ? 0 : kc * 32 % CIN; }
    // false for the padding lanes of the last slice
    __device__ bool valid(int kc) const { return !(hi && tap(kc) + 1 >= 27); }
};

// element offset of tap t = 9 dz + 3 dy + dx in a box of HY x HX voxels per
// plane, CP bf16 per voxel; in_plane leaves dz out
template <int HY, int HX, int CP>
__device__ __forceinline__ constexpr int tap_off(int t) {
    return (((t / 9) * HY + (t / 3) % 3) * HX + t % 3) * CP;
}
template <int HX, int CP>
__device__ __forceinline__ constexpr int in_plane(int t) {
    return (((t / 3) % 3) * HX + t % 3) * CP;
}

// weight fragment (slice j of input-channel chunk c, N tile n) of this lane:
// w[16 n + row][k .. k+7] with k = tap * cin + channel; wl = w + row * K +
// weight_lane (for chunks of 16 channels, lane groups 2-3 sit one tap
// further).  Zeros for the padding lanes (their address is clamped into the
// row).
template <int CIN, int CH>
__device__ __forceinline__ int weight_lane(int grp) {
    return CH == 16 ? (grp >= 2 ? CIN : 0) + (grp & 1) * 8 : grp * 8;
}
template <int CIN, int CH, int K>
__device__ __forceinline__ bf16x8 weight_frag(const bf16_t* wl, const SliceLanes<CH>& sl, int c, int j, int n) {
    const bool ok = sl.valid(j);
    const int kofs =
        (CH == 16 ? 2 * j * CIN : SliceLanes<CH>::tap(j) * CIN + SliceLanes<CH>::cbase(j)) + c * CH;
    const u32x4 v = *(const u32x4*)(wl + n * 16 * K + (ok ? kofs : 0));
    const u32x4 zero = {0u, 0u, 0u, 0u};
    return __builtin_bit_cast(bf16x8, ok ? v : zero);
}

// output channels 16 n + 4 grp .. +3 of voxel vrow + row (vrow: first voxel
// of the line); N tiles below o0 go to y0, the others to y1
template <int COUT>
__device__ __forceinline__ void store_quad(const ConvArgs& a, size_t vrow, int row, int n, int grp, floatx4 v) {
    if (a.relu) {
#pragma unroll
        for (int i = 0; i < 4; ++i) v[i] = fmaxf(v[i], 0.0f);
    }
    u32x2 q;
    q[0] = f2bf2(v[0], v[1]);
    q[1] = f2bf2(v[2], v[3]);
    if (n * 16 < a.o0) {
        bf16_t* p = a.y0 + vrow * a.o0;
        *(u32x2*)(p + row * a.o0 + n * 16 + grp * 4) = q;
    } else {
        const int o1 = COUT - a.o0;
        bf16_t* p = a.y1 + vrow * o1;
        *(u32x2*)(p + row * o1 + n * 16 - a.o0 + grp * 4) = q;
    }
}

// The head's logit b + sum_c w[c] x[c] over 16 channels is summed in a fixed
// order shared by the head kernel and the fused form: four partials of four
// consecutive channels, p_g = fma chain over channels 4g .. 4g+3, then
// ((p0 + p1) + (p2 + p3)) + b.
__device__ __forceinline__ float head_partial(const float* w4, const float* x4) {
    float p = w4[0] * x4[0];
    p = fmaf(w4[1], x4[1], p);
    p = fmaf(w4[2], x4[2], p);
    return fmaf(w4[3], x4[3], p);
}
// sigmoid on the transcendental unit (v_exp_f32, v_rcp_f32: about 1 ulp each)
__device__ __forceinline__ float head_sigmoid(float lg) { return __builtin_amdgcn_rcpf(1.0f + __expf(-lg)); }

// Fused head (inference, cout = 16): lane group grp holds channels 4 grp ..
// +3 of voxel vrow + row; rounded to bf16 as the stored activation would be,
// they give partial p_grp, and two butterflies over the lane groups form the
// logit in the shared order (fp addition is commutative, so every lane of the
// row holds the same value).
__device__ __forceinline__ void head_quad(const ConvArgs& a, size_t vrow, int row, int grp, floatx4 v) {
    const floatx4 w = *(const floatx4*)(a.head + grp * 4);  // the lane group's 4 head weights
    float xr[4], w4[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const float r = a.relu ? fmaxf(v[i], 0.0f) : v[i];
        xr[i] = __uint_as_float((unsigned)f2bf(r) << 16);
        w4[i] = w[i];
    }
    const float p = head_partial(w4, xr);
    const float s1 = p + __shfl_xor(p, 16, 64);     // p0 + p1 (groups 0, 1), p2 + p3 (groups 2, 3)
    const float s2 = s1 + __shfl_xor(s1, 32, 64);   // (p0 + p1) + (p2 + p3)
    const float prob = head_sigmoid(s2 + a.head[16]);
    if (grp == 0) a.prob[vrow + row] = prob;
}

// bias of output channels 16 n + 4 grp .. +3
template <int NT>
__device__ __forceinline__ void load_bias(const ConvArgs& a, int b, int grp, floatx4 (&bv)[NT]) {
#pragma unroll
    for (int n = 0; n < NT; ++n)
        bv[n] = a.bias ? *(const floatx4*)(a.bias + b * a.bias_stride + n * 16 + grp * 4) : floatx4{0.f, 0.f, 0.f, 0.f};
}

// One halo tile per block: TZ x TY lines of 16 voxels, input with a one-voxel
// halo in LDS; each of the four waves owns MW lines of one z-plane.  With
// CH < CIN the tile holds CH input channels at a time and the K loop walks
// the channel chunks: a smaller tile, so more lines per block and more MFMAs
// per weight fragment fetched from L2.
template <int CIN, int COUT, int TZ, int TY, int CH = CIN>
__global__ __launch_bounds__(256) void k_conv3_mfma(ConvArgs a) {
    constexpr int NCH = CIN / CH;
    constexpr int HZ = TZ + 2, HY = TY + 2, HX = 18, CP = CH + 8;
    constexpr int MT = TZ * TY, MW = MT / 4, NT = COUT / 16;
    constexpr int K = 27 * CIN, KC = (27 * CH + 31) / 32;  // K slices per chunk
    constexpr int CPR = HX * (CH / 8);                      // 16-byte chunks per halo row
    constexpr int NR = HZ * HY, NP = (CPR + 63) / 64, RW = NR / 4;
    static_assert(CIN % CH == 0 && (CH == 16 || CH % 32 == 0), "channel chunks");
    static_assert(MT % 4 == 0 && TY % MW == 0, "a wave's lines lie in one z-plane");
    static_assert(NR % 4 == 0, "the waves share the halo rows");
    static_assert(COUT % 16 == 0, "channel counts");
    __shared__ __attribute__((aligned(16))) bf16_t tile[HZ * HY * HX * CP];
    const int D = a.D;
    const int nyt = D / TY;
    const int x0 = blockIdx.x * 16;
    const int y0 = (blockIdx.y % nyt) * TY, z0 = (blockIdx.y / nyt) * TZ;
    const int b = blockIdx.z;
    const int c0 = a.c0, c1 = CIN - c0;
    const size_t vb = (size_t)b * D * D * D;
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int row = lane & 15, grp = lane >> 4;
    const u32x4 zero = {0u, 0u, 0u, 0u};

    // halo tile of input channels cb .. cb+CH-1: halo rows (hz, hy) go to the
    // waves in turn; lane p < CPR of a wave loads 16-byte chunk p of the row
    // (voxel hx = p / (CH/8), channels ch .. ch+7), so only the row offset
    // changes between the rows.  All rows are in flight before the LDS
    // writes; lanes past CPR load a copy of chunk CPR-1 and write nothing
    // (their writes to one address would serialise in the LDS).
    auto load_tile = [&](int cb) {
        const bf16_t* lsrc[NP];
        const bf16_t* lmask[NP];
        int lmul[NP], llds[NP];
        bool lw[NP];
#pragma unroll
        for (int j = 0; j < NP; ++j) {
            lw[j] = lane + 64 * j < CPR;
            const int p = min(lane + 64 * j, CPR - 1);
            const int hx = p / (CH / 8), ch = (p % (CH / 8)) * 8, gch = cb + ch;
            const unsigned gx = wrap(x0 + hx - 1, D);
            if (gch < c0) {
                lsrc[j] = a.x0 + (vb + gx) * c0 + gch;
                lmul[j] = c0;
            } else {
                lsrc[j] = a.x1 + (vb + gx) * c1 + (gch - c0);
                lmul[j] = c1;
            }
            lmask[j] = a.mask ? a.mask + (vb + gx) * c0 + gch : nullptr;
            llds[j] = hx * CP + ch;
        }
        auto row_off = [&](int it) -> unsigned {
            const int r = wave + 4 * it;
            const int hz = r / HY, hy = r - hz * HY;
            return ((unsigned)wrap(z0 + hz - 1, D) * D + wrap(y0 + hy - 1, D)) * D;
        };
        if (!a.mask) {
            u32x4 v[RW][NP];
#pragma unroll
            for (int it = 0; it < RW; ++it) {
                const unsigned rowv = row_off(it);
#pragma unroll
                for (int j = 0; j < NP; ++j) v[it][j] = *(const u32x4*)(lsrc[j] + (size_t)rowv * lmul[j]);
            }
#pragma unroll
            for (int it = 0; it < RW; ++it)
#pragma unroll
                for (int j = 0; j < NP; ++j)
                    if (lw[j]) *(u32x4*)(tile + (wave + 4 * it) * HX * CP + llds[j]) = v[it][j];
        } else {
            for (int it = 0; it < RW; ++it) {
                const unsigned rowv = row_off(it);
#pragma unroll
                for (int j = 0; j < NP; ++j) {
                    const u32x4 v = *(const u32x4*)(lsrc[j] + (size_t)rowv * lmul[j]);
                    const u32x4 m = *(const u32x4*)(lmask[j] + (size_t)rowv * c0);
                    if (lw[j]) *(u32x4*)(tile + (wave + 4 * it) * HX * CP + llds[j]) = mask8(v, m);
                }
            }
        }
    };

    const SliceLanes<CH> sl(grp);
    const int mt0 = wave * MW, lz = mt0 / TY, ly0 = mt0 % TY;
    const bf16_t* abase = tile + ((lz * HY + ly0) * HX + row) * CP + sl.ci;
    const bf16_t* wl = a.w + (size_t)row * K + weight_lane<CIN, CH>(grp);
    auto load_b = [&](int c, int j, bf16x8 (&bw)[NT]) {
#pragma unroll
        for (int n = 0; n < NT; ++n) bw[n] = weight_frag<CIN, CH, K>(wl, sl, c, j, n);
    };
    auto load_a = [&](int j, bf16x8 (&av)[MW]) {
        const int t = SliceLanes<CH>::tap(j);
        const bool ok = sl.valid(j);
        const int off = (sl.hi && ok) ? tap_off<HY, HX, CP>(t + 1) : tap_off<HY, HX, CP>(t) + SliceLanes<CH>::cbase(j);
#pragma unroll
        for (int m = 0; m < MW; ++m) {
            const u32x4 v = *(const u32x4*)(abase + off + m * HX * CP);
            av[m] = __builtin_bit_cast(bf16x8, ok ? v : zero);
        }
    };
    floatx4 acc[MW][NT];
#pragma unroll
    for (int m = 0; m < MW; ++m)
#pragma unroll
        for (int n = 0; n < NT; ++n) acc[m][n] = floatx4{0.0f, 0.0f, 0.0f, 0.0f};
    // software pipeline: the weight fragments (from L2) run PD slices ahead
    // in a ring (more when a slice has few MFMAs to cover the latency), the
    // LDS fragments one slice ahead
    constexpr int PDC = 8 / NT < 2 ? 2 : (8 / NT > 6 ? 6 : 8 / NT);  // ring of at most 32 VGPRs
    constexpr int PDW = 24 / (MW * NT) < 2 ? 2 : (24 / (MW * NT) > PDC ? PDC : 24 / (MW * NT));
    constexpr int PD = KC < PDW ? KC : PDW;
    // dy reuse (16-channel chunks, each wave owning a whole z-plane of
    // lines): the K slices are regrouped into five families of tap pairs
    // whose members differ only in dy,
    //   f = 0, 1, 2: taps (dz=f, dy, dx=0) | (f, dy, 1)
    //   f = 3:       taps (0, dy, 2)       | (1, dy, 2)
    //   f = 4:       tap  (2, dy, 2)       | zero padding,
    // (lane groups 0-1 take the first tap, 2-3 the second), so one LDS
    // fragment of input line L feeds output lines L, L-1 and L-2 with the
    // weights of dy = 0, 1, 2: 5 x (MW + 2) LDS reads per chunk instead of
    // 14 x MW, for one padded slice more of MFMAs.
    constexpr bool REUSE = CH == 16 && NT <= 2 && MW == TY;
    const bool hi = grp >= 2;
    const int lane_a = ((lz * HY) * HX + row) * CP + (grp & 1) * 8;
    const int hi_ax = hi ? CP : 0, hi_az = hi ? HY * HX * CP : 0;
    const bf16_t* wl0 = a.w + (size_t)row * K + (grp & 1) * 8;
    const int hi_wx = hi ? CIN : 0, hi_wz = hi ? 9 * CIN : 0;
    auto fam_w = [&](int c, int f, bf16x8 (&wq)[3][NT]) {
        const int dz = f < 3 ? f : (f == 3 ? 0 : 2), dx = f < 3 ? 0 : 2;
        const bool ok = !(f == 4 && hi);
        const int hw = f < 3 ? hi_wx : (f == 3 ? hi_wz : 0);
#pragma unroll
        for (int dy = 0; dy < 3; ++dy)
#pragma unroll
            for (int n = 0; n < NT; ++n) {
                const int tap = dz * 9 + dy * 3 + dx;
                const u32x4 v = *(const u32x4*)(wl0 + n * 16 * K + tap * CIN + c * CH + hw);
                wq[dy][n] = __builtin_bit_cast(bf16x8, ok ? v : zero);
            }
    };
    auto fam_a = [&](int f, int L) -> bf16x8 {
        const int dz = f < 3 ? f : (f == 3 ? 0 : 2), dx = f < 3 ? 0 : 2;
        const bool ok = !(f == 4 && hi);
        const int ha = f < 3 ? hi_ax : (f == 3 ? hi_az : 0);
        const u32x4 v = *(const u32x4*)(tile + lane_a + ((dz * HY + L) * HX + dx) * CP + ha);
        return __builtin_bit_cast(bf16x8, ok ? v : zero);
    };
#pragma unroll
    for (int c = 0; c < NCH; ++c) {
        if constexpr (REUSE) {
            bf16x8 wq[2][3][NT];
            fam_w(c, 0, wq[0]);  // in flight during the tile load
            if (c > 0) __syncthreads();
            load_tile(c * CH);
            __syncthreads();
            // steps k = (family f, input line L); the LDS fragments run two
            // steps ahead (across the family boundaries), since the one to
            // three MFMAs of a step are shorter than the LDS latency
            constexpr int NL = MW + 2, NS = 5 * NL;
            bf16x8 ar[3];
            ar[0] = fam_a(0, 0);
            ar[1] = fam_a(0, 1);
#pragma unroll
            for (int k = 0; k < NS; ++k) {
                const int f = k / NL, L = k % NL;
                if (L == 0 && f + 1 < 5) fam_w(c, f + 1, wq[(f + 1) & 1]);
                if (k + 2 < NS) ar[(k + 2) % 3] = fam_a((k + 2) / NL, (k + 2) % NL);
                const bf16x8 acur = ar[k % 3];
#pragma unroll
                for (int dy = 0; dy < 3; ++dy) {
                    const int m = L - dy;
                    if (m >= 0 && m < MW) {
#pragma unroll
                        for (int n = 0; n < NT; ++n)
                            acc[m][n] =
                                __builtin_amdgcn_mfma_f32_16x16x32_bf16(wq[f & 1][dy][n], acur, acc[m][n], 0, 0, 0);
                    }
                }
                __builtin_amdgcn_sched_barrier(0);
            }
            continue;
        }
        bf16x8 bq[PD][NT], acur[MW], anext[MW];
#pragma unroll
        for (int j = 0; j < PD; ++j) load_b(c, j, bq[j]);  // in flight during the tile load
        if (c > 0) __syncthreads();                       // every wave is done with the previous chunk
        load_tile(c * CH);
        __syncthreads();
        load_a(0, acur);
#pragma unroll
        for (int j = 0; j < KC; ++j) {
            if (j + 1 < KC) load_a(j + 1, anext);
#pragma unroll
            for (int m = 0; m < MW; ++m)
#pragma unroll
                for (int n = 0; n < NT; ++n)
                    acc[m][n] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bq[j % PD][n], acur[m], acc[m][n], 0, 0, 0);
            if (j + PD < KC) load_b(c, j + PD, bq[j % PD]);
#pragma unroll
            for (int m = 0; m < MW; ++m) acur[m] = anext[m];
            // keep each slice's loads in its own slice: the scheduler would
            // otherwise sink them next to their use and expose the L2 latency
            __builtin_amdgcn_sched_barrier(0);
        }
    }
    floatx4 bv[NT];
    load_bias<NT>(a, b, grp, bv);
    if constexpr (COUT == 16) {
        if (a.head) {
#pragma unroll
            for (int m = 0; m < MW; ++m) {
                const size_t vrow = vb + ((size_t)(z0 + lz) * D + (y0 + ly0 + m)) * D + x0;
                head_quad(a, vrow, row, grp, acc[m][0] + bv[0]);
            }
            return;
        }
    }
#pragma unroll
    for (int m = 0; m < MW; ++m) {
        const size_t vrow = vb + ((size_t)(z0 + lz) * D + (y0 + ly0 + m)) * D + x0;
#pragma unroll
        for (int n = 0; n < NT; ++n) store_quad<COUT>(a, vrow, row, n, grp, acc[m][n] + bv[n]);
    }
}

// Sliding-window form for large boxes: a block owns 16 x TY output voxels of
// NZ consecutive z-planes and walks z with a ring of four input planes (each
// (TY+2) x 18 voxels with halo) in LDS: one new plane is fetched per output
// plane, into registers while the previous plane computes.  Layers with at
// most 14 weight fragments (K slices x N tiles) keep every B fragment in
// registers for the whole walk; the others stream them per K slice,
// double-buffered.
//
// POOL (16 -> 16 channels, two lines per wave): the 2x2x2 mean of the stored
// output goes to a.pool as well.  Lanes of rows 2i and 2i+1 hold an x-pair,
// a wave's two lines the y-pair, consecutive planes the z-pair; the sum runs
// in the order of k_pool_fwd (z, y, x, from 0), so the pooled box equals that
// kernel's output bit for bit.
template <int CIN, int COUT, int TY, int NZ, bool HEAD = false, bool POOL = false>
__global__ __launch_bounds__(256, POOL ? 4 : 1) void k_conv3_slide(ConvArgs a) {
    constexpr int HY = TY + 2, HX = 18, CP = CIN >= 32 ? CIN + 8 : CIN;
    constexpr int PLANE = HY * HX * CP;               // bf16 per ring slot
    constexpr int MW = TY / 4, NT = COUT / 16;
    constexpr int K = 27 * CIN, KC = (K + 31) / 32;
    constexpr bool HOLD = KC * NT <= 14;
    constexpr int CHUNKS = HY * HX * (CIN / 8);       // 16-byte chunks per plane
    constexpr int NPF = (CHUNKS + 255) / 256;
    static_assert(TY % 4 == 0, "four waves share the lines");
    static_assert(CIN % 8 == 0 && COUT % 16 == 0, "channel counts");
    __shared__ __attribute__((aligned(16))) bf16_t ring[4 * PLANE];
    const int D = a.D;
    const int nyt = D / TY;
    const int x0 = blockIdx.x * 16;
    const int y0 = (blockIdx.y % nyt) * TY, z0 = (blockIdx.y / nyt) * NZ;
    const int b = blockIdx.z;
    const int c0 = a.c0, c1 = CIN - c0;
    const int t = threadIdx.x;
    const u32x4 zero = {0u, 0u, 0u, 0u};
    const size_t vb = (size_t)b * D * D * D;
    const bf16_t* xs0 = a.x0 + vb * c0;
    const bf16_t* xs1 = c1 ? a.x1 + vb * c1 : nullptr;
    const bf16_t* ms = a.mask ? a.mask + vb * c0 : nullptr;

    // global -> registers for halo plane hz (global z = z0 - 1 + hz)
    auto fetch = [&](int hz, u32x4 (&pf)[NPF]) {
        const unsigned gzD = (unsigned)wrap(z0 - 1 + hz, D) * D;
#pragma unroll
        for (int q = 0; q < NPF; ++q) {
            const int i = t + q * 256;
            u32x4 v = zero;
            if (i < CHUNKS) {
                const int ch = (i % (CIN / 8)) * 8;
                const int vox = i / (CIN / 8);
                const int hx = vox % HX, hy = vox / HX;
                const unsigned g = (gzD + wrap(y0 + hy - 1, D)) * D + wrap(x0 + hx - 1, D);
                if (ch < c0) {
                    v = *(const u32x4*)(xs0 + (size_t)g * c0 + ch);
                    if (ms) v = mask8(v, *(const u32x4*)(ms + (size_t)g * c0 + ch));
                } else {
                    v = *(const u32x4*)(xs1 + (size_t)g * c1 + (ch - c0));
                }
            }
            pf[q] = v;
        }
    };
    auto store = [&](int hz, const u32x4 (&pf)[NPF]) {
        bf16_t* slot = ring + (hz & 3) * PLANE;
#pragma unroll
        for (int q = 0; q < NPF; ++q) {
            const int i = t + q * 256;
            if (i < CHUNKS) {
                const int ch = (i % (CIN / 8)) * 8;
                const int vox = i / (CIN / 8);
                *(u32x4*)(slot + vox * CP + ch) = pf[q];
            }
        }
    };
    {
        u32x4 pf[NPF];
        for (int hz = 0; hz < 3; ++hz) {
            fetch(hz, pf);
            store(hz, pf);
        }
    }
    const int lane = t & 63;
    const int wave = __builtin_amdgcn_readfirstlane(t >> 6);
    const int row = lane & 15, grp = lane >> 4;
    const SliceLanes<CIN> sl(grp);
    const bf16_t* wl = a.w + (size_t)row * K + weight_lane<CIN, CIN>(grp);
    const int lane_off = ((wave * MW) * HX + row) * CP + sl.ci;
    bf16x8 bhold[HOLD ? KC : 1][NT];
    if constexpr (HOLD) {
#pragma unroll
        for (int kc = 0; kc < KC; ++kc)
#pragma unroll
            for (int n = 0; n < NT; ++n) bhold[kc][n] = weight_frag<CIN, CIN, K>(wl, sl, 0, kc, n);
    }
    floatx4 bv[NT];
    load_bias<NT>(a, b, grp, bv);
    static_assert(!POOL || (COUT == 16 && MW == 2 && NZ % 2 == 0 && !HEAD), "pool pairs");
    // pool sums of channels 4 grp .. +3, carried from an even plane to the next
    __shared__ __attribute__((aligned(16))) float psum[POOL ? 256 * 4 : 4];
    __syncthreads();
    for (int oz = 0; oz < NZ; ++oz) {
        u32x4 pf[NPF];
        const bool more = oz + 3 <= NZ + 1;
        if (more) fetch(oz + 3, pf);
        floatx4 acc[MW][NT];
#pragma unroll
        for (int m = 0; m < MW; ++m)
#pragma unroll
            for (int n = 0; n < NT; ++n) acc[m][n] = floatx4{0.0f, 0.0f, 0.0f, 0.0f};
        // ring slots of the three planes this output plane reads
        const int s0 = (oz & 3) * PLANE, s1 = ((oz + 1) & 3) * PLANE, s2 = ((oz + 2) & 3) * PLANE;
        auto slot_of = [&](int dz) { return dz == 0 ? s0 : (dz == 1 ? s1 : s2); };
        // one K slice: A fragments of this wave's lines from the ring, MFMAs
        auto slice = [&](int kc, const bf16x8 (&bw)[NT]) {
            const int tp = SliceLanes<CIN>::tap(kc);
            const bool ok = sl.valid(kc);
            const int lo = slot_of(tp / 9) + in_plane<HX, CP>(tp) + SliceLanes<CIN>::cbase(kc);
            const int up = slot_of(((tp + 1) / 9) % 3) + in_plane<HX, CP>(tp + 1);
            const int off = (sl.hi && ok) ? up : lo;
            bf16x8 av[MW];
#pragma unroll
            for (int m = 0; m < MW; ++m) {
                const u32x4 v = *(const u32x4*)(ring + lane_off + off + m * HX * CP);
                av[m] = __builtin_bit_cast(bf16x8, ok ? v : zero);
            }
#pragma unroll
            for (int m = 0; m < MW; ++m)
#pragma unroll
                for (int n = 0; n < NT; ++n)
                    acc[m][n] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bw[n], av[m], acc[m][n], 0, 0, 0);
        };
        if constexpr (HOLD) {
#pragma unroll
            for (int kc = 0; kc < KC; ++kc) slice(kc, bhold[kc]);
        } else {
            // weight fragments PD slices ahead in a ring
            constexpr int PD = 3;
            bf16x8 bq[PD][NT];
#pragma unroll
            for (int j = 0; j < PD; ++j)
#pragma unroll
                for (int n = 0; n < NT; ++n) bq[j][n] = weight_frag<CIN, CIN, K>(wl, sl, 0, j, n);
#pragma unroll
            for (int kc = 0; kc < KC; ++kc) {
                slice(kc, bq[kc % PD]);
                if (kc + PD < KC) {
#pragma unroll
                    for (int n = 0; n < NT; ++n) bq[kc % PD][n] = weight_frag<CIN, CIN, K>(wl, sl, 0, kc + PD, n);
                }
                __builtin_amdgcn_sched_barrier(0);
            }
        }
        // the fused head is its own instance: its epilogue registers would
        // cost the plain instance a wave per SIMD
        if constexpr (HEAD) {
            static_assert(COUT == 16, "the fused head reads 16 channels");
#pragma unroll
            for (int m = 0; m < MW; ++m) {
                const size_t vrow = vb + ((size_t)(z0 + oz) * D + (y0 + wave * MW + m)) * D + x0;
                head_quad(a, vrow, row, grp, acc[m][0] + bv[0]);
            }
        } else {
#pragma unroll
            for (int m = 0; m < MW; ++m) {
                const size_t vrow = vb + ((size_t)(z0 + oz) * D + (y0 + wave * MW + m)) * D + x0;
#pragma unroll
                for (int n = 0; n < NT; ++n) store_quad<COUT>(a, vrow, row, n, grp, acc[m][n] + bv[n]);
            }
        }
        if (more) store(oz + 3, pf);
        // (after the ring store, so that the prefetch registers are free)
        if constexpr (POOL) {
            // per channel: the stored (bf16) values of lines 0 and 1 and of
            // the x-neighbour row, added to the sums of the even plane (kept
            // in the lane's own LDS slot, not in registers)
            const bool odd_plane = oz & 1;
            floatx4 o;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const float r0 = acc[0][0][i] + bv[0][i], r1 = acc[1][0][i] + bv[0][i];
                const float s0 = __uint_as_float((unsigned)f2bf(a.relu ? fmaxf(r0, 0.0f) : r0) << 16);
                const float s1 = __uint_as_float((unsigned)f2bf(a.relu ? fmaxf(r1, 0.0f) : r1) << 16);
                const float n0 = __shfl_xor(s0, 1, 64), n1 = __shfl_xor(s1, 1, 64);
                o[i] = (((odd_plane ? psum[t * 4 + i] : 0.0f) + s0) + n0) + s1 + n1;
            }
            if (!odd_plane) {
                *(floatx4*)(psum + t * 4) = o;
            } else {
#pragma unroll
                for (int i = 0; i < 4; ++i) o[i] *= 0.125f;
                if ((row & 1) == 0) {
                    const int Dh = D / 2;
                    const size_t pv =
                        (((size_t)b * Dh + (z0 + oz) / 2) * Dh + (y0 + wave * MW) / 2) * Dh + (x0 + row) / 2;
                    u32x2 q;
                    q[0] = f2bf2(o[0], o[1]);
                    q[1] = f2bf2(o[2], o[3]);
                    *(u32x2*)(a.pool + pv * 16 + grp * 4) = q;
                }
            }
        }
        __syncthreads();
    }
}

// ------------------------------------------------------------------ 3^3 conv weight gradient

struct WgradArgs {
    int B, D, c0;
    const bf16_t* x0;
    const bf16_t* x1;
    const bf16_t* g;
    const bf16_t* mask;
    float* partial;  // [gridDim.x][cout][27][cin]
    int tiles_per_block;
};

// Voxel tile 2 x 2 x 16 (64 voxels); taps in groups of 3 (grid.y = 9).  Each
// thread owns 4 output x 4 input channels for the group's 3 taps; thread sets
// split the voxels and are added in LDS at the end.
template <int CIN, int COUT>
__global__ __launch_bounds__(256) void k_conv3_wgrad(WgradArgs a) {
    constexpr int TZ = 2, TY = 2, HZ = 4, HY = 4, HX = 18, NV = TZ * TY * 16;
    constexpr int PAIRS = (COUT / 4) * (CIN / 4);
    constexpr int SETS = 256 / PAIRS;
    static_assert(PAIRS <= 256 && 256 % PAIRS == 0, "pairs per block");
    __shared__ __attribute__((aligned(16))) bf16_t xt[HZ * HY * HX * CIN];
    __shared__ float gt[NV * COUT];
    __shared__ float red[(SETS > 1 ? 256 : 1) * 16];
    const int D = a.D;
    const int tg = blockIdx.y;  // taps 3*tg .. 3*tg+2
    const int t = threadIdx.x;
    const int pair = t % PAIRS, set = t / PAIRS;
    const int co0 = (pair % (COUT / 4)) * 4, ci0 = (pair / (COUT / 4)) * 4;
    const int c1 = CIN - a.c0;
    const int nx = D / 16, ny = D / TY, nz = D / TZ;
    const int ntiles = a.B * nz * ny * nx;
    float acc[3][4][4];
#pragma unroll
    for (int q = 0; q < 3; ++q)
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j) acc[q][i][j] = 0.0f;
    for (int s = 0; s < a.tiles_per_block; ++s) {
        const int tile = blockIdx.x * a.tiles_per_block + s;
        if (tile >= ntiles) break;  // uniform
        const int xt0 = (tile % nx) * 16;
        const int yt0 = ((tile / nx) % ny) * TY;
        const int zt0 = ((tile / (nx * ny)) % nz) * TZ;
        const int b = tile / (nx * ny * nz);
        __syncthreads();  // previous tile fully consumed
        for (int i = t; i < HZ * HY * HX * (CIN / 8); i += 256) {
            const int ch = (i % (CIN / 8)) * 8;
            const int vox = i / (CIN / 8);
            const int hx = vox % HX, hy = (vox / HX) % HY, hz = vox / (HX * HY);
            const size_t g = (((size_t)b * D + wrap(zt0 + hz - 1, D)) * D + wrap(yt0 + hy - 1, D)) * D +
                             wrap(xt0 + hx - 1, D);
            const u32x4 v = ch < a.c0 ? *(const u32x4*)(a.x0 + g * a.c0 + ch)
                                      : *(const u32x4*)(a.x1 + g * c1 + (ch - a.c0));
            *(u32x4*)(xt + vox * CIN + ch) = v;
        }
        for (int i = t; i < NV * COUT; i += 256) {
            const int co = i % COUT, v = i / COUT;
            const int lx = v % 16, ly = (v / 16) % TY, lz = v / (16 * TY);
            const size_t g = (((size_t)b * D + zt0 + lz) * D + yt0 + ly) * D + xt0 + lx;
            float gv = bf2f(a.g[g * COUT + co]);
            if (a.mask && !pos_bf(a.mask[g * COUT + co])) gv = 0.0f;
            gt[i] = gv;
        }
        __syncthreads();
        for (int v = set; v < NV; v += SETS) {
            const int lx = v % 16, ly = (v / 16) % TY, lz = v / (16 * TY);
            float gv[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) gv[i] = gt[v * COUT + co0 + i];
#pragma unroll
            for (int q = 0; q < 3; ++q) {
                const int tap = tg * 3 + q;
                const int dz = tap / 9, dy = (tap / 3) % 3, dx = tap % 3;
                const bf16_t* xp = xt + (((lz + dz) * HY + (ly + dy)) * HX + (lx + dx)) * CIN + ci0;
                float xv[4];
#pragma unroll
                for (int j = 0; j < 4; ++j) xv[j] = bf2f(xp[j]);
#pragma unroll
                for (int i = 0; i < 4; ++i)
#pragma unroll
                    for (int j = 0; j < 4; ++j) acc[q][i][j] = fmaf(gv[i], xv[j], acc[q][i][j]);
            }
        }
    }
    // add the sets (fixed order) and write this block's partial
    float* out = a.partial + (size_t)blockIdx.x * COUT * 27 * CIN;
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        const int tap = tg * 3 + q;
        if (SETS > 1) {
            __syncthreads();
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j) red[t * 16 + i * 4 + j] = acc[q][i][j];
            __syncthreads();
            if (set == 0) {
                for (int s2 = 1; s2 < SETS; ++s2)
#pragma unroll
                    for (int i = 0; i < 4; ++i)
#pragma unroll
                        for (int j = 0; j < 4; ++j) acc[q][i][j] += red[(s2 * PAIRS + pair) * 16 + i * 4 + j];
            }
        }
        if (set == 0) {
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j) out[((size_t)(co0 + i) * 27 + tap) * CIN + ci0 + j] = acc[q][i][j];
        }
    }
}

// Matrix-core weight gradient, dW[co][tap][ci] = sum_v g'[v][co] x[v + off(tap)][ci]
// (g' = g where the forward output is positive), on v_mfma_f32_16x16x32_bf16
// with K = 32 voxels of an x-line.  Grid: (voxel-slab blocks, 16-input-channel
// chunk x dz).  Voxel tiles of 2 x 2 lines of 32 voxels; the masked gradient
// (all cout channels) and the chunk's input (the 2 planes at dz, 4 rows, 34
// voxels) are written to LDS channel-major, the input in three copies shifted
// by dx = 0, 1, 2, so that every fragment (8 consecutive voxels of one
// channel) is one aligned 16-byte read.  Wave w owns line w of every tile and
// keeps the 9 (dy, dx) taps x cout/16 accumulator tiles over all of the
// block's tiles; the four waves are added in a fixed order at the end, and
// the blocks by k_sum_blocks.
template <int CIN, int COUT>
__global__ __launch_bounds__(256) void k_conv3_wgrad_mfma(WgradArgs a) {
    constexpr int NTC = COUT / 16, NTI = CIN / 16, XT = 32, XH = XT + 2, HY = 4;
    constexpr int GS = 4 * XT + 8;           // gs channel stride (bf16), padded against bank conflicts
    constexpr int XS = 2 * HY * XT + 8;      // xs channel stride
    static_assert(CIN % 16 == 0 && COUT % 16 == 0, "channel counts");
    __shared__ __attribute__((aligned(16))) bf16_t xs[3 * 16 * XS];
    __shared__ __attribute__((aligned(16))) bf16_t gs[COUT * GS];
    const int D = a.D;
    const int cit = blockIdx.y % NTI, dz = blockIdx.y / NTI;
    const int t = threadIdx.x, lane = t & 63;
    const int wave = __builtin_amdgcn_readfirstlane(t >> 6);
    const int row = lane & 15, grp = lane >> 4;
    const int lz = wave >> 1, ly = wave & 1;
    const int nx = D / XT, ny = D / 2, nz = D / 2;
    const int ntiles = a.B * nz * ny * nx;
    // the chunk's channels come from x0 or x1 (c0 is a multiple of 16)
    const int cb = cit * 16;
    const bool from0 = cb < a.c0;
    const bf16_t* xsrc = from0 ? a.x0 : a.x1;
    const int xstride = from0 ? a.c0 : CIN - a.c0;
    const int xoff = from0 ? cb : cb - a.c0;
    floatx4 acc[9][NTC];
#pragma unroll
    for (int q = 0; q < 9; ++q)
#pragma unroll
        for (int n = 0; n < NTC; ++n) acc[q][n] = floatx4{0.0f, 0.0f, 0.0f, 0.0f};
    for (int s = 0; s < a.tiles_per_block; ++s) {
        const int tile = blockIdx.x * a.tiles_per_block + s;
        if (tile >= ntiles) break;  // uniform
        const int x0 = (tile % nx) * XT;
        const int y0 = ((tile / nx) % ny) * 2;
        const int z0 = ((tile / (nx * ny)) % nz) * 2;
        const int b = tile / (nx * ny * nz);
        const size_t vb = (size_t)b * D * D * D;
        __syncthreads();  // the previous tile is consumed
        // masked output gradient: gs[co][line][x], line = 2 lz + ly
        for (int i = t; i < 4 * XT * (COUT / 8); i += 256) {
            const int ch = (i % (COUT / 8)) * 8, v = i / (COUT / 8);
            const int x = v % XT, l = v / XT;
            const size_t g = vb + ((size_t)(z0 + (l >> 1)) * D + (y0 + (l & 1))) * D + x0 + x;
            u32x4 q = *(const u32x4*)(a.g + g * COUT + ch);
            if (a.mask) q = mask8(q, *(const u32x4*)(a.mask + g * COUT + ch));
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                gs[(ch + 2 * k) * GS + l * XT + x] = (bf16_t)(q[k] & 0xffffu);
                gs[(ch + 2 * k + 1) * GS + l * XT + x] = (bf16_t)(q[k] >> 16);
            }
        }
        // the chunk's input: xs[dx][ci][pz][hy][xd] = x(z0 + pz + dz - 1, y0 + hy - 1, x0 + xd + dx - 1)
        for (int i = t; i < 2 * HY * XH * 2; i += 256) {
            const int half = i & 1, v = i >> 1;
            const int xi = v % XH, r = v / XH;
            const int pz = r / HY, hy = r % HY;
            const size_t g =
                vb + ((size_t)wrap(z0 + pz + dz - 1, D) * D + wrap(y0 + hy - 1, D)) * D + wrap(x0 + xi - 1, D);
            const u32x4 q = *(const u32x4*)(xsrc + g * xstride + xoff + half * 8);
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const int xd = xi - c;
                if (xd >= 0 && xd < XT) {
#pragma unroll
                    for (int k = 0; k < 4; ++k) {
                        const int ci = half * 8 + 2 * k;
                        xs[(c * 16 + ci) * XS + r * XT + xd] = (bf16_t)(q[k] & 0xffffu);
                        xs[(c * 16 + ci + 1) * XS + r * XT + xd] = (bf16_t)(q[k] >> 16);
                    }
                }
            }
        }
        __syncthreads();
        bf16x8 ga[NTC];
#pragma unroll
        for (int n = 0; n < NTC; ++n)
            ga[n] = __builtin_bit_cast(bf16x8, *(const u32x4*)(gs + (n * 16 + row) * GS + wave * XT + grp * 8));
#pragma unroll
        for (int dy = 0; dy < 3; ++dy)
#pragma unroll
            for (int dx = 0; dx < 3; ++dx) {
                const bf16x8 xb = __builtin_bit_cast(
                    bf16x8, *(const u32x4*)(xs + (dx * 16 + row) * XS + (lz * HY + ly + dy) * XT + grp * 8));
#pragma unroll
                for (int n = 0; n < NTC; ++n)
                    acc[dy * 3 + dx][n] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ga[n], xb, acc[dy * 3 + dx][n], 0, 0, 0);
            }
    }
    // fixed-order sum of the four waves, one tap at a time, into this block's
    // partial (lane: co = 16 n + 4 grp + i, ci = cb + row)
    float* red = (float*)xs;
    float* out = a.partial + (size_t)blockIdx.x * COUT * 27 * CIN;
#pragma unroll
    for (int q = 0; q < 9; ++q) {
        __syncthreads();
        if (wave > 0) {
#pragma unroll
            for (int n = 0; n < NTC; ++n) *(floatx4*)(red + (((wave - 1) * NTC + n) * 64 + lane) * 4) = acc[q][n];
        }
        __syncthreads();
        if (wave == 0) {
            const int tap = dz * 9 + q;
#pragma unroll
            for (int n = 0; n < NTC; ++n) {
                floatx4 sum = acc[q][n];
#pragma unroll
                for (int w = 0; w < 3; ++w) sum += *(const floatx4*)(red + ((w * NTC + n) * 64 + lane) * 4);
#pragma unroll
                for (int i = 0; i < 4; ++i) out[((size_t)(n * 16 + grp * 4 + i) * 27 + tap) * CIN + cb + row] = sum[i];
            }
        }
    }
}

// Fixed-order sum over blocks of n-float partials: out[i] = sum_b partial[b][i].
// A block of 256 threads covers 32 outputs; thread group g (of 8) adds the
// partials b = g, g + 8, ... in turn, then the 8 group sums are added in
// order (deterministic, and 8 short dependent chains instead of one long one).
constexpr int kSumOut = 32;
__global__ __launch_bounds__(256) void k_sum_blocks(const float* __restrict__ partial, int nblk, int n,
                                                    float* __restrict__ out) {
    __shared__ float red[8][kSumOut];
    const int j = threadIdx.x % kSumOut, g = threadIdx.x / kSumOut;
    const int i = blockIdx.x * kSumOut + j;
    float s = 0.0f;
    if (i < n)
        for (int b = g; b < nblk; b += 8) s += partial[(size_t)b * n + i];
    red[g][j] = s;
    __syncthreads();
    if (g == 0 && i < n) {
        float r = red[0][j];
        for (int q = 1; q < 8; ++q) r += red[q][j];
        out[i] = r;
    }
}
inline unsigned sum_blocks_grid(int n) { return (unsigned)((n + kSumOut - 1) / kSumOut); }

// Per-sample channel sums of a (masked) gradient over a chunk of one
// sample's voxels per block: partial[b][blk][c].  Each thread reads 16-byte
// vectors of 8 channels; lanes with the same vector slot are added in a
// fixed order at the end.
__global__ __launch_bounds__(256) void k_channel_partials(const bf16_t* __restrict__ g,
                                                          const bf16_t* __restrict__ mask, int C,
                                                          size_t vox_per_sample, float* __restrict__ partial) {
    __shared__ float red[256 * 8];
    const int b = blockIdx.y;
    const int vc = C / 8;  // 16-byte vectors per voxel (C in {16, 32, 64})
    const int t = threadIdx.x;
    const int cv = t % vc, lanes = 256 / vc, l = t / vc;
    const size_t per = (vox_per_sample + gridDim.x - 1) / gridDim.x;
    const size_t v0 = (size_t)blockIdx.x * per;
    const size_t v1 = v0 + per < vox_per_sample ? v0 + per : vox_per_sample;
    float s[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (size_t v = v0 + l; v < v1; v += lanes) {
        const size_t idx = ((size_t)b * vox_per_sample + v) * C + cv * 8;
        u32x4 q = *(const u32x4*)(g + idx);
        if (mask) q = mask8(q, *(const u32x4*)(mask + idx));
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            s[2 * k] += __uint_as_float(q[k] << 16);
            s[2 * k + 1] += __uint_as_float(q[k] & 0xffff0000u);
        }
    }
#pragma unroll
    for (int k = 0; k < 8; ++k) red[t * 8 + k] = s[k];
    __syncthreads();
    if (t < C) {
        const int c8 = t / 8, k = t % 8;
        float acc = 0.0f;
        for (int q = 0; q < lanes; ++q) acc += red[(q * vc + c8) * 8 + k];
        partial[((size_t)b * gridDim.x + blockIdx.x) * C + t] = acc;
    }
}

// per_sample[b][c] = sum_blk partial[b][blk][c]; total[c] = sum_b per_sample[b][c]
// (either output may be NULL).
// One block per channel, one thread per partial block (nblk <= 128): for each
// sample the 128 partials are added by a fixed butterfly within each wave and
// the two wave sums in order; the samples in order.
__global__ __launch_bounds__(128) void k_channel_final(const float* __restrict__ partial, int B, int nblk, int C,
                                                       float* __restrict__ per_sample, float* __restrict__ total) {
    __shared__ float ws[2];
    const int c = blockIdx.x, k = threadIdx.x, lane = k & 63, w = k >> 6;
    float tot = 0.0f;
    for (int b = 0; b < B; ++b) {
        float v = k < nblk ? partial[((size_t)b * nblk + k) * C + c] : 0.0f;
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
        if (lane == 0) ws[w] = v;
        __syncthreads();
        const float s = ws[0] + ws[1];
        __syncthreads();
        if (k == 0 && per_sample) per_sample[b * C + c] = s;
        tot += s;
    }
    if (k == 0 && total) total[c] = tot;
}

// ------------------------------------------------------------------ pooling

// One thread per (output voxel, 8 channels): eight 16-byte loads, the mean
// in fp32 (the eight inputs in z, y, x order), one 16-byte store.  Grid:
// (x and channel chunks of a row, output y, output z * B).
__global__ __launch_bounds__(64) void k_pool_fwd(const bf16_t* __restrict__ x, bf16_t* __restrict__ y, int D, int C) {
    const int Dh = D / 2, C8 = C / 8;
    const int i = blockIdx.x * 64 + threadIdx.x;
    if (i >= Dh * C8) return;
    const int ox = i / C8, cc = (i - ox * C8) * 8;
    const int oy = blockIdx.y, oz = blockIdx.z % Dh, b = blockIdx.z / Dh;
    const size_t in0 = ((((size_t)b * D + 2 * oz) * D + 2 * oy) * D + 2 * ox) * C + cc;
    const size_t dy = (size_t)D * C, dz = (size_t)D * D * C;
    float s[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        const u32x4 v = *(const u32x4*)(x + in0 + (q >> 2) * dz + ((q >> 1) & 1) * dy + (q & 1) * C);
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            s[2 * k] += __uint_as_float(v[k] << 16);
            s[2 * k + 1] += __uint_as_float(v[k] & 0xffff0000u);
        }
    }
    u32x4 o;
#pragma unroll
    for (int k = 0; k < 4; ++k) o[k] = f2bf2(s[2 * k] * 0.125f, s[2 * k + 1] * 0.125f);
    *(u32x4*)(y + ((((size_t)b * Dh + oz) * Dh + oy) * Dh + ox) * C + cc) = o;
}

__global__ void k_pool_bwd(const bf16_t* __restrict__ dy, bf16_t* __restrict__ dx, int B, int D, int C, int add) {
    const int Dh = D / 2;
    const size_t n = (size_t)B * D * D * D * C;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const int c = i % C;
        size_t v = i / C;
        const int x = v % D, y = (v / D) % D, z = (v / ((size_t)D * D)) % D;
        const int b = v / ((size_t)D * D * D);
        const size_t u = (((size_t)b * Dh + z / 2) * Dh + y / 2) * Dh + x / 2;
        float g = bf2f(dy[u * C + c]) * 0.125f;
        if (add) g += bf2f(dx[i]);
        dx[i] = f2bf(g);
    }
}

// ------------------------------------------------------------------ 2^3 stride-2 transposed conv

// Matrix-core forms of the transposed convolution.  Forward: for 16
// consecutive input voxels (M) and n = s * COUT + co (N), K = CIN:
// y[2v+s][co] = bias[co] + sum_ci wn[n][ci] x[v][ci], with wn bf16 [8*COUT][CIN].
// Input gradient: K = (s, co), N = ci: dx[v][ci] = sum_k wk[ci][k] dy[2v+s][co],
// with wk bf16 [CIN][8*COUT].  A and B fragments are 16-byte loads of 8
// consecutive channels; each wave walks its 16-voxel row segments.
constexpr int kUpSegs = 4;  // 16-voxel input segments per wave of the transposed conv

template <int CIN, int COUT>
__global__ __launch_bounds__(256) void k_up_fwd_mfma(const bf16_t* __restrict__ x, const bf16_t* __restrict__ wn,
                                                     const float* __restrict__ bias, bf16_t* __restrict__ y, int B,
                                                     int D) {
    // weights are the A operand (N = s * COUT + co), voxels the B operand:
    // lane (row, grp) ends with outputs n = 16 j + 4 grp .. +3 of input voxel
    // v0 + row, i.e. 4 consecutive channels of one output voxel (16 j lies
    // inside one sub-position s because COUT is a multiple of 16).  The tiles
    // of sub-positions s (dx = 0) and s + 1 (dx = 1) are formed together and
    // lane groups 2g and 2g + 1 swap halves, so that each lane stores 8
    // consecutive channels of one output voxel (16 bytes) and the four lanes
    // of a row cover both output voxels of the x-pair.  Voxel indices are
    // 32-bit (the host checks B D^3 < 2^32).
    constexpr int KC = CIN / 32, TPS = COUT / 16;  // N tiles per sub-position
    constexpr bool KEEPW = 8 * TPS * KC <= 8;      // every weight fragment stays in registers
    static_assert(CIN % 32 == 0 && COUT % 16 == 0, "channel counts");
    const int lane = threadIdx.x & 63, row = lane & 15, grp = lane >> 4;
    const bool odd = grp & 1;
    const unsigned nseg = (unsigned)B * D * D * D / 16;
    const unsigned seg0 = (blockIdx.x * 4 + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6)) * kUpSegs;
    auto wfrag = [&](int j, int k) {
        return __builtin_bit_cast(bf16x8, *(const u32x4*)(wn + (j * 16 + row) * CIN + k * 32 + grp * 8));
    };
    bf16x8 wk[KEEPW ? 8 * TPS * KC : 1];
    if constexpr (KEEPW) {
#pragma unroll
        for (int j = 0; j < 8 * TPS; ++j)
#pragma unroll
            for (int k = 0; k < KC; ++k) wk[j * KC + k] = wfrag(j, k);
    }
    const unsigned ud = (unsigned)D;
    const size_t D2 = 2 * (size_t)D;
    for (int sg = 0; sg < kUpSegs; ++sg) {
        const unsigned seg = seg0 + sg;
        if (seg >= nseg) return;  // whole wave
        const unsigned v = seg * 16 + row;  // this lane's input voxel
        bf16x8 av[KC];
#pragma unroll
        for (int k = 0; k < KC; ++k)
            av[k] = __builtin_bit_cast(bf16x8, *(const u32x4*)(x + (size_t)v * CIN + k * 32 + grp * 8));
        const unsigned xx = v % ud, r1 = v / ud, yy = r1 % ud, r2 = r1 / ud, zz = r2 % ud, b = r2 / ud;
        // output voxel of the lane's input voxel at sub-position (0, 0, dx = odd)
        const size_t u0 = (((size_t)b * D2 + 2 * zz) * D2 + 2 * yy) * D2 + 2 * xx + (odd ? 1 : 0);
#pragma unroll
        for (int s = 0; s < 8; s += 2) {
#pragma unroll
            for (int h = 0; h < TPS; ++h) {
                const int j0 = s * TPS + h, j1 = (s + 1) * TPS + h;
                floatx4 a0 = {0.0f, 0.0f, 0.0f, 0.0f}, a1 = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
                for (int k = 0; k < KC; ++k) {
                    a0 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(KEEPW ? wk[j0 * KC + k] : wfrag(j0, k), av[k], a0,
                                                                 0, 0, 0);
                    a1 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(KEEPW ? wk[j1 * KC + k] : wfrag(j1, k), av[k], a1,
                                                                 0, 0, 0);
                }
                const floatx4 bv = *(const floatx4*)(bias + h * 16 + grp * 4);
                a0 += bv;
                a1 += bv;
                const unsigned p0 = f2bf2(a0[0], a0[1]), p1 = f2bf2(a0[2], a0[3]);
                const unsigned q0 = f2bf2(a1[0], a1[1]), q1 = f2bf2(a1[2], a1[3]);
                // even lane groups pass on their dx = 1 channels, odd ones their dx = 0 channels
                const unsigned r0 = __shfl_xor(odd ? p0 : q0, 16, 64), r1 = __shfl_xor(odd ? p1 : q1, 16, 64);
                u32x4 o;
                o[0] = odd ? r0 : p0;
                o[1] = odd ? r1 : p1;
                o[2] = odd ? q0 : r0;
                o[3] = odd ? q1 : r1;
                const size_t u = u0 + ((s >> 2) * D2 + ((s >> 1) & 1)) * D2;
                *(u32x4*)(y + u * COUT + h * 16 + (grp & 2) * 4) = o;
            }
        }
    }
}

template <int CIN, int COUT>
__global__ __launch_bounds__(256) void k_up_bwd_mfma(const bf16_t* __restrict__ dy, const bf16_t* __restrict__ wk,
                                                     bf16_t* __restrict__ dx, int B, int D) {
    constexpr int K = 8 * COUT, KC = K / 32, NT = CIN / 16;
    static_assert(COUT % 4 == 0 && CIN % 16 == 0 && K % 32 == 0, "channel counts");
    const int lane = threadIdx.x & 63, row = lane & 15, grp = lane >> 4;
    const size_t nseg = (size_t)B * D * D * D / 16;
    const size_t seg = (size_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (seg >= nseg) return;  // whole wave
    const size_t v = seg * 16 + row;  // this lane's A row (input voxel)
    const int xx = v % D, yy = (v / D) % D, zz = (v / ((size_t)D * D)) % D;
    const int b = v / ((size_t)D * D * D);
    const int D2 = 2 * D;
    floatx4 acc[NT];
#pragma unroll
    for (int n = 0; n < NT; ++n) acc[n] = floatx4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll 2
    for (int kc = 0; kc < KC; ++kc) {
        const int k0 = kc * 32 + grp * 8;  // 8 consecutive (s, co) with one s
        const int s = k0 / COUT, co = k0 % COUT;
        const size_t u = (((size_t)b * D2 + 2 * zz + (s >> 2)) * D2 + 2 * yy + ((s >> 1) & 1)) * D2 + 2 * xx + (s & 1);
        const bf16x8 a = __builtin_bit_cast(bf16x8, *(const u32x4*)(dy + u * COUT + co));
#pragma unroll
        for (int n = 0; n < NT; ++n) {
            const bf16x8 bw = __builtin_bit_cast(bf16x8, *(const u32x4*)(wk + (size_t)(n * 16 + row) * K + k0));
            acc[n] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, bw, acc[n], 0, 0, 0);
        }
    }
#pragma unroll
    for (int n = 0; n < NT; ++n)
#pragma unroll
        for (int i = 0; i < 4; ++i) dx[(seg * 16 + grp * 4 + i) * CIN + n * 16 + row] = f2bf(acc[n][i]);
}

// bf16 copies of transposed-conv weights w [cout][8][cin]:
// wn[s*cout+co][ci] (forward) and wk[ci][s*cout+co] (input gradient).
__global__ void k_pack_up(const float* __restrict__ w, bf16_t* __restrict__ wn, bf16_t* __restrict__ wk, int cin,
                          int cout) {
    const int n = cout * 8 * cin;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int ci = i % cin, s = (i / cin) % 8, co = i / (8 * cin);
    const bf16_t h = f2bf(w[i]);
    const int k = s * cout + co;
    wn[(size_t)k * cin + ci] = h;
    wk[(size_t)ci * 8 * cout + k] = h;
}

// dW[co][s][ci] = sum_v dy[2v+s][co] x[v][ci]: block partials over voxel
// ranges; each thread owns 4 x 4 (co, ci) for all 8 s; sets split voxels.
// Matrix-core weight gradient of the transposed conv on
// v_mfma_f32_16x16x4f32: dW[co][s][ci] = sum_v dy[2v+s][co] x[v][ci].  K = 4
// voxels per product, one per lane group, so both operands come straight
// from the channels-last tensors (bf16 -> fp32 is exact): lane (row, grp)
// supplies dy[out(voxel grp, s)][16 cot + row] and x[voxel grp][16 cit + row].
// Grid: (blocks over x-lines, groups of SPW sub-positions); wave w takes the
// 4-voxel steps 4 w, 4 w + 16, ... of each line and keeps SPW x cout/16 x
// cin/16 accumulator tiles; the waves are added in a fixed order at the end.
template <int CIN, int COUT, int SPW>
__global__ __launch_bounds__(256) void k_up_wgrad_mfma(const bf16_t* __restrict__ x, const bf16_t* __restrict__ dy,
                                                       float* __restrict__ partial, int B, int D,
                                                       int lines_per_block) {
    constexpr int NTC = COUT / 16, NTI = CIN / 16;
    static_assert(CIN % 16 == 0 && COUT % 16 == 0 && 8 % SPW == 0, "shapes");
    __shared__ floatx4 red[3][64];
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int row = lane & 15, grp = lane >> 4;
    const int s0 = blockIdx.y * SPW;
    const size_t D2 = 2 * (size_t)D;
    floatx4 acc[SPW][NTC][NTI];
#pragma unroll
    for (int a = 0; a < SPW; ++a)
#pragma unroll
        for (int n = 0; n < NTC; ++n)
#pragma unroll
            for (int m = 0; m < NTI; ++m) acc[a][n][m] = floatx4{0.f, 0.f, 0.f, 0.f};
    const int nlines = B * D * D;
    const int L0 = blockIdx.x * lines_per_block;
    const int L1 = L0 + lines_per_block < nlines ? L0 + lines_per_block : nlines;
    for (int L = L0; L < L1; ++L) {
        const size_t yy = L % D, zz = (L / D) % D, b = L / (D * D);
        size_t ob[SPW];
#pragma unroll
        for (int a = 0; a < SPW; ++a) {
            const int s = s0 + a;
            ob[a] = ((b * D2 + 2 * zz + (s >> 2)) * D2 + 2 * yy + ((s >> 1) & 1)) * D2 + (s & 1);
        }
        for (int xb = 4 * wave; xb < D; xb += 16) {
            const int xx = xb + grp;
            float bx[NTI];
#pragma unroll
            for (int m = 0; m < NTI; ++m) bx[m] = bf2f(x[((size_t)L * D + xx) * CIN + m * 16 + row]);
#pragma unroll
            for (int a = 0; a < SPW; ++a) {
                const size_t u = ob[a] + 2 * (size_t)xx;
#pragma unroll
                for (int n = 0; n < NTC; ++n) {
                    const float av = bf2f(dy[u * COUT + n * 16 + row]);
#pragma unroll
                    for (int m = 0; m < NTI; ++m)
                        acc[a][n][m] = __builtin_amdgcn_mfma_f32_16x16x4f32(av, bx[m], acc[a][n][m], 0, 0, 0);
                }
            }
        }
    }
    float* out = partial + (size_t)blockIdx.x * COUT * 8 * CIN;
#pragma unroll
    for (int a = 0; a < SPW; ++a)
#pragma unroll
        for (int n = 0; n < NTC; ++n)
#pragma unroll
            for (int m = 0; m < NTI; ++m) {
                __syncthreads();
                if (wave > 0) red[wave - 1][lane] = acc[a][n][m];
                __syncthreads();
                if (wave == 0) {
                    floatx4 sum = acc[a][n][m];
#pragma unroll
                    for (int w = 0; w < 3; ++w) sum += red[w][lane];
#pragma unroll
                    for (int i = 0; i < 4; ++i)
                        out[((size_t)(n * 16 + grp * 4 + i) * 8 + s0 + a) * CIN + m * 16 + row] = sum[i];
                }
            }
}

template <int CIN, int COUT>
__global__ __launch_bounds__(256) void k_up_wgrad(const bf16_t* __restrict__ x, const bf16_t* __restrict__ dy,
                                                  float* __restrict__ partial, int B, int D, int vox_per_block) {
    constexpr int PAIRS = (COUT / 4) * (CIN / 4);
    constexpr int SETS = 256 / PAIRS;
    static_assert(PAIRS <= 256 && 256 % PAIRS == 0, "pairs per block");
    __shared__ float red[(SETS > 1 ? 256 : 1) * 16];
    const int t = threadIdx.x;
    const int pair = t % PAIRS, set = t / PAIRS;
    const int co0 = (pair % (COUT / 4)) * 4, ci0 = (pair / (COUT / 4)) * 4;
    const size_t nv = (size_t)B * D * D * D;
    const size_t v0 = (size_t)blockIdx.x * vox_per_block;
    const size_t v1 = v0 + vox_per_block < nv ? v0 + vox_per_block : nv;
    const int D2 = 2 * D;
    float* out = partial + (size_t)blockIdx.x * COUT * 8 * CIN;
    for (int s = 0; s < 8; ++s) {
        float acc[4][4];
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j) acc[i][j] = 0.0f;
        for (size_t v = v0 + set; v < v1; v += SETS) {
            const int xx = v % D, yy = (v / D) % D, zz = (v / ((size_t)D * D)) % D;
            const int b = v / ((size_t)D * D * D);
            const size_t u =
                (((size_t)b * D2 + 2 * zz + (s >> 2)) * D2 + 2 * yy + ((s >> 1) & 1)) * D2 + 2 * xx + (s & 1);
            // 4 consecutive bf16 = one 8-byte load each
            const uint2 gq = *(const uint2*)(dy + u * COUT + co0);
            const uint2 xq = *(const uint2*)(x + v * CIN + ci0);
            const float gv[4] = {__uint_as_float(gq.x << 16), __uint_as_float(gq.x & 0xffff0000u),
                                 __uint_as_float(gq.y << 16), __uint_as_float(gq.y & 0xffff0000u)};
            const float xv[4] = {__uint_as_float(xq.x << 16), __uint_as_float(xq.x & 0xffff0000u),
                                 __uint_as_float(xq.y << 16), __uint_as_float(xq.y & 0xffff0000u)};
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j) acc[i][j] = fmaf(gv[i], xv[j], acc[i][j]);
        }
        if (SETS > 1) {
            __syncthreads();
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j) red[t * 16 + i * 4 + j] = acc[i][j];
            __syncthreads();
            if (set == 0)
                for (int s2 = 1; s2 < SETS; ++s2)
#pragma unroll
                    for (int i = 0; i < 4; ++i)
#pragma unroll
                        for (int j = 0; j < 4; ++j) acc[i][j] += red[(s2 * PAIRS + pair) * 16 + i * 4 + j];
        }
        if (set == 0) {
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j) out[((size_t)(co0 + i) * 8 + s) * CIN + ci0 + j] = acc[i][j];
        }
    }
}

// ------------------------------------------------------------------ input layer (density + conditioning)

// beff[b][co] = bias[co] + cond[b][0] sum_t w[co][t][1] + cond[b][1] sum_t w[co][t][2]
__global__ void k_in_bias(const float* __restrict__ w, const float* __restrict__ bias, const float* __restrict__ cond,
                          float* __restrict__ beff, int B, int COUT) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= B * COUT) return;
    const int b = i / COUT, co = i % COUT;
    float s1 = 0.0f, s2 = 0.0f;
    for (int t = 0; t < 27; ++t) {
        s1 += w[(co * 27 + t) * 3 + 1];
        s2 += w[(co * 27 + t) * 3 + 2];
    }
    beff[i] = bias[co] + cond[b * 2 + 0] * s1 + cond[b * 2 + 1] * s2;
}

// y[v][co] = relu(beff[b][co] + sum_t w[co][t][0] * scale * delta[v + off_t])
// for the 16 output channels, on the fp32 matrix cores
// (v_mfma_f32_16x16x4f32): A = scaled weights (16 channels x 4 taps, lane
// (row, grp) holds channel row, tap 4 s + grp), B = densities (4 taps x the
// 16 voxels of a line), D = channels x voxels, so a lane ends with channels
// 4 grp .. +3 of voxel row.  Seven products cover the 27 taps (tap 27 has
// zero weight).  A block owns TZ x TY lines of 16 voxels with the density
// halo in LDS; each wave MW lines of one z-plane.
template <int TZ, int TY>
__global__ __launch_bounds__(256) void k_in_fwd(const float* __restrict__ delta, const float* __restrict__ w,
                                                const float* __restrict__ beff, bf16_t* __restrict__ y, int D,
                                                float scale) {
    constexpr int HZ = TZ + 2, HY = TY + 2, HX = 18, MW = TZ * TY / 4;
    static_assert(TY % MW == 0, "a wave's lines lie in one z-plane");
    __shared__ float tile[HZ * HY * HX];
    const int nyt = D / TY;
    const int x0 = blockIdx.x * 16;
    const int y0 = (blockIdx.y % nyt) * TY, z0 = (blockIdx.y / nyt) * TZ;
    const int b = blockIdx.z;
    const size_t vb = (size_t)b * D * D * D;
    const float* db = delta + vb;  // this box; 32-bit offsets inside it (D^3 < 2^31)
    // halo rows (hz, hy) of HX voxels: thread t loads column t % HX of rows
    // t / HX + RPI it (threads past RPI HX idle), every load in flight
    // before the LDS writes
    constexpr int NR = HZ * HY, RPI = 256 / HX, NIT = (NR + RPI - 1) / RPI;
    const int hx = threadIdx.x % HX, hr0 = threadIdx.x / HX;
    const unsigned gx = (unsigned)wrap(x0 + hx - 1, D);
    float hv[NIT];
#pragma unroll
    for (int it = 0; it < NIT; ++it) {
        const int r = hr0 + it * RPI;
        if (hr0 < RPI && r < NR) {
            const int hz = r / HY, hy = r - hz * HY;
            hv[it] = db[((unsigned)wrap(z0 + hz - 1, D) * D + (unsigned)wrap(y0 + hy - 1, D)) * D + gx];
        }
    }
#pragma unroll
    for (int it = 0; it < NIT; ++it) {
        const int r = hr0 + it * RPI;
        if (hr0 < RPI && r < NR) tile[r * HX + hx] = hv[it];
    }
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int row = lane & 15, grp = lane >> 4;
    float wa[7];
    int toff[7];
#pragma unroll
    for (int s = 0; s < 7; ++s) {
        const int t = 4 * s + grp;
        wa[s] = t < 27 ? w[(row * 27 + t) * 3] * scale : 0.0f;
        // tile offset of tap t: one of four compile-time constants
        constexpr auto off = [](int tp) { return tp < 27 ? tap_off<HY, HX, 1>(tp) : 0; };
        toff[s] = grp == 0 ? off(4 * s) : (grp == 1 ? off(4 * s + 1) : (grp == 2 ? off(4 * s + 2) : off(4 * s + 3)));
    }
    const floatx4 bias = *(const floatx4*)(beff + b * 16 + grp * 4);
    __syncthreads();
    const int mt0 = wave * MW, lz = mt0 / TY, ly0 = mt0 % TY;
#pragma unroll
    for (int m = 0; m < MW; ++m) {
        const float* lt = tile + (lz * HY + ly0 + m) * HX + row;
        floatx4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int s = 0; s < 7; ++s) acc = __builtin_amdgcn_mfma_f32_16x16x4f32(wa[s], lt[toff[s]], acc, 0, 0, 0);
        acc += bias;
        u32x2 q;
        q[0] = f2bf2(fmaxf(acc[0], 0.0f), fmaxf(acc[1], 0.0f));
        q[1] = f2bf2(fmaxf(acc[2], 0.0f), fmaxf(acc[3], 0.0f));
        const size_t vrow = vb + ((size_t)(z0 + lz) * D + (y0 + ly0 + m)) * D + x0;
        *(u32x2*)(y + vrow * 16 + row * 16 + grp * 4) = q;
    }
}

// Input-layer weight gradient of the density channel on the fp32 matrix
// cores (v_mfma_f32_16x16x4f32): dwd[co][t] = sum_v g'[v][co] scale
// delta[v + off(t)] (g' = g where the layer's output is positive).  K = 4
// voxels per product, one per lane group, so both operands come straight from
// memory: lane (row, grp) supplies g'[voxel grp][channel row] and the scaled
// density at voxel grp shifted by tap 16 j + row (two N tiles cover the 27
// taps; tap rows fetched per x-line).  Grid: blocks over x-lines; wave w takes
// the 4-voxel steps 4 w, 4 w + 16, ... of each line; the waves are added in a
// fixed order at the end.
__global__ __launch_bounds__(256) void k_in_wgrad_mfma(const float* __restrict__ delta, const bf16_t* __restrict__ g,
                                                       const bf16_t* __restrict__ y, float* __restrict__ partial, int B,
                                                       int D, float scale, int lines_per_block) {
    __shared__ floatx4 red[3][2][64];
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int row = lane & 15, grp = lane >> 4;
    int tdz[2], tdy[2], tdx[2];
    bool tok[2];
#pragma unroll
    for (int j = 0; j < 2; ++j) {
        const int t = 16 * j + row;
        tok[j] = t < 27;
        const int tt = tok[j] ? t : 0;
        tdz[j] = tt / 9 - 1;
        tdy[j] = (tt / 3) % 3 - 1;
        tdx[j] = tt % 3 - 1;
    }
    floatx4 acc[2] = {floatx4{0.f, 0.f, 0.f, 0.f}, floatx4{0.f, 0.f, 0.f, 0.f}};
    const int nlines = B * D * D;
    const int L0 = blockIdx.x * lines_per_block;
    const int L1 = L0 + lines_per_block < nlines ? L0 + lines_per_block : nlines;
    for (int L = L0; L < L1; ++L) {
        const int yy = L % D, zz = (L / D) % D, b = L / (D * D);
        const float* drow[2];
#pragma unroll
        for (int j = 0; j < 2; ++j)
            drow[j] = delta + (((size_t)b * D + wrap(zz + tdz[j], D)) * D + wrap(yy + tdy[j], D)) * D;
        const size_t lbase = (size_t)L * D;
        for (int xb = 4 * wave; xb < D; xb += 16) {
            const int xx = xb + grp;
            const size_t idx = (lbase + xx) * 16 + row;
            const float av = pos_bf(y[idx]) ? bf2f(g[idx]) : 0.0f;
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                const float bv = tok[j] ? scale * drow[j][wrap(xx + tdx[j], D)] : 0.0f;
                acc[j] = __builtin_amdgcn_mfma_f32_16x16x4f32(av, bv, acc[j], 0, 0, 0);
            }
        }
    }
    if (wave > 0) {
#pragma unroll
        for (int j = 0; j < 2; ++j) red[wave - 1][j][lane] = acc[j];
    }
    __syncthreads();
    if (wave == 0) {
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            floatx4 sum = acc[j];
#pragma unroll
            for (int w = 0; w < 3; ++w) sum += red[w][j][lane];
            const int tap = 16 * j + row;
            if (tap < 27) {
#pragma unroll
                for (int i = 0; i < 4; ++i) partial[(size_t)blockIdx.x * 16 * 27 + (grp * 4 + i) * 27 + tap] = sum[i];
            }
        }
    }
}

// Full input-layer weight gradient [co][t][3] and bias gradient from the
// density part (dwd[co][t]) and the per-sample gradient sums gsum[b][co]:
// the conditioning channels are constant, so their gradient is the same for
// every tap: sum_b cond[b][k] gsum[b][co].
__global__ void k_in_finish(const float* __restrict__ dwd, const float* __restrict__ gsum,
                            const float* __restrict__ cond, float* __restrict__ dw, float* __restrict__ db, int B,
                            int COUT) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= COUT * 27) return;
    const int co = i / 27;
    float s1 = 0.0f, s2 = 0.0f, s0 = 0.0f;
    for (int b = 0; b < B; ++b) {
        s0 += gsum[b * COUT + co];
        s1 += cond[b * 2 + 0] * gsum[b * COUT + co];
        s2 += cond[b * 2 + 1] * gsum[b * COUT + co];
    }
    dw[i * 3 + 0] = dwd[i];
    dw[i * 3 + 1] = s1;
    dw[i * 3 + 2] = s2;
    if (i % 27 == 0) db[co] = s0;
}

// ------------------------------------------------------------------ head: 1^3 conv, sigmoid, BCE

// logit = b + sum_c w[c] x[v][c]; p = sigmoid(logit).  With a target:
// loss += -(t log p + (1-t) log(1-p)) (FP64 partial per block),
// dlogit = (p - t) * inv_n, dx[v][c] = dlogit * w[c],
// partial dw[c] += dlogit x[v][c], db += dlogit.
template <int C>
__global__ __launch_bounds__(256) void k_head(const bf16_t* __restrict__ x, const float* __restrict__ w,
                                              const float* __restrict__ bias, const float* __restrict__ target,
                                              float* __restrict__ prob, bf16_t* __restrict__ dx,
                                              float* __restrict__ partial, double* __restrict__ lpart, size_t nv,
                                              float inv_n) {
    __shared__ float red[256 * (C + 1)];
    __shared__ double lred[256];
    const int t = threadIdx.x;
    float wv[C];
#pragma unroll
    for (int c = 0; c < C; ++c) wv[c] = w[c];
    const float b0 = bias[0];
    float gw[C + 1];
#pragma unroll
    for (int c = 0; c <= C; ++c) gw[c] = 0.0f;
    double loss = 0.0;
    for (size_t v = (size_t)blockIdx.x * 256 + t; v < nv; v += (size_t)gridDim.x * 256) {
        float xv[C];
        const u32x4* xp = (const u32x4*)(x + v * C);
#pragma unroll
        for (int q = 0; q < C / 8; ++q) {
            const u32x4 r = xp[q];
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                xv[q * 8 + 2 * k] = __uint_as_float(r[k] << 16);
                xv[q * 8 + 2 * k + 1] = __uint_as_float(r[k] & 0xffff0000u);
            }
        }
        static_assert(C == 16, "the shared summation order is defined for 16 channels");
        const float p0 = head_partial(wv, xv), p1 = head_partial(wv + 4, xv + 4);
        const float p2 = head_partial(wv + 8, xv + 8), p3 = head_partial(wv + 12, xv + 12);
        const float lg = ((p0 + p1) + (p2 + p3)) + b0;
        const float p = head_sigmoid(lg);
        if (prob) prob[v] = p;
        if (target) {
            const float tv = target[v];
            // log p = -softplus(-lg), log(1-p) = -softplus(lg), stable for large |lg|
            const double lp = -(double)(fmaxf(-lg, 0.0f) + log1pf(expf(-fabsf(lg))));
            const double lq = -(double)(fmaxf(lg, 0.0f) + log1pf(expf(-fabsf(lg))));
            loss -= (double)tv * lp + (1.0 - (double)tv) * lq;
            const float dl = (p - tv) * inv_n;
            if (dx) {
#pragma unroll
                for (int c = 0; c < C; ++c) dx[v * C + c] = f2bf(dl * wv[c]);
            }
#pragma unroll
            for (int c = 0; c < C; ++c) gw[c] = fmaf(dl, xv[c], gw[c]);
            gw[C] += dl;
        }
    }
    if (!target) return;
#pragma unroll
    for (int c = 0; c <= C; ++c) red[c * 256 + t] = gw[c];
    lred[t] = loss;
    __syncthreads();
    if (t <= C) {
        float s = 0.0f;
        for (int q = 0; q < 256; ++q) s += red[t * 256 + q];
        partial[(size_t)blockIdx.x * (C + 1) + t] = s;
    }
    if (t == 0) {
        double s = 0.0;
        for (int q = 0; q < 256; ++q) s += lred[q];
        lpart[blockIdx.x] = s;
    }
}

__global__ void k_sum_f64(const double* __restrict__ p, int n, double scale, double* __restrict__ out) {
    double s = 0.0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) s += p[i];
    s = block_sum(s);
    if (threadIdx.x == 0) *out = s * scale;
}

// ------------------------------------------------------------------ packing and AdamW

// wb[co][t][ci] = bf16(w[co][t][ci]);  wt[ci][t'][co] = bf16(w[co][26-t'][ci])
__global__ void k_pack(const float* __restrict__ w, bf16_t* __restrict__ wb, bf16_t* __restrict__ wt, int cin,
                       int cout) {
    const int n = cout * 27 * cin;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int ci = i % cin, t = (i / cin) % 27, co = i / (27 * cin);
    const bf16_t h = f2bf(w[i]);
    if (wb) wb[i] = h;
    if (wt) wt[((size_t)ci * 27 + (26 - t)) * cout + co] = h;
}

__global__ void k_adamw(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                        float* __restrict__ v, size_t n, float lr, float b1, float b2, float eps, float wd, float c1,
                        float c2) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const float gi = g[i];
        const float mi = b1 * m[i] + (1.0f - b1) * gi;
        const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
        m[i] = mi;
        v[i] = vi;
        const float mh = mi / c1, vh = vi / c2;
        p[i] = p[i] - lr * (mh / (sqrtf(vh) + eps) + wd * p[i]);
    }
}

// ------------------------------------------------------------------ dispatch

using ConvFn = void (*)(const ConvArgs&, hipStream_t);
template <int CIN, int COUT, int TZ, int TY, int CH = CIN>
void launch_conv(const ConvArgs& a, hipStream_t st) {
    const dim3 grid(a.D / 16, (a.D / TZ) * (a.D / TY), a.B);
    k_conv3_mfma<CIN, COUT, TZ, TY, CH><<<grid, 256, 0, st>>>(a);
}
template <int CIN, int COUT, int TY, int NZ, bool HEAD = false, bool POOL = false>
void launch_slide(const ConvArgs& a, hipStream_t st) {
    const dim3 grid(a.D / 16, (a.D / TY) * (a.D / NZ), a.B);
    k_conv3_slide<CIN, COUT, TY, NZ, HEAD, POOL><<<grid, 256, 0, st>>>(a);
}
// sliding form with the pool of its output fused (16 -> 16 channels, D >= 64)
ConvFn slide_pool(int cin, int cout, int D) {
    if (cin != 16 || cout != 16 || D < 64) return nullptr;
    return D >= 128 ? launch_slide<16, 16, 8, 16, false, true> : launch_slide<16, 16, 8, 8, false, true>;
}

// Tiled form: 4 x 4 x 16 (cin <= 32) or 4 x 2 x 16 (cin = 64) voxels per
// block for D >= 64, 2 x 2 x 16 below (enough blocks to fill the chip).
// Sliding form: 8 (cin <= 32) or 4 lines of 16 voxels over 16 planes
// (D >= 128) or 8 planes (D = 64).  Chunked tile: 16 input channels at a
// time, 4 x 8 x 16 voxels per block (2 x 8 x 16 for 64 output channels).
struct ConvShape {
    int cin, cout;
    ConvFn tile_big, tile_small, slide128, slide64, chunked;
    ConvFn slide128_head, slide64_head;  // sliding form with the fused head (cout = 16)
};

#define CM21_CONV(CI, CO)                                                                             \
    {CI, CO, launch_conv<CI, CO, 4, (CI >= 64 ? 2 : 4)>, launch_conv<CI, CO, 2, 2>,                    \
     launch_slide<CI, CO, (CI >= 64 ? 4 : 8), 16>, launch_slide<CI, CO, (CI >= 64 ? 4 : 8), 8>,       \
     launch_conv<CI, CO, (CO >= 64 ? 2 : 4), 8, 16>,                                                  \
     launch_slide<CI, CO, (CI >= 64 ? 4 : 8), 16, CO == 16>, launch_slide<CI, CO, (CI >= 64 ? 4 : 8), 8, CO == 16>}
const ConvShape kConvShapes[] = {
    CM21_CONV(16, 16), CM21_CONV(16, 32), CM21_CONV(32, 16), CM21_CONV(32, 32),
    CM21_CONV(32, 64), CM21_CONV(64, 32), CM21_CONV(64, 64),
};
#undef CM21_CONV

const ConvShape* conv_shape(int cin, int cout) {
    for (const auto& s : kConvShapes)
        if (s.cin == cin && s.cout == cout) return &s;
    return nullptr;
}

using WgradFn = void (*)(const WgradArgs&, int nblk, hipStream_t);
template <int CIN, int COUT>
void launch_wgrad(const WgradArgs& a, int nblk, hipStream_t st) {
    k_conv3_wgrad<CIN, COUT><<<dim3(nblk, 9), 256, 0, st>>>(a);
}
template <int CIN, int COUT>
void launch_wgrad_mfma(const WgradArgs& a, int nblk, hipStream_t st) {
    k_conv3_wgrad_mfma<CIN, COUT><<<dim3(nblk, 3 * (CIN / 16)), 256, 0, st>>>(a);
}
struct WgradShape {
    int cin, cout;
    WgradFn fn;    // fp32 VALU, voxel tiles 2 x 2 x 16 (any D)
    WgradFn mfma;  // matrix cores, voxel tiles 2 x 2 x 32 (D a multiple of 32, c0 a multiple of 16)
};
#define CM21_WGRAD(CI, CO) {CI, CO, launch_wgrad<CI, CO>, launch_wgrad_mfma<CI, CO>}
const WgradShape kWgradShapes[] = {
    CM21_WGRAD(16, 16), CM21_WGRAD(16, 32), CM21_WGRAD(32, 16), CM21_WGRAD(32, 32),
    CM21_WGRAD(32, 64), CM21_WGRAD(64, 32), CM21_WGRAD(64, 64),
};
#undef CM21_WGRAD
const WgradShape* wgrad_shape(int cin, int cout) {
    for (const auto& s : kWgradShapes)
        if (s.cin == cin && s.cout == cout) return &s;
    return nullptr;
}

constexpr int kWgradBlocks = 256;   // voxel-slab blocks of the weight gradient

constexpr int kChanBlocks = 128;    // per-sample blocks of the channel sums

int channel_sums(const void* g, const void* mask, int C, size_t vps, int B, float* cpart, float* total,
                 float* per_sample, hipStream_t st) {
    k_channel_partials<<<dim3(kChanBlocks, B), 256, 0, st>>>((const bf16_t*)g, (const bf16_t*)mask, C, vps, cpart);
    CM21_LAUNCH_CHECK("k_channel_partials");
    static_assert(kChanBlocks <= 128, "one thread per partial block");
    k_channel_final<<<C, 128, 0, st>>>(cpart, B, kChanBlocks, C, per_sample, total);
    CM21_LAUNCH_CHECK("k_channel_final");
    return CM21_OK;
}
constexpr int kHeadBlocks = 1024;
constexpr int kUpBlocksMax = 1024;  // line blocks of the transposed-conv weight gradient
constexpr int kInBlocks = 1024;     // line blocks of the input-layer weight gradient

bool valid_unet_dim(int D) { return D >= 16 && D <= 1024 && D % 16 == 0; }

// 2x2x2 mean pool, C a multiple of 8, B * D / 2 < 65536 (grid z)
int pool_fwd_launch(const bf16_t* x, bf16_t* y, int B, int D, int C, hipStream_t st) {
    const int Dh = D / 2;
    k_pool_fwd<<<dim3((Dh * (C / 8) + 63) / 64, Dh, B * Dh), 64, 0, st>>>(x, y, D, C);
    CM21_LAUNCH_CHECK("k_pool_fwd");
    return CM21_OK;
}

}  // namespace
}  // namespace cm21

using namespace cm21;

extern "C" int cm21_unet_conv3_fwd(const cm21_unet_conv* c, void* stream) {
    CM21_REQUIRE(c, "cm21_unet_conv3_fwd: NULL descriptor");
    const ConvShape* s = conv_shape(c->cin, c->cout);
    CM21_REQUIRE(s, "cm21_unet_conv3_fwd: no kernel for cin=%d cout=%d", c->cin, c->cout);
    CM21_REQUIRE(c->B >= 1 && valid_unet_dim(c->D), "cm21_unet_conv3_fwd: B=%d D=%d", c->B, c->D);
    CM21_REQUIRE(c->c0 > 0 && c->c0 <= c->cin && c->c0 % 8 == 0, "cm21_unet_conv3_fwd: c0=%d", c->c0);
    CM21_REQUIRE(c->o0 > 0 && c->o0 <= c->cout && c->o0 % 16 == 0, "cm21_unet_conv3_fwd: o0=%d", c->o0);
    CM21_REQUIRE(c->x0 && c->w && (c->y0 || c->head), "cm21_unet_conv3_fwd: NULL x0, w or y0");
    CM21_REQUIRE(!c->head || (c->cout == 16 && c->o0 == 16 && c->prob),
                 "cm21_unet_conv3_fwd: the fused head needs cout = 16, o0 = 16 and prob");
    CM21_REQUIRE(c->c0 == c->cin || c->x1, "cm21_unet_conv3_fwd: x1 is NULL with c0 < cin");
    CM21_REQUIRE(c->o0 == c->cout || c->y1, "cm21_unet_conv3_fwd: y1 is NULL with o0 < cout");
    CM21_REQUIRE(!c->mask || c->c0 == c->cin, "cm21_unet_conv3_fwd: mask needs a single input");
    CM21_REQUIRE(!c->pool || (!c->head && c->o0 == c->cout && (size_t)c->B * (c->D / 2) < 65536),
                 "cm21_unet_conv3_fwd: pool needs every output channel in y0 and no head");
    ConvArgs a{};
    a.B = c->B;
    a.D = c->D;
    a.c0 = c->c0;
    a.o0 = c->o0;
    a.x0 = (const bf16_t*)c->x0;
    a.x1 = (const bf16_t*)c->x1;
    a.mask = (const bf16_t*)c->mask;
    a.w = (const bf16_t*)c->w;
    a.bias = c->bias;
    a.bias_stride = c->bias_per_sample ? c->cout : 0;
    a.relu = c->relu;
    a.y0 = (bf16_t*)c->y0;
    a.y1 = (bf16_t*)c->y1;
    a.head = c->head;
    a.prob = c->prob;
    hipStream_t st = (hipStream_t)stream;
    CM21_REQUIRE(c->algo >= 0 && c->algo <= 3, "cm21_unet_conv3_fwd: algo=%d", c->algo);
    CM21_REQUIRE(c->algo != 2 || c->D >= 64, "cm21_unet_conv3_fwd: algo 2 needs D >= 64");
    // automatic choice from per-layer timings on the MI355X: the sliding
    // window wins for 16 -> 16 channels at D >= 128, the tile of 16 input
    // channels at a time elsewhere (up to 3x faster than the whole-channel
    // tile for 64 input channels, whose weights would otherwise be fetched
    // per pair of lines)
    // (the fused head goes on the chunked tile, whose occupancy is set by
    // its LDS, so the head's epilogue registers cost nothing)
    const int algo =
        c->algo ? c->algo : ((c->cin == 16 && c->cout == 16 && c->D >= 128 && !c->head) ? 2 : 3);
    ConvFn fn = algo == 3 ? s->chunked
                          : (c->D < 64 ? s->tile_small
                                       : (algo == 1 ? s->tile_big : (c->D >= 128 ? s->slide128 : s->slide64)));
    if (c->head && algo == 2) fn = c->D >= 128 ? s->slide128_head : s->slide64_head;
    // the pool is fused into the sliding form of 16 -> 16 channels; the other
    // forms pool the stored output afterwards (the same sums, bit for bit)
    const ConvFn fused_pool = c->pool && algo == 2 ? slide_pool(c->cin, c->cout, c->D) : nullptr;
    if (fused_pool) {
        a.pool = (bf16_t*)c->pool;
        fn = fused_pool;
    }
    fn(a, st);
    CM21_LAUNCH_CHECK("k_conv3_mfma");
    if (c->pool && !fused_pool) return pool_fwd_launch(a.y0, (bf16_t*)c->pool, c->B, c->D, c->cout, st);
    return CM21_OK;
}

extern "C" size_t cm21_unet_wgrad_workspace_bytes(int B, int D, int cin, int cout) {
    if (B < 1 || !valid_unet_dim(D) || cin < 1 || cout < 1) return 0;
    const size_t w = (size_t)kWgradBlocks * cout * 27 * cin * sizeof(float);
    const size_t c = (size_t)B * kChanBlocks * cout * sizeof(float);
    return align_up(w, 256) + align_up(c, 256);
}

extern "C" int cm21_unet_conv3_wgrad(const cm21_unet_wgrad* w, void* stream) {
    CM21_REQUIRE(w, "cm21_unet_conv3_wgrad: NULL descriptor");
    const WgradShape* s = wgrad_shape(w->cin, w->cout);
    CM21_REQUIRE(s, "cm21_unet_conv3_wgrad: no kernel for cin=%d cout=%d", w->cin, w->cout);
    CM21_REQUIRE(w->B >= 1 && valid_unet_dim(w->D), "cm21_unet_conv3_wgrad: B=%d D=%d", w->B, w->D);
    CM21_REQUIRE(w->c0 > 0 && w->c0 <= w->cin && w->c0 % 8 == 0, "cm21_unet_conv3_wgrad: c0=%d", w->c0);
    CM21_REQUIRE(w->x0 && w->g && w->dw && w->work, "cm21_unet_conv3_wgrad: NULL argument");
    CM21_REQUIRE(w->c0 == w->cin || w->x1, "cm21_unet_conv3_wgrad: x1 is NULL with c0 < cin");
    const size_t need = cm21_unet_wgrad_workspace_bytes(w->B, w->D, w->cin, w->cout);
    if (w->work_bytes < need)
        return fail(CM21_ERR_NOMEM, "cm21_unet_conv3_wgrad: workspace %zu < %zu", w->work_bytes, need);
    hipStream_t st = (hipStream_t)stream;
    float* partial = (float*)w->work;
    float* cpart = (float*)((char*)w->work + align_up((size_t)kWgradBlocks * w->cout * 27 * w->cin * sizeof(float), 256));
    // the matrix-core form needs x-lines of 32 voxels and whole 16-channel
    // chunks per input tensor; the VALU form takes the rest
    const bool mfma = w->D % 32 == 0 && w->c0 % 16 == 0;
    const int ntiles = w->B * (w->D / 2) * (w->D / 2) * (w->D / (mfma ? 32 : 16));
    const int tpb = (ntiles + kWgradBlocks - 1) / kWgradBlocks;
    const int nblk = (ntiles + tpb - 1) / tpb;
    WgradArgs a{};
    a.B = w->B;
    a.D = w->D;
    a.c0 = w->c0;
    a.x0 = (const bf16_t*)w->x0;
    a.x1 = (const bf16_t*)w->x1;
    a.g = (const bf16_t*)w->g;
    a.mask = (const bf16_t*)w->mask;
    a.partial = partial;
    a.tiles_per_block = tpb;
    (mfma ? s->mfma : s->fn)(a, nblk, st);
    CM21_LAUNCH_CHECK("k_conv3_wgrad");
    const int n = w->cout * 27 * w->cin;
    k_sum_blocks<<<sum_blocks_grid(n), 256, 0, st>>>(partial, nblk, n, w->dw);
    CM21_LAUNCH_CHECK("k_sum_blocks");
    if (w->db) CM21_TRY(channel_sums(w->g, w->mask, w->cout, (size_t)w->D * w->D * w->D, w->B, cpart, w->db, nullptr, st));
    return CM21_OK;
}

extern "C" int cm21_unet_pool_fwd(const void* x, void* y, int B, int D, int C, void* stream) {
    CM21_REQUIRE(x && y && B >= 1 && D >= 2 && D % 2 == 0 && C >= 8 && C % 8 == 0 && (size_t)B * (D / 2) < 65536,
                 "cm21_unet_pool_fwd: bad arguments");
    return pool_fwd_launch((const bf16_t*)x, (bf16_t*)y, B, D, C, (hipStream_t)stream);
}

extern "C" int cm21_unet_pool_bwd(const void* dy, void* dx, int B, int D, int C, int accumulate, void* stream) {
    CM21_REQUIRE(dy && dx && B >= 1 && D >= 2 && D % 2 == 0 && C >= 1, "cm21_unet_pool_bwd: bad arguments");
    const size_t n = (size_t)B * D * D * D * C;
    k_pool_bwd<<<grid_for(n), kThreads, 0, (hipStream_t)stream>>>((const bf16_t*)dy, (bf16_t*)dx, B, D, C,
                                                                    accumulate ? 1 : 0);
    CM21_LAUNCH_CHECK("k_pool_bwd");
    return CM21_OK;
}

namespace {
template <int CIN, int COUT>
int up_fwd_t(const void* x, const void* wn, const float* b, void* y, int B, int D, hipStream_t st) {
    const size_t nseg = (size_t)B * D * D * D / 16;
    k_up_fwd_mfma<CIN, COUT><<<(unsigned)((nseg + 4 * kUpSegs - 1) / (4 * kUpSegs)), 256, 0, st>>>(
        (const bf16_t*)x, (const bf16_t*)wn, b, (bf16_t*)y, B, D);
    CM21_LAUNCH_CHECK("k_up_fwd_mfma");
    return CM21_OK;
}
template <int CIN, int COUT>
int up_bwd_t(const void* x, const void* dy, const void* wk, void* dx, float* dw, float* db, int B, int D, float* part,
             float* cpart, hipStream_t st) {
    const size_t nv = (size_t)B * D * D * D;
    if (dx) {
        k_up_bwd_mfma<CIN, COUT><<<(unsigned)((nv / 16 + 3) / 4), 256, 0, st>>>((const bf16_t*)dy, (const bf16_t*)wk,
                                                                               (bf16_t*)dx, B, D);
        CM21_LAUNCH_CHECK("k_up_bwd_mfma");
    }
    int nblk;
    if (D % 4 == 0) {
        // matrix cores; SPW sub-positions per block keep 64 accumulator registers
        constexpr int SPW = (COUT / 16) * (CIN / 16) <= 2 ? 8 : 2;
        const int nlines = B * D * D;
        const int lpb = (nlines + kUpBlocksMax - 1) / kUpBlocksMax;
        nblk = (nlines + lpb - 1) / lpb;
        k_up_wgrad_mfma<CIN, COUT, SPW><<<dim3(nblk, 8 / SPW), 256, 0, st>>>((const bf16_t*)x, (const bf16_t*)dy,
                                                                             part, B, D, lpb);
    } else {
        const int vpb = (int)((nv + kUpBlocksMax - 1) / kUpBlocksMax);
        nblk = (int)((nv + vpb - 1) / vpb);
        k_up_wgrad<CIN, COUT><<<nblk, 256, 0, st>>>((const bf16_t*)x, (const bf16_t*)dy, part, B, D, vpb);
    }
    CM21_LAUNCH_CHECK("k_up_wgrad");
    const int n = COUT * 8 * CIN;
    k_sum_blocks<<<sum_blocks_grid(n), 256, 0, st>>>(part, nblk, n, dw);
    CM21_LAUNCH_CHECK("k_sum_blocks");
    const size_t vps = (size_t)8 * D * D * D;
    CM21_TRY(channel_sums(dy, nullptr, COUT, vps, B, cpart, db, nullptr, st));
    return CM21_OK;
}
}  // namespace

extern "C" int cm21_unet_up_fwd(const void* x, const void* wn, const float* bias, void* y, int B, int D, int cin,
                                int cout, void* stream) {
    CM21_REQUIRE(x && wn && bias && y && B >= 1 && D >= 2 && ((size_t)B * D * D * D) % 16 == 0 &&
                     (size_t)B * D * D * D < ((size_t)1 << 32),
                 "cm21_unet_up_fwd: bad arguments");
    hipStream_t st = (hipStream_t)stream;
    if (cin == 64 && cout == 32) return up_fwd_t<64, 32>(x, wn, bias, y, B, D, st);
    if (cin == 32 && cout == 16) return up_fwd_t<32, 16>(x, wn, bias, y, B, D, st);
    return fail(CM21_ERR_INVALID, "cm21_unet_up_fwd: no kernel for cin=%d cout=%d", cin, cout);
}

extern "C" size_t cm21_unet_up_workspace_bytes(int B, int D, int cin, int cout) {
    if (B < 1 || D < 2 || cin < 1 || cout < 1) return 0;
    return align_up((size_t)kUpBlocksMax * cout * 8 * cin * sizeof(float), 256) +
           align_up((size_t)B * kChanBlocks * cout * sizeof(float), 256);
}

extern "C" int cm21_unet_up_bwd(const void* x, const void* dy, const void* wk, void* dx, float* dw, float* db, int B,
                                int D, int cin, int cout, void* work, size_t work_bytes, void* stream) {
    CM21_REQUIRE(x && dy && dw && db && work && B >= 1 && D >= 2 && ((size_t)B * D * D * D) % 16 == 0,
                 "cm21_unet_up_bwd: bad arguments");
    CM21_REQUIRE(!dx || wk, "cm21_unet_up_bwd: dx needs the packed weights wk");
    const size_t need = cm21_unet_up_workspace_bytes(B, D, cin, cout);
    if (work_bytes < need) return fail(CM21_ERR_NOMEM, "cm21_unet_up_bwd: workspace %zu < %zu", work_bytes, need);
    float* part = (float*)work;
    float* cpart = (float*)((char*)work + align_up((size_t)kUpBlocksMax * cout * 8 * cin * sizeof(float), 256));
    hipStream_t st = (hipStream_t)stream;
    if (cin == 64 && cout == 32) return up_bwd_t<64, 32>(x, dy, wk, dx, dw, db, B, D, part, cpart, st);
    if (cin == 32 && cout == 16) return up_bwd_t<32, 16>(x, dy, wk, dx, dw, db, B, D, part, cpart, st);
    return fail(CM21_ERR_INVALID, "cm21_unet_up_bwd: no kernel for cin=%d cout=%d", cin, cout);
}

extern "C" int cm21_unet_pack_up(const float* w, void* wn, void* wk, int cin, int cout, void* stream) {
    CM21_REQUIRE(w && wn && wk && cin >= 1 && cout >= 1, "cm21_unet_pack_up: bad arguments");
    const int n = cout * 8 * cin;
    k_pack_up<<<(n + 255) / 256, 256, 0, (hipStream_t)stream>>>(w, (bf16_t*)wn, (bf16_t*)wk, cin, cout);
    CM21_LAUNCH_CHECK("k_pack_up");
    return CM21_OK;
}

extern "C" int cm21_unet_in_fwd(const float* delta, const float* cond, const float* w, const float* bias, float* beff,
                                void* y, int B, int D, float delta_scale, void* stream) {
    CM21_REQUIRE(delta && cond && w && bias && beff && y && B >= 1 && valid_unet_dim(D),
                 "cm21_unet_in_fwd: bad arguments");
    hipStream_t st = (hipStream_t)stream;
    k_in_bias<<<(B * 16 + 63) / 64, 64, 0, st>>>(w, bias, cond, beff, B, 16);
    CM21_LAUNCH_CHECK("k_in_bias");
    k_in_fwd<4, 8><<<dim3(D / 16, (D / 4) * (D / 8), B), 256, 0, st>>>(delta, w, beff, (bf16_t*)y, D, delta_scale);
    CM21_LAUNCH_CHECK("k_in_fwd");
    return CM21_OK;
}

extern "C" size_t cm21_unet_in_workspace_bytes(int B, int D) {
    if (B < 1 || !valid_unet_dim(D)) return 0;
    return align_up((size_t)kInBlocks * 16 * 27 * sizeof(float), 256) +
           align_up((size_t)B * kChanBlocks * 16 * sizeof(float), 256) + align_up((size_t)B * 16 * sizeof(float), 256) +
           align_up((size_t)16 * 27 * sizeof(float), 256);
}

extern "C" int cm21_unet_in_wgrad(const float* delta, const float* cond, const void* g, const void* y, float* dw,
                                  float* db, int B, int D, float delta_scale, void* work, size_t work_bytes,
                                  void* stream) {
    CM21_REQUIRE(delta && cond && g && y && dw && db && work && B >= 1 && valid_unet_dim(D),
                 "cm21_unet_in_wgrad: bad arguments");
    const size_t need = cm21_unet_in_workspace_bytes(B, D);
    if (work_bytes < need) return fail(CM21_ERR_NOMEM, "cm21_unet_in_wgrad: workspace %zu < %zu", work_bytes, need);
    hipStream_t st = (hipStream_t)stream;
    char* p = (char*)work;
    float* part = (float*)p;
    p += align_up((size_t)kInBlocks * 16 * 27 * sizeof(float), 256);
    float* cpart = (float*)p;
    p += align_up((size_t)B * kChanBlocks * 16 * sizeof(float), 256);
    float* gsum = (float*)p;
    p += align_up((size_t)B * 16 * sizeof(float), 256);
    float* dwd = (float*)p;
    const int nlines = B * D * D;
    const int lpb = (nlines + kInBlocks - 1) / kInBlocks;
    const int nblk = (nlines + lpb - 1) / lpb;
    k_in_wgrad_mfma<<<nblk, 256, 0, st>>>(delta, (const bf16_t*)g, (const bf16_t*)y, part, B, D, delta_scale, lpb);
    CM21_LAUNCH_CHECK("k_in_wgrad_mfma");
    k_sum_blocks<<<sum_blocks_grid(16 * 27), 256, 0, st>>>(part, nblk, 16 * 27, dwd);
    CM21_LAUNCH_CHECK("k_sum_blocks");
    CM21_TRY(channel_sums(g, y, 16, (size_t)D * D * D, B, cpart, nullptr, gsum, st));
    k_in_finish<<<(16 * 27 + 255) / 256, 256, 0, st>>>(dwd, gsum, cond, dw, db, B, 16);
    CM21_LAUNCH_CHECK("k_in_finish");
    return CM21_OK;
}

extern "C" size_t cm21_unet_head_workspace_bytes(void) {
    return align_up((size_t)kHeadBlocks * 17 * sizeof(float), 256) + align_up((size_t)kHeadBlocks * sizeof(double), 256);
}

extern "C" int cm21_unet_head(const void* x, const float* w, const float* bias, const float* target, float* prob,
                              void* dx, float* dwb, double* loss, int B, int D, void* work, size_t work_bytes,
                              void* stream) {
    CM21_REQUIRE(x && w && bias && B >= 1 && valid_unet_dim(D), "cm21_unet_head: bad arguments");
    CM21_REQUIRE(!target || (dwb && loss && work), "cm21_unet_head: a target needs dwb, loss and work");
    if (target && work_bytes < cm21_unet_head_workspace_bytes())
        return fail(CM21_ERR_NOMEM, "cm21_unet_head: workspace too small");
    hipStream_t st = (hipStream_t)stream;
    const size_t nv = (size_t)B * D * D * D;
    float* part = (float*)work;
    double* lpart = work ? (double*)((char*)work + align_up((size_t)kHeadBlocks * 17 * sizeof(float), 256)) : nullptr;
    k_head<16><<<kHeadBlocks, 256, 0, st>>>((const bf16_t*)x, w, bias, target, prob, (bf16_t*)dx, part, lpart, nv,
                                            (float)(1.0 / (double)nv));
    CM21_LAUNCH_CHECK("k_head");
    if (target) {
        k_sum_blocks<<<sum_blocks_grid(17), 256, 0, st>>>(part, kHeadBlocks, 17, dwb);
        CM21_LAUNCH_CHECK("k_sum_blocks");
        k_sum_f64<<<1, 256, 0, st>>>(lpart, kHeadBlocks, 1.0 / (double)nv, loss);
        CM21_LAUNCH_CHECK("k_sum_f64");
    }
    return CM21_OK;
}

extern "C" int cm21_unet_pack(const float* w, void* wb, void* wt, int cin, int cout, void* stream) {
    CM21_REQUIRE(w && (wb || wt) && cin >= 1 && cout >= 1, "cm21_unet_pack: bad arguments");
    const int n = cout * 27 * cin;
    k_pack<<<(n + 255) / 256, 256, 0, (hipStream_t)stream>>>(w, (bf16_t*)wb, (bf16_t*)wt, cin, cout);
    CM21_LAUNCH_CHECK("k_pack");
    return CM21_OK;
}

extern "C" int cm21_adamw(float* p, const float* g, float* m, float* v, size_t n, float lr, float beta1, float beta2,
                          float eps, float weight_decay, int step, void* stream) {
    CM21_REQUIRE(p && g && m && v && n > 0 && step >= 1, "cm21_adamw: bad arguments");
    const float c1 = (float)(1.0 - std::pow((double)beta1, (double)step));
    const float c2 = (float)(1.0 - std::pow((double)beta2, (double)step));
    k_adamw<<<grid_for(n), kThreads, 0, (hipStream_t)stream>>>(p, g, m, v, n, lr, beta1, beta2, eps, weight_decay, c1,
                                                               c2);
    CM21_LAUNCH_CHECK("k_adamw");
    return CM21_OK;
}
