// Initial conditions on the GPU: the Gaussian random field written directly
// as a Hermitian r2c half spectrum, the nearest-pick subsample and the
// Zel'dovich displacement ("velocity") spectra.
//
// Reference: Programs/initialize.cu:58-95 (init_kernel), :133-146 (subsample),
// :148-190 (set_velocity); Programs/initialize.py:33-111 (driver).
#include <cmath>

#include "cm21_internal.hpp"
#include "cm21_window.hpp"

using namespace cm21;

namespace {

// Philox4x32-10 (Salmon et al., SC'11): ten rounds of two 32x32->64-bit
// multiplies; the key is bumped by the Weyl constants before rounds 2..10.
__device__ inline uint4 philox4x32_10(uint4 c, uint2 k) {
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        if (r) {
            k.x += 0x9E3779B9u;
            k.y += 0xBB67AE85u;
        }
        const uint32_t hi0 = __umulhi(0xD2511F53u, c.x), lo0 = 0xD2511F53u * c.x;
        const uint32_t hi1 = __umulhi(0xCD9E8D57u, c.z), lo1 = 0xCD9E8D57u * c.z;
        c = make_uint4(hi1 ^ c.y ^ k.x, lo1, hi0 ^ c.w ^ k.y, lo0);
    }
    return c;
}

// Box-Muller in float32, r = sqrt(-2 ln a), theta = 2 pi b, with the
// uniforms a = ((u0 >> 8) + 0.5) 2^-24 and b likewise, handled so that no
// argument is rounded: a = m 2^-25 with m = 2 (u0 >> 8) + 1 is an exact float
// when a <= 1/2, and 1 - a is one otherwise (ln a = log1p(-(1 - a))); 2b =
// m' 2^-24 is exact below 1, and 2b - 1 is exact above (sin and cos change
// sign), so sincospi sees an exact argument.  Each result is within a few
// float32 ulp of the FP64 values.
__device__ inline void box_muller(uint32_t u0, uint32_t u1, float& g1, float& g2) {
    const uint32_t m0 = 2u * (u0 >> 8) + 1u;
    const float lna = m0 < (1u << 24) ? logf(ldexpf((float)m0, -25))
                                      : log1pf(-ldexpf((float)((1u << 25) - m0), -25));
    const float r = sqrtf(-2.0f * lna);
    const uint32_t m1 = 2u * (u1 >> 8) + 1u;
    float s, c;
    if (m1 < (1u << 24)) {
        sincospif(ldexpf((float)m1, -24), &s, &c);
    } else {
        sincospif(ldexpf((float)(m1 - (1u << 24)), -24), &s, &c);
        s = -s;
        c = -c;
    }
    g1 = r * c;
    g2 = r * s;
}

struct GrfParams {
    float dkf;              // 2 pi / box_len
    float amp_over_volume;  // P(k) normalisation / box_len^3
    float n_indexf;         // spectral slope
    float inv_k_cut;        // 1 / Gaussian cut-off scale
    float inv_volume;       // 1 / box_len^3
    uint32_t seed, box_id;
    const float* ktab;  // tabulated P(k): increasing K (NULL = power law)
    const float* ptab;
    int ntab;
};

constexpr int kMaxTable = 2048;

// Where to start looking for k in a table: the index of k if the table were
// spaced uniformly in log k between its end points (the P(k) tables are),
// from log2 K[0] and (n-1) / log2(K[n-1]/K[0]); any table works, the search
// then only walks further.
struct TableGuess {
    float l0, inv;
};

__device__ inline TableGuess table_guess(const float* K, int n) {
    TableGuess g{0.0f, 0.0f};
    if (K[0] > 0.0f && K[n - 1] > K[0]) {
        g.l0 = __log2f(K[0]);
        g.inv = (float)(n - 1) / (__log2f(K[n - 1]) - g.l0);
    }
    return g;
}

// P(k) from the table by the rule of initialize.cu:87-90: i is the first
// index with K[i] >= k and P is interpolated linearly between i-1 and i in
// float32, as init_kernel does; modes outside [K[0], K[n-1]] get 0.  The
// index is found from the log-uniform guess by stepping down while K[i-1] >= k
// and up while K[i] < k: the same first index as a bisection, in one or two
// steps on a log-spaced table.
__device__ inline float table_pk(float k, const float* K, const float* Pt, int n, const TableGuess& tg) {
    if (!(k >= K[0]) || k > K[n - 1]) return 0.0f;
    const float x = (__log2f(k) - tg.l0) * tg.inv + 1.0f;
    int hi = (int)fminf(fmaxf(x, 1.0f), (float)(n - 1));  // in [1, n-1]
    while (hi > 1 && K[hi - 1] >= k) --hi;
    while (K[hi] < k) ++hi;  // stops at n-1 at the latest: K[n-1] >= k
    const int lo = hi - 1;
    if (K[lo] >= k) return Pt[lo];  // k == K[0]
    return Pt[lo] + (Pt[hi] - Pt[lo]) * ((k - K[lo]) / (K[hi] - K[lo]));
}

// One mode: its Philox draw, amplitude and Hermitian placement.
__device__ __forceinline__ void grf_mode(float2* __restrict__ spec, int D, int a, int b, int c, const GrfParams& gp,
                                         const float* sK, const float* sP, const TableGuess& tg) {
    const int H = D / 2 + 1;
    const size_t idx = ((size_t)a * D + b) * H + c;
    if (a == 0 && b == 0 && c == 0) {
        spec[idx] = make_float2(0.0f, 0.0f);
        return;
    }
    // Modes on the c = 0 and c = D/2 planes pair with (-a, -b) in the same
    // plane; the pair shares the random numbers of its lower linear index.
    const bool edge = (c == 0) || (c == D / 2);
    const int ma = (D - a) % D, mb = (D - b) % D;
    const long lin_self = (long)a * D + b, lin_mirror = (long)ma * D + mb;
    const bool use_mirror = edge && lin_mirror < lin_self;
    const bool selfconj = edge && ma == a && mb == b;
    const long ra = use_mirror ? ma : a, rb = use_mirror ? mb : b;
    const unsigned long long rep = ((unsigned long long)(ra * D + rb)) * (unsigned long long)H + (unsigned long long)c;
    const uint4 u = philox4x32_10(make_uint4((uint32_t)rep, (uint32_t)(rep >> 32), 0u, 0u), make_uint2(gp.seed, gp.box_id));
    float g1, g2;
    box_muller(u.x, u.y, g1, g2);

    // |n|^2 is an exact integer; k and the amplitude in float32
    const int na = mode_of(a, D), nb = mode_of(b, D);
    const float k = gp.dkf * sqrtf((float)(na * na + nb * nb + c * c));
    float amp;
    if (gp.ntab > 0) {
        amp = sqrtf(table_pk(k, sK, sP, gp.ntab, tg) * gp.inv_volume);
    } else {
        const float x = k * gp.inv_k_cut;
        // k^n as exp2(n log2 k) (k > 0 here): about 1e-6 relative, a third of
        // the instructions of powf
        amp = sqrtf(gp.amp_over_volume * (exp2f(gp.n_indexf * log2f(k)) * expf(-(x * x))));
    }
    if (selfconj) {
        spec[idx] = make_float2(amp * g1, 0.0f);
    } else {
        const float h = amp * 0.70710678118654752f;
        spec[idx] = make_float2(h * g1, use_mirror ? -(h * g2) : h * g2);
    }
}

// Blocks of 64 x 4 threads; block (0, y, z) covers a = z and the kGrfRows
// rows b from y * kGrfRows, its 256 threads walking the rows' modes as one
// run (no lane idles at the end of each row of D/2+1 modes), so the P(k)
// table is loaded into LDS once per kGrfRows * H modes.
constexpr int kGrfRows = 16;

__global__ __launch_bounds__(kModeBlockX* kModeBlockY) void k_grf(float2* __restrict__ spec, int D, GrfParams gp) {
    __shared__ float sK[kMaxTable], sP[kMaxTable];
    if (gp.ntab > 0) {
        const int t = threadIdx.y * kModeBlockX + threadIdx.x;
        for (int i = t; i < gp.ntab; i += kModeBlockX * kModeBlockY) {
            sK[i] = gp.ktab[i];
            sP[i] = gp.ptab[i];
        }
        __syncthreads();
    }
    const TableGuess tg = gp.ntab > 0 ? table_guess(sK, gp.ntab) : TableGuess{0.0f, 0.0f};
    const int H = D / 2 + 1;
    const int a = blockIdx.z;
    const int b0 = blockIdx.y * kGrfRows;
    const int n = (min(D, b0 + kGrfRows) - b0) * H;  // the block's modes, rows b0.. of plane a
    for (int q = threadIdx.y * kModeBlockX + threadIdx.x; q < n; q += kModeBlockX * kModeBlockY) {
        const int rb = q / H;
        grf_mode(spec, D, a, b0 + rb, q - rb * H, gp, sK, sP, tg);
    }
}

__global__ void k_subsample(float* __restrict__ small, const float* __restrict__ large, int D, int sw, float pf) {
    const int k = blockIdx.x * kModeBlockX + threadIdx.x, j = blockIdx.y * kModeBlockY + threadIdx.y;
    const int i = blockIdx.z;
    if (k >= sw || j >= sw) return;
    const size_t p = ((size_t)i * sw + j) * sw + k;
    // floor(k*pixel_factor + 0.5): product in float, + 0.5 in double.
    const int lk = (int)floor((double)__fmul_rn((float)k, pf) + 0.5);
    const int lj = (int)floor((double)__fmul_rn((float)j, pf) + 0.5);
    const int li = (int)floor((double)__fmul_rn((float)i, pf) + 0.5);
    if (li >= D || lj >= D || lk >= D) return;  // never taken for pf = D/sw
    small[p] = large[((size_t)li * D + lj) * D + lk];
}

__global__ void k_velocity(float2* __restrict__ out, const float2* __restrict__ in, int D, float dk, int comp,
                           int filter_type, float R) {
    const int H = D / 2 + 1;
    const int c = blockIdx.x * kModeBlockX + threadIdx.x, b = blockIdx.y * kModeBlockY + threadIdx.y;
    const int a = blockIdx.z;
    if (c >= H || b >= D) return;
    const size_t idx = ((size_t)a * D + b) * H + c;
    const float k0 = __fmul_rn((float)mode_of(a, D), dk);
    const float k1 = __fmul_rn((float)mode_of(b, D), dk);
    const float k2 = __fmul_rn((float)c, dk);
    // initialize.cu:162: k_sq = k_x^2 + k_y^2 + k_z^2 with x the slowest axis.
    const float ksq = __fadd_rn(__fadd_rn(__fmul_rn(k0, k0), __fmul_rn(k1, k1)), __fmul_rn(k2, k2));
    const int idx_c = comp == 0 ? a : (comp == 1 ? b : c);
    const float kc = comp == 0 ? k0 : (comp == 1 ? k1 : k2);
    if (ksq == 0.0f || idx_c == D / 2) {
        out[idx] = make_float2(0.0f, 0.0f);
        return;
    }
    const float f = __fdiv_rn(kc, ksq);
    const float2 v = in[idx];
    float2 r = make_float2(-__fmul_rn(f, v.y), __fmul_rn(f, v.x));  // (i f) * v
    if (filter_type >= 0) {
        WindowParams wp{dk, R, filter_type, 1, 1};
        float w0, w2;
        bool cut;
        window_of(sqrt_rn(ksq), wp, w0, cut, w2);
        r = apply_window(r, w0, cut, w2, 1.0f);
    }
    out[idx] = r;
}

bool valid_dim(int D) { return D >= 2 && D <= 2048 && D % 2 == 0; }

dim3 grf_grid(int D) { return dim3(1u, (unsigned)((D + kGrfRows - 1) / kGrfRows), (unsigned)D); }

}  // namespace

extern "C" int cm21_grf_init(void* spec_out, int D, double box_len, double amp, double n_index, double k_cut,
                             uint32_t seed, uint32_t box_id, void* stream) {
    CM21_REQUIRE(spec_out, "cm21_grf_init: NULL spectrum");
    CM21_REQUIRE(valid_dim(D), "cm21_grf_init: D=%d must be even in [2,2048]", D);
    CM21_REQUIRE(box_len > 0.0 && k_cut > 0.0 && amp >= 0.0, "cm21_grf_init: box_len, k_cut must be > 0, amp >= 0");
    const double vol = box_len * box_len * box_len;
    GrfParams gp{(float)(2.0 * M_PI / box_len), (float)(amp / vol), (float)n_index, (float)(1.0 / k_cut),
                 (float)(1.0 / vol), seed, box_id, nullptr, nullptr, 0};
    k_grf<<<grf_grid(D), mode_block(), 0, (hipStream_t)stream>>>((float2*)spec_out, D, gp);
    CM21_LAUNCH_CHECK("k_grf");
    return CM21_OK;
}

extern "C" int cm21_grf_init_table(void* spec_out, int D, double box_len, const float* k_table,
                                   const float* p_table, int n_table, uint32_t seed, uint32_t box_id,
                                   void* stream) {
    CM21_REQUIRE(spec_out && k_table && p_table, "cm21_grf_init_table: NULL argument");
    CM21_REQUIRE(valid_dim(D), "cm21_grf_init_table: D=%d must be even in [2,2048]", D);
    CM21_REQUIRE(box_len > 0.0, "cm21_grf_init_table: box_len must be > 0");
    CM21_REQUIRE(n_table >= 2 && n_table <= kMaxTable, "cm21_grf_init_table: n_table=%d not in [2,%d]", n_table,
                 kMaxTable);
    GrfParams gp{(float)(2.0 * M_PI / box_len), 0.0f, 0.0f, 1.0f, (float)(1.0 / (box_len * box_len * box_len)),
                 seed, box_id, k_table, p_table, n_table};
    k_grf<<<grf_grid(D), mode_block(), 0, (hipStream_t)stream>>>((float2*)spec_out, D, gp);
    CM21_LAUNCH_CHECK("k_grf");
    return CM21_OK;
}

extern "C" int cm21_subsample(float* small, const float* large, int D, int sw, float pixel_factor, void* stream) {
    CM21_REQUIRE(small && large, "cm21_subsample: NULL argument");
    CM21_REQUIRE(D >= 1 && sw >= 1 && sw <= D && D <= 4096, "cm21_subsample: D=%d sw=%d", D, sw);
    // The largest pick is floor((sw-1)*pf + 0.5); it must stay inside the large box.
    const double last = std::floor((double)((float)(sw - 1) * pixel_factor) + 0.5);
    CM21_REQUIRE(pixel_factor > 0.0f && last < (double)D, "cm21_subsample: pixel_factor=%g reads past D=%d",
                 (double)pixel_factor, D);
    k_subsample<<<mode_grid(sw, sw), mode_block(), 0, (hipStream_t)stream>>>(small, large, D, sw, pixel_factor);
    CM21_LAUNCH_CHECK("k_subsample");
    return CM21_OK;
}

extern "C" int cm21_velocity(void* spec_out, const void* spec_in, int D, float delta_k, int comp, int filter_type,
                             float R, void* stream) {
    CM21_REQUIRE(spec_out && spec_in, "cm21_velocity: NULL spectrum");
    CM21_REQUIRE(valid_dim(D), "cm21_velocity: D=%d must be even in [2,2048]", D);
    CM21_REQUIRE(comp >= 0 && comp <= 2, "cm21_velocity: comp=%d not in {0,1,2}", comp);
    CM21_REQUIRE(filter_type <= 2, "cm21_velocity: filter_type=%d", filter_type);
    k_velocity<<<mode_grid(D, D / 2 + 1), mode_block(), 0, (hipStream_t)stream>>>(
        (float2*)spec_out, (const float2*)spec_in, D, delta_k, comp, filter_type, R);
    CM21_LAUNCH_CHECK("k_velocity");
    return CM21_OK;
}
