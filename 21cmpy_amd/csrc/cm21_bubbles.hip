// find_HII_bubbles on the GPU: k-space filter, f_coll with its FP64 box sum,
// threshold / final paint, sphere paint and the radius-loop driver.
//
// Reference: Programs/find_bubbles.cu:12-158 (kernels) and
// Programs/find_HII_bubbles.py:17-193 (loop).  Differences by design:
//  * the forward transform runs once per box (the reference re-uploads and
//    re-transforms the same box every radius, find_HII_bubbles.py:141-144);
//  * the 1/N of the inverse transform and the growth factor of nonlinear
//    input are folded into the filter multiply;
//  * the box mean of f_coll is an FP64 device scalar read by the threshold
//    kernel, so the loop never synchronises with the host.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <unordered_map>
#include <vector>

#include "cm21_internal.hpp"
#include "cm21_window.hpp"
#include "cm21_fused_api.hpp"

using namespace cm21;

namespace {

__global__ void k_filter(float2* __restrict__ out, const float2* __restrict__ in, int D, WindowParams wp,
                         float scale) {
    const int H = D / 2 + 1;
    const int c = blockIdx.x * kModeBlockX + threadIdx.x, b = blockIdx.y * kModeBlockY + threadIdx.y;
    const int a = blockIdx.z;
    if (c >= H || b >= D) return;
    const size_t idx = ((size_t)a * D + b) * H + c;
    float w0, w2;
    bool cut;
    window_of(kmag_of(a, b, c, D, wp.dk, wp.korder), wp, w0, cut, w2);
    out[idx] = apply_window(in[idx], w0, cut, w2, scale);
}

__global__ void k_fill(float* __restrict__ p, size_t n, float v) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__device__ inline float fcoll_of(float smoothed, float deltac, float denom) {
    return erfcf(__fdiv_rn(__fsub_rn(deltac, smoothed), denom));
}

// Per-block FP64 partial sums of f_coll over a fixed grid of kReduceBlocks.
__global__ void k_fcoll_partial(const float* __restrict__ sm, size_t N, float deltac, float denom,
                                double* __restrict__ partial) {
    double acc = 0.0;
    for (size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x; p < N; p += (size_t)gridDim.x * blockDim.x)
        acc += (double)fcoll_of(sm[p], deltac, denom);
    acc = block_sum(acc);
    if (threadIdx.x == 0) partial[blockIdx.x] = acc;
}

// Sums `n` partials in a fixed order into *out (one block).
__global__ void k_sum_partials(const double* __restrict__ partial, int n, double* __restrict__ out) {
    double acc = 0.0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) acc += partial[i];
    acc = block_sum(acc);
    if (threadIdx.x == 0) *out = acc;
}

__device__ inline float norm_scale(const double* sum, size_t N, float fc_mean) {
    if (!sum) return 1.0f;
    const double mean = *sum / (double)N;
    return (float)((double)fc_mean / mean);
}

__global__ void k_threshold(float* __restrict__ ion, const float* __restrict__ sm, size_t N, float deltac,
                            float denom, double zeta, double inv_zeta, const double* __restrict__ sum,
                            float fc_mean, int final_step) {
    const size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= N) return;
    const float cur = ion[p];
    if (cur == 1.0f) return;
    float f = fcoll_of(sm[p], deltac, denom);
    if (sum) f = __fmul_rn(f, norm_scale(sum, N, fc_mean));
    if ((double)f >= inv_zeta)
        ion[p] = 1.0f;
    else if (final_step)
        ion[p] = (float)((double)f * zeta);
}

// Sphere paint as a separable squared-distance transform.  d2 holds, per
// cell, a squared distance to the nearest ionizing centre (kFar when none is
// within range).  Only offsets with o^2 < R^2 matter, because the paint test
// is d^2 < R^2 (Programs/find_bubbles.cu:97-98).
constexpr uint32_t kFar = 0x3FFFFFFFu;

__global__ void k_sphere_seed(uint32_t* __restrict__ d2, const float* __restrict__ sm, size_t N, float deltac,
                              float denom, double inv_zeta, const double* __restrict__ sum, float fc_mean) {
    const size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= N) return;
    float f = fcoll_of(sm[p], deltac, denom);
    if (sum) f = __fmul_rn(f, norm_scale(sum, N, fc_mean));
    d2[p] = ((double)f >= inv_zeta) ? 0u : kFar;
}

__global__ void k_sphere_pass(uint32_t* __restrict__ out, const uint32_t* __restrict__ in, int D, int axis,
                              int omax, int periodic) {
    const size_t N = (size_t)D * D * D;
    const size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= N) return;
    const size_t stride = axis == 2 ? 1 : (axis == 1 ? (size_t)D : (size_t)D * D);
    const int pos = (int)((p / stride) % D);
    const size_t base = p - (size_t)pos * stride;
    uint32_t best = kFar;
    for (int o = -omax; o <= omax; ++o) {
        int q = pos + o;
        if (periodic) {
            q %= D;
            if (q < 0) q += D;
        } else if (q < 0 || q >= D) {
            continue;
        }
        const uint32_t v = in[base + (size_t)q * stride];
        if (v < kFar) {
            const uint32_t c = v + (uint32_t)(o * o);
            best = c < best ? c : best;
        }
    }
    out[p] = best;
}

__global__ void k_sphere_apply(float* __restrict__ ion, const uint32_t* __restrict__ d2, size_t N, float R2) {
    const size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= N) return;
    const uint32_t v = d2[p];
    if (v < kFar && (float)v < R2) ion[p] = 1.0f;
}

int launch_filter(void* spec_out, const void* spec_in, int D, const WindowParams& wp, float scale,
                  hipStream_t st) {
    k_filter<<<mode_grid(D, D / 2 + 1), mode_block(), 0, st>>>((float2*)spec_out, (const float2*)spec_in, D, wp,
                                                               scale);
    CM21_LAUNCH_CHECK("k_filter");
    return CM21_OK;
}

int launch_fcoll_sum(const float* sm, int D, float deltac, float denom, double* dev_sum, void* work,
                     hipStream_t st) {
    const size_t N = (size_t)D * D * D;
    double* partial = (double*)work;
    k_fcoll_partial<<<kReduceBlocks, kThreads, 0, st>>>(sm, N, deltac, denom, partial);
    CM21_LAUNCH_CHECK("k_fcoll_partial");
    k_sum_partials<<<1, kThreads, 0, st>>>(partial, kReduceBlocks, dev_sum);
    CM21_LAUNCH_CHECK("k_sum_partials");
    return CM21_OK;
}

int launch_threshold(float* ion, const float* sm, int D, float deltac, float denom, double zeta,
                     const double* sum, float fc_mean, int final_step, hipStream_t st) {
    const size_t N = (size_t)D * D * D;
    k_threshold<<<grid_for(N), kThreads, 0, st>>>(ion, sm, N, deltac, denom, zeta, 1.0 / zeta, sum, fc_mean,
                                                  final_step);
    CM21_LAUNCH_CHECK("k_threshold");
    return CM21_OK;
}

int launch_sphere(float* ion, const float* sm, int D, float deltac, float denom, double zeta, const double* sum,
                  float fc_mean, float R_cells, int periodic, void* work, hipStream_t st) {
    const size_t N = (size_t)D * D * D;
    uint32_t* a = (uint32_t*)work;
    uint32_t* b = a + N;
    const float R2 = R_cells * R_cells;  // float32(R*R), find_bubbles.cu:98
    int omax = (int)std::ceil((double)R_cells);
    while (omax > 0 && (float)(omax * omax) >= R2) --omax;  // largest o with o^2 < R^2
    if (omax >= D) omax = periodic ? D / 2 : D - 1;
    k_sphere_seed<<<grid_for(N), kThreads, 0, st>>>(a, sm, N, deltac, denom, 1.0 / zeta, sum, fc_mean);
    CM21_LAUNCH_CHECK("k_sphere_seed");
    k_sphere_pass<<<grid_for(N), kThreads, 0, st>>>(b, a, D, 2, omax, periodic);
    CM21_LAUNCH_CHECK("k_sphere_pass");
    k_sphere_pass<<<grid_for(N), kThreads, 0, st>>>(a, b, D, 1, omax, periodic);
    CM21_LAUNCH_CHECK("k_sphere_pass");
    k_sphere_pass<<<grid_for(N), kThreads, 0, st>>>(b, a, D, 0, omax, periodic);
    CM21_LAUNCH_CHECK("k_sphere_pass");
    k_sphere_apply<<<grid_for(N), kThreads, 0, st>>>(ion, b, N, R2);
    CM21_LAUNCH_CHECK("k_sphere_apply");
    return CM21_OK;
}

bool valid_dim(int D) { return D >= 2 && D <= 2048 && D % 2 == 0; }

}  // namespace

extern "C" int cm21_hii_filter(void* spec_out, const void* spec_in, int D, float delta_k, int filter_type,
                               int ref_fallthrough, int korder, float R, float scale, void* stream) {
    CM21_REQUIRE(spec_out && spec_in, "cm21_hii_filter: NULL spectrum");
    CM21_REQUIRE(valid_dim(D), "cm21_hii_filter: D=%d must be even in [2,2048]", D);
    CM21_REQUIRE(filter_type >= 0 && filter_type <= 2, "cm21_hii_filter: filter_type=%d not in {0,1,2}", filter_type);
    CM21_REQUIRE(korder == 0 || korder == 1, "cm21_hii_filter: korder=%d not in {0,1}", korder);
    WindowParams wp{delta_k, R, filter_type, ref_fallthrough ? 1 : 0, korder};
    return launch_filter(spec_out, spec_in, D, wp, scale, (hipStream_t)stream);
}

extern "C" size_t cm21_reduce_workspace_bytes(void) { return align_up(sizeof(double) * kReduceBlocks, 256); }

extern "C" int cm21_fcoll_sum(const float* smoothed, int D, float deltac, float denom, double* dev_sum,
                              void* work, void* stream) {
    CM21_REQUIRE(smoothed && dev_sum && work, "cm21_fcoll_sum: NULL argument");
    CM21_REQUIRE(valid_dim(D), "cm21_fcoll_sum: D=%d must be even in [2,2048]", D);
    return launch_fcoll_sum(smoothed, D, deltac, denom, dev_sum, work, (hipStream_t)stream);
}

extern "C" int cm21_fcoll_threshold(float* ionized, const float* smoothed, int D, float deltac, float denom,
                                    double zeta, const double* dev_sum, float fc_mean, int final_step,
                                    void* stream) {
    CM21_REQUIRE(ionized && smoothed, "cm21_fcoll_threshold: NULL argument");
    CM21_REQUIRE(valid_dim(D), "cm21_fcoll_threshold: D=%d must be even in [2,2048]", D);
    CM21_REQUIRE(zeta > 0.0, "cm21_fcoll_threshold: zeta must be positive");
    return launch_threshold(ionized, smoothed, D, deltac, denom, zeta, dev_sum, fc_mean, final_step,
                            (hipStream_t)stream);
}

extern "C" size_t cm21_sphere_workspace_bytes(int D) {
    return align_up(sizeof(uint32_t) * 2 * (size_t)D * D * D, 256);
}

extern "C" int cm21_update_sphere(float* ionized, const float* smoothed, int D, float deltac, float denom,
                                  double zeta, const double* dev_sum, float fc_mean, float R_cells, int periodic,
                                  void* work, void* stream) {
    CM21_REQUIRE(ionized && smoothed && work, "cm21_update_sphere: NULL argument");
    CM21_REQUIRE(valid_dim(D), "cm21_update_sphere: D=%d must be even in [2,2048]", D);
    CM21_REQUIRE(zeta > 0.0, "cm21_update_sphere: zeta must be positive");
    CM21_REQUIRE(R_cells >= 0.0f && R_cells <= 1024.0f, "cm21_update_sphere: R_cells=%g not in [0,1024]",
                 (double)R_cells);
    return launch_sphere(ionized, smoothed, D, deltac, denom, zeta, dev_sum, fc_mean, R_cells, periodic ? 1 : 0,
                         work, (hipStream_t)stream);
}

// Workspace carve-up of the loop driver.
namespace {
// Pipelined fused loop: radius r uses intermediate r % kInter and the small
// per-radius buffers r % kSlots; a launch touches radii r-2 .. r+2.
constexpr int kInter = 5, kSlots = 3;

struct LoopWork {
    float2* spec;
    float2* filt;             // generic path: filtered spectrum
    float* smoothed;          // generic path: smoothed box
    float2* inter[kInter];    // fused path: intermediates (inter[0] is filt,
                              // inter[1] shares smoothed's region)
    double* partial;
    double* sums;
    void* sphere;
    unsigned long long* mask;
    double* lpart[kSlots];    // box-sum partials
    float* bmax[kSlots];      // per line group maxima
    int* active[kSlots];      // line groups the threshold pass visits, and their number
    int* nactive[kSlots];
    ScaledThreshold* scaled;  // per radius
    size_t bytes;
};

LoopWork carve(void* base, int D, int n_radii, int update) {
    const size_t N = (size_t)D * D * D;
    const size_t Nc = (size_t)D * D * (D / 2 + 1);
    const size_t Nt = (size_t)D * D * padded_row(D / 2 + 1);  // an intermediate (padded rows)
    char* p = (char*)base;
    LoopWork w{};
    size_t off = 0;
    // the spectrum: the r2c layout (stride D/2+1) for the paint path, the
    // fused transform's b-major layout with padded rows otherwise
    w.spec = (float2*)(p + off);
    off = align_up(off + sizeof(float2) * (update ? Nc : Nt), 256);
    w.filt = w.inter[0] = (float2*)(p + off);
    off = align_up(off + sizeof(float2) * Nt, 256);
    // 8Nt >= 4N bytes
    w.smoothed = (float*)(p + off);
    w.inter[1] = (float2*)(p + off);
    off = align_up(off + std::max(sizeof(float) * N, sizeof(float2) * Nt), 256);
    // the fused path (update == 0 only) needs three more intermediates
    for (int k = 2; k < kInter && !update; ++k) {
        w.inter[k] = (float2*)(p + off);
        off = align_up(off + sizeof(float2) * Nt, 256);
    }
    w.partial = (double*)(p + off);
    off = align_up(off + cm21_reduce_workspace_bytes(), 256);
    w.sums = (double*)(p + off);
    off = align_up(off + sizeof(double) * (size_t)(n_radii > 0 ? n_radii : 1), 256);
    w.sphere = update ? (void*)(p + off) : nullptr;
    if (update) off = align_up(off + cm21_sphere_workspace_bytes(D), 256);
    w.mask = (unsigned long long*)(p + off);
    off = align_up(off + sizeof(unsigned long long) * ((N + 63) / 64), 256);
    for (int k = 0; k < kSlots; ++k) {
        w.lpart[k] = (double*)(p + off);
        off = align_up(off + sizeof(double) * (fused_partials(D) + 1), 256);
        w.bmax[k] = (float*)(p + off);
        off = align_up(off + sizeof(float) * (fused_partials(D) + 1), 256);
        w.active[k] = (int*)(p + off);
        off = align_up(off + sizeof(int) * (fused_partials(D) + 1), 256);
        w.nactive[k] = (int*)(p + off);
        off = align_up(off + sizeof(int), 256);
    }
    w.scaled = (ScaledThreshold*)(p + off);
    off = align_up(off + sizeof(ScaledThreshold) * (size_t)(n_radii > 0 ? n_radii : 1), 256);
    w.bytes = off;
    return w;
}
}  // namespace

namespace {
int run_loop(cm21_plan* plan, const float* delta, float* ionized, const float* radii, const float* denoms,
             int n_radii, const cm21_bubble_params* prm, const LoopWork& w, hipStream_t st);
int graph_loop(cm21_plan* plan, const float* delta, float* ionized, const float* radii, const float* denoms,
               int n_radii, const cm21_bubble_params* prm, const LoopWork& w, hipStream_t st);
}  // namespace

extern "C" size_t cm21_find_bubbles_workspace_bytes(int D, int n_radii, int update) {
    if (!valid_dim(D)) return 0;
    return carve(nullptr, D, n_radii, update).bytes;
}

namespace {
// Argument checks shared by cm21_find_bubbles and cm21_find_bubbles_noalloc.
int check_loop_args(const char* fn, cm21_plan* plan, const float* delta, float* ionized, const float* radii,
                    const float* denoms, int n_radii, const cm21_bubble_params* prm) {
    CM21_REQUIRE(plan && delta && ionized && radii && denoms && prm, "%s: NULL argument", fn);
    CM21_REQUIRE(plan_batch(plan) == 1, "%s: plan must have batch 1", fn);
    CM21_REQUIRE(n_radii >= 1 && n_radii <= 4096, "%s: n_radii=%d out of range", fn, n_radii);
    CM21_REQUIRE(prm->filter_type >= 0 && prm->filter_type <= 2, "%s: filter_type=%d", fn, prm->filter_type);
    CM21_REQUIRE(prm->zeta > 0.0, "%s: zeta must be positive", fn);
    CM21_REQUIRE(prm->box_len > 0.0f, "%s: box_len must be positive", fn);
    for (int r = 0; r < n_radii; ++r) {
        if (!(denoms[r] > 0.0f) || !std::isfinite(denoms[r]))
            return cm21::fail(CM21_ERR_DOMAIN,
                              "%s: denominator %g at R=%g is not positive and finite (sigma_min^2 <= S(R))", fn,
                              (double)denoms[r], (double)radii[r]);
        if (prm->update && !(radii[r] * prm->sphere_r_scale <= 1024.0f))
            return cm21::fail(CM21_ERR_INVALID, "%s: sphere radius %g cells exceeds 1024", fn,
                              (double)(radii[r] * prm->sphere_r_scale));
    }
    return CM21_OK;
}
}  // namespace

extern "C" int cm21_find_bubbles(cm21_plan* plan, const float* delta, float* ionized, const float* radii,
                                 const float* denoms, int n_radii, const cm21_bubble_params* prm, void* work,
                                 size_t work_bytes, void* stream) {
    CM21_TRY(check_loop_args("cm21_find_bubbles", plan, delta, ionized, radii, denoms, n_radii, prm));
    CM21_REQUIRE(work, "cm21_find_bubbles: NULL argument");
    const int D = plan_dim(plan);
    LoopWork w = carve(work, D, n_radii, prm->update);
    if (work_bytes < w.bytes)
        return cm21::fail(CM21_ERR_NOMEM, "cm21_find_bubbles: workspace %zu bytes < required %zu", work_bytes,
                          w.bytes);
    hipStream_t st = (hipStream_t)stream;
    if (prm->graph && st != nullptr && !prof_any()) return graph_loop(plan, delta, ionized, radii, denoms, n_radii, prm, w, st);
    return run_loop(plan, delta, ionized, radii, denoms, n_radii, prm, w, st);
}

namespace {
int run_loop(cm21_plan* plan, const float* delta, float* ionized, const float* radii, const float* denoms,
             int n_radii, const cm21_bubble_params* prm, const LoopWork& w, hipStream_t st) {
    void* stream = (void*)st;
    const int D = plan_dim(plan);
    const size_t N = (size_t)D * D * D;

    const float dk = (float)(2.0 * M_PI / (double)prm->box_len);
    const float scale = prm->linear_input ? (float)(1.0 / (double)N) : (float)((double)prm->growth_inv / (double)N);
    const FusedShape* sh = fused_shape(D);
    if (prm->fused && sh && prm->update == 0 && (prm->filter_type == 1 || prm->filter_type == 2)) {
        const float2* tw = plan_twiddles(plan);
        // the forward transform forms only the modes the widest window keeps,
        // through the intermediate that radius 4 uses first
        FusedParams widest = fused_params(D, dk, radii[0], prm->filter_type, prm->ref_fallthrough, scale);
        for (int r = 1; r < n_radii; ++r) {
            const FusedParams f = fused_params(D, dk, radii[r], prm->filter_type, prm->ref_fallthrough, scale);
            if (f.m2max > widest.m2max) widest = f;
        }
        CM21_TRY(fused_forward(*sh, delta, w.inter[kInter - 1], w.spec, tw, widest, st));
        CM21_TRY(fused_mask_clear(w.mask, (N + 63) / 64, st));
        // One launch per radius on the caller's stream, each with up to five
        // independent roles over neighbouring radii (k_fused_radius): launch
        // j holds the columns pass of radius j+2, the rows pass of radius j+1,
        // the box-sum pass of radius j, the box-sum reduction of radius j-1
        // and the threshold pass of radius j-2 (nonlinear input), or the
        // threshold pass of radius j (linear input, threshold from the host).
        // In-order launches carry every dependency (rows after columns, sum
        // after rows, reduction after sum, threshold after reduction), so the
        // loop needs no events, and the pruned transforms of later radii
        // share the chip with the box sums instead of waiting for them.
        // The threshold pass recomputes the last-axis transform in the line
        // groups whose maximum can reach the threshold rather than storing and
        // re-reading the smoothed field; the final paint is one more launch.
        const int nblk = (int)fused_partials(D);
        const bool nl = !prm->linear_input;
        auto params = [&](int r) {
            return fused_params(D, dk, radii[r], prm->filter_type, prm->ref_fallthrough, scale);
        };
        auto threshold_epilogue = [&](int r) {
            const int q = r % kSlots;
            LineEpilogue e = line_epilogue(prm->deltac, denoms[r], prm->zeta, nl ? w.sums + r : nullptr,
                                           prm->fc_mean, w.mask, ionized, w.lpart[q], (double)N,
                                           nl ? w.bmax[q] : nullptr);
            if (nl) {
                e.scaled = w.scaled + r;
                if (r < n_radii - 1) {
                    e.active = w.active[q];
                    e.nactive = w.nactive[q];
                }
            } else {
                e.th = density_threshold_host(prm->deltac, denoms[r], prm->zeta, 1.0f);
            }
            return e;
        };
        const int last = n_radii - 1;
        // launches j = -2 .. j_end-1: the last one holds the reduction of the
        // last radius and the threshold pass of the one before (nonlinear), or
        // the threshold pass of the radius before the last (linear)
        const int j_end = nl ? n_radii + 1 : n_radii - 1;
        for (int j = -2; j < j_end; ++j) {
            RadiusRoles a{};
            if (j + 2 <= last) {
                const int r = j + 2;
                a.S = w.spec;
                a.Tcols = w.inter[r % kInter];
                a.fp_cols = params(r);
                int nb = 0;
                column_groups(*sh, a.fp_cols, &a.ncg_cols, &nb);
                a.n_cols = a.ncg_cols * nb;
            }
            a.n_mem = a.n_cols;
            if (j + 1 >= 0 && j + 1 <= last) {
                const int r = j + 1;
                a.Trows = w.inter[r % kInter];
                a.fp_rows = params(r);
                int nb = 0;
                column_groups(*sh, a.fp_rows, &a.ncg_rows, &nb);
                a.n_mem += a.ncg_rows * D;
            }
            if (nl && j - 1 >= 0 && j - 1 <= last) {
                const int r = j - 1, q = r % kSlots;
                a.partial = w.lpart[q];
                a.npart = nblk;
                a.total = w.sums + r;
                a.sa = SumScaleArgs{w.scaled + r, prm->fc_mean, prm->deltac, denoms[r], (double)N, prm->zeta,
                                    w.bmax[q], r < last ? w.active[q] : nullptr, r < last ? w.nactive[q] : nullptr};
                a.n_red = 1;
            }
            const int rt = nl ? j - 2 : j;  // the threshold pass (the last radius is painted below)
            if (rt >= 0 && rt < last) {
                a.Uthr = w.inter[rt % kInter];
                a.fp_thr = params(rt);
                a.ep_thr = threshold_epilogue(rt);
                a.n_thr = nblk;
            }
            if (nl && j >= 0 && j <= last) {
                const int r = j, q = r % kSlots;
                a.Usum = w.inter[r % kInter];
                a.fp_sum = params(r);
                a.ep_sum = line_epilogue(prm->deltac, denoms[r], prm->zeta, nullptr, prm->fc_mean, w.mask, ionized,
                                         w.lpart[q], (double)N, w.bmax[q]);
                a.n_sum = nblk;
            }
            CM21_TRY(fused_radius_launch(*sh, a, tw, st));
        }
        // mode 2: the final paint of the last radius
        CM21_TRY(fused_lines(*sh, 2, w.inter[last % kInter], tw, params(last), threshold_epilogue(last), st,
                             CM21_STAGE_FUSED_LINES_THRESHOLD));
        return CM21_OK;
    }
    {
        ProfToken t = prof_begin(CM21_STAGE_R2C, st);
        CM21_TRY(cm21_r2c(plan, delta, w.spec, stream));
        prof_end(t, st);
    }
    k_fill<<<grid_for(N), kThreads, 0, st>>>(ionized, N, 0.0f);
    CM21_LAUNCH_CHECK("k_fill");
    for (int r = 0; r < n_radii; ++r) {
        const bool final_step = r == n_radii - 1;
        WindowParams wp{dk, radii[r], prm->filter_type, prm->ref_fallthrough ? 1 : 0, 0};
        ProfToken t = prof_begin(CM21_STAGE_FILTER, st);
        CM21_TRY(launch_filter(w.filt, w.spec, D, wp, scale, st));
        prof_end(t, st);
        t = prof_begin(CM21_STAGE_C2R, st);
        CM21_TRY(cm21_c2r(plan, w.filt, w.smoothed, stream));
        prof_end(t, st);
        const double* sum = nullptr;
        if (!prm->linear_input) {
            t = prof_begin(CM21_STAGE_FCOLL_SUM, st);
            CM21_TRY(launch_fcoll_sum(w.smoothed, D, prm->deltac, denoms[r], w.sums + r, w.partial, st));
            prof_end(t, st);
            sum = w.sums + r;
        }
        if (prm->update && !final_step) {
            t = prof_begin(CM21_STAGE_SPHERE, st);
            CM21_TRY(launch_sphere(ionized, w.smoothed, D, prm->deltac, denoms[r], prm->zeta, sum, prm->fc_mean,
                                   radii[r] * prm->sphere_r_scale, prm->sphere_periodic, w.sphere, st));
            prof_end(t, st);
        } else {
            t = prof_begin(CM21_STAGE_THRESHOLD, st);
            CM21_TRY(launch_threshold(ionized, w.smoothed, D, prm->deltac, denoms[r], prm->zeta, sum, prm->fc_mean,
                                      final_step ? 1 : 0, st));
            prof_end(t, st);
        }
    }
    return CM21_OK;
}

// Graph replay of the whole loop.  The first call with a given set of
// arguments (pointers, parameters, radii, denominators, stream) captures the
// launches into a HIP graph; identical later calls replay it, which removes
// the host launch cost of the ~5 kernels per radius.  One graph is kept per
// plan.  A capture that fails falls back to eager launches.
struct GraphKey {
    const void* delta;
    const void* ionized;
    const void* work;
    hipStream_t stream;
    cm21_bubble_params prm;
    std::vector<float> radii, denoms;
    bool operator==(const GraphKey& o) const {
        return delta == o.delta && ionized == o.ionized && work == o.work && stream == o.stream &&
               std::memcmp(&prm, &o.prm, sizeof(prm)) == 0 && radii == o.radii && denoms == o.denoms;
    }
};
struct GraphEntry {
    GraphKey key;
    hipGraphExec_t exec = nullptr;
};
std::mutex g_graph_mu;
std::unordered_map<const cm21_plan*, GraphEntry> g_graphs;

int graph_loop(cm21_plan* plan, const float* delta, float* ionized, const float* radii, const float* denoms,
               int n_radii, const cm21_bubble_params* prm, const LoopWork& w, hipStream_t st) {
    GraphKey key{delta, ionized, w.spec, st, *prm, std::vector<float>(radii, radii + n_radii),
                 std::vector<float>(denoms, denoms + n_radii)};
    std::lock_guard<std::mutex> lk(g_graph_mu);
    auto it = g_graphs.find(plan);
    if (it != g_graphs.end() && it->second.exec && it->second.key == key) {
        CM21_HIP_CHECK(hipGraphLaunch(it->second.exec, st));
        return CM21_OK;
    }
    if (it != g_graphs.end() && it->second.exec) {
        (void)hipGraphExecDestroy(it->second.exec);
        it->second.exec = nullptr;
    }
    if (hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal) != hipSuccess) {
        (void)hipGetLastError();
        return run_loop(plan, delta, ionized, radii, denoms, n_radii, prm, w, st);
    }
    const int rc = run_loop(plan, delta, ionized, radii, denoms, n_radii, prm, w, st);
    hipGraph_t graph = nullptr;
    const hipError_t ec = hipStreamEndCapture(st, &graph);
    hipGraphExec_t exec = nullptr;
    bool ok = rc == CM21_OK && ec == hipSuccess && graph;
    if (ok) ok = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) == hipSuccess;
    if (graph) (void)hipGraphDestroy(graph);
    if (!ok) {
        (void)hipGetLastError();
        if (rc != CM21_OK) return rc;
        return run_loop(plan, delta, ionized, radii, denoms, n_radii, prm, w, st);
    }
    g_graphs[plan] = GraphEntry{std::move(key), exec};
    CM21_HIP_CHECK(hipGraphLaunch(exec, st));
    return CM21_OK;
}
}  // namespace

// ------------------------------------------------------------------ no-alloc loop
// cm21_find_bubbles_noalloc runs the loop of cm21_find_bubbles with its
// workspace in a file-scope device array instead of a caller's buffer.  Each
// GPU holds its own copy of the array (CM21_NOALLOC_SCRATCH_BYTES) from the
// library's load until the process ends; host code reaches it only through
// the address hipGetSymbolAddress returns on that device, taken once per
// device at the first call.  Calls on one GPU share the array.
namespace {
constexpr size_t kScratchBytes = (size_t)CM21_NOALLOC_SCRATCH_BYTES;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch[kMaxDevices] = {};

int scratch_on_current_device(void** out) {
    int dev = 0;
    CM21_HIP_CHECK(hipGetDevice(&dev));
    CM21_REQUIRE(dev >= 0 && dev < kMaxDevices, "cm21_find_bubbles_noalloc: device %d out of range", dev);
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || p == nullptr)
            return cm21::fail(CM21_ERR_HIP, "cm21_find_bubbles_noalloc: hipGetSymbolAddress(scratch) on device %d: %s",
                              dev, hipGetErrorString(e));
        g_scratch[dev] = p;
    }
    *out = g_scratch[dev];
    return CM21_OK;
}
}  // namespace

extern "C" int cm21_find_bubbles_noalloc(cm21_plan* plan, const float* delta, float* ionized, const float* radii,
                                         const float* denoms, int n_radii, const cm21_bubble_params* prm,
                                         void* stream) {
    CM21_TRY(check_loop_args("cm21_find_bubbles_noalloc", plan, delta, ionized, radii, denoms, n_radii, prm));
    const int D = plan_dim(plan);
    const size_t need = carve(nullptr, D, n_radii, prm->update).bytes;
    if (need > kScratchBytes)
        return cm21::fail(CM21_ERR_NOMEM,
                          "cm21_find_bubbles_noalloc: D=%d with %d radii (update %d) needs %zu bytes of workspace, "
                          "more than the %zu-byte static scratch; use cm21_find_bubbles with a caller workspace",
                          D, n_radii, prm->update, need, kScratchBytes);
    void* base = nullptr;
    CM21_TRY(scratch_on_current_device(&base));
    LoopWork w = carve(base, D, n_radii, prm->update);
    // the launches go straight to the stream (params->graph is ignored), so a
    // caller that captures the stream records them into its own graph
    return run_loop(plan, delta, ionized, radii, denoms, n_radii, prm, w, (hipStream_t)stream);
}

namespace cm21 {
void forget_plan_graphs(const cm21_plan* plan) {
    std::lock_guard<std::mutex> lk(g_graph_mu);
    auto it = g_graphs.find(plan);
    if (it != g_graphs.end()) {
        if (it->second.exec) (void)hipGraphExecDestroy(it->second.exec);
        g_graphs.erase(it);
    }
}
}  // namespace cm21
