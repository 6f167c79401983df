// Internal interface of the fused, pruned transform passes (cm21_fused.hip)
// used by the radius-loop driver.
#pragma once

#include "cm21_internal.hpp"

namespace cm21 {

struct FusedParams {
    int m2max;      // modes with |n|^2 > m2max are zero
    int m2lo;       // modes with |n|^2 >= m2lo take the exact float32 cut test
    int bm;         // largest |b| with b^2 <= m2max (capped at D/2)
    int cmax;       // largest c with c^2 <= m2max (capped at D/2)
    int exact_cut;  // 1 when the k-space cut applies
    float dk, R;    // for the exact test
    float gcoef;    // Gaussian factor exp2(m2 * gcoef); 0 for none
    float scale;    // 1/N times the growth factor of nonlinear input
};

// Density-space form of the centre-pixel threshold: f = erfcf((deltac -
// d)/denom) * scale >= 1/zeta holds exactly when d >= dstar, up to the float32
// rounding of the left side; cells within `band` of dstar take the exact
// float32 test, the others are decided by comparing d with dstar.
struct DensityThreshold {
    float dstar;  // decided by d > dstar outside the band
    float band;   // +inf: every cell takes the exact test
};

// Box-mean scale of nonlinear input and the threshold it implies, written
// once per radius by the reduction of the box sum.
struct ScaledThreshold {
    float scale;
    DensityThreshold th;
    float pad;
};

// The threshold for linear input (scale 1), computed on the host.
DensityThreshold density_threshold_host(float deltac, float denom, double zeta, float scale);

struct LineEpilogue {
    float deltac, denom;
    double zeta, inv_zeta;
    const double* sum;  // FP64 box sum of f_coll (nonlinear input), or NULL
    float fc_mean;
    unsigned long long* mask;  // ionization state, one bit per cell
    float* out;                // final float box
    double* partial;           // per-block FP64 partial sums
    double ncells;
    float* bmax;               // per-block maximum of the smoothed field: written by
                               // the box-sum pass, read by the threshold pass to skip
                               // blocks no cell of which can reach the threshold; or NULL
    DensityThreshold th;       // threshold when `scaled` is NULL (linear input)
    const ScaledThreshold* scaled;  // nonlinear input: scale and threshold of the radius
    const int* active;         // threshold pass: the line groups to visit, or NULL (all)
    const int* nactive;        //   and their number
};

struct FusedShape {
    int D;
    int (*columns)(const float2*, float2*, const float2*, const FusedParams&, int ncg, int nb, hipStream_t);
    int (*rows)(float2*, const float2*, const FusedParams&, int ncg, hipStream_t);
    int (*lines)(int mode, const float2*, const float2*, const FusedParams&, const LineEpilogue&, hipStream_t);
    int (*pair)(const float2*, const FusedParams&, const LineEpilogue&, const float2*, const FusedParams&,
                const LineEpilogue&, const float2*, hipStream_t);
    int W, L;
};

const FusedShape* fused_shape(int D);
size_t fused_partials(int D);
FusedParams fused_params(int D, float dk, float R, int filter_type, int ref_fallthrough, float scale);
LineEpilogue line_epilogue(float deltac, float denom, double zeta, const double* sum, float fc_mean,
                           unsigned long long* mask, float* out, double* partial, double ncells,
                           float* bmax = nullptr);
int fused_radius(const FusedShape& sh, const float2* S, float2* T, const float2* tw, const FusedParams& fp,
                 hipStream_t st);
int fused_lines(const FusedShape& sh, int mode, const float2* U, const float2* tw, const FusedParams& fp,
                const LineEpilogue& ep, hipStream_t st, int stage = CM21_STAGE_FUSED_LINES);
// Threshold pass of one radius (intermediate Uthr, epilogue ept with its
// active list) and box-sum pass of the next (Usum, eps) in one launch.
int fused_lines_pair(const FusedShape& sh, const float2* Uthr, const FusedParams& fpt, const LineEpilogue& ept,
                     const float2* Usum, const FusedParams& fps, const LineEpilogue& eps, const float2* tw,
                     hipStream_t st);
int fused_mask_clear(unsigned long long* mask, size_t words, hipStream_t st);
// Fixed-order FP64 sum of the box-sum partials into *out; with `scaled`
// also the box-mean scale float32(fc_mean / (sum/ncells)) and its threshold.
// With bmax, active and nactive it also lists the line groups whose maximum
// can reach the threshold (for the threshold pass).
int fused_sum(const double* partial, int n, double* out, hipStream_t st, ScaledThreshold* scaled = nullptr,
              float fc_mean = 0.0f, double ncells = 1.0, float deltac = 0.0f, float denom = 1.0f,
              double zeta = 1.0, const float* bmax = nullptr, int* active = nullptr, int* nactive = nullptr);

}  // namespace cm21
