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

// Arguments of the box-sum reduction of one radius: the FP64 total, and for
// nonlinear input the box-mean scale float32(fc_mean / (sum / ncells)), its
// density threshold and the list of line groups whose maximum can reach it.
struct SumScaleArgs {
    ScaledThreshold* scaled;
    float fc_mean, deltac, denom;
    double ncells, zeta;
    const float* bmax;
    int* active;
    int* nactive;
};

// The roles of one launch of the pipelined radius loop (k_fused_radius) and
// their block counts; a count of 0 is a role the launch does not have.
struct RadiusRoles {
    int n_red;               // box-sum reduction: 0 or 1 block
    int n_thr;               // threshold blocks (one line group each)
    int n_cols, n_mem;       // columns blocks; columns + rows blocks
    int n_sum;               // box-sum blocks (one line group each)
    int ncg_cols, ncg_rows;  // column groups of the columns and rows roles
    const float2* S;         // columns: pristine spectrum -> Tcols
    float2* Tcols;
    FusedParams fp_cols;
    float2* Trows;           // rows: in place
    FusedParams fp_rows;
    const double* partial;   // reduction: npart box-sum partials -> total
    int npart;
    double* total;
    SumScaleArgs sa;
    const float2* Uthr;      // threshold pass
    FusedParams fp_thr;
    LineEpilogue ep_thr;
    const float2* Usum;      // box-sum pass
    FusedParams fp_sum;
    LineEpilogue ep_sum;
};

struct FusedShape {
    int D;
    int (*lines)(int mode, const float2*, const float2*, const FusedParams&, const LineEpilogue&, hipStream_t);
    int (*radius)(const RadiusRoles&, const float2* tw, hipStream_t);
    int (*forward)(const float* delta, float2* T0, float2* S, const float2* tw, const FusedParams& fp, hipStream_t);
    int W, L;  // columns per column group, lines per line group
};

const FusedShape* fused_shape(int D);
size_t fused_partials(int D);
FusedParams fused_params(int D, float dk, float R, int filter_type, int ref_fallthrough, float scale);
LineEpilogue line_epilogue(float deltac, float denom, double zeta, const double* sum, float fc_mean,
                           unsigned long long* mask, float* out, double* partial, double ncells,
                           float* bmax = nullptr);
// Column groups and surviving b rows of the columns role at one radius
// (the rows role uses the same column groups over all D rows p).
void column_groups(const FusedShape& sh, const FusedParams& fp, int* ncg, int* nb);
int fused_radius_launch(const FusedShape& sh, const RadiusRoles& a, const float2* tw, hipStream_t st);
// The r2c of delta into the half spectrum S for the fused loop, stored
// b-major (S[b][a][c]), formed only for the modes the widest window fp keeps
// (|n|^2 <= fp.m2max), through the intermediate T0 (rows padded to 16
// modes).  S entries outside that sphere are left untouched: no radius reads
// them.
int fused_forward(const FusedShape& sh, const float* delta, float2* T0, float2* S, const float2* tw,
                  const FusedParams& fp, hipStream_t st);
int fused_lines(const FusedShape& sh, int mode, const float2* U, const float2* tw, const FusedParams& fp,
                const LineEpilogue& ep, hipStream_t st, int stage = CM21_STAGE_FUSED_LINES);
int fused_mask_clear(unsigned long long* mask, size_t words, hipStream_t st);

}  // namespace cm21
