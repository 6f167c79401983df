/*
 * cm21.h - C ABI of the MI355X (gfx950) semi-numerical reionization hot path.
 *
 * Every entry point is asynchronous on the HIP stream it is given (a
 * hipStream_t passed as void*, NULL = the null stream), takes plain device
 * pointers owned by the caller, and returns an int status: CM21_OK (0) or a
 * negative CM21_ERR_* code with a message in cm21_last_error() (thread-local).
 * No exception or abort crosses this boundary.  The library owns only opaque
 * plan handles (rocFFT plans and their work buffers).
 *
 * Layouts.  A real box is float32 (D,D,D), C order, axis 2 fastest.  A
 * spectrum is the r2c half spectrum: float2 (re,im) interleaved, (D,D,D/2+1),
 * axis 2 fastest.  Spectra are "c2r-ready": the unnormalised inverse real
 * transform (cm21_c2r) of a spectrum gives the real-space field.  cm21_r2c is
 * unnormalised (a forward transform of a field gives N times its c2r-ready
 * spectrum).
 *
 * Each declaration names the reference interface it replaces (file:line in
 * yunfanz/21cmpy).  Host-side scalars that the reference computes in Python
 * (the radius list, S(R), denominators, the spectrum normalisation) are
 * passed in by the caller.
 */
#ifndef CM21_H
#define CM21_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CM21_ABI_VERSION 1

#define CM21_OK 0
#define CM21_ERR_INVALID (-1) /* bad argument (sizes, null pointers, flags) */
#define CM21_ERR_HIP (-2)     /* HIP runtime or kernel launch error */
#define CM21_ERR_FFT (-3)     /* rocFFT error */
#define CM21_ERR_DOMAIN (-4)  /* physics domain error, e.g. sigma_min^2 <= S(R) */
#define CM21_ERR_NOMEM (-5)   /* workspace too small or allocation failed */

typedef struct cm21_plan cm21_plan;

/* Version and last error (no reference counterpart: PyCUDA raised Python
 * exceptions, Programs/find_HII_bubbles.py:116-118). */
int cm21_abi_version(void);
const char* cm21_last_error(void);

/* ------------------------------------------------------------------ FFT
 * Replaces pyfft.cuda.Plan(shape, complex64) and Plan.execute(buf,
 * inverse=...) (Programs/find_HII_bubbles.py:91,144,150; initialize.py:79-80;
 * delta_T.py:84-86).  The reference runs complex-to-complex transforms of real
 * data; these are the equivalent real transforms on the half spectrum.
 * A plan serves `batch` boxes of side D laid out back to back.  cm21_c2r may
 * overwrite its input spectrum. */
int cm21_plan_create(cm21_plan** out, int D, int batch);
int cm21_plan_destroy(cm21_plan* plan);
int cm21_r2c(cm21_plan* plan, const float* in, void* spec_out, void* stream);
int cm21_c2r(cm21_plan* plan, void* spec_in, float* out, void* stream);

/* ------------------------------------------------------------------ filter
 * Replaces HII_filter (Programs/find_bubbles.cu:12-42, initialize.cu:10-40).
 * spec_out = window(|k| R) * spec_in * scale on the half spectrum (in place
 * allowed).  filter_type 0/1/2 as INIT_PARAMS.HII_FILTER; ref_fallthrough=1
 * reproduces the reference switch without breaks (0 -> top-hat * k-cut *
 * Gaussian(0.2659 kR), 1 -> k-cut * Gaussian(0.2659 kR), 2 -> Gaussian(0.643
 * kR)); 0 gives the textbook windows.  korder selects the order of the |k|^2
 * sum: 0 = fastest axis first (find_bubbles.cu:26), 1 = slowest first
 * (initialize.cu:24).  delta_k = 2 pi / box_len in float32. */
int cm21_hii_filter(void* spec_out, const void* spec_in, int D, float delta_k,
                    int filter_type, int ref_fallthrough, int korder, float R,
                    float scale, void* stream);

/* ------------------------------------------------------------------ f_coll
 * Replaces fcoll_kernel + gpuarray.sum (Programs/find_bubbles.cu:43-58;
 * find_HII_bubbles.py:156-159): dev_sum[0] = sum over the box of
 * erfcf((deltac - smoothed)/denom), accumulated in FP64 with a fixed
 * reduction order (deterministic).  `work` holds at least
 * cm21_reduce_workspace_bytes() bytes. */
size_t cm21_reduce_workspace_bytes(void);
int cm21_fcoll_sum(const float* smoothed, int D, float deltac, float denom,
                   double* dev_sum, void* work, void* stream);

/* Replaces fcoll_kernel + update_kernel (final_step = 0) or final_kernel
 * (final_step = 1) (Programs/find_bubbles.cu:43-71,142-158;
 * find_HII_bubbles.py:155-175).  f = erfcf((deltac - smoothed)/denom),
 * optionally normalised by float32(fc_mean / (dev_sum[0]/N)) when dev_sum is
 * not NULL (nonlinear input); cells with ionized == 1 are skipped, cells with
 * f >= 1/zeta (compared in double) become 1, and on the final step the rest
 * become f*zeta. */
int cm21_fcoll_threshold(float* ionized, const float* smoothed, int D,
                         float deltac, float denom, double zeta,
                         const double* dev_sum, float fc_mean, int final_step,
                         void* stream);

/* Replaces update_sphere_kernel (Programs/find_bubbles.cu:73-140): every cell
 * whose integer squared distance to a cell meeting the threshold is below
 * float32(R_cells^2) becomes 1.  periodic = 0 clips at the box faces as the
 * reference does; 1 wraps.  Computed as three separable min-plus passes
 * (O(N R)) instead of the reference's O(R^3) loop per ionizing cell.  `work`
 * holds at least cm21_sphere_workspace_bytes(D) bytes.  R_cells <= 1024. */
size_t cm21_sphere_workspace_bytes(int D);
int cm21_update_sphere(float* ionized, const float* smoothed, int D,
                       float deltac, float denom, double zeta,
                       const double* dev_sum, float fc_mean, float R_cells,
                       int periodic, void* work, void* stream);

/* ------------------------------------------------------------------ loop
 * Replaces conv_bubbles (Programs/find_HII_bubbles.py:17-193): one forward
 * transform of delta, then for each radius the filter, inverse transform,
 * optional FP64 box mean and threshold/paint, all on `stream` with no host
 * synchronisation.  radii/denoms are host arrays of n_radii float32 values
 * (find_HII_bubbles.py:121-128); the last radius is the final step.
 * `ionized` receives the ionized-fraction box (1 - xH).  `work` holds at
 * least cm21_find_bubbles_workspace_bytes(D, n_radii, update) bytes. */
typedef struct cm21_bubble_params {
    float box_len;        /* Mpc; DELTA_K = float32(2 pi / box_len) */
    float deltac;         /* barrier DELTAC (find_HII_bubbles.py:57,66) */
    double zeta;          /* ionizing efficiency ZETA */
    float growth_inv;     /* fgrowth = deltac/1.686, applied when !linear_input */
    float fc_mean;        /* fc_mean_ps (find_HII_bubbles.py:60) */
    int linear_input;     /* LE: 1 = linear z=0 input, no normalisation */
    int filter_type;      /* HII_FILTER: 0, 1 or 2 */
    int ref_fallthrough;  /* 1 = reference window chain */
    int update;           /* 0 = centre pixel, 1 = sphere paint */
    int sphere_periodic;  /* sphere paint wraps (1) or clips (0) */
    float sphere_r_scale; /* R_cells = R * sphere_r_scale (1 = reference) */
    int fused;            /* 1 = fused pruned transforms where supported (update 0,
                             filter_type 1 or 2, D in {64,96,128,160,256,512}):
                             one launch per radius, every launch on `stream`;
                             0 = rocFFT c2r + separate filter/threshold kernels */
    int graph;            /* 1 = capture the loop into a HIP graph on the first call
                             and replay it for identical later calls (same
                             pointers, parameters, radii, denominators and a
                             non-null stream); ignored while profiling */
} cm21_bubble_params;

size_t cm21_find_bubbles_workspace_bytes(int D, int n_radii, int update);
int cm21_find_bubbles(cm21_plan* plan, const float* delta, float* ionized,
                      const float* radii, const float* denoms, int n_radii,
                      const cm21_bubble_params* params, void* work,
                      size_t work_bytes, void* stream);

/* No counterpart in the reference: a variant of cm21_find_bubbles for callers
 * that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  It allocates nothing: the
 * workspace is a static device array of CM21_NOALLOC_SCRATCH_BYTES that the
 * library holds on each GPU it runs on until the process ends, and the call
 * makes no hipMalloc, hipFree or allocating library call (the plan, created
 * beforehand, brings its own rocFFT work buffer).  Boxes whose workspace
 * (cm21_find_bubbles_workspace_bytes) exceeds the array are rejected with
 * CM21_ERR_NOMEM: every mode fits for D <= 128.  params->graph is ignored (the
 * launches go to `stream`, where a capturing caller records them).  Calls on
 * one GPU share the array: calls that overlap in time on one GPU race, so
 * order them on one stream. */
#define CM21_NOALLOC_SCRATCH_BYTES (64u << 20)
int cm21_find_bubbles_noalloc(cm21_plan* plan, const float* delta, float* ionized,
                              const float* radii, const float* denoms, int n_radii,
                              const cm21_bubble_params* params, void* stream);

/* ------------------------------------------------------------------ init
 * Replaces init_kernel + MRG32k3a normals + complex assembly
 * (Programs/initialize.cu:58-95; initialize.py:63-72): writes the c2r-ready
 * Hermitian half spectrum of a Gaussian random field with
 * P(k) = amp * k^n_index * exp(-(k/k_cut)^2), P(0) = 0, from Philox4x32-10
 * keyed by (seed, box_id).  Paired modes get (g1 + i g2) sqrt(P/V)/sqrt(2),
 * self-conjugate modes g1 sqrt(P/V), the DC mode 0. */
int cm21_grf_init(void* spec_out, int D, double box_len, double amp,
                  double n_index, double k_cut, uint32_t seed, uint32_t box_id,
                  void* stream);

/* Replaces init_kernel with its tabulated spectrum (Programs/initialize.cu:
 * 58-95; table of initialize.py:27-31): as cm21_grf_init, with P(k) (Mpc^3)
 * interpolated from n_table (K, P) pairs in device memory (float32, K
 * increasing, 2 <= n_table <= 2048): i is the first index with K[i] >= k and
 * P = P[i-1] + (P[i]-P[i-1]) (k-K[i-1])/(K[i]-K[i-1]); modes outside
 * [K[0], K[n_table-1]] get 0.  k = 2 pi |n| / box_len in FP64. */
int cm21_grf_init_table(void* spec_out, int D, double box_len, const float* k_table, const float* p_table,
                        int n_table, uint32_t seed, uint32_t box_id, void* stream);

/* Replaces subsample (Programs/initialize.cu:133-146): small[i,j,k] =
 * large[l(i),l(j),l(k)], l(i) = floor(float32(i*pixel_factor) + 0.5). */
int cm21_subsample(float* small, const float* large, int D, int sw,
                   float pixel_factor, void* stream);

/* Replaces set_velocity followed by HII_filter (Programs/initialize.cu:148-190,
 * initialize.py:103-104): spec_out = i k_comp / k^2 * spec_in, zero at k = 0
 * and on the Nyquist plane of axis comp, then the window of filter_type at R
 * (filter_type < 0: no window).  comp = 0, 1, 2 is the array axis. */
int cm21_velocity(void* spec_out, const void* spec_in, int D, float delta_k,
                  int comp, int filter_type, float R, void* stream);

/* ------------------------------------------------------------------ perturb
 * Replaces move_mass + the renormalisation (Programs/perturb_field.cu:90-133;
 * perturb_field.py:93-105): the DIM^3 particles at cell centres move by
 * v*disp_scale (box units; v at the HII cell floor(i/pixel_factor)) on the
 * sw^3 grid and deposit float(1 + growth_init*delta); scheme 0 = NGP
 * (reference truncation), 1 = cloud-in-cell.  updated = sum/MASS_FACTOR - 1. */
int cm21_move_mass(float* updated, const float* deltax, const float* vx,
                   const float* vy, const float* vz, int DIM, int sw,
                   float disp_scale, float growth_init, int scheme,
                   void* stream);

/* ------------------------------------------------------------------ delta_T
 * Replaces the host NumPy brightness temperature (Programs/delta_T.py:73-76):
 * dT = (tb_const*xH)*(1+delta) in float32; dev_sum[0] = FP64 sum of dT. */
int cm21_delta_T(float* dT, const float* xH, const float* delta, int D,
                 float tb_const, double* dev_sum, void* work, void* stream);

/* Replaces pbox_kernel and its host driver (Programs/delta_T.cu:5-29,
 * delta_T.py:80-98).  Transforms (field/mean - 1) V/N (times mean when
 * dimensional) and bins k^3 |d_k|^2 / (2 pi^2 V) by bin_lut[|n|^2] (an int32
 * table of 3(D/2)^2 + 1 entries: bin index, -1 or >= nbins to drop).  Writes
 * FP64 sums ps_sum[nbins], k_sum[nbins] and exact counts[nbins] (uint64) on
 * the device, deterministically.  1 <= nbins <= 512.  field_sum is the FP64 device sum of field.
 * `work` holds at least cm21_pk_workspace_bytes(D, nbins) bytes. */
size_t cm21_pk_workspace_bytes(int D, int nbins);
int cm21_power_spectrum(cm21_plan* plan, const float* field,
                        const double* field_sum, int D, double box_len,
                        const int32_t* bin_lut, int nbins, int dimensional,
                        double* ps_sum, double* k_sum, uint64_t* counts,
                        void* work, size_t work_bytes, void* stream);

/* ------------------------------------------------------------------ profiler
 * No reference counterpart (the reference records CUDA events it never reads,
 * Programs/find_HII_bubbles.py:133-177).  cm21_profile_enable(mask) makes the
 * launches of every stage s with bit s of mask set record a HIP event pair on
 * their stream (mask 0 turns profiling off, -1 enables all stages); collect
 * waits for the recorded events and returns per-stage total milliseconds and
 * launch counts (arrays of nstages entries, indexed by CM21_STAGE_*), then
 * clears the record. */
#define CM21_STAGE_R2C 0
#define CM21_STAGE_FILTER 1
#define CM21_STAGE_C2R 2
#define CM21_STAGE_FCOLL_SUM 3
#define CM21_STAGE_THRESHOLD 4
#define CM21_STAGE_SPHERE 5
#define CM21_STAGE_FUSED_COLUMNS 6         /* batched: columns pass of every (radius, box) */
#define CM21_STAGE_FUSED_ROWS 7            /* batched: rows pass */
#define CM21_STAGE_FUSED_LINES 8           /* batched: persistent lines pass */
#define CM21_STAGE_FUSED_LINES_THRESHOLD 9 /* single box: the final paint */
#define CM21_STAGE_FUSED_RADIUS 10         /* single box: one launch per radius (columns, rows,
                                              box sum, reduction and threshold roles of
                                              neighbouring radii) */
#define CM21_NSTAGES 11
int cm21_profile_enable(int stage_mask);
int cm21_profile_collect(double* ms, int64_t* count, int nstages);

/* ---- Batched find_HII_bubbles.  No reference counterpart: the reference
 * runs one box per conv_bubbles call (Programs/find_HII_bubbles.py:17-193,
 * driven box by box from Programs/Drives.py:7-14).  Boxes are grouped by side
 * D; every box of a group shares D, box_len and the radius list, and has its
 * own barrier, efficiency and denominators.  The per-cell arithmetic is that
 * of cm21_find_bubbles with params->fused = 1 (centre-pixel update, filter 1
 * or 2). */
typedef struct cm21_batch_box {
    float deltac;     /* barrier of this box (1.686 / D(z) for linear z=0 input) */
    float growth_inv; /* nonlinear input: deltac / 1.686 */
    double zeta;      /* ionizing efficiency */
    float fc_mean;    /* nonlinear input: mean collapsed fraction */
    int reserved;
} cm21_batch_box;

typedef struct cm21_batch_group {
    int D;                       /* 64, 96, 128, 160, 256 or 512 */
    int nbox;                    /* boxes in the group */
    float box_len;               /* Mpc, shared */
    int n_radii;                 /* shared radius list, host float32 */
    const float* radii;
    const float* denoms;         /* host, nbox x n_radii float32 */
    const cm21_batch_box* boxes; /* host, nbox */
    const float* delta;          /* device, nbox boxes of D^3 float32, contiguous */
    float* ionized;              /* device, nbox boxes */
    cm21_plan* plan;             /* cm21_plan_create(&p, D, nbox) */
    int tile_columns;            /* candidate index, -1 = default */
    int tile_lines;              /* candidate index, -1 = default */
} cm21_batch_group;

/* params: filter_type, ref_fallthrough, linear_input, update (must be 0) and
 * graph are used; the per-box fields come from the groups.  With graph = 1
 * and a non-null stream the launches after the table upload are captured
 * into a HIP graph on the first call and replayed for identical calls.
 * Returns 0 from the size query for invalid groups. */
size_t cm21_find_bubbles_batched_workspace_bytes(const cm21_batch_group* groups, int ngroups,
                                                 const cm21_bubble_params* params);
int cm21_find_bubbles_batched(const cm21_batch_group* groups, int ngroups, const cm21_bubble_params* params,
                              void* work, size_t work_bytes, void* stream);

/* Tile candidates per D: stage 0 = columns/rows passes, 1 = lines pass.
 * cm21_batch_tile_count returns -1 for an unsupported D. */
int cm21_batch_tile_count(int D, int stage);
const char* cm21_batch_tile_name(int D, int stage, int index);

/* Times every candidate of both stages on the caller's data (reps timed runs
 * each, after one warm run) and writes the fastest into each group's
 * tile_columns / tile_lines.  ms_out (optional) receives the mean time of
 * each candidate, group by group, columns candidates first.  Leaves the
 * outputs unspecified; synchronises the stream. */
int cm21_batch_autotune(cm21_batch_group* groups, int ngroups, const cm21_bubble_params* params, void* work,
                        size_t work_bytes, int reps, double* ms_out, void* stream);

/* ---- 3D U-Net surrogate layers.  No reference counterpart: a learned
 * stand-in for the excursion-set loop (Programs/find_HII_bubbles.py:17-193),
 * trained on its outputs.  Activations are bf16 (uint16 bit patterns),
 * channels-last (B, D, D, D, C), periodic boundaries; master weights fp32,
 * accumulation fp32.  3^3 conv weights are [cout][27][cin] with tap
 * t = 9 (dz+1) + 3 (dy+1) + (dx+1). */
typedef struct cm21_unet_conv {
    int B, D;            /* batch, spatial side (input = output) */
    int cin, cout;       /* channels; cin = c0 + (channels of x1) */
    int c0;              /* input channels read from x0 (the rest from x1) */
    int o0;              /* output channels written to y0 (the rest to y1) */
    const void* x0;      /* bf16 (B,D,D,D,c0) */
    const void* x1;      /* bf16 (B,D,D,D,cin-c0) or NULL */
    const void* mask;    /* bf16 like x0 or NULL: x0 * (mask > 0) (single input only) */
    const void* w;       /* bf16 [cout][27][cin] */
    const float* bias;   /* fp32 [cout] ([B][cout] if bias_per_sample) or NULL */
    int bias_per_sample;
    int relu;            /* 1: y = max(0, conv + bias) */
    void* y0;            /* bf16 (B,D,D,D,o0) */
    void* y1;            /* bf16 (B,D,D,D,cout-o0) or NULL */
    int algo;            /* 0 = automatic, 1 = one halo tile per block,
                            2 = sliding window along z (D >= 64),
                            3 = halo tile of 16 input channels at a time */
    const float* head;   /* fp32 w[16], b of a 1^3 head fused into this conv
                            (cout = 16, inference): prob receives
                            sigmoid(b + w . y) and y0 is not written; or NULL */
    float* prob;         /* fp32 (B,D,D,D) when head is set */
    void* pool;          /* bf16 (B,D/2,D/2,D/2,cout) or NULL: also receives the
                            2x2x2 mean of y0, equal to cm21_unet_pool_fwd of y0
                            (fused into the sliding form for 16 -> 16 channels;
                            needs o0 = cout and no head) */
} cm21_unet_conv;
/* y = conv(x) on the matrix cores; (cin, cout) in {(16,16), (16,32), (32,16),
 * (32,32), (32,64), (64,32), (64,64)}, D a multiple of 16.  The input
 * gradient of a conv is this call on the masked output gradient with the
 * weights from cm21_unet_pack's transposed copy. */
int cm21_unet_conv3_fwd(const cm21_unet_conv* c, void* stream);

typedef struct cm21_unet_wgrad {
    int B, D, cin, cout, c0;
    const void* x0;      /* forward input (bf16), c0 channels */
    const void* x1;      /* forward input, cin - c0 channels, or NULL */
    const void* g;       /* output gradient (bf16, cout channels) */
    const void* mask;    /* forward output (bf16) or NULL: g * (mask > 0) */
    float* dw;           /* fp32 [cout][27][cin], overwritten */
    float* db;           /* fp32 [cout] or NULL, overwritten */
    void* work;
    size_t work_bytes;
} cm21_unet_wgrad;
size_t cm21_unet_wgrad_workspace_bytes(int B, int D, int cin, int cout);
int cm21_unet_conv3_wgrad(const cm21_unet_wgrad* w, void* stream);

/* 2x2x2 mean pooling: x (B,D,D,D,C) -> y (B,D/2,D/2,D/2,C), C a multiple
 * of 8 for the forward; backward dx (+)= dy/8 broadcast (accumulate = 1 adds
 * to dx). */
int cm21_unet_pool_fwd(const void* x, void* y, int B, int D, int C, void* stream);
int cm21_unet_pool_bwd(const void* dy, void* dx, int B, int D, int C, int accumulate, void* stream);

/* 2x2x2 stride-2 transposed conv on the matrix cores, x (B,D,D,D,cin) ->
 * y (B,2D,2D,2D,cout); master weights fp32 w [cout][8][cin] with sub-voxel
 * s = 4a + 2b + c, packed by cm21_unet_pack_up into wn (bf16 [8 cout][cin],
 * forward) and wk (bf16 [cin][8 cout], input gradient); (cin, cout) in
 * {(64,32), (32,16)}.  Backward: dx (or NULL), dw and db overwritten. */
int cm21_unet_pack_up(const float* w, void* wn, void* wk, int cin, int cout, void* stream);
int cm21_unet_up_fwd(const void* x, const void* wn, const float* bias, void* y, int B, int D, int cin, int cout,
                     void* stream);
size_t cm21_unet_up_workspace_bytes(int B, int D, int cin, int cout);
int cm21_unet_up_bwd(const void* x, const void* dy, const void* wk, void* dx, float* dw, float* db, int B, int D,
                     int cin, int cout, void* work, size_t work_bytes, void* stream);

/* Input layer: 16 channels from fp32 density delta (B,D,D,D) scaled by
 * delta_scale and two constant conditioning channels cond[B][2]
 * (w fp32 [16][27][3]); beff[B][16] receives the per-sample bias. */
int cm21_unet_in_fwd(const float* delta, const float* cond, const float* w, const float* bias, float* beff,
                     void* y, int B, int D, float delta_scale, void* stream);
size_t cm21_unet_in_workspace_bytes(int B, int D);
int cm21_unet_in_wgrad(const float* delta, const float* cond, const void* g, const void* y, float* dw, float* db,
                       int B, int D, float delta_scale, void* work, size_t work_bytes, void* stream);

/* Head: p = sigmoid(b + w . x) (x bf16 (B,D,D,D,16)) into prob (fp32, or
 * NULL).  With a target (fp32 in [0,1]): loss = mean binary cross entropy
 * (FP64, *loss), dx = dloss/dx (bf16, or NULL), dwb[0..15] = dloss/dw,
 * dwb[16] = dloss/db. */
size_t cm21_unet_head_workspace_bytes(void);
int cm21_unet_head(const void* x, const float* w, const float* bias, const float* target, float* prob, void* dx,
                   float* dwb, double* loss, int B, int D, void* work, size_t work_bytes, void* stream);

/* bf16 copies of fp32 conv weights w [cout][27][cin]: wb same layout, wt the
 * transposed, flipped kernel [cin][27][cout] (input gradient); either NULL. */
int cm21_unet_pack(const float* w, void* wb, void* wt, int cin, int cout, void* stream);

/* AdamW (decoupled weight decay) over n fp32 parameters, step >= 1. */
int cm21_adamw(float* p, const float* g, float* m, float* v, size_t n, float lr, float beta1, float beta2, float eps,
               float weight_decay, int step, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* CM21_H */
