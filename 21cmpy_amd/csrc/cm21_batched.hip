// Batched find_HII_bubbles: many boxes, all radii, few launches.
//
// Boxes are grouped by side D (one group per D).  For a group of B boxes and
// n_R radii the whole excursion-set loop is
//   r2c        one batched rocFFT forward transform of the B boxes;
//   columns    ONE launch covering every (radius, box): window + inverse DFT
//              along axis 0 of the surviving columns of that radius;
//   rows       ONE launch covering every (radius, box): inverse DFT along
//              axis 1, in place;
//   [sum]      nonlinear input only: one launch of the last-axis transform of
//              every (radius, box) with FP64 block partials of f_coll, then a
//              fixed-order reduction per (box, radius);
//   thresholds one thread per (box, radius): box-mean scale and the
//              density-space threshold;
//   lines      ONE persistent launch: each block owns L lines of one box and
//              walks all radii, transforming its lines back to real space per
//              radius and OR-ing the centre-pixel decisions into an
//              ionization mask that stays in LDS; the final radius paints the
//              float output.  A block whose cells are all ionized stops
//              transforming and writes 1.
// Per radius r only the prefix c <= cmax_r of each line can be non-zero after
// the k-space top-hat, so the intermediate of radius r is stored compactly as
// T_r[p][q][c] with row length hs_r = cmax_r + 1; all radii of all boxes stay
// resident in HBM between the passes.
//
// The arithmetic per cell is that of the single-box loop (same device bodies,
// cm21_fused_dev.hpp), so for equal tile factorisations a batched box equals
// the same box run alone up to the forward transform (a batched rocFFT plan).
//
// Tile shapes (columns/rows: D1 x D2 factorisation and W columns per block;
// lines: N1 x N2 factorisation and L lines per block) come from a candidate
// table per D; cm21_batch_autotune times the candidates on the caller's data
// and returns the fastest.
#include <algorithm>
#include <cstring>
#include <mutex>
#include <unordered_map>
#include <vector>

#include "cm21.h"
#include "cm21_fused_dev.hpp"

namespace cm21 {
namespace {

struct BatchRadius {
    FusedParams fp;
    unsigned long long toff;  // float2 offset of the radius' compact intermediate in a box's region
    int hs;                   // row length of the compact intermediate (cmax + 1)
    int ncg;                  // column groups of the columns/rows passes (depends on W)
    int pad[2];
};

struct BatchCell {  // one (box, radius)
    float deltac, denom, fc_mean, scale;
    double zeta, inv_zeta;
    DensityThreshold th;
    int use_scale;  // nonlinear input: f_coll is scaled by the box-mean factor
    int pad;
};

struct BatchLaunch {
    const float2* S;
    float2* T;
    const float2* tw;
    const BatchRadius* rad;
    const int* cblk;  // nr+1 first flattened blocks of the columns launch
    const int* rblk;  // nr+1 first flattened blocks of the rows launch
    BatchCell* cells;
    const float* bscale;  // per box: window scale (1/N, times the growth factor of nonlinear input)
    double* partial;
    double* sums;
    float* out;
    unsigned long long nc;    // half-spectrum entries per box
    unsigned long long tbox;  // float2 entries of a box's intermediate (all radii)
    int nr, nbox, cblocks, rblocks, nblk;
};

__device__ __forceinline__ int find_radius(const int* __restrict__ blk, int nr, int x) {
    int lo = 0, hi = nr - 1;  // last r with blk[r] <= x
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (blk[mid] <= x)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

template <int D1, int D2, int W>
__global__ __launch_bounds__(W* D1) void k_batch_columns(BatchLaunch a) {
    constexpr int D = D1 * D2;
    __shared__ float2 lds[W * (D + 1)];
    const int box = blockIdx.y;
    const int x = blockIdx.x;
    const int r = find_radius(a.cblk, a.nr, x);
    const int local = x - a.cblk[r];
    const BatchRadius& R = a.rad[r];
    const int cg = local % R.ncg, slot = local / R.ncg;
    columns_body<D1, D2, W>(a.S + box * a.nc, a.T + box * a.tbox + R.toff, R.hs, a.tw, R.fp, a.bscale[box], slot,
                            cg, lds);
}

template <int D1, int D2, int W>
__global__ __launch_bounds__(W* D1) void k_batch_rows(BatchLaunch a) {
    constexpr int D = D1 * D2;
    __shared__ float2 lds[W * (D + 1)];
    const int box = blockIdx.y;
    const int x = blockIdx.x;
    const int r = find_radius(a.rblk, a.nr, x);
    const int local = x - a.rblk[r];
    const BatchRadius& R = a.rad[r];
    const int cg = local % R.ncg, p = local / R.ncg;
    rows_body<D1, D2, W>(a.T + box * a.tbox + R.toff, R.hs, a.tw, R.fp, p, cg, lds);
}

// Nonlinear input: FP64 block partials of f_coll for every (box, radius);
// grid (D*D/L, nr, nbox).
template <int N1, int N2, int L>
__global__ __launch_bounds__(L* N1) void k_batch_lines_sum(BatchLaunch a) {
    using C = LinesCore<N1, N2, L>;
    constexpr int D = C::D, NT = C::NT;
    __shared__ float2 buf[L * C::HP];
    __shared__ float2 stw[D];
    const int t = threadIdx.x;
    const int r = blockIdx.y, box = blockIdx.z;
    const size_t line0 = (size_t)blockIdx.x * L;
    const BatchRadius& R = a.rad[r];
    float2 tmp[C::LOADS];
    C::load(a.T + box * a.tbox + R.toff + line0 * R.hs, R.hs, R.hs, tmp);
    for (int q = t; q < D; q += NT) stw[q] = a.tw[q];
    C::store(tmp, buf);
    __syncthreads();
    C::transform(buf, stw);
    const BatchCell cell = a.cells[box * a.nr + r];
    const float* xrf = reinterpret_cast<const float*>(buf);
    double acc = 0.0;
    const float inv_denom = 1.0f / cell.denom;
#pragma unroll 4
    for (int f = t; f < L * D; f += NT) {
        const int i = f / D, rr = f - i * D;
        acc += (double)fcoll_for_sum(xrf[i * 2 * C::HP + rr], cell.deltac, cell.denom, inv_denom);
    }
    acc = block_sum(acc);
    if (t == 0) a.partial[((size_t)box * a.nr + r) * a.nblk + blockIdx.x] = acc;
}

// Fixed-order sum of the partials of one (box, radius) per block.
__global__ void k_batch_sum(BatchLaunch a) {
    const double* p = a.partial + (size_t)blockIdx.x * a.nblk;
    double acc = 0.0;
    for (int i = threadIdx.x; i < a.nblk; i += blockDim.x) acc += p[i];
    acc = block_sum(acc);
    if (threadIdx.x == 0) a.sums[blockIdx.x] = acc;
}

// Box-mean scale (nonlinear input) and density-space threshold per (box, radius).
__global__ void k_batch_thresholds(BatchLaunch a, double ncells) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.nbox * a.nr) return;
    BatchCell c = a.cells[i];
    float scale = 1.0f;
    if (c.use_scale) scale = (float)((double)c.fc_mean / (a.sums[i] / ncells));
    c.scale = scale;
    c.th = density_threshold(c.deltac, c.denom, c.zeta, scale);
    a.cells[i] = c;
}

// Persistent threshold pass: grid (D*D/L, nbox); each block walks all radii
// for its L lines with the ionization mask of its L*D cells in LDS.
template <int N1, int N2, int L>
__device__ __forceinline__ void batch_line_transform(const BatchLaunch& a, const float2* Tb, size_t line0, int r,
                                                     float2* buf, const float2* stw) {
    using C = LinesCore<N1, N2, L>;
    const BatchRadius& R = a.rad[r];
    float2 tmp[C::LOADS];
    C::load(Tb + R.toff + line0 * R.hs, R.hs, R.hs, tmp);
    C::store(tmp, buf);
    __syncthreads();
    C::transform(buf, stw);
}

template <int N1, int N2, int L>
__global__ __launch_bounds__(L* N1) void k_batch_lines(BatchLaunch a) {
    using C = LinesCore<N1, N2, L>;
    constexpr int D = C::D, NT = C::NT, WORDS = C::WORDS;
    __shared__ float2 buf[L * C::HP];
    __shared__ float2 stw[D];
    __shared__ unsigned long long smask[WORDS];
    const int t = threadIdx.x;
    const int box = blockIdx.y;
    const size_t line0 = (size_t)blockIdx.x * L;
    const float2* Tb = a.T + box * a.tbox;
    const BatchCell* cb = a.cells + (size_t)box * a.nr;
    float* ob = a.out + (size_t)box * D * D * D + line0 * D;
    for (int q = t; q < D; q += NT) stw[q] = a.tw[q];
    if (t < WORDS) smask[t] = 0ull;
    const int lane = t & 63;
    const float* xrf = reinterpret_cast<const float*>(buf);
    for (int r = 0; r < a.nr - 1; ++r) {
        batch_line_transform<N1, N2, L>(a, Tb, line0, r, buf, stw);
        const BatchCell& cell = cb[r];
        const bool use_scale = cell.use_scale != 0;
#pragma unroll 4
        for (int f = t; f < L * D; f += NT) {
            const int i = f / D, rr = f - i * D;
            const float delta = xrf[i * 2 * C::HP + rr];
            const bool hit = cell_hit(delta, cell.th, cell.deltac, cell.denom, use_scale, cell.scale, cell.inv_zeta);
            const unsigned long long bits = __ballot(hit);
            if (lane == 0) smask[f >> 6] |= bits;
        }
        __syncthreads();  // mask words complete, buf free for the next radius
        // every cell of the block ionized: the final paint is 1 everywhere
        if (__syncthreads_and(t >= WORDS || smask[t] == ~0ull)) {
            for (int f = t; f < L * D; f += NT) ob[f] = 1.0f;
            return;
        }
    }
    const int r = a.nr - 1;
    batch_line_transform<N1, N2, L>(a, Tb, line0, r, buf, stw);
    const BatchCell& cell = cb[r];
    const bool use_scale = cell.use_scale != 0;
#pragma unroll 4
    for (int f = t; f < L * D; f += NT) {
        const int i = f / D, rr = f - i * D;
        const float delta = xrf[i * 2 * C::HP + rr];
        const bool was = (smask[f >> 6] >> (f & 63)) & 1ull;
        float val = 1.0f;
        if (!was && !cell_hit(delta, cell.th, cell.deltac, cell.denom, use_scale, cell.scale, cell.inv_zeta))
            val = final_value(delta, cell.deltac, cell.denom, use_scale, cell.scale, cell.zeta, cell.inv_zeta);
        ob[f] = val;
    }
}

// ------------------------------------------------------------------ tiles

using LaunchFn = int (*)(const BatchLaunch&, hipStream_t);

template <int D1, int D2, int W>
int launch_bcols(const BatchLaunch& a, hipStream_t st) {
    k_batch_columns<D1, D2, W><<<dim3(a.cblocks, a.nbox), W * D1, 0, st>>>(a);
    CM21_LAUNCH_CHECK("k_batch_columns");
    return CM21_OK;
}
template <int D1, int D2, int W>
int launch_brows(const BatchLaunch& a, hipStream_t st) {
    k_batch_rows<D1, D2, W><<<dim3(a.rblocks, a.nbox), W * D1, 0, st>>>(a);
    CM21_LAUNCH_CHECK("k_batch_rows");
    return CM21_OK;
}
template <int N1, int N2, int L>
int launch_blines(const BatchLaunch& a, hipStream_t st) {
    constexpr int D = 2 * N1 * N2;
    k_batch_lines<N1, N2, L><<<dim3(D * D / L, a.nbox), L * N1, 0, st>>>(a);
    CM21_LAUNCH_CHECK("k_batch_lines");
    return CM21_OK;
}
template <int N1, int N2, int L>
int launch_bsum(const BatchLaunch& a, hipStream_t st) {
    constexpr int D = 2 * N1 * N2;
    k_batch_lines_sum<N1, N2, L><<<dim3(D * D / L, a.nr, a.nbox), L * N1, 0, st>>>(a);
    CM21_LAUNCH_CHECK("k_batch_lines_sum");
    return CM21_OK;
}

struct ColTile {
    int W;
    LaunchFn cols, rows;
    const char* name;
};
struct LineTile {
    int L;
    LaunchFn lines, sum;
    const char* name;
};
struct BatchShape {
    int D;
    ColTile cols[3];
    int ncols;
    LineTile lines[4];
    int nlines;
};

#define CM21_COL(D1, D2, W) {W, launch_bcols<D1, D2, W>, launch_brows<D1, D2, W>, #D1 "x" #D2 "/W" #W}
#define CM21_LINE(N1, N2, L) {L, launch_blines<N1, N2, L>, launch_bsum<N1, N2, L>, #N1 "x" #N2 "/L" #L}

// Candidate tiles per D; the first of each list is the default (the shape of
// the single-box loop, so that defaults reproduce it exactly).
const BatchShape kBatchShapes[] = {
    {64, {CM21_COL(8, 8, 32), CM21_COL(8, 8, 16), CM21_COL(16, 4, 16)}, 3,
     {CM21_LINE(4, 8, 64), CM21_LINE(8, 4, 32), CM21_LINE(4, 8, 32)}, 3},
    {96, {CM21_COL(8, 12, 32), CM21_COL(8, 12, 16), CM21_COL(16, 6, 16)}, 3,
     {CM21_LINE(4, 12, 64), CM21_LINE(8, 6, 32), CM21_LINE(4, 12, 32), CM21_LINE(6, 8, 32)}, 4},
    {128, {CM21_COL(8, 16, 32), CM21_COL(16, 8, 16), CM21_COL(8, 16, 16)}, 3,
     {CM21_LINE(8, 8, 32), CM21_LINE(4, 16, 64), CM21_LINE(8, 8, 16), CM21_LINE(16, 4, 16)}, 4},
    {160, {CM21_COL(16, 10, 16), CM21_COL(8, 20, 32), CM21_COL(16, 10, 8)}, 3,
     {CM21_LINE(8, 10, 32), CM21_LINE(4, 20, 64), CM21_LINE(8, 10, 16), CM21_LINE(16, 5, 16)}, 4},
    {256, {CM21_COL(16, 16, 16), CM21_COL(16, 16, 8)}, 2, {CM21_LINE(16, 8, 16), CM21_LINE(8, 16, 32)}, 2},
    {512, {CM21_COL(16, 32, 8)}, 1, {CM21_LINE(16, 16, 8)}, 1},
};

#undef CM21_COL
#undef CM21_LINE

const BatchShape* batch_shape(int D) {
    for (const auto& s : kBatchShapes)
        if (s.D == D) return &s;
    return nullptr;
}

int min_lines_per_block(const BatchShape& s) {
    int m = s.lines[0].L;
    for (int i = 1; i < s.nlines; ++i) m = std::min(m, s.lines[i].L);
    return m;
}

// ------------------------------------------------------------------ host plan

size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

struct GroupHost {
    const cm21_batch_group* g;
    const BatchShape* shape;
    int D, nbox, nr;
    size_t N, Nc, tbox;
    std::vector<BatchRadius> rad;
    // table offsets inside the blob
    size_t o_rad, o_cblk, o_rblk, o_cells, o_bscale;
    // device buffers inside the work region
    size_t o_S, o_T, o_partial, o_sums;
    int nblk_max;
};

struct BatchLayout {
    std::vector<GroupHost> groups;
    size_t blob_bytes = 0;  // tables, at the start of the workspace
    size_t bytes = 0;       // whole workspace
};

int build_layout(const cm21_batch_group* groups, int ngroups, const cm21_bubble_params* prm, BatchLayout& lay,
                 bool check_arrays) {
    CM21_REQUIRE(groups && prm, "cm21_find_bubbles_batched: NULL argument");
    CM21_REQUIRE(ngroups >= 1 && ngroups <= 64, "cm21_find_bubbles_batched: ngroups=%d out of range", ngroups);
    CM21_REQUIRE(prm->update == 0, "cm21_find_bubbles_batched: only the centre-pixel update (0) is batched");
    CM21_REQUIRE(prm->filter_type == 1 || prm->filter_type == 2,
                 "cm21_find_bubbles_batched: filter_type=%d (1 or 2 are batched)", prm->filter_type);
    lay.groups.clear();
    size_t blob = 0;
    for (int gi = 0; gi < ngroups; ++gi) {
        const cm21_batch_group& g = groups[gi];
        GroupHost h{};
        h.g = &g;
        h.shape = batch_shape(g.D);
        CM21_REQUIRE(h.shape, "cm21_find_bubbles_batched: group %d: D=%d has no batched tiles", gi, g.D);
        CM21_REQUIRE(g.nbox >= 1 && g.nbox <= 4096, "cm21_find_bubbles_batched: group %d: nbox=%d", gi, g.nbox);
        CM21_REQUIRE(g.n_radii >= 1 && g.n_radii <= 1024, "cm21_find_bubbles_batched: group %d: n_radii=%d", gi,
                     g.n_radii);
        CM21_REQUIRE(g.box_len > 0.0f, "cm21_find_bubbles_batched: group %d: box_len must be positive", gi);
        CM21_REQUIRE(g.radii, "cm21_find_bubbles_batched: group %d: radii is NULL", gi);
        CM21_REQUIRE(g.tile_columns >= -1 && g.tile_columns < h.shape->ncols,
                     "cm21_find_bubbles_batched: group %d: tile_columns=%d out of range", gi, g.tile_columns);
        CM21_REQUIRE(g.tile_lines >= -1 && g.tile_lines < h.shape->nlines,
                     "cm21_find_bubbles_batched: group %d: tile_lines=%d out of range", gi, g.tile_lines);
        if (check_arrays) {
            CM21_REQUIRE(g.delta && g.ionized && g.denoms && g.boxes && g.plan,
                         "cm21_find_bubbles_batched: group %d: NULL pointer", gi);
            CM21_REQUIRE(plan_dim(g.plan) == g.D && plan_batch(g.plan) == g.nbox,
                         "cm21_find_bubbles_batched: group %d: plan is (D=%d, batch=%d), need (D=%d, batch=%d)", gi,
                         plan_dim(g.plan), plan_batch(g.plan), g.D, g.nbox);
            for (int b = 0; b < g.nbox; ++b) {
                CM21_REQUIRE(g.boxes[b].zeta > 0.0, "cm21_find_bubbles_batched: group %d box %d: zeta <= 0", gi, b);
                for (int r = 0; r < g.n_radii; ++r) {
                    const float dn = g.denoms[(size_t)b * g.n_radii + r];
                    if (!(dn > 0.0f) || !std::isfinite(dn))
                        return fail(CM21_ERR_DOMAIN,
                                    "cm21_find_bubbles_batched: group %d box %d: denominator %g at R=%g is not "
                                    "positive and finite (sigma_min^2 <= S(R))",
                                    gi, b, (double)dn, (double)g.radii[r]);
                }
            }
        }
        h.D = g.D;
        h.nbox = g.nbox;
        h.nr = g.n_radii;
        h.N = (size_t)g.D * g.D * g.D;
        h.Nc = (size_t)g.D * g.D * (g.D / 2 + 1);
        const float dk = (float)(2.0 * M_PI / (double)g.box_len);
        h.rad.resize(h.nr);
        size_t toff = 0;
        for (int r = 0; r < h.nr; ++r) {
            BatchRadius& R = h.rad[r];
            std::memset(&R, 0, sizeof(R));
            R.fp = fused_params(g.D, dk, g.radii[r], prm->filter_type, prm->ref_fallthrough, 1.0f);
            R.hs = R.fp.cmax + 1;
            R.toff = toff;
            toff += (size_t)g.D * g.D * R.hs;
        }
        h.tbox = toff;
        h.nblk_max = g.D * g.D / min_lines_per_block(*h.shape);
        h.o_rad = blob;
        blob = align256(blob + sizeof(BatchRadius) * h.nr);
        h.o_cblk = blob;
        blob = align256(blob + sizeof(int) * (h.nr + 1));
        h.o_rblk = blob;
        blob = align256(blob + sizeof(int) * (h.nr + 1));
        h.o_cells = blob;
        blob = align256(blob + sizeof(BatchCell) * (size_t)h.nbox * h.nr);
        h.o_bscale = blob;
        blob = align256(blob + sizeof(float) * h.nbox);
        lay.groups.push_back(std::move(h));
    }
    lay.blob_bytes = blob;
    size_t off = blob;
    for (auto& h : lay.groups) {
        h.o_S = off;
        off = align256(off + sizeof(float2) * h.Nc * h.nbox);
        h.o_T = off;
        off = align256(off + sizeof(float2) * h.tbox * h.nbox);
        h.o_sums = off;
        off = align256(off + sizeof(double) * (size_t)h.nbox * h.nr);
        h.o_partial = off;
        if (!prm->linear_input) off = align256(off + sizeof(double) * (size_t)h.nbox * h.nr * h.nblk_max);
    }
    lay.bytes = off;
    return CM21_OK;
}

// Host tables of one group for the given columns tile, written into the blob.
void fill_tables(const GroupHost& h, const cm21_bubble_params* prm, int col_tile, std::vector<char>& blob,
                 int* cblocks, int* rblocks) {
    const cm21_batch_group& g = *h.g;
    const int W = h.shape->cols[col_tile].W;
    const int D = h.D, H = D / 2 + 1;
    BatchRadius* rad = (BatchRadius*)(blob.data() + h.o_rad);
    int* cblk = (int*)(blob.data() + h.o_cblk);
    int* rblk = (int*)(blob.data() + h.o_rblk);
    int ctot = 0, rtot = 0;
    for (int r = 0; r < h.nr; ++r) {
        BatchRadius R = h.rad[r];
        int ncg = R.fp.cmax / W + 1;
        const int ncg_all = (H + W - 1) / W;
        if (ncg > ncg_all) ncg = ncg_all;
        R.ncg = ncg;
        const int bpos = R.fp.bm < D / 2 ? R.fp.bm : D / 2;
        const int bneg = R.fp.bm < D / 2 - 1 ? R.fp.bm : D / 2 - 1;
        const int nb = bpos + 1 + bneg;
        rad[r] = R;
        cblk[r] = ctot;
        rblk[r] = rtot;
        ctot += ncg * nb;
        rtot += ncg * D;
    }
    cblk[h.nr] = ctot;
    rblk[h.nr] = rtot;
    *cblocks = ctot;
    *rblocks = rtot;
    BatchCell* cells = (BatchCell*)(blob.data() + h.o_cells);
    float* bscale = (float*)(blob.data() + h.o_bscale);
    for (int b = 0; b < h.nbox; ++b) {
        const cm21_batch_box& bx = g.boxes[b];
        bscale[b] = prm->linear_input ? (float)(1.0 / (double)h.N) : (float)((double)bx.growth_inv / (double)h.N);
        for (int r = 0; r < h.nr; ++r) {
            BatchCell c{};
            c.deltac = bx.deltac;
            c.denom = g.denoms[(size_t)b * h.nr + r];
            c.fc_mean = bx.fc_mean;
            c.scale = 1.0f;
            c.zeta = bx.zeta;
            c.inv_zeta = 1.0 / bx.zeta;
            c.use_scale = prm->linear_input ? 0 : 1;
            cells[(size_t)b * h.nr + r] = c;
        }
    }
}

BatchLaunch make_launch(const GroupHost& h, char* work, int cblocks, int rblocks, int nblk) {
    BatchLaunch a{};
    a.S = (const float2*)(work + h.o_S);
    a.T = (float2*)(work + h.o_T);
    a.tw = plan_twiddles(h.g->plan);
    a.rad = (const BatchRadius*)(work + h.o_rad);
    a.cblk = (const int*)(work + h.o_cblk);
    a.rblk = (const int*)(work + h.o_rblk);
    a.cells = (BatchCell*)(work + h.o_cells);
    a.bscale = (const float*)(work + h.o_bscale);
    a.partial = (double*)(work + h.o_partial);
    a.sums = (double*)(work + h.o_sums);
    a.out = h.g->ionized;
    a.nc = h.Nc;
    a.tbox = h.tbox;
    a.nr = h.nr;
    a.nbox = h.nbox;
    a.cblocks = cblocks;
    a.rblocks = rblocks;
    a.nblk = nblk;
    return a;
}

int tile_or_default(int t) { return t < 0 ? 0 : t; }

// Stages of one group after the tables are in place.
int run_forward(const GroupHost& h, hipStream_t st, char* work) {
    return cm21_r2c(h.g->plan, h.g->delta, work + h.o_S, (void*)st);
}

int run_transforms(const GroupHost& h, const BatchLaunch& a, int col_tile, hipStream_t st) {
    const ColTile& ct = h.shape->cols[col_tile];
    ProfToken t = prof_begin(CM21_STAGE_FUSED_COLUMNS, st);
    CM21_TRY(ct.cols(a, st));
    prof_end(t, st);
    t = prof_begin(CM21_STAGE_FUSED_ROWS, st);
    CM21_TRY(ct.rows(a, st));
    prof_end(t, st);
    return CM21_OK;
}

int run_threshold(const GroupHost& h, const BatchLaunch& a, int line_tile, const cm21_bubble_params* prm,
                  hipStream_t st) {
    const LineTile& lt = h.shape->lines[line_tile];
    if (!prm->linear_input) {
        ProfToken t = prof_begin(CM21_STAGE_FCOLL_SUM, st);
        CM21_TRY(lt.sum(a, st));
        k_batch_sum<<<a.nbox * a.nr, kThreads, 0, st>>>(a);
        CM21_LAUNCH_CHECK("k_batch_sum");
        prof_end(t, st);
    }
    const int n = a.nbox * a.nr;
    k_batch_thresholds<<<(n + 63) / 64, 64, 0, st>>>(a, (double)h.N);
    CM21_LAUNCH_CHECK("k_batch_thresholds");
    ProfToken t = prof_begin(CM21_STAGE_FUSED_LINES, st);
    CM21_TRY(lt.lines(a, st));
    prof_end(t, st);
    return CM21_OK;
}

struct Prepared {
    BatchLayout lay;
    std::vector<char> blob;
    std::vector<BatchLaunch> launches;
    std::vector<int> col_tile, line_tile;
};

int prepare(const cm21_batch_group* groups, int ngroups, const cm21_bubble_params* prm, void* work,
            size_t work_bytes, Prepared& p) {
    CM21_TRY(build_layout(groups, ngroups, prm, p.lay, true));
    CM21_REQUIRE(work, "cm21_find_bubbles_batched: work is NULL");
    if (work_bytes < p.lay.bytes)
        return fail(CM21_ERR_NOMEM, "cm21_find_bubbles_batched: workspace %zu bytes < required %zu", work_bytes,
                    p.lay.bytes);
    p.blob.assign(p.lay.blob_bytes, 0);
    p.launches.clear();
    p.col_tile.clear();
    p.line_tile.clear();
    for (const auto& h : p.lay.groups) {
        const int ct = tile_or_default(h.g->tile_columns);
        const int lt = tile_or_default(h.g->tile_lines);
        int cb = 0, rb = 0;
        fill_tables(h, prm, ct, p.blob, &cb, &rb);
        p.launches.push_back(make_launch(h, (char*)work, cb, rb, h.D * h.D / h.shape->lines[lt].L));
        p.col_tile.push_back(ct);
        p.line_tile.push_back(lt);
    }
    return CM21_OK;
}

int upload(const Prepared& p, void* work, hipStream_t st) {
    CM21_HIP_CHECK(hipMemcpyAsync(work, p.blob.data(), p.blob.size(), hipMemcpyHostToDevice, st));
    return CM21_OK;
}

int run_group(const Prepared& p, size_t i, const cm21_bubble_params* prm, void* work, hipStream_t gs) {
    const GroupHost& h = p.lay.groups[i];
    ProfToken t = prof_begin(CM21_STAGE_R2C, gs);
    CM21_TRY(run_forward(h, gs, (char*)work));
    prof_end(t, gs);
    CM21_TRY(run_transforms(h, p.launches[i], p.col_tile[i], gs));
    CM21_TRY(run_threshold(h, p.launches[i], p.line_tile[i], prm, gs));
    return CM21_OK;
}

// Groups are independent: the first runs on the caller's stream, the others
// on library helper streams forked after the table upload and joined back at
// the end, so the small passes of one group overlap the persistent passes of
// another.  Groups that share a plan (same D and box count, different
// box_len) share its rocFFT work buffer, so they go to the same stream and
// their forward transforms run one after the other.
int run_all(const Prepared& p, const cm21_bubble_params* prm, void* work, hipStream_t st) {
    const int G = (int)p.lay.groups.size();
    if (G == 1) return run_group(p, 0, prm, work, st);
    SideStreams* ss = nullptr;
    CM21_TRY(side_streams(&ss));
    std::lock_guard<std::mutex> lk(ss->mu);
    // lane[i]: 0 = the caller's stream, k >= 1 = helper stream k-1
    std::vector<int> lane(G, 0);
    int next = 0;
    for (int i = 1; i < G; ++i) {
        int shared = -1;
        for (int j = 0; j < i && shared < 0; ++j)
            if (p.lay.groups[j].g->plan == p.lay.groups[i].g->plan) shared = j;
        if (shared >= 0) {
            lane[i] = lane[shared];
        } else {
            lane[i] = 1 + next % SideStreams::kStreams;
            ++next;
        }
    }
    const int nside = std::min(next, SideStreams::kStreams);
    CM21_HIP_CHECK(hipEventRecord(ss->fork, st));
    for (int k = 0; k < nside; ++k) CM21_HIP_CHECK(hipStreamWaitEvent(ss->s[k], ss->fork, 0));
    for (int i = 0; i < G; ++i) CM21_TRY(run_group(p, i, prm, work, lane[i] == 0 ? st : ss->s[lane[i] - 1]));
    for (int k = 0; k < nside; ++k) {
        CM21_HIP_CHECK(hipEventRecord(ss->join[k], ss->s[k]));
        CM21_HIP_CHECK(hipStreamWaitEvent(st, ss->join[k], 0));
    }
    return CM21_OK;
}

// Graph cache: the launches after the table upload, keyed by everything that
// reaches a kernel argument or a table.
struct BatchGraph {
    std::vector<char> key;
    hipGraphExec_t exec = nullptr;
};
std::mutex g_bgraph_mu;
std::unordered_map<hipStream_t, BatchGraph> g_bgraphs;

void append(std::vector<char>& k, const void* p, size_t n) {
    const char* c = (const char*)p;
    k.insert(k.end(), c, c + n);
}

std::vector<char> graph_key(const Prepared& p, const cm21_batch_group* groups, int ngroups,
                            const cm21_bubble_params* prm, void* work) {
    std::vector<char> k;
    append(k, &work, sizeof(work));
    append(k, prm, sizeof(*prm));
    append(k, &ngroups, sizeof(ngroups));
    for (int i = 0; i < ngroups; ++i) append(k, &groups[i], sizeof(groups[i]));
    append(k, p.blob.data(), p.blob.size());
    return k;
}

int graph_run(const Prepared& p, const cm21_batch_group* groups, int ngroups, const cm21_bubble_params* prm,
              void* work, hipStream_t st) {
    std::vector<char> key = graph_key(p, groups, ngroups, prm, work);
    std::lock_guard<std::mutex> lk(g_bgraph_mu);
    BatchGraph& e = g_bgraphs[st];
    if (e.exec && e.key == key) {
        CM21_HIP_CHECK(hipGraphLaunch(e.exec, st));
        return CM21_OK;
    }
    if (e.exec) {
        (void)hipGraphExecDestroy(e.exec);
        e.exec = nullptr;
    }
    if (hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal) != hipSuccess) {
        (void)hipGetLastError();
        return run_all(p, prm, work, st);
    }
    const int rc = run_all(p, prm, work, st);
    hipGraph_t graph = nullptr;
    const hipError_t ec = hipStreamEndCapture(st, &graph);
    hipGraphExec_t exec = nullptr;
    bool ok = rc == CM21_OK && ec == hipSuccess && graph;
    if (ok) ok = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) == hipSuccess;
    if (graph) (void)hipGraphDestroy(graph);
    if (!ok) {
        (void)hipGetLastError();
        if (rc != CM21_OK) return rc;
        return run_all(p, prm, work, st);
    }
    e.key = std::move(key);
    e.exec = exec;
    CM21_HIP_CHECK(hipGraphLaunch(exec, st));
    return CM21_OK;
}

}  // namespace

void forget_batch_graphs() {
    std::lock_guard<std::mutex> lk(g_bgraph_mu);
    for (auto& kv : g_bgraphs)
        if (kv.second.exec) (void)hipGraphExecDestroy(kv.second.exec);
    g_bgraphs.clear();
}

}  // namespace cm21

using namespace cm21;

extern "C" size_t cm21_find_bubbles_batched_workspace_bytes(const cm21_batch_group* groups, int ngroups,
                                                             const cm21_bubble_params* params) {
    BatchLayout lay;
    if (build_layout(groups, ngroups, params, lay, false) != CM21_OK) return 0;
    return lay.bytes;
}

extern "C" int cm21_find_bubbles_batched(const cm21_batch_group* groups, int ngroups,
                                         const cm21_bubble_params* params, void* work, size_t work_bytes,
                                         void* stream) {
    Prepared p;
    CM21_TRY(prepare(groups, ngroups, params, work, work_bytes, p));
    hipStream_t st = (hipStream_t)stream;
    CM21_TRY(upload(p, work, st));
    if (params->graph && st != nullptr && !prof_any()) return graph_run(p, groups, ngroups, params, work, st);
    return run_all(p, params, work, st);
}

extern "C" int cm21_batch_tile_count(int D, int stage) {
    const BatchShape* s = batch_shape(D);
    if (!s || stage < 0 || stage > 1) return -1;
    return stage == 0 ? s->ncols : s->nlines;
}

extern "C" const char* cm21_batch_tile_name(int D, int stage, int index) {
    const BatchShape* s = batch_shape(D);
    if (!s || stage < 0 || stage > 1) return nullptr;
    if (stage == 0) return (index >= 0 && index < s->ncols) ? s->cols[index].name : nullptr;
    return (index >= 0 && index < s->nlines) ? s->lines[index].name : nullptr;
}

extern "C" int cm21_batch_autotune(cm21_batch_group* groups, int ngroups, const cm21_bubble_params* params,
                                   void* work, size_t work_bytes, int reps, double* ms_out, void* stream) {
    CM21_REQUIRE(reps >= 1 && reps <= 1000, "cm21_batch_autotune: reps=%d out of range", reps);
    CM21_REQUIRE(stream != nullptr, "cm21_batch_autotune: needs a non-null stream");
    CM21_REQUIRE(!prof_any(), "cm21_batch_autotune: disable the stage profiler first");
    hipStream_t st = (hipStream_t)stream;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    CM21_HIP_CHECK(hipEventCreate(&e0));
    if (hipEventCreate(&e1) != hipSuccess) {
        (void)hipEventDestroy(e0);
        return fail(CM21_ERR_HIP, "cm21_batch_autotune: hipEventCreate failed");
    }
    auto elapsed = [&](float* ms) -> int {
        CM21_HIP_CHECK(hipEventSynchronize(e1));
        CM21_HIP_CHECK(hipEventElapsedTime(ms, e0, e1));
        return CM21_OK;
    };
    int rc = CM21_OK;
    int slot = 0;
    for (int gi = 0; gi < ngroups && rc == CM21_OK; ++gi) {
        // tune this group alone, with the others' choices fixed
        Prepared p;
        rc = prepare(groups + gi, 1, params, work, work_bytes, p);
        if (rc != CM21_OK) break;
        const GroupHost& h = p.lay.groups[0];
        rc = upload(p, work, st);
        if (rc == CM21_OK) rc = run_forward(h, st, (char*)work);
        // columns/rows candidates (idempotent: S -> T)
        double best = 1e300;
        int best_c = 0;
        for (int c = 0; c < h.shape->ncols && rc == CM21_OK; ++c) {
            int cb = 0, rb = 0;
            fill_tables(h, params, c, p.blob, &cb, &rb);
            BatchLaunch a = make_launch(h, (char*)work, cb, rb, p.launches[0].nblk);
            rc = upload(p, work, st);
            if (rc == CM21_OK) rc = run_transforms(h, a, c, st);  // warm
            if (rc != CM21_OK) break;
            if (hipEventRecord(e0, st) != hipSuccess) rc = fail(CM21_ERR_HIP, "hipEventRecord failed");
            for (int i = 0; i < reps && rc == CM21_OK; ++i) rc = run_transforms(h, a, c, st);
            if (rc == CM21_OK && hipEventRecord(e1, st) != hipSuccess) rc = fail(CM21_ERR_HIP, "hipEventRecord failed");
            float ms = 0.0f;
            if (rc == CM21_OK) rc = elapsed(&ms);
            if (ms_out) ms_out[slot++] = ms / reps;
            if (rc == CM21_OK && ms < best) {
                best = ms;
                best_c = c;
            }
        }
        if (rc != CM21_OK) break;
        // leave T from the chosen columns tile, then time the threshold stage
        int cb = 0, rb = 0;
        fill_tables(h, params, best_c, p.blob, &cb, &rb);
        rc = upload(p, work, st);
        BatchLaunch a = make_launch(h, (char*)work, cb, rb, p.launches[0].nblk);
        if (rc == CM21_OK) rc = run_transforms(h, a, best_c, st);
        best = 1e300;
        int best_l = 0;
        for (int l = 0; l < h.shape->nlines && rc == CM21_OK; ++l) {
            a.nblk = h.D * h.D / h.shape->lines[l].L;
            rc = run_threshold(h, a, l, params, st);  // warm
            if (rc != CM21_OK) break;
            if (hipEventRecord(e0, st) != hipSuccess) rc = fail(CM21_ERR_HIP, "hipEventRecord failed");
            for (int i = 0; i < reps && rc == CM21_OK; ++i) rc = run_threshold(h, a, l, params, st);
            if (rc == CM21_OK && hipEventRecord(e1, st) != hipSuccess) rc = fail(CM21_ERR_HIP, "hipEventRecord failed");
            float ms = 0.0f;
            if (rc == CM21_OK) rc = elapsed(&ms);
            if (ms_out) ms_out[slot++] = ms / reps;
            if (rc == CM21_OK && ms < best) {
                best = ms;
                best_l = l;
            }
        }
        if (rc == CM21_OK) {
            groups[gi].tile_columns = best_c;
            groups[gi].tile_lines = best_l;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return rc;
}
