// Zel'dovich mass move on the GPU (perturb_field stage).
//
// Reference: Programs/perturb_field.cu:90-133 (move_mass) and
// Programs/perturb_field.py:93-105.  One thread per DIM^3 particle, in the
// reference's float32 position arithmetic.  NGP deposits the whole mass in the
// truncated cell (bit-exact indices); CIC shares it over the 8 nearest cell
// centres.  Deposits are float atomics: the summation order varies from run
// to run at the 1e-7 relative level, the target cells never do.
//
// Particles of a block come from a 16^3 patch of the DIM grid whose
// displacements are small, so the block first accumulates into an LDS tile
// around the patch's HII footprint and flushes it once; deposits that fall
// outside the tile go straight to global memory.
#include <cmath>

#include "cm21_internal.hpp"

using namespace cm21;

namespace {

struct MoveParams {
    int DIM, sw;
    float disp_scale;   // (D(z) - D_init)/BOX_LEN
    float growth_init;  // D_init
    double pf;          // PIXEL_FACTOR = DIM / sw
};

// Position of particle index i on one axis after the wrap loops of
// perturb_field.cu:102-118.  Returns false for a non-finite position.
__device__ inline bool particle_coord(int i, float v, const MoveParams& mp, float& f) {
    const float swf = (float)mp.sw;
    f = (float)(((double)i + 0.5) / (double)mp.DIM);
    f = __fadd_rn(f, __fmul_rn(v, mp.disp_scale));
    f = __fmul_rn(f, swf);
    if (!isfinite(f)) return false;
    // The reference loops until in range; bound the loop and finish with an
    // exact floor-based wrap for displacements of many box lengths.
    int it = 0;
    while (f >= swf && it < 64) { f = __fsub_rn(f, swf); ++it; }
    while (f < 0.0f && it < 128) { f = __fadd_rn(f, swf); ++it; }
    if (f >= swf || f < 0.0f) {
        f = f - swf * floorf(f / swf);
        if (f >= swf) f = 0.0f;
    }
    return true;
}

__device__ inline int hii_index(int i, double pf) { return (int)floor((double)i / pf); }

__device__ inline bool particle_pos(size_t p, const float* vx, const float* vy, const float* vz,
                                    const MoveParams& mp, float& xf, float& yf, float& zf) {
    const int DIM = mp.DIM;
    const int k = (int)(p % DIM);
    const int j = (int)((p / DIM) % DIM);
    const int i = (int)(p / ((size_t)DIM * DIM));
    const int hi = hii_index(i, mp.pf), hj = hii_index(j, mp.pf), hk = hii_index(k, mp.pf);
    const size_t h = ((size_t)hi * mp.sw + hj) * mp.sw + hk;  // C_INDEX(HII_k, HII_j, HII_i, sw)
    return particle_coord(i, vx[h], mp, xf) && particle_coord(j, vy[h], mp, yf) && particle_coord(k, vz[h], mp, zf);
}

__device__ inline float particle_mass(float delta, float growth_init) {
    return (float)(1.0 + (double)__fmul_rn(growth_init, delta));
}

__device__ inline int wrap_cell(int x, int sw) {
    if (x >= sw) x -= sw;
    if (x < 0) x += sw;
    return x;
}

__global__ void k_move_ngp(float* __restrict__ acc, const float* __restrict__ delta, const float* __restrict__ vx,
                           const float* __restrict__ vy, const float* __restrict__ vz, MoveParams mp) {
    const size_t n = (size_t)mp.DIM * mp.DIM * mp.DIM;
    const size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    float xf, yf, zf;
    if (!particle_pos(p, vx, vy, vz, mp, xf, yf, zf)) return;
    const int sw = mp.sw;
    const int xi = wrap_cell((int)xf, sw), yi = wrap_cell((int)yf, sw), zi = wrap_cell((int)zf, sw);
    if ((unsigned)xi >= (unsigned)sw || (unsigned)yi >= (unsigned)sw || (unsigned)zi >= (unsigned)sw) return;
    atomicAdd(&acc[((size_t)xi * sw + yi) * sw + zi], particle_mass(delta[p], mp.growth_init));
}

// Tiled deposit.  A block of 256 threads moves the 16^3 particles of one
// patch of the DIM grid (thread t owns (j,k) = (t/16, t%16) and walks 16
// values of i).  Their HII targets lie near the patch's HII footprint, so the
// block accumulates into an LDS tile covering the footprint plus a halo of
// kHalo cells (plus one for the CIC neighbour) and flushes each non-zero tile
// cell with one global atomic; deposits outside the tile go straight to
// global memory.  scheme 0 = NGP, 1 = CIC.
//
// CIC: along a thread's walk in i the particles advance by 1/PIXEL_FACTOR
// HII cells in x and share (y, z) cells while their displacement cells do, so
// the thread keeps the deposits of the two x cells {base, base+1} times the
// four (y, z) cells in registers and writes a cell to the tile only when the
// walk leaves it: about 40 LDS atomics per 16 particles instead of 128.
constexpr int kPatch = 16;
constexpr int kHalo = 4;

__device__ inline int mod_sw(int x, int sw) {
    x %= sw;
    return x < 0 ? x + sw : x;
}

// One deposit into the block's tile, or into global memory when the cell
// lies outside the tile.  x, y, z are unwrapped HII cells.
__device__ __forceinline__ void deposit(float* tile, float* acc, int x, int y, int z, float w, int ox, int oy,
                                        int oz, int T, int sw, bool use_tile) {
    if (w == 0.0f) return;
    const int lx = mod_sw(x - ox, sw), ly = mod_sw(y - oy, sw), lz = mod_sw(z - oz, sw);
    if (use_tile && lx < T && ly < T && lz < T)
        atomicAdd(&tile[(lx * T + ly) * T + lz], w);
    else
        atomicAdd(&acc[((size_t)mod_sw(x, sw) * sw + mod_sw(y, sw)) * sw + mod_sw(z, sw)], w);
}

// The register accumulator of one thread's CIC walk: cells (base + a, y0 + b,
// z0 + c) for a, b, c in {0, 1}.
struct CicRun {
    int base, y0, z0;
    bool live;
    float v[2][4];  // [a][2b + c]

    __device__ __forceinline__ void flush_x(int a, float* tile, float* acc, int ox, int oy, int oz, int T, int sw,
                                            bool use_tile) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            deposit(tile, acc, base + a, y0 + (q >> 1), z0 + (q & 1), v[a][q], ox, oy, oz, T, sw, use_tile);
            v[a][q] = 0.0f;
        }
    }
};

template <int SCHEME>
__global__ void k_move_tile(float* __restrict__ acc, const float* __restrict__ delta, const float* __restrict__ vx,
                            const float* __restrict__ vy, const float* __restrict__ vz, MoveParams mp, int T) {
    extern __shared__ float tile[];
    const int DIM = mp.DIM, sw = mp.sw;
    const int np = DIM / kPatch;
    const int bk = blockIdx.x % np, bj = (blockIdx.x / np) % np, bi = blockIdx.x / (np * np);
    const int t = threadIdx.x;
    const int j = bj * kPatch + (t >> 4), k = bk * kPatch + (t & 15);
    // Tile origin in unwrapped HII cells: footprint start minus the halo.
    const int ox = hii_index(bi * kPatch, mp.pf) - kHalo;
    const int oy = hii_index(bj * kPatch, mp.pf) - kHalo;
    const int oz = hii_index(bk * kPatch, mp.pf) - kHalo;
    const bool use_tile = sw >= T;
    const int T3 = T * T * T;
    if (use_tile) {
        for (int q = t; q < T3; q += blockDim.x) tile[q] = 0.0f;
        __syncthreads();
    }
    CicRun run;
    run.live = false;
    run.base = run.y0 = run.z0 = 0;
#pragma unroll
    for (int q = 0; q < 4; ++q) run.v[0][q] = run.v[1][q] = 0.0f;
    for (int ii = 0; ii < kPatch; ++ii) {
        const int i = bi * kPatch + ii;
        const size_t p = ((size_t)i * DIM + j) * DIM + k;
        float xf, yf, zf;
        if (!particle_pos(p, vx, vy, vz, mp, xf, yf, zf)) continue;
        const float m = particle_mass(delta[p], mp.growth_init);
        if (SCHEME == 0) {
            const int xi = wrap_cell((int)xf, sw), yi = wrap_cell((int)yf, sw), zi = wrap_cell((int)zf, sw);
            if ((unsigned)xi >= (unsigned)sw || (unsigned)yi >= (unsigned)sw || (unsigned)zi >= (unsigned)sw) continue;
            const int lx = mod_sw(xi - ox, sw), ly = mod_sw(yi - oy, sw), lz = mod_sw(zi - oz, sw);
            if (use_tile && lx < T && ly < T && lz < T)
                atomicAdd(&tile[(lx * T + ly) * T + lz], m);
            else
                atomicAdd(&acc[((size_t)xi * sw + yi) * sw + zi], m);
        } else {
            const float cx = __fsub_rn(xf, 0.5f), cy = __fsub_rn(yf, 0.5f), cz = __fsub_rn(zf, 0.5f);
            const float fx = floorf(cx), fy = floorf(cy), fz = floorf(cz);
            const float tx = __fsub_rn(cx, fx), ty = __fsub_rn(cy, fy), tz = __fsub_rn(cz, fz);
            const int x0 = (int)fx, y0 = (int)fy, z0 = (int)fz;
            if (run.live && (y0 != run.y0 || z0 != run.z0 || x0 < run.base || x0 > run.base + 1)) {
                run.flush_x(0, tile, acc, ox, oy, oz, T, sw, use_tile);
                run.flush_x(1, tile, acc, ox, oy, oz, T, sw, use_tile);
                run.live = false;
            }
            if (!run.live) {
                run.base = x0;
                run.y0 = y0;
                run.z0 = z0;
                run.live = true;
            } else if (x0 == run.base + 1) {
                // the walk leaves cell `base`: write it, slide the window
                run.flush_x(0, tile, acc, ox, oy, oz, T, sw, use_tile);
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    run.v[0][q] = run.v[1][q];
                    run.v[1][q] = 0.0f;
                }
                run.base = x0;
            }
            const float wx[2] = {__fsub_rn(1.0f, tx), tx};
            const float wy[2] = {__fsub_rn(1.0f, ty), ty};
            const float wz[2] = {__fsub_rn(1.0f, tz), tz};
#pragma unroll
            for (int dx = 0; dx < 2; ++dx)
#pragma unroll
                for (int q = 0; q < 4; ++q)
                    run.v[dx][q] += __fmul_rn(__fmul_rn(__fmul_rn(m, wx[dx]), wy[q >> 1]), wz[q & 1]);
        }
    }
    if (SCHEME == 1 && run.live) {
        run.flush_x(0, tile, acc, ox, oy, oz, T, sw, use_tile);
        run.flush_x(1, tile, acc, ox, oy, oz, T, sw, use_tile);
    }
    if (!use_tile) return;
    __syncthreads();
    for (int q = t; q < T3; q += blockDim.x) {
        const float v = tile[q];
        if (v != 0.0f) {
            const int lz = q % T, ly = (q / T) % T, lx = q / (T * T);
            atomicAdd(&acc[((size_t)mod_sw(ox + lx, sw) * sw + mod_sw(oy + ly, sw)) * sw + mod_sw(oz + lz, sw)], v);
        }
    }
}

__global__ void k_zero(float* __restrict__ p, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0f;
}

__global__ void k_renorm(float* __restrict__ p, size_t n, float mass_factor) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = __fsub_rn(__fdiv_rn(p[i], mass_factor), 1.0f);
}

}  // namespace

extern "C" int cm21_move_mass(float* updated, const float* deltax, const float* vx, const float* vy,
                              const float* vz, int DIM, int sw, float disp_scale, float growth_init, int scheme,
                              void* stream) {
    CM21_REQUIRE(updated && deltax && vx && vy && vz, "cm21_move_mass: NULL argument");
    CM21_REQUIRE(DIM >= 8 && DIM <= 2048 && DIM % 8 == 0, "cm21_move_mass: DIM=%d must be a multiple of 8 in [8,2048]",
                 DIM);
    CM21_REQUIRE(sw >= 1 && sw <= DIM && DIM % sw == 0, "cm21_move_mass: sw=%d must divide DIM=%d", sw, DIM);
    CM21_REQUIRE(scheme == 0 || scheme == 1, "cm21_move_mass: scheme=%d not in {0 (NGP), 1 (CIC)}", scheme);
    hipStream_t st = (hipStream_t)stream;
    const size_t n_small = (size_t)sw * sw * sw;
    const size_t n_part = (size_t)DIM * DIM * DIM;
    const float pff = (float)DIM / (float)sw;  // PIXEL_FACTOR (INIT_PARAMS.py:7)
    MoveParams mp{DIM, sw, disp_scale, growth_init, (double)pff};
    // Tile edge: HII footprint of a patch (ceil(16/pf)) plus halo on both
    // sides plus the CIC neighbour.
    const int foot = (int)std::ceil((double)kPatch / (double)pff);
    const int T = foot + 2 * kHalo + 1;
    const size_t lds = sizeof(float) * (size_t)T * T * T;
    const unsigned blocks = (unsigned)((DIM / kPatch) * (DIM / kPatch) * (DIM / kPatch));
    if (scheme == 1 && (DIM % kPatch != 0 || lds > 64 * 1024))
        return cm21::fail(CM21_ERR_INVALID, "cm21_move_mass: CIC needs DIM %% 16 == 0 and DIM/sw <= 16");
    k_zero<<<grid_for(n_small), kThreads, 0, st>>>(updated, n_small);
    CM21_LAUNCH_CHECK("k_zero");
    if (DIM % kPatch != 0 || lds > 64 * 1024) {
        k_move_ngp<<<grid_for(n_part), kThreads, 0, st>>>(updated, deltax, vx, vy, vz, mp);
        CM21_LAUNCH_CHECK("k_move_ngp");
    } else if (scheme == 0) {
        k_move_tile<0><<<blocks, kThreads, lds, st>>>(updated, deltax, vx, vy, vz, mp, T);
        CM21_LAUNCH_CHECK("k_move_tile<NGP>");
    } else {
        k_move_tile<1><<<blocks, kThreads, lds, st>>>(updated, deltax, vx, vy, vz, mp, T);
        CM21_LAUNCH_CHECK("k_move_tile<CIC>");
    }
    const float mass_factor = (float)std::pow((double)pff, 3.0);  // MASS_FACTOR (INIT_PARAMS.py:8)
    k_renorm<<<grid_for(n_small), kThreads, 0, st>>>(updated, n_small, mass_factor);
    CM21_LAUNCH_CHECK("k_renorm");
    return CM21_OK;
}
