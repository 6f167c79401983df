// 21-cm brightness temperature and its spherically binned power spectrum.
//
// Reference: Programs/delta_T.py:73-98 (host NumPy dT, transform, binning
// driver) and Programs/delta_T.cu:5-29 (pbox_kernel).  Bin sums are FP64 and
// counts are exact integers; the reduction order is fixed, so results are
// bitwise reproducible (the reference kernel races on `+=`, delta_T.cu:27-29).
#include <cmath>

#include "cm21_internal.hpp"

using namespace cm21;

namespace {

__global__ void k_delta_T(float* __restrict__ dT, const float* __restrict__ xH, const float* __restrict__ delta,
                          size_t N, float c, double* __restrict__ partial) {
    double acc = 0.0;
    for (size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x; p < N; p += (size_t)gridDim.x * blockDim.x) {
        const float v = __fmul_rn(__fmul_rn(c, xH[p]), __fadd_rn(1.0f, delta[p]));
        dT[p] = v;
        acc += (double)v;
    }
    acc = block_sum(acc);
    if (threadIdx.x == 0) partial[blockIdx.x] = acc;
}

__global__ void k_sum_partials_pk(const double* __restrict__ partial, int n, double* __restrict__ out) {
    double acc = 0.0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) acc += partial[i];
    acc = block_sum(acc);
    if (threadIdx.x == 0) *out = acc;
}

// (field/mean - 1) * V/N [* mean], the array transformed at delta_T.py:80-86.
__global__ void k_pk_prep(float* __restrict__ out, const float* __restrict__ f, size_t N,
                          const double* __restrict__ sum, float vol_over_n, int dimensional) {
    const size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= N) return;
    const float ave = (float)(*sum / (double)N);
    float v = __fmul_rn(__fsub_rn(__fdiv_rn(f[p], ave), 1.0f), vol_over_n);
    if (dimensional) v = __fmul_rn(v, ave);
    out[p] = v;
}

// Bins the half spectrum.  Each wavefront keeps private per-bin slots in LDS
// and adds to them only from lane 0, once per distinct bin present in the
// wavefront at each step, with a fixed-order wavefront sum: the result does not
// depend on scheduling.  Blocks then write per-block partials.
__global__ void k_pk_bin(const float2* __restrict__ F, int D, size_t Nc, const int32_t* __restrict__ lut, int nbins,
                         double dk, double norm, double* __restrict__ part_ps, double* __restrict__ part_k,
                         unsigned long long* __restrict__ part_n) {
    extern __shared__ unsigned char smem[];
    double* s_ps = (double*)smem;
    double* s_k = s_ps + kWaves * nbins;
    unsigned long long* s_n = (unsigned long long*)(s_k + kWaves * nbins);
    for (int q = threadIdx.x; q < kWaves * nbins; q += blockDim.x) {
        s_ps[q] = 0.0;
        s_k[q] = 0.0;
        s_n[q] = 0ull;
    }
    __syncthreads();
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int H = D / 2 + 1;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t base = (size_t)blockIdx.x * blockDim.x; base < Nc; base += stride) {
        const size_t idx = base + threadIdx.x;
        int bin = -1;
        double w = 0.0, kk = 0.0, contrib = 0.0;
        if (idx < Nc) {
            const int c = (int)(idx % H);
            const size_t ab = idx / H;
            const int b = (int)(ab % D), a = (int)(ab / D);
            const long na = mode_of(a, D), nb = mode_of(b, D);
            const long m2 = na * na + nb * nb + (long)c * c;
            const int bb = lut[m2];
            if (bb >= 0 && bb < nbins) {
                bin = bb;
                w = (c == 0 || c == D / 2) ? 1.0 : 2.0;
                kk = dk * sqrt((double)m2);
                const float2 v = F[idx];
                const double p2 = (double)v.x * (double)v.x + (double)v.y * (double)v.y;
                contrib = kk * kk * kk * p2 * norm;
            }
        }
        bool pending = bin >= 0;
        while (__any(pending)) {
            const unsigned long long mask = __ballot(pending);
            const int leader = __ffsll((long long)mask) - 1;
            const int b0 = __shfl(bin, leader, 64);
            const bool mine = pending && bin == b0;
            const double sp = wave_sum(mine ? w * contrib : 0.0);
            const double sk = wave_sum(mine ? w * kk : 0.0);
            const unsigned long long sn = wave_sum_u64(mine ? (unsigned long long)w : 0ull);
            if (lane == 0) {
                s_ps[wv * nbins + b0] += sp;
                s_k[wv * nbins + b0] += sk;
                s_n[wv * nbins + b0] += sn;
            }
            pending = pending && !mine;
        }
    }
    __syncthreads();
    for (int q = threadIdx.x; q < nbins; q += blockDim.x) {
        double ps = 0.0, ks = 0.0;
        unsigned long long n = 0;
        for (int v = 0; v < kWaves; ++v) {
            ps += s_ps[v * nbins + q];
            ks += s_k[v * nbins + q];
            n += s_n[v * nbins + q];
        }
        part_ps[(size_t)blockIdx.x * nbins + q] = ps;
        part_k[(size_t)blockIdx.x * nbins + q] = ks;
        part_n[(size_t)blockIdx.x * nbins + q] = n;
    }
}

// One block per bin: each thread adds a strided subset of the per-block
// partials in a fixed order, then a fixed tree over lanes and waves
// (deterministic).
__global__ __launch_bounds__(kThreads) void k_pk_final(const double* __restrict__ part_ps,
                                                       const double* __restrict__ part_k,
                                                       const unsigned long long* __restrict__ part_n, int nblocks,
                                                       int nbins, double* __restrict__ ps, double* __restrict__ ks,
                                                       unsigned long long* __restrict__ n) {
    __shared__ double wa[kWaves], wb[kWaves];
    __shared__ unsigned long long wc[kWaves];
    const int q = blockIdx.x;
    double a = 0.0, b = 0.0;
    unsigned long long c = 0;
    for (int i = threadIdx.x; i < nblocks; i += kThreads) {
        a += part_ps[(size_t)i * nbins + q];
        b += part_k[(size_t)i * nbins + q];
        c += part_n[(size_t)i * nbins + q];
    }
    a = wave_sum(a);
    b = wave_sum(b);
    c = wave_sum_u64(c);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) {
        wa[w] = a;
        wb[w] = b;
        wc[w] = c;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double ta = 0.0, tb = 0.0;
        unsigned long long tc = 0;
        for (int v = 0; v < kWaves; ++v) {
            ta += wa[v];
            tb += wb[v];
            tc += wc[v];
        }
        ps[q] = ta;
        ks[q] = tb;
        n[q] = tc;
    }
}

constexpr int kPkBlocks = 512;

struct PkWork {
    float* field;
    float2* spec;
    double* part_ps;
    double* part_k;
    unsigned long long* part_n;
    size_t bytes;
};

PkWork carve_pk(void* base, int D, int nbins) {
    const size_t N = (size_t)D * D * D, Nc = (size_t)D * D * (D / 2 + 1);
    char* p = (char*)base;
    PkWork w{};
    size_t off = 0;
    w.field = (float*)(p + off);
    off = align_up(off + sizeof(float) * N, 256);
    w.spec = (float2*)(p + off);
    off = align_up(off + sizeof(float2) * Nc, 256);
    w.part_ps = (double*)(p + off);
    off = align_up(off + sizeof(double) * kPkBlocks * (size_t)nbins, 256);
    w.part_k = (double*)(p + off);
    off = align_up(off + sizeof(double) * kPkBlocks * (size_t)nbins, 256);
    w.part_n = (unsigned long long*)(p + off);
    off = align_up(off + sizeof(unsigned long long) * kPkBlocks * (size_t)nbins, 256);
    w.bytes = off;
    return w;
}

}  // namespace

namespace cm21 {
int plan_dim(const cm21_plan* p);
int plan_batch(const cm21_plan* p);
}

extern "C" int cm21_delta_T(float* dT, const float* xH, const float* delta, int D, float tb_const, double* dev_sum,
                            void* work, void* stream) {
    CM21_REQUIRE(dT && xH && delta && dev_sum && work, "cm21_delta_T: NULL argument");
    CM21_REQUIRE(D >= 1 && D <= 2048, "cm21_delta_T: D=%d out of range", D);
    hipStream_t st = (hipStream_t)stream;
    const size_t N = (size_t)D * D * D;
    double* partial = (double*)work;
    k_delta_T<<<kReduceBlocks, kThreads, 0, st>>>(dT, xH, delta, N, tb_const, partial);
    CM21_LAUNCH_CHECK("k_delta_T");
    k_sum_partials_pk<<<1, kThreads, 0, st>>>(partial, kReduceBlocks, dev_sum);
    CM21_LAUNCH_CHECK("k_sum_partials_pk");
    return CM21_OK;
}

extern "C" size_t cm21_pk_workspace_bytes(int D, int nbins) {
    if (D < 2 || D > 2048 || D % 2 || nbins < 1) return 0;
    return carve_pk(nullptr, D, nbins).bytes;
}

extern "C" int cm21_power_spectrum(cm21_plan* plan, const float* field, const double* field_sum, int D,
                                   double box_len, const int32_t* bin_lut, int nbins, int dimensional, double* ps_sum,
                                   double* k_sum, uint64_t* counts, void* work, size_t work_bytes, void* stream) {
    CM21_REQUIRE(plan && field && field_sum && bin_lut && ps_sum && k_sum && counts && work,
                 "cm21_power_spectrum: NULL argument");
    CM21_REQUIRE(plan_dim(plan) == D && plan_batch(plan) == 1, "cm21_power_spectrum: plan is not a batch-1 D=%d plan",
                 D);
    CM21_REQUIRE(nbins >= 1 && nbins <= 512, "cm21_power_spectrum: nbins=%d not in [1,512]", nbins);
    CM21_REQUIRE(box_len > 0.0, "cm21_power_spectrum: box_len must be positive");
    PkWork w = carve_pk(work, D, nbins);
    if (work_bytes < w.bytes)
        return cm21::fail(CM21_ERR_NOMEM, "cm21_power_spectrum: workspace %zu bytes < required %zu", work_bytes,
                          w.bytes);
    hipStream_t st = (hipStream_t)stream;
    const size_t N = (size_t)D * D * D, Nc = (size_t)D * D * (D / 2 + 1);
    const double V = box_len * box_len * box_len;
    k_pk_prep<<<grid_for(N), kThreads, 0, st>>>(w.field, field, N, field_sum, (float)(V / (double)N),
                                                dimensional ? 1 : 0);
    CM21_LAUNCH_CHECK("k_pk_prep");
    CM21_TRY(cm21_r2c(plan, w.field, w.spec, stream));
    const size_t lds = (sizeof(double) * 2 + sizeof(unsigned long long)) * kWaves * (size_t)nbins;
    k_pk_bin<<<kPkBlocks, kThreads, lds, st>>>(w.spec, D, Nc, bin_lut, nbins, 2.0 * M_PI / box_len,
                                               1.0 / (2.0 * M_PI * M_PI * V), w.part_ps, w.part_k, w.part_n);
    CM21_LAUNCH_CHECK("k_pk_bin");
    k_pk_final<<<nbins, kThreads, 0, st>>>(w.part_ps, w.part_k, w.part_n, kPkBlocks, nbins, ps_sum, k_sum,
                                           (unsigned long long*)counts);
    CM21_LAUNCH_CHECK("k_pk_final");
    return CM21_OK;
}
