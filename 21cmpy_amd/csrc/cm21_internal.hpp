// Internal helpers shared by the cm21 HIP sources: error reporting across the
// C ABI, launch geometry and deterministic wavefront/block reductions.
#pragma once

#include <hip/hip_runtime.h>
#include <cstddef>
#include <cstdint>
#include <mutex>

#include "cm21.h"

namespace cm21 {

// Record a formatted message for cm21_last_error() and return `code`.
int fail(int code, const char* fmt, ...) __attribute__((format(printf, 2, 3)));
void clear_error();

constexpr int kThreads = 256;          // threads per block for streaming kernels
constexpr int kWaves = kThreads / 64;  // wavefronts per block (wave64 on CDNA)
constexpr int kReduceBlocks = 1024;    // fixed grid for deterministic reductions

inline size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

// Row length (complex elements) of the fused loop's intermediates: the D/2+1
// modes of a half-spectrum line padded to whole 128-byte lines.
__host__ __device__ constexpr int padded_row(int h) { return (h + 15) / 16 * 16; }

inline unsigned grid_for(size_t n, int threads = kThreads) {
    return (unsigned)((n + threads - 1) / threads);
}

// Sum over the 64 lanes of a wavefront; the tree order is fixed, so the
// result is bitwise reproducible.  Lane 0 holds the total.
__device__ inline double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    return v;
}

__device__ inline unsigned long long wave_sum_u64(unsigned long long v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    return v;
}

// Block-wide FP64 sum in a fixed order (blocks of up to 1024 threads, a
// multiple of 64); thread 0 receives the total.
__device__ inline double block_sum(double v) {
    __shared__ double part[16];
    v = wave_sum(v);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) part[w] = v;
    __syncthreads();
    double r = 0.0;
    if (threadIdx.x == 0) {
        const int nw = (int)(blockDim.x >> 6);
        for (int i = 0; i < nw; ++i) r += part[i];
    }
    return r;
}

// Event-pair timing of instrumented launches (no-op unless enabled with
// cm21_profile_enable).
struct ProfToken {
    int stage;
    void* start;
};
ProfToken prof_begin(int stage, hipStream_t st);
void prof_end(const ProfToken& t, hipStream_t st);
bool prof_any();  // any stage instrumented

// Plan accessors (cm21_core.hip).
int plan_dim(const ::cm21_plan* p);
int plan_batch(const ::cm21_plan* p);
const float2* plan_twiddles(const ::cm21_plan* p);

// Library-owned helper streams of one device (default priority, created on
// first use, kept until the process ends) and events to fork work off a
// caller's stream and join it back (the fork/join becomes graph edges when
// the caller's stream is capturing).  `mu` serialises host threads that
// launch through the same set.
struct SideStreams {
    static constexpr int kStreams = 4;
    std::mutex mu;
    hipStream_t s[kStreams] = {};
    hipEvent_t fork = nullptr;
    hipEvent_t join[kStreams] = {};
    hipEvent_t ev[6] = {};  // free for the loop's own dependencies
};
int side_streams(SideStreams** out);

// Drop captured graphs that reference a plan about to be destroyed.
void forget_plan_graphs(const ::cm21_plan* plan);
void forget_batch_graphs();

// Launch geometry of the per-mode kernels over a half spectrum (D, D, H):
// blocks of 64 x 4 threads, x over the halved axis c, y over b, z over a, so
// a thread finds its mode without any 64-bit division.
constexpr int kModeBlockX = 64, kModeBlockY = 4;
inline dim3 mode_grid(int D, int H) {
    return dim3((unsigned)((H + kModeBlockX - 1) / kModeBlockX), (unsigned)((D + kModeBlockY - 1) / kModeBlockY),
                (unsigned)D);
}
inline dim3 mode_block() { return dim3(kModeBlockX, kModeBlockY, 1); }

// Signed mode number of FFT index i on an axis of length D; index D/2 maps
// to +D/2 (Programs/find_bubbles.cu:22-24).
__device__ inline int mode_of(int i, int D) { return (i > D / 2) ? i - D : i; }

}  // namespace cm21

#define CM21_HIP_CHECK(call)                                                        \
    do {                                                                            \
        hipError_t e_ = (call);                                                     \
        if (e_ != hipSuccess)                                                       \
            return cm21::fail(CM21_ERR_HIP, "%s failed at %s:%d: %s", #call,        \
                              __FILE__, __LINE__, hipGetErrorString(e_));           \
    } while (0)

#define CM21_LAUNCH_CHECK(name)                                                     \
    do {                                                                            \
        hipError_t e_ = hipGetLastError();                                          \
        if (e_ != hipSuccess)                                                       \
            return cm21::fail(CM21_ERR_HIP, "launch of %s failed: %s", name,       \
                              hipGetErrorString(e_));                               \
    } while (0)

#define CM21_REQUIRE(cond, ...)                                                     \
    do {                                                                            \
        if (!(cond)) return cm21::fail(CM21_ERR_INVALID, __VA_ARGS__);              \
    } while (0)

#define CM21_TRY(call)                                                              \
    do {                                                                            \
        int s_ = (call);                                                            \
        if (s_ != CM21_OK) return s_;                                               \
    } while (0)
