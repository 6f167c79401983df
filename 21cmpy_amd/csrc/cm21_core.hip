// Error reporting, ABI version and the rocFFT plans behind the C ABI.
//
// Plans replace pyfft.cuda.Plan (Programs/find_HII_bubbles.py:91): one plan
// object holds a real-forward and a real-inverse rocFFT plan for `batch`
// packed D^3 boxes, a preallocated work buffer and an execution-info object
// whose stream is rebound on every call.
#include <cstdarg>
#include <cstdio>
#include <mutex>
#include <unordered_map>
#include <string>
#include <vector>

#include <rocfft/rocfft.h>

#include "cm21_dft.hpp"
#include "cm21_internal.hpp"

namespace {
thread_local std::string g_last_error;
std::once_flag g_rocfft_once;
rocfft_status g_rocfft_setup_status = rocfft_status_success;
}  // namespace

namespace cm21 {
int fail(int code, const char* fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return code;
}
void clear_error() { g_last_error.clear(); }
}  // namespace cm21

struct cm21_plan {
    int D = 0;
    int batch = 0;
    rocfft_plan fwd = nullptr;
    rocfft_plan inv = nullptr;
    rocfft_execution_info info = nullptr;
    void* work = nullptr;
    size_t work_bytes = 0;
    float2* twiddles = nullptr;  // w_D^m = exp(+2 pi i m/D), m < D, for the fused transforms
};

extern "C" int cm21_abi_version(void) { return CM21_ABI_VERSION; }

extern "C" const char* cm21_last_error(void) { return g_last_error.c_str(); }

#define CM21_FFT_CHECK(call)                                                        \
    do {                                                                            \
        rocfft_status st_ = (call);                                                 \
        if (st_ != rocfft_status_success)                                           \
            return cm21::fail(CM21_ERR_FFT, "%s failed (rocfft status %d)", #call,  \
                              (int)st_);                                            \
    } while (0)

static void plan_release(cm21_plan* p) {
    if (!p) return;
    cm21::forget_plan_graphs(p);
    cm21::forget_batch_graphs();
    if (p->fwd) rocfft_plan_destroy(p->fwd);
    if (p->inv) rocfft_plan_destroy(p->inv);
    if (p->info) rocfft_execution_info_destroy(p->info);
    if (p->work) (void)hipFree(p->work);
    if (p->twiddles) (void)hipFree(p->twiddles);
    delete p;
}

extern "C" int cm21_plan_create(cm21_plan** out, int D, int batch) {
    CM21_REQUIRE(out != nullptr, "cm21_plan_create: out is NULL");
    *out = nullptr;
    CM21_REQUIRE(D >= 2 && D <= 2048 && (D % 2) == 0, "cm21_plan_create: D=%d must be even in [2,2048]", D);
    CM21_REQUIRE(batch >= 1 && batch <= 4096, "cm21_plan_create: batch=%d out of range", batch);
    std::call_once(g_rocfft_once, [] { g_rocfft_setup_status = rocfft_setup(); });
    if (g_rocfft_setup_status != rocfft_status_success)
        return cm21::fail(CM21_ERR_FFT, "rocfft_setup failed (status %d)", (int)g_rocfft_setup_status);

    cm21_plan* p = new cm21_plan();
    p->D = D;
    p->batch = batch;
    const size_t len[3] = {(size_t)D, (size_t)D, (size_t)D};
    rocfft_status st = rocfft_plan_create(&p->fwd, rocfft_placement_notinplace,
                                          rocfft_transform_type_real_forward, rocfft_precision_single,
                                          3, len, (size_t)batch, nullptr);
    if (st == rocfft_status_success)
        st = rocfft_plan_create(&p->inv, rocfft_placement_notinplace,
                                rocfft_transform_type_real_inverse, rocfft_precision_single, 3, len,
                                (size_t)batch, nullptr);
    if (st != rocfft_status_success) {
        plan_release(p);
        return cm21::fail(CM21_ERR_FFT, "rocfft_plan_create(D=%d, batch=%d) failed (status %d)", D, batch, (int)st);
    }
    size_t w1 = 0, w2 = 0;
    rocfft_plan_get_work_buffer_size(p->fwd, &w1);
    rocfft_plan_get_work_buffer_size(p->inv, &w2);
    p->work_bytes = w1 > w2 ? w1 : w2;
    st = rocfft_execution_info_create(&p->info);
    if (st != rocfft_status_success) {
        plan_release(p);
        return cm21::fail(CM21_ERR_FFT, "rocfft_execution_info_create failed (status %d)", (int)st);
    }
    if (p->work_bytes) {
        hipError_t e = hipMalloc(&p->work, p->work_bytes);
        if (e != hipSuccess) {
            plan_release(p);
            return cm21::fail(CM21_ERR_NOMEM, "hipMalloc of %zu-byte FFT work buffer failed: %s", p->work_bytes,
                              hipGetErrorString(e));
        }
        st = rocfft_execution_info_set_work_buffer(p->info, p->work, p->work_bytes);
        if (st != rocfft_status_success) {
            plan_release(p);
            return cm21::fail(CM21_ERR_FFT, "rocfft_execution_info_set_work_buffer failed (status %d)", (int)st);
        }
    }
    {
        std::vector<float2> tw((size_t)D);
        for (int m = 0; m < D; ++m) tw[m] = cm21::twiddle(m, D);
        hipError_t e = hipMalloc(&p->twiddles, sizeof(float2) * (size_t)D);
        if (e == hipSuccess) e = hipMemcpy(p->twiddles, tw.data(), sizeof(float2) * (size_t)D, hipMemcpyHostToDevice);
        if (e != hipSuccess) {
            plan_release(p);
            return cm21::fail(CM21_ERR_NOMEM, "twiddle table for D=%d: %s", D, hipGetErrorString(e));
        }
    }
    *out = p;
    return CM21_OK;
}

extern "C" int cm21_plan_destroy(cm21_plan* plan) {
    plan_release(plan);
    return CM21_OK;
}

extern "C" int cm21_r2c(cm21_plan* plan, const float* in, void* spec_out, void* stream) {
    CM21_REQUIRE(plan && in && spec_out, "cm21_r2c: NULL argument");
    CM21_FFT_CHECK(rocfft_execution_info_set_stream(plan->info, stream));
    void* ib[1] = {const_cast<float*>(in)};
    void* ob[1] = {spec_out};
    CM21_FFT_CHECK(rocfft_execute(plan->fwd, ib, ob, plan->info));
    return CM21_OK;
}

extern "C" int cm21_c2r(cm21_plan* plan, void* spec_in, float* out, void* stream) {
    CM21_REQUIRE(plan && spec_in && out, "cm21_c2r: NULL argument");
    CM21_FFT_CHECK(rocfft_execution_info_set_stream(plan->info, stream));
    void* ib[1] = {spec_in};
    void* ob[1] = {out};
    CM21_FFT_CHECK(rocfft_execute(plan->inv, ib, ob, plan->info));
    return CM21_OK;
}

// Plan dimension accessor used by the loop driver in other translation units.
namespace cm21 {
int side_streams(SideStreams** out) {
    static std::mutex mu;
    static std::unordered_map<int, SideStreams> sets;
    int dev = 0;
    CM21_HIP_CHECK(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(mu);
    SideStreams& o = sets[dev];
    if (!o.fork) {
        for (int i = 0; i < SideStreams::kStreams; ++i) {
            CM21_HIP_CHECK(hipStreamCreateWithFlags(&o.s[i], hipStreamNonBlocking));
            CM21_HIP_CHECK(hipEventCreateWithFlags(&o.join[i], hipEventDisableTiming));
        }
        for (auto& e : o.ev) CM21_HIP_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        CM21_HIP_CHECK(hipEventCreateWithFlags(&o.fork, hipEventDisableTiming));
    }
    *out = &o;
    return CM21_OK;
}

int plan_dim(const cm21_plan* p) { return p ? p->D : 0; }
int plan_batch(const cm21_plan* p) { return p ? p->batch : 0; }
const float2* plan_twiddles(const cm21_plan* p) { return p ? p->twiddles : nullptr; }
}  // namespace cm21

// ------------------------------------------------------------------ profiler
// Per-stage HIP event timing.  While enabled, every instrumented launch
// records an event pair on its own stream; cm21_profile_collect() waits for
// the recorded events, adds their elapsed times to per-stage totals and
// returns them.  Events come from a pool and are reused.
namespace {
struct ProfPair {
    int stage;
    hipEvent_t a, b;
};
std::mutex g_prof_mu;
unsigned g_prof_mask = 0u;  // bit s enables stage s
std::vector<ProfPair> g_prof_pending;
std::vector<hipEvent_t> g_prof_pool;

hipEvent_t pool_get() {
    if (!g_prof_pool.empty()) {
        hipEvent_t e = g_prof_pool.back();
        g_prof_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}
}  // namespace

namespace cm21 {
ProfToken prof_begin(int stage, hipStream_t st) {
    ProfToken t{-1, nullptr};
    std::lock_guard<std::mutex> lk(g_prof_mu);
    if (stage < 0 || stage >= CM21_NSTAGES || !((g_prof_mask >> stage) & 1u)) return t;
    hipEvent_t a = pool_get();
    if (!a) return t;
    if (hipEventRecord(a, st) != hipSuccess) {
        g_prof_pool.push_back(a);
        return t;
    }
    t.stage = stage;
    t.start = a;
    return t;
}

bool prof_any() {
    std::lock_guard<std::mutex> lk(g_prof_mu);
    return g_prof_mask != 0u;
}

void prof_end(const ProfToken& t, hipStream_t st) {
    if (t.stage < 0) return;
    std::lock_guard<std::mutex> lk(g_prof_mu);
    hipEvent_t b = pool_get();
    if (!b || hipEventRecord(b, st) != hipSuccess) {
        g_prof_pool.push_back((hipEvent_t)t.start);
        if (b) g_prof_pool.push_back(b);
        return;
    }
    g_prof_pending.push_back({t.stage, (hipEvent_t)t.start, b});
}
}  // namespace cm21

extern "C" int cm21_profile_enable(int stage_mask) {
    std::lock_guard<std::mutex> lk(g_prof_mu);
    g_prof_mask = (unsigned)stage_mask;
    return CM21_OK;
}

extern "C" int cm21_profile_collect(double* ms, int64_t* count, int nstages) {
    CM21_REQUIRE(ms && count && nstages >= 1, "cm21_profile_collect: NULL argument");
    std::vector<ProfPair> pend;
    {
        std::lock_guard<std::mutex> lk(g_prof_mu);
        pend.swap(g_prof_pending);
    }
    for (int i = 0; i < nstages; ++i) {
        ms[i] = 0.0;
        count[i] = 0;
    }
    int status = CM21_OK;
    for (auto& p : pend) {
        float t = 0.0f;
        hipError_t e = hipEventSynchronize(p.b);
        if (e == hipSuccess) e = hipEventElapsedTime(&t, p.a, p.b);
        if (e != hipSuccess && status == CM21_OK)
            status = cm21::fail(CM21_ERR_HIP, "cm21_profile_collect: %s", hipGetErrorString(e));
        if (p.stage < nstages) {
            ms[p.stage] += (double)t;
            count[p.stage] += 1;
        }
    }
    std::lock_guard<std::mutex> lk(g_prof_mu);
    for (auto& p : pend) {
        g_prof_pool.push_back(p.a);
        g_prof_pool.push_back(p.b);
    }
    return status;
}
