// The HII filter window (Programs/find_bubbles.cu:20-41), evaluated per mode
// in the reference's float32 operation order.  Shared by the filter, velocity
// and fused-transform kernels.
#pragma once

#include "cm21_internal.hpp"

namespace cm21 {

struct WindowParams {
    float dk;             // DELTA_K = float32(2 pi / box_len)
    float R;              // filter radius (Mpc)
    int filter_type;      // 0 real-space top-hat, 1 k-space top-hat, 2 Gaussian
    int ref_fallthrough;  // 1: switch without breaks as in the reference
    int korder;           // 0: |k|^2 summed fastest axis first, 1: slowest first
};

constexpr float kE = 2.7182818284f;  // find_bubbles.cu:3, `(float)(2.7182818284)`

// Correctly rounded float32 square root: the FP64 root of a float32 value
// rounded once to float32 (53 >= 2*24+2 bits, so the double rounding is
// innocuous).  The cut-off decision kR*0.413566994 > 1 depends on the last
// bit of |k|.
__device__ inline float sqrt_rn(float s) { return (float)__dsqrt_rn((double)s); }

// |k| of half-spectrum element (a, b, c); axis 2 (c) is the halved axis.
__device__ inline float kmag_of(int a, int b, int c, int D, float dk, int korder) {
    const float kz = __fmul_rn((float)mode_of(a, D), dk);
    const float ky = __fmul_rn((float)mode_of(b, D), dk);
    const float kx = __fmul_rn((float)c, dk);
    const float sx = __fmul_rn(kx, kx), sy = __fmul_rn(ky, ky), sz = __fmul_rn(kz, kz);
    const float s = korder == 0 ? __fadd_rn(__fadd_rn(sx, sy), sz) : __fadd_rn(__fadd_rn(sz, sy), sx);
    return sqrt_rn(s);
}

// Evaluates the window for |k| = kmag.  The caller multiplies the spectrum
// value by w0 and then by w2 (each product rounded), or zeroes it when cut.
__device__ inline void window_of(float kmag, const WindowParams& p, float& w0, bool& cut, float& w2) {
    float kR = __fmul_rn(kmag, p.R);
    w0 = 1.0f;
    cut = false;
    w2 = 1.0f;
    if (p.filter_type == 0 && (double)kR > 1e-4) {
        // sin, cos and the powers are rounded once from double: the float32
        // difference below cancels catastrophically for small kR, so both the
        // GPU and the CPU oracle take correctly rounded inputs.
        const double x = (double)kR;
        const float s = (float)sin(x), c = (float)cos(x);
        const float p3 = (float)(x * x * x), p2 = (float)(x * x);
        const float inner = __fsub_rn(__fdiv_rn(s, p3), __fdiv_rn(c, p2));
        w0 = (float)(3.0 * (double)inner);
    }
    if (!p.ref_fallthrough) {
        if (p.filter_type == 1) {
            cut = (float)((double)kR * 0.413566994) > 1.0f;
        } else if (p.filter_type == 2) {
            const float kg = (float)((double)kR * 0.643);
            w2 = powf(kE, -__fmul_rn(kg, kg) * 0.5f);
        }
        return;
    }
    if (p.filter_type <= 1) {  // case 1 (reached from case 0 by fall-through)
        kR = (float)((double)kR * 0.413566994);
        cut = kR > 1.0f;
    }
    kR = (float)((double)kR * 0.643);  // case 2
    w2 = powf(kE, -__fmul_rn(kR, kR) * 0.5f);
}

__device__ inline float2 apply_window(float2 v, float w0, bool cut, float w2, float scale) {
    if (cut) return make_float2(0.0f, 0.0f);
    v.x = __fmul_rn(__fmul_rn(__fmul_rn(v.x, w0), w2), scale);
    v.y = __fmul_rn(__fmul_rn(__fmul_rn(v.y, w0), w2), scale);
    return v;
}

}  // namespace cm21
