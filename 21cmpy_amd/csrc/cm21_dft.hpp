// Small inverse DFTs held in registers, for the fused transform kernels.
//
// dft<N>(x) replaces x[k] by sum_n x[n] exp(+2 pi i n k / N) (unnormalised
// inverse transform) for N with prime factors 2, 3 and 5.  Composite sizes are
// split by decimation in time (radix 4 first, then 2, 3, 5); every loop has a
// compile-time trip count, so after unrolling all indices and twiddles are
// constants and the arrays live in registers.  Twiddles are evaluated in
// double from an integer octant reduction and a short Taylor series (exact to
// ~1e-15 for |angle| <= pi/4), then rounded to float.
#pragma once

#include <hip/hip_runtime.h>

namespace cm21 {

__host__ __device__ constexpr double sin_taylor(double x) {  // |x| <= pi/4
    const double x2 = x * x;
    return x * (1.0 - x2 / 6.0 * (1.0 - x2 / 20.0 * (1.0 - x2 / 42.0 * (1.0 - x2 / 72.0 *
                (1.0 - x2 / 110.0 * (1.0 - x2 / 156.0 * (1.0 - x2 / 210.0)))))));
}

__host__ __device__ constexpr double cos_taylor(double x) {  // |x| <= pi/4
    const double x2 = x * x;
    return 1.0 - x2 / 2.0 * (1.0 - x2 / 12.0 * (1.0 - x2 / 30.0 * (1.0 - x2 / 56.0 *
                (1.0 - x2 / 90.0 * (1.0 - x2 / 132.0 * (1.0 - x2 / 182.0))))));
}

// (cos, sin) of 2 pi j / n.
__host__ __device__ constexpr void unit_root(long j, long n, double& c, double& s) {
    constexpr double kQuarterPi = 0.78539816339744830962;
    j %= n;
    if (j < 0) j += n;
    const long o = (8 * j) / n;           // octant 0..7
    const long r = 8 * j - o * n;         // 0 <= r < n
    const double th = kQuarterPi * (double)r / (double)n;           // angle inside the octant
    const double ph = kQuarterPi * (double)(n - r) / (double)n;     // complement to the next multiple of pi/4
    const double st = sin_taylor(th), ct = cos_taylor(th), sp = sin_taylor(ph), cp = cos_taylor(ph);
    switch (o) {
        case 0: c = ct; s = st; break;
        case 1: c = sp; s = cp; break;
        case 2: c = -st; s = ct; break;
        case 3: c = -cp; s = sp; break;
        case 4: c = -ct; s = -st; break;
        case 5: c = -sp; s = -cp; break;
        case 6: c = st; s = -ct; break;
        default: c = cp; s = -sp; break;
    }
}

__host__ __device__ inline float2 twiddle(long j, long n) {
    double c = 0.0, s = 0.0;
    unit_root(j, n, c, s);
    return make_float2((float)c, (float)s);
}

__host__ __device__ inline float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__host__ __device__ inline float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__host__ __device__ inline float2 cmul(float2 a, float2 b) {
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__host__ __device__ inline float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }
__host__ __device__ inline float2 times_i(float2 a) { return make_float2(-a.y, a.x); }  // i*a
__host__ __device__ inline float2 cconj(float2 a) { return make_float2(a.x, -a.y); }

__host__ __device__ constexpr int dft_radix(int n) {
    return (n % 4 == 0 && n != 4) ? 4 : (n % 2 == 0) ? 2 : (n % 3 == 0) ? 3 : 5;
}

template <int N>
__host__ __device__ inline void dft(float2 (&x)[N]) {
    static_assert(N >= 1, "dft size");
    if constexpr (N == 1) {
        return;
    } else if constexpr (N == 2) {
        const float2 a = x[0], b = x[1];
        x[0] = cadd(a, b);
        x[1] = csub(a, b);
    } else if constexpr (N == 3) {
        const float h = 0.86602540378443864676f;  // sin(2 pi/3)
        const float2 t1 = cadd(x[1], x[2]);
        const float2 t2 = csub(x[0], cscale(t1, 0.5f));
        const float2 t3 = times_i(cscale(csub(x[1], x[2]), h));
        x[0] = cadd(x[0], t1);
        x[1] = cadd(t2, t3);
        x[2] = csub(t2, t3);
    } else if constexpr (N == 4) {
        const float2 a = cadd(x[0], x[2]), b = csub(x[0], x[2]);
        const float2 c = cadd(x[1], x[3]), d = times_i(csub(x[1], x[3]));
        x[0] = cadd(a, c);
        x[2] = csub(a, c);
        x[1] = cadd(b, d);
        x[3] = csub(b, d);
    } else if constexpr (N == 5) {
        const float c1 = 0.30901699437494742410f, c2 = -0.80901699437494742410f;  // cos(2pi/5), cos(4pi/5)
        const float s1 = 0.95105651629515357212f, s2 = 0.58778525229247312917f;   // sin(2pi/5), sin(4pi/5)
        const float2 a1 = cadd(x[1], x[4]), b1 = csub(x[1], x[4]);
        const float2 a2 = cadd(x[2], x[3]), b2 = csub(x[2], x[3]);
        const float2 r1 = cadd(x[0], cadd(cscale(a1, c1), cscale(a2, c2)));
        const float2 r2 = cadd(x[0], cadd(cscale(a1, c2), cscale(a2, c1)));
        const float2 i1 = times_i(cadd(cscale(b1, s1), cscale(b2, s2)));
        const float2 i2 = times_i(csub(cscale(b1, s2), cscale(b2, s1)));
        x[0] = cadd(x[0], cadd(a1, a2));
        x[1] = cadd(r1, i1);
        x[4] = csub(r1, i1);
        x[2] = cadd(r2, i2);
        x[3] = csub(r2, i2);
    } else {
        constexpr int R = dft_radix(N);
        constexpr int M = N / R;
        static_assert(R * M == N, "dft: size must factor into 2, 3 and 5");
        float2 sub[R][M];
#pragma unroll
        for (int r = 0; r < R; ++r)
#pragma unroll
            for (int m = 0; m < M; ++m) sub[r][m] = x[R * m + r];
#pragma unroll
        for (int r = 0; r < R; ++r) dft<M>(sub[r]);
#pragma unroll
        for (int k = 0; k < M; ++k) {
            float2 t[R];
#pragma unroll
            for (int r = 0; r < R; ++r) t[r] = (r * k == 0) ? sub[r][k] : cmul(sub[r][k], twiddle(r * k, N));
            dft<R>(t);
#pragma unroll
            for (int q = 0; q < R; ++q) x[k + M * q] = t[q];
        }
    }
}

}  // namespace cm21
