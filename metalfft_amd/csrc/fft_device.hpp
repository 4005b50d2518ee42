// Device-side building blocks for the gfx950 FFT kernels: complex helpers,
// in-register DFT-2/4/8/16 butterflies, padded LDS indexing and the compile-
// time radix schedule of the LDS-resident Stockham transform.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mifft {

// ------------------------------------------------------------------ complex
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// a * conj(b)
__device__ __forceinline__ float2 cmulc(float2 a, float2 b) {
    return make_float2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
__device__ __forceinline__ float2 conjf2(float2 a) { return make_float2(a.x, -a.y); }
// multiply by -i
__device__ __forceinline__ float2 mul_mi(float2 a) { return make_float2(a.y, -a.x); }

// exp(-2 pi i m / 16) constants, m = 1..3 (all other powers used by the
// 4x4 DFT-16 reduce to sign/swap of these).
#define MIFFT_C16 0.92387953251128675613f   // cos(pi/8)
#define MIFFT_S16 0.38268343236508977173f   // sin(pi/8)
#define MIFFT_SQRT1_2 0.70710678118654752440f     // sqrt(2)/2

// -------------------------------------------------------------- butterflies
// Forward DFT of length R on registers, natural order in and out:
//   y[k] = sum_j x[j] exp(-2 pi i j k / R)
__device__ __forceinline__ void dft2(float2& a, float2& b) {
    float2 t = a;
    a = cadd(t, b);
    b = csub(t, b);
}

__device__ __forceinline__ void dft4(float2& x0, float2& x1, float2& x2, float2& x3) {
    float2 t0 = cadd(x0, x2), t1 = cadd(x1, x3);
    float2 t2 = csub(x0, x2), t3 = mul_mi(csub(x1, x3));
    x0 = cadd(t0, t1);
    x2 = csub(t0, t1);
    x1 = cadd(t2, t3);
    x3 = csub(t2, t3);
}

// DFT-8 = even/odd split into two DFT-4 and a W8 combine.
__device__ __forceinline__ void dft8(float2* x) {
    float2 e0 = x[0], e1 = x[2], e2 = x[4], e3 = x[6];
    float2 o0 = x[1], o1 = x[3], o2 = x[5], o3 = x[7];
    dft4(e0, e1, e2, e3);
    dft4(o0, o1, o2, o3);
    // W8^1 = (h, -h), W8^2 = -i, W8^3 = (-h, -h)
    float2 w1o = make_float2(MIFFT_SQRT1_2 * (o1.x + o1.y), MIFFT_SQRT1_2 * (o1.y - o1.x));
    float2 w2o = mul_mi(o2);
    float2 w3o = make_float2(MIFFT_SQRT1_2 * (o3.y - o3.x), -MIFFT_SQRT1_2 * (o3.x + o3.y));
    x[0] = cadd(e0, o0); x[4] = csub(e0, o0);
    x[1] = cadd(e1, w1o); x[5] = csub(e1, w1o);
    x[2] = cadd(e2, w2o); x[6] = csub(e2, w2o);
    x[3] = cadd(e3, w3o); x[7] = csub(e3, w3o);
}

// DFT-16 as 4 x 4: column DFT-4s over j1 (x[4 j1 + j2]), internal twiddle
// W16^{j2 k1}, row DFT-4s over j2; output index k1 + 4 k2.
__device__ __forceinline__ void dft16(float2* x) {
    float2 a[4][4];
#pragma unroll
    for (int j2 = 0; j2 < 4; ++j2) {
        float2 p0 = x[j2], p1 = x[4 + j2], p2 = x[8 + j2], p3 = x[12 + j2];
        dft4(p0, p1, p2, p3);
        a[j2][0] = p0; a[j2][1] = p1; a[j2][2] = p2; a[j2][3] = p3;
    }
    // twiddles W16^{j2*k1}: (j2,k1) with j2,k1 >= 1
    const float2 w1 = make_float2(MIFFT_C16, -MIFFT_S16);
    const float2 w3 = make_float2(MIFFT_S16, -MIFFT_C16);
    const float2 w9 = make_float2(-MIFFT_C16, MIFFT_S16);
    a[1][1] = cmul(a[1][1], w1);
    a[1][2] = make_float2(MIFFT_SQRT1_2 * (a[1][2].x + a[1][2].y), MIFFT_SQRT1_2 * (a[1][2].y - a[1][2].x));  // W^2
    a[1][3] = cmul(a[1][3], w3);
    a[2][1] = make_float2(MIFFT_SQRT1_2 * (a[2][1].x + a[2][1].y), MIFFT_SQRT1_2 * (a[2][1].y - a[2][1].x));  // W^2
    a[2][2] = mul_mi(a[2][2]);                                                                  // W^4
    a[2][3] = make_float2(MIFFT_SQRT1_2 * (a[2][3].y - a[2][3].x), -MIFFT_SQRT1_2 * (a[2][3].x + a[2][3].y)); // W^6
    a[3][1] = cmul(a[3][1], w3);
    a[3][2] = make_float2(MIFFT_SQRT1_2 * (a[3][2].y - a[3][2].x), -MIFFT_SQRT1_2 * (a[3][2].x + a[3][2].y)); // W^6
    a[3][3] = cmul(a[3][3], w9);
#pragma unroll
    for (int k1 = 0; k1 < 4; ++k1) {
        float2 q0 = a[0][k1], q1 = a[1][k1], q2 = a[2][k1], q3 = a[3][k1];
        dft4(q0, q1, q2, q3);
        x[k1] = q0; x[k1 + 4] = q1; x[k1 + 8] = q2; x[k1 + 12] = q3;
    }
}

template <int R>
__device__ __forceinline__ void dft(float2* x) {
    if constexpr (R == 2) {
        dft2(x[0], x[1]);
    } else if constexpr (R == 4) {
        dft4(x[0], x[1], x[2], x[3]);
    } else if constexpr (R == 8) {
        dft8(x);
    } else if constexpr (R == 16) {
        dft16(x);
    } else {
        static_assert(R == 2 || R == 4 || R == 8 || R == 16, "unsupported radix");
    }
}

// ------------------------------------------------------------ radix schedule
constexpr int ilog2c(int n) { return n <= 1 ? 0 : 1 + ilog2c(n / 2); }

// Stockham schedule for N = 2^L with at most E elements per thread: one
// smaller radix pass first for the remainder, then full radix-E passes.
// Pass 0 needs no twiddles, so the small radix there (E/R butterflies per
// thread) costs none, and every twiddled pass is one radix-E butterfly per
// thread with log2 E table entries (fewer registers and table rows than
// with the remainder last: 16384 = 4*16*16*16 has 1092 packed rows, not
// 4368).
// E = 32 keeps radix 16 (two butterflies per thread and pass).
template <int N, int E>
struct Schedule {
    static constexpr int RM = E < 16 ? E : 16;  // largest radix
    static constexpr int L = ilog2c(N);
    static constexpr int e = ilog2c(RM);
    static constexpr int full = L / e;
    static constexpr int rem = L % e;
    static constexpr int passes = full + (rem ? 1 : 0);
    static constexpr int T = N / E;  // threads per transform
    static constexpr int radix(int p) { return (rem && p == 0) ? (1 << rem) : RM; }
    static constexpr int stride(int p) { return p == 0 ? 1 : stride(p - 1) * radix(p - 1); }
};

// Padded LDS index: one spare complex every 16 keeps radix-16/8 scatter
// writes (stride 16 or 8 complex across lanes) off a single bank group.
__device__ __forceinline__ int lds_idx(int i) { return i + (i >> 4); }
constexpr int lds_len(int n) { return n + n / 16; }

}  // namespace mifft
