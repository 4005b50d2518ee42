// Device-side building blocks for the gfx950 FFT kernels: complex helpers,
// in-register DFT-2/4/8/16 butterflies, padded LDS indexing and the compile-
// time radix schedule of the LDS-resident Stockham transform.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mifft {

// ------------------------------------------------------------------ complex
#define MIFFT_HD __host__ __device__ __forceinline__

MIFFT_HD float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
MIFFT_HD float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
MIFFT_HD float2 cmul(float2 a, float2 b) {
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// a * conj(b)
MIFFT_HD float2 cmulc(float2 a, float2 b) {
    return make_float2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
MIFFT_HD float2 conjf2(float2 a) { return make_float2(a.x, -a.y); }
// multiply by -i
MIFFT_HD float2 mul_mi(float2 a) { return make_float2(a.y, -a.x); }

// exp(-2 pi i m / 16) constants, m = 1..3 (all other powers used by the
// 4x4 DFT-16 reduce to sign/swap of these).
#define MIFFT_C16 0.92387953251128675613f   // cos(pi/8)
#define MIFFT_S16 0.38268343236508977173f   // sin(pi/8)
#define MIFFT_SQRT1_2 0.70710678118654752440f     // sqrt(2)/2

// -------------------------------------------------------------- butterflies
// Forward DFT of length R on registers, natural order in and out:
//   y[k] = sum_j x[j] exp(-2 pi i j k / R)
MIFFT_HD void dft2(float2& a, float2& b) {
    float2 t = a;
    a = cadd(t, b);
    b = csub(t, b);
}

MIFFT_HD void dft4(float2& x0, float2& x1, float2& x2, float2& x3) {
    float2 t0 = cadd(x0, x2), t1 = cadd(x1, x3);
    float2 t2 = csub(x0, x2), t3 = mul_mi(csub(x1, x3));
    x0 = cadd(t0, t1);
    x2 = csub(t0, t1);
    x1 = cadd(t2, t3);
    x3 = csub(t2, t3);
}

// DFT-8 = even/odd split into two DFT-4 and a W8 combine.
MIFFT_HD void dft8(float2* x) {
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

// y + w x and 2 y - s as fused multiply-adds.  Folding a twiddle product
// into the butterfly add that consumes it takes the pair (y + w x, y - w x)
// from 4 (product) + 4 (add, subtract) instructions to 4 + 2: the sum is
// y + w x in two FMAs per component, the difference 2 y - (y + w x) in one.
MIFFT_HD float2 cfma(float2 y, float2 w, float2 x) {
    return make_float2(__builtin_fmaf(w.x, x.x, __builtin_fmaf(-w.y, x.y, y.x)),
                       __builtin_fmaf(w.x, x.y, __builtin_fmaf(w.y, x.x, y.y)));
}
MIFFT_HD float2 cdbl_sub(float2 y, float2 s) {
    return make_float2(__builtin_fmaf(2.f, y.x, -s.x), __builtin_fmaf(2.f, y.y, -s.y));
}
// y + W16^2 a and y + W16^6 a, W16^2 = h (1 - i), W16^6 = -h (1 + i),
// h = sqrt(2)/2: one add and one FMA per component
MIFFT_HD float2 add_w2(float2 y, float2 a) {
    return make_float2(__builtin_fmaf(MIFFT_SQRT1_2, a.x + a.y, y.x), __builtin_fmaf(MIFFT_SQRT1_2, a.y - a.x, y.y));
}
MIFFT_HD float2 add_w6(float2 y, float2 a) {
    return make_float2(__builtin_fmaf(MIFFT_SQRT1_2, a.y - a.x, y.x), __builtin_fmaf(-MIFFT_SQRT1_2, a.x + a.y, y.y));
}
MIFFT_HD float2 mul_w2(float2 a) {
    return make_float2(MIFFT_SQRT1_2 * (a.x + a.y), MIFFT_SQRT1_2 * (a.y - a.x));
}

// DFT-4 of (p0, p1, p2, p3) from s02 = p0 + p2, d02 = p0 - p2,
// s13 = p1 + p3, d13 = p1 - p3: outputs in p0 .. p3.
MIFFT_HD void dft4_from(float2 s02, float2 d02, float2 s13, float2 d13, float2& p0, float2& p1, float2& p2,
                        float2& p3) {
    p0 = cadd(s02, s13);
    p2 = csub(s02, s13);
    const float2 t3 = mul_mi(d13);
    p1 = cadd(d02, t3);
    p3 = csub(d02, t3);
}

// DFT-16 as 4 x 4: column DFT-4s over j1 (x[4 j1 + j2]), internal twiddle
// W16^{j2 k1}, row DFT-4s over j2; output index k1 + 4 k2.  The row DFT-4s
// take their twiddled inputs through cfma / add_w2 / add_w6 (the twiddle
// product folded into the first add of the DFT-4).  TW: the inputs are
// first multiplied by w[j] (j >= 1, the Stockham pass twiddles), folded the
// same way into the column DFT-4s.
template <bool TW>
MIFFT_HD void dft16_fused(float2* x, const float2* w) {
    float2 a[4][4];
#pragma unroll
    for (int j2 = 0; j2 < 4; ++j2) {
        float2 s02, d02, s13, d13;
        if constexpr (TW) {
            // p_i = w[4i + j2] x[4i + j2]; p_0 untwiddled at j2 = 0
            const float2 y0 = j2 == 0 ? x[0] : cmul(x[j2], w[j2]);
            s02 = cfma(y0, w[8 + j2], x[8 + j2]);
            d02 = cdbl_sub(y0, s02);
            const float2 y1 = cmul(x[4 + j2], w[4 + j2]);
            s13 = cfma(y1, w[12 + j2], x[12 + j2]);
            d13 = cdbl_sub(y1, s13);
        } else {
            s02 = cadd(x[j2], x[8 + j2]);
            d02 = csub(x[j2], x[8 + j2]);
            s13 = cadd(x[4 + j2], x[12 + j2]);
            d13 = csub(x[4 + j2], x[12 + j2]);
        }
        dft4_from(s02, d02, s13, d13, a[j2][0], a[j2][1], a[j2][2], a[j2][3]);
    }
    const float2 w1 = make_float2(MIFFT_C16, -MIFFT_S16);
    const float2 w3 = make_float2(MIFFT_S16, -MIFFT_C16);
    const float2 w9 = make_float2(-MIFFT_C16, MIFFT_S16);
    // row k1: q_j2 = a[j2][k1] W16^{j2 k1}
    {  // k1 = 0: no twiddles
        float2 q0, q1, q2, q3;
        dft4_from(cadd(a[0][0], a[2][0]), csub(a[0][0], a[2][0]), cadd(a[1][0], a[3][0]),
                  csub(a[1][0], a[3][0]), q0, q1, q2, q3);
        x[0] = q0; x[4] = q1; x[8] = q2; x[12] = q3;
    }
    {  // k1 = 1: W^1, W^2, W^3
        const float2 s02 = add_w2(a[0][1], a[2][1]);
        const float2 d02 = cdbl_sub(a[0][1], s02);
        const float2 y1 = cmul(a[1][1], w1);
        const float2 s13 = cfma(y1, w3, a[3][1]);
        const float2 d13 = cdbl_sub(y1, s13);
        float2 q0, q1, q2, q3;
        dft4_from(s02, d02, s13, d13, q0, q1, q2, q3);
        x[1] = q0; x[5] = q1; x[9] = q2; x[13] = q3;
    }
    {  // k1 = 2: W^2, W^4 = -i, W^6
        const float2 t = mul_mi(a[2][2]);
        const float2 s02 = cadd(a[0][2], t);
        const float2 d02 = csub(a[0][2], t);
        const float2 y1 = mul_w2(a[1][2]);
        const float2 s13 = add_w6(y1, a[3][2]);
        const float2 d13 = cdbl_sub(y1, s13);
        float2 q0, q1, q2, q3;
        dft4_from(s02, d02, s13, d13, q0, q1, q2, q3);
        x[2] = q0; x[6] = q1; x[10] = q2; x[14] = q3;
    }
    {  // k1 = 3: W^3, W^6, W^9
        const float2 s02 = add_w6(a[0][3], a[2][3]);
        const float2 d02 = cdbl_sub(a[0][3], s02);
        const float2 y1 = cmul(a[1][3], w3);
        const float2 s13 = cfma(y1, w9, a[3][3]);
        const float2 d13 = cdbl_sub(y1, s13);
        float2 q0, q1, q2, q3;
        dft4_from(s02, d02, s13, d13, q0, q1, q2, q3);
        x[3] = q0; x[7] = q1; x[11] = q2; x[15] = q3;
    }
}

MIFFT_HD void dft16(float2* x) { dft16_fused<false>(x, nullptr); }

template <int R>
MIFFT_HD void dft(float2* x) {
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

// x[j] *= w[j] for j >= 1 (w[0] is not read), then the DFT-R; at R = 16
// the products are folded into the butterfly's first adds (dft16_fused).
template <int R>
MIFFT_HD void dft_tw(float2* x, const float2* w) {
    if constexpr (R == 16) {
        dft16_fused<true>(x, w);
    } else {
#pragma unroll
        for (int j = 1; j < R; ++j) x[j] = cmul(x[j], w[j]);
        dft<R>(x);
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
