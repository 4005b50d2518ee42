// FP16 arithmetic for the N = 4096 radix-16 Stockham transform in the three
// precision modes of the reference's FP16 kernels
// (Sources/MetalFFT/Resources/fft_fused_convolve_fp16.metal:160-498,
// src/fft_sar_fused_fp16.metal:10-34):
//   PURE     butterflies, twiddle powers, twiddle and filter products in FP16
//            (packed f16 VALU), FP16 LDS exchange;
//   STORAGE  every LDS exchange stored as FP16, all arithmetic in FP32;
//   MIXED    twiddle powers, twiddle and filter products in FP16,
//            butterflies in FP32, FP16 LDS exchange.
// Shared by the fused convolver and the SAR rows.
#pragma once
#include <hip/hip_runtime.h>

#include <type_traits>

#include "stockham_core.hpp"

namespace mifft {
namespace fp16 {

typedef _Float16 h2 __attribute__((ext_vector_type(2)));

enum { kConvPure = 1, kConvStorage = 2, kConvMixed = 3 };

__device__ __forceinline__ h2 to_h2(float2 a) { return (h2){(_Float16)a.x, (_Float16)a.y}; }
__device__ __forceinline__ float2 to_f2(h2 a) { return make_float2((float)a.x, (float)a.y); }
__device__ __forceinline__ float2 round_h(float2 a) { return to_f2(to_h2(a)); }
// Complex product as two packed FP16 operations: a.x * (b.x, b.y) +
// a.y * (-b.y, b.x).  The broadcasts and the swapped, negated operand are
// op_sel / neg modifiers of v_pk_mul_f16 and v_pk_fma_f16, so no scalar
// f16 multiply or repacking is issued (written component-wise, the compiler
// scalarised every product).
__device__ __forceinline__ h2 hmul(h2 a, h2 b) {
    const h2 ax = __builtin_shufflevector(a, a, 0, 0);
    const h2 ay = __builtin_shufflevector(a, a, 1, 1);
    const h2 bs = __builtin_shufflevector(-b, b, 1, 2);  // (-b.y, b.x)
    return __builtin_elementwise_fma(ay, bs, ax * b);
}
// multiply by -i: (a.y, -a.x)
__device__ __forceinline__ h2 hmul_mi(h2 a) { return __builtin_shufflevector(a, -a, 1, 2); }
__device__ __forceinline__ h2 hconj(h2 a) { return __builtin_shufflevector(a, -a, 0, 3); }

__device__ __forceinline__ void hdft4(h2& x0, h2& x1, h2& x2, h2& x3) {
    const h2 t0 = x0 + x2, t1 = x1 + x3, t2 = x0 - x2, t3 = hmul_mi(x1 - x3);
    x0 = t0 + t1;
    x2 = t0 - t1;
    x1 = t2 + t3;
    x3 = t2 - t3;
}

// DFT-16 in FP16: the 4 x 4 split of dft16 (fft_device.hpp), every operation
// rounded to FP16.
__device__ __forceinline__ void hdft16(h2* x) {
    h2 a[4][4];
#pragma unroll
    for (int j2 = 0; j2 < 4; ++j2) {
        h2 p0 = x[j2], p1 = x[4 + j2], p2 = x[8 + j2], p3 = x[12 + j2];
        hdft4(p0, p1, p2, p3);
        a[j2][0] = p0; a[j2][1] = p1; a[j2][2] = p2; a[j2][3] = p3;
    }
    const _Float16 c = (_Float16)MIFFT_C16, sn = (_Float16)MIFFT_S16, h = (_Float16)MIFFT_SQRT1_2;
    const h2 w1 = (h2){c, -sn}, w3 = (h2){sn, -c}, w9 = (h2){-c, sn};
    const h2 hh = (h2){h, h};
    // W16^2 v = h (v.x + v.y, v.y - v.x) = h (v + (v.y, -v.x));
    // W16^6 v = h (v.y - v.x, -v.x - v.y) = h ((v.y, -v.x) - v)
    auto w2 = [&](h2 v) { return (v + hmul_mi(v)) * hh; };
    auto w6 = [&](h2 v) { return (hmul_mi(v) - v) * hh; };
    a[1][1] = hmul(a[1][1], w1);
    a[1][2] = w2(a[1][2]);
    a[1][3] = hmul(a[1][3], w3);
    a[2][1] = w2(a[2][1]);
    a[2][2] = hmul_mi(a[2][2]);
    a[2][3] = w6(a[2][3]);
    a[3][1] = hmul(a[3][1], w3);
    a[3][2] = w6(a[3][2]);
    a[3][3] = hmul(a[3][3], w9);
#pragma unroll
    for (int k1 = 0; k1 < 4; ++k1) {
        h2 q0 = a[0][k1], q1 = a[1][k1], q2 = a[2][k1], q3 = a[3][k1];
        hdft4(q0, q1, q2, q3);
        x[k1] = q0; x[k1 + 4] = q1; x[k1 + 8] = q2; x[k1 + 12] = q3;
    }
}

// w[j] = W^{jk}, j < 16, from the FP32 bases W^k, W^2k, W^4k, W^8k, the
// products in FP16 (the reference forms w2 = w1*w1, w3 = w2*w1 in half).
__device__ __forceinline__ void hpowers16(h2 (&w)[16], const float2* b) {
    w[0] = (h2){(_Float16)1.f, (_Float16)0.f};
    w[1] = to_h2(b[0]);
    w[2] = to_h2(b[1]);
    w[3] = hmul(w[1], w[2]);
    w[4] = to_h2(b[2]);
    w[5] = hmul(w[1], w[4]);
    w[6] = hmul(w[2], w[4]);
    w[7] = hmul(w[3], w[4]);
    w[8] = to_h2(b[3]);
#pragma unroll
    for (int j = 9; j < 16; ++j) w[j] = hmul(w[j - 8], w[8]);
}

template <int PM>
using ConvReg = typename std::conditional<PM == kConvPure, h2, float2>::type;

// Pass P of the N = 4096 radix-16 Stockham transform (stockham_regs) with
// the mode's arithmetic; pb: the thread's FP32 table entries (PreBases).
// SCALE: every pass's outputs times 1/16 (exact in FP16 and FP32 while
// normal), the transform divided by N without growing past its input.
template <int PM, int P, bool SYNC0, bool SCALE = false>
__device__ __forceinline__ void conv_passes_h(ConvReg<PM> (&v)[16], h2* buf, int t,
                                              const float2 (*pb)[16]) {
    using S = Schedule<4096, 16>;
    constexpr int SD = S::stride(P);
    constexpr int T = S::T;
    if constexpr (SD > 1) {
        if constexpr (PM == kConvStorage) {
            float2 w[16];
            powers_from_bases<16>(w, pb[P - 1]);
#pragma unroll
            for (int j = 1; j < 16; ++j) v[j] = cmul(v[j], w[j]);
        } else {
            h2 w[16];
            hpowers16(w, pb[P - 1]);
#pragma unroll
            for (int j = 1; j < 16; ++j) {
                if constexpr (PM == kConvPure) v[j] = hmul(v[j], w[j]);
                else v[j] = to_f2(hmul(to_h2(v[j]), w[j]));
            }
        }
    }
    if constexpr (PM == kConvPure) hdft16(v); else dft<16>(v);
    if constexpr (SCALE) {
#pragma unroll
        for (int j = 0; j < 16; ++j) {
            if constexpr (PM == kConvPure) v[j] = v[j] * (_Float16)0.0625f;
            else v[j] = make_float2(v[j].x * 0.0625f, v[j].y * 0.0625f);
        }
    }
    if constexpr (P < S::passes - 1) {
        if constexpr (P > 0 || SYNC0) lds_barrier();
        const int base = (t / SD) * (16 * SD) + (t & (SD - 1));
#pragma unroll
        for (int j = 0; j < 16; ++j) {
            if constexpr (PM == kConvPure) buf[LdsPad::idx(base + j * SD)] = v[j];
            else buf[LdsPad::idx(base + j * SD)] = to_h2(v[j]);
        }
        lds_barrier();
#pragma unroll
        for (int m = 0; m < 16; ++m) {
            if constexpr (PM == kConvPure) v[m] = buf[LdsPad::idx(t + T * m)];
            else v[m] = to_f2(buf[LdsPad::idx(t + T * m)]);
        }
        conv_passes_h<PM, P + 1, SYNC0, SCALE>(v, buf, t, pb);
    }
}

}  // namespace fp16
}  // namespace mifft
