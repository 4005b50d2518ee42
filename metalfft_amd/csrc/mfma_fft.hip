// N = 4096 complex FFT with every radix-16 stage done as a GEMM on the
// matrix cores (v_mfma_f32_16x16x32_{f16,bf16}, FP32 accumulate).
//
// Reference counterpart: the CT-MMA kernel (src/fft_4096_ct_mma.metal:88-206),
// which runs radix-8 CT-DIF stages as complex 8x8 GEMMs Y = F8 * X on Apple
// simdgroup_float8x8 tiles (Yr = Fr Xr + (-Fi) Xi, Yi = Fr Xi + Fi Xr,
// :147-151) with a host twiddle table.  Here, for gfx950:
//
//  * 4096 = 16 x 16 x 16: three self-sorting Stockham passes (the same pass
//    structure as the FP32 LDS kernels).  Pass p's 256 radix-16 butterflies
//    are a GEMM Y[k][b] = sum_j F16[k][j] x_j[b] over 16 tiles of 16
//    butterflies.
//  * Complex arithmetic folded into K = 32: B holds (re, im) of x_j
//    interleaved (K index 2j + c), A_re[k][2j] = cos, A_re[k][2j+1] = sin,
//    A_im[k][2j] = -sin, A_im[k][2j+1] = cos (angle 2 pi j k / 16), so one
//    16x16x32 MFMA gives Re(Y) and another Im(Y) for a 16-butterfly tile.
//  * Stage twiddles W_4096^{j*pos*N/(16 s)} are applied in FP32 on the VALU
//    before the operands are rounded to FP16/BF16; results stay FP32 in LDS.
//  * hi/lo split: x = x_hi + x_lo, F = F_hi + F_lo (each rounded to the MFMA
//    type), Y ~ F_lo x_hi + F_hi x_lo + F_hi x_hi accumulated in FP32 (three
//    MFMAs per half-tile), recovering roughly twice the mantissa bits.
//  * Persistent workgroups: the next transform streams HBM -> LDS by DMA
//    (global_load_lds) while the current one computes; the stage twiddles
//    are loaded into registers once per workgroup.
//  * Per-transform power-of-two scaling: the input is multiplied by 2^-e
//    (max |x| 2^-e < 1, a workgroup max-reduction), so FP16 operands stay in
//    range through the three passes (growth <= 16 per pass); the output is
//    multiplied back by 2^e, which is exact.
//
//  * Radix-64 form (R64): 4096 = 64 x 64, two self-sorting passes of 64
//    radix-64 butterflies, each radix-64 butterfly B = 16 w + c done as four
//    DFT-16s on the matrix cores (tile t4 = j4 takes X[j4 + 4 j16], column
//    c, so one lane holds the four sub-transforms' outputs at the same k16)
//    followed, in FP32 registers, by the twiddle W_64^{j4 k16} and a DFT-4
//    across the tiles:  Y[k16 + 16 k4] = sum_j4 W_4^{j4 k4} W_64^{j4 k16}
//    Z_j4[k16].  One LDS exchange and two operand splits per transform
//    instead of two and three, the same MFMA tiles per pass, two passes
//    instead of three.
//
// Lane layout of v_mfma_f32_16x16x32 (checked on the device): A lane l holds
// row l%16, K = 8*(l/16) .. +7; B lane l holds column l%16, the same K range;
// D lane l holds column l%16, rows 4*(l/16) .. +3.  So lane (c = l%16,
// g = l/16) supplies x_j for j = 4g .. 4g+3 of butterfly 16*tile + c and
// receives y_k for k = 4g .. 4g+3.
#include <hip/hip_runtime.h>

#include "fft_device.hpp"
#include "kernels.hpp"
#include "mfma_fft.hpp"
#include "stockham_core.hpp"

namespace mifft {

namespace {

typedef _Float16 h8 __attribute__((ext_vector_type(8)));
typedef __bf16 b8 __attribute__((ext_vector_type(8)));
typedef float f4 __attribute__((ext_vector_type(4)));
typedef _Float16 h2 __attribute__((ext_vector_type(2)));

template <bool BF>
struct MT;
template <>
struct MT<false> {
    using v8 = h8;
    using e = _Float16;
    static __device__ __forceinline__ f4 mma(v8 a, v8 b, f4 c) {
        return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
    }
};
template <>
struct MT<true> {
    using v8 = b8;
    using e = __bf16;
    static __device__ __forceinline__ f4 mma(v8 a, v8 b, f4 c) {
        return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
    }
};

// complex64 or complex-FP16 element in HBM
struct IoC64 {
    using T = float2;
    static __device__ __forceinline__ float2 ld(const T* p) { return *p; }
    static __device__ __forceinline__ void st(T* p, float2 v) { *p = v; }
};
struct IoC32h {
    using T = h2;
    static __device__ __forceinline__ float2 ld(const T* p) {
        h2 v = *p;
        return make_float2((float)v.x, (float)v.y);
    }
    static __device__ __forceinline__ void st(T* p, float2 v) {
        h2 o = {(_Float16)v.x, (_Float16)v.y};
        *p = o;
    }
};

// LDS index: +1 complex per 16 and +16 per 1024 spread the stride-16 and
// stride-1024 lane groups of the MFMA operand layout over the banks:
// midx(i) = i + (i >> 4) + ((i >> 10) << 4).  For lane (w, c, g), tile t4
// and element i (butterfly b = 16*(4w + t4) + c, j = 4g + i) the pass
// addresses reduce to a per-lane base plus constants:
//   read  midx(b + 256 j)                 = 68w + c + 1104g + 17 t4 + 272 i
//   write midx(16 b + j)         (s = 1)  = 1104w + 17c + 4g + 272 t4 + i
//   write midx(256(b/16) + b%16 + 16 j)   = 1104w + c + 68g + 272 t4 + 17 i
// (b < 256, so the >> 10 term is g or w).  The FP16 split form, whose lane
// coordinates come from an opaque copy of the thread index (so the compiler
// cannot bound them), uses these forms directly: the t4 and i terms become
// immediate offsets of the LDS instructions (fp16 split, complex-FP16 I/O:
// 1082 -> 888 VALU instructions per thread, 44.4-44.8 -> 41.8-42.0 us per
// 4096 x 4096 batch).  The other forms keep midx(): there the compiler
// already simplifies it, and the closed forms measured 1-3 % slower.
__device__ __forceinline__ int midx(int i) { return i + (i >> 4) + ((i >> 10) << 4); }
constexpr int kMfmaLds = 4096 + 256 + 64;

// Radix-64 form's LDS index: bits 6-9 of a spread onto bits 0, 1, 3, 4 and
// bit 10 onto bit 4 (XOR, a bijection on [0, 4096)).  Pass 0 writes
// 64 (16 w + c) + 4 g + i + 16 k4 and pass 1 reads 16 w + c + 64 (t4 + 16 g
// + 4 i); in both, the 32 lanes of a half-wave land on 32 distinct 8-byte
// slots (checked on the CPU for every lane and element).
__device__ __forceinline__ int r64idx(int a) {
    const int m = ((a >> 6) & 3) | (((a >> 8) & 1) << 3) | ((((a >> 9) ^ (a >> 10)) & 1) << 4);
    return a ^ m;
}

template <bool BF, bool SPLIT>
__device__ __forceinline__ void split8(const float (&x)[8], typename MT<BF>::v8& hi,
                                       typename MT<BF>::v8& lo) {
    using E = typename MT<BF>::e;
    if constexpr (SPLIT && !BF) {
        // FP16 hi/lo with the packed round-toward-zero convert (two values
        // per instruction): hi keeps the top 11 bits of x, x - hi is exact in
        // FP32 and lo keeps its top 11, about 22 bits in all
#pragma unroll
        for (int i = 0; i < 8; i += 2) {
            const h2 h = __builtin_bit_cast(h2, __builtin_amdgcn_cvt_pkrtz(x[i], x[i + 1]));
            const h2 l = __builtin_bit_cast(
                h2, __builtin_amdgcn_cvt_pkrtz(x[i] - (float)h.x, x[i + 1] - (float)h.y));
            hi[i] = h.x;
            hi[i + 1] = h.y;
            lo[i] = l.x;
            lo[i + 1] = l.y;
        }
    } else {
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            E h = (E)x[i];
            hi[i] = h;
            if constexpr (SPLIT) lo[i] = (E)(x[i] - (float)h);
        }
    }
}

template <bool BF, bool SPLIT>
__device__ __forceinline__ f4 gemm_half(typename MT<BF>::v8 a_hi, typename MT<BF>::v8 a_lo,
                                        typename MT<BF>::v8 b_hi, typename MT<BF>::v8 b_lo) {
    f4 acc = {0.f, 0.f, 0.f, 0.f};
    if constexpr (SPLIT) {
        acc = MT<BF>::mma(a_lo, b_hi, acc);  // small terms first
        acc = MT<BF>::mma(a_hi, b_lo, acc);
    }
    return MT<BF>::mma(a_hi, b_hi, acc);
}

// STAGE: persistent workgroups, the next transform staged HBM -> LDS by DMA
// during the current one (register budget for three resident workgroups
// per CU); otherwise one transform per workgroup loaded straight into
// registers, no stage buffer (35 KiB of LDS: four workgroups per CU with
// complex64 data, where the staged form fits two).
template <bool BF, bool SPLIT, bool INV, class IO, bool STAGE, bool R64>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(STAGE || !(SPLIT && !BF) ? 3 : 4, 8)))
k_fft4096_mfma(const typename IO::T* in, typename IO::T* out, const float2* __restrict__ tw,
               int64_t batch, float out_scale) {
    constexpr int N = 4096;
    constexpr int PASSES = R64 ? 2 : 3;
    // stage: the next transform as it sits in HBM (DMA target, natural order);
    // lds: FP32 exchange buffer of the passes.
    // SHARED (staged radix-64 FP16 hi/lo form, complex64 data): the stage
    // buffer is also the exchange buffer.  The next transform's DMA is issued
    // once pass 1 has read the exchange (overlapping pass 1's MFMAs and
    // stores instead of the whole transform), for 32 KiB of LDS instead of
    // 64: three workgroups per CU instead of two (53 against 57.5 us per
    // 4096 x 4096).  The other precisions need more than the 168 registers of
    // three waves per SIMD there and spill, and keep two buffers.
    constexpr bool SHARED = STAGE && R64 && SPLIT && !BF && sizeof(typename IO::T) == sizeof(float2);
    __shared__ typename IO::T stage[STAGE ? N : 1];
    __shared__ float2 lds_own[SHARED ? 1 : (R64 ? N : kMfmaLds)];
    static_assert(!SHARED || sizeof(stage) >= N * sizeof(float2), "the shared buffer holds the exchange");
    float2* const lds = SHARED ? reinterpret_cast<float2*>(stage) : lds_own;
    __shared__ float red[4];
    const int lane = threadIdx.x & 63;
    const int w = threadIdx.x >> 6;
    const int c = lane & 15;
    const int g = lane >> 4;
    using V8 = typename MT<BF>::v8;

    // DFT-16 operands of this lane: row k = c, columns j = 4g .. 4g+3
    V8 are_hi, are_lo, aim_hi, aim_lo;
    {
        float fr[8], fi[8];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int j = 4 * g + i;
            const float2 W = tw[(256 * j * c) & (N - 1)];  // exp(-2 pi i j c / 16)
            fr[2 * i] = W.x;
            fr[2 * i + 1] = -W.y;
            fi[2 * i] = W.y;
            fi[2 * i + 1] = W.x;
        }
        split8<BF, SPLIT>(fr, are_hi, are_lo);
        split8<BF, SPLIT>(fi, aim_hi, aim_lo);
    }
    // stage twiddles W_4096^{j*pos*N/(16 s)}: pass 1 (s = 16) has pos = c for
    // every tile, pass 2 (s = 256) pos = 16*(4w + tile) + c.
    // R64: pass 1's W_4096^{J B} (J = t4 + 16 g + 4 i, B = 16 w + c) as
    // W^{B (16 g + 4 i)} (tw1[i]) times W^{B t4} (tw2[0][t4]); the in-butterfly
    // W_64^{t4 k16} (k16 = 4 g + i) as tw2[1][i] = W_64^{k16}, squared and
    // multiplied at use.
    float2 tw1[4], tw2[4][4];
    if constexpr (R64) {
        const int B = 16 * w + c;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            tw1[i] = tw[B * (16 * g + 4 * i)];
            tw2[0][i] = tw[B * i];
            tw2[1][i] = tw[64 * (4 * g + i)];
        }
    } else {
#pragma unroll
        for (int i = 0; i < 4; ++i) tw1[i] = tw[(4 * g + i) * c * 16];
#pragma unroll
        for (int t4 = 0; t4 < 4; ++t4) {
            const int pos = 16 * (4 * w + t4) + c;
#pragma unroll
            for (int i = 0; i < 4; ++i) tw2[t4][i] = tw[(4 * g + i) * pos];
        }
    }

    constexpr int kElemBytes = sizeof(typename IO::T);
    constexpr int kCount = N * kElemBytes / 8;  // transform size in float2 units (DMA granule)
    int64_t tr = blockIdx.x;
    if (STAGE && tr < batch)
        dma_transform(reinterpret_cast<float2*>(stage),
                      reinterpret_cast<const float2*>(in + tr * N), kCount, threadIdx.x, 256);
    for (; tr < batch; tr += gridDim.x) {  // uniform per workgroup
        // FP16 split: lane coordinates from an opaque copy of the thread
        // index, so the LDS and HBM index arithmetic below is recomputed per
        // transform instead of hoisted out of the loop and held in registers
        // across it (which spills that form); the others are faster with it
        // hoisted
        const int tid = SPLIT && !BF ? fresh((int)threadIdx.x) : (int)threadIdx.x;
        const int lane = tid & 63;
        const int w = tid >> 6;
        const int c = lane & 15;
        const int g = lane >> 4;
        typename IO::T* dst = out + tr * N;
        if constexpr (STAGE) {
            __builtin_amdgcn_s_waitcnt(0);  // this wave's DMA (and earlier stores) done
            ds_barrier();                   // every wave's DMA landed; previous transform done
        }
        // the transform as it sits in LDS (staged) or in HBM: 16 lanes of a
        // row read 16 consecutive elements.  One-shot radix-64: the transform
        // is first copied HBM -> LDS through registers with lane-contiguous
        // loads (a wave reads 64 consecutive elements per instruction, where
        // the operand layout would read four runs of 16 elements 1024 apart)
        // into the exchange buffer, which pass 0 overwrites only after the
        // max-reduction barrier below
        constexpr bool VIA_LDS = !STAGE && R64;
        if constexpr (VIA_LDS) {
            typename IO::T x[16];
            const typename IO::T* g = in + tr * N;
#pragma unroll
            for (int m = 0; m < 16; ++m) x[m] = g[tid + 256 * m];
            typename IO::T* st = reinterpret_cast<typename IO::T*>(lds_own);
#pragma unroll
            for (int m = 0; m < 16; ++m) st[tid + 256 * m] = x[m];
            ds_barrier();
        }
        const typename IO::T* src = STAGE     ? static_cast<const typename IO::T*>(stage)
                                    : VIA_LDS ? reinterpret_cast<const typename IO::T*>(lds_own)
                                              : in + tr * N;
        float2 v[4][4];  // [tile][i]: x_{4g+i} of butterfly 16*(4w+tile) + c
        float m = 0.f;
#pragma unroll
        for (int t4 = 0; t4 < 4; ++t4)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                // R64: X_J of butterfly 16 w + c, J = t4 + 16 g + 4 i
                float2 x = IO::ld(src + (R64 ? 16 * w + c + 64 * t4 + 1024 * g + 256 * i
                                             : 16 * (4 * w + t4) + c + 256 * (4 * g + i)));
                if constexpr (INV) x.y = -x.y;
                v[t4][i] = x;
                m = fmaxf(m, fmaxf(fabsf(x.x), fabsf(x.y)));
            }
        // workgroup max -> power-of-two scale
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) m = fmaxf(m, __shfl_xor(m, off));
        if (lane == 0) red[w] = m;
        ds_barrier();  // stage fully read, red[] written
        m = fmaxf(fmaxf(red[0], red[1]), fmaxf(red[2], red[3]));
        if (STAGE && !SHARED && tr + gridDim.x < batch)
            dma_transform(reinterpret_cast<float2*>(stage),
                          reinterpret_cast<const float2*>(in + (tr + gridDim.x) * N), kCount,
                          tid, 256);
        int e = 0;
        if (m > 0.f) (void)frexpf(m, &e);  // m = f 2^e, f in [0.5, 1)
        // a subnormal maximum (below 2^-126) would give 2^-e = inf and turn
        // zero entries into NaN: the scale stops at 2^125 (max |x| 2^-e < 1
        // still holds, the data are just not normalised all the way)
        if (e < -125) e = -125;
        const float sc = ldexpf(1.f, -e);
#pragma unroll
        for (int t4 = 0; t4 < 4; ++t4)
#pragma unroll
            for (int i = 0; i < 4; ++i) v[t4][i] = make_float2(v[t4][i].x * sc, v[t4][i].y * sc);

        constexpr bool CLOSED = SPLIT && !BF;        // closed-form pass addresses
        const int rd0 = 68 * w + c + 1104 * g;      // pass reads
        const int wr0 = 1104 * w + 17 * c + 4 * g;  // pass 0 writes
        const int wr1 = 1104 * w + c + 68 * g;      // pass 1 writes
#pragma unroll
        for (int p = 0; p < PASSES; ++p) {
            const int s = p == 0 ? 1 : (p == 1 ? 16 : 256);
            if (p > 0) {
#pragma unroll
                for (int t4 = 0; t4 < 4; ++t4) {
                    const int b = 16 * (4 * w + t4) + c;
#pragma unroll
                    for (int i = 0; i < 4; ++i) {
                        if constexpr (R64) {
                            const float2 wj = t4 == 0 ? tw1[i] : cmul(tw1[i], tw2[0][t4]);
                            v[t4][i] = cmul(lds[r64idx(16 * w + c + 64 * (t4 + 16 * g + 4 * i))], wj);
                        } else {
                            const int a = CLOSED ? rd0 + 17 * t4 + 272 * i : midx(b + 256 * (4 * g + i));
                            v[t4][i] = cmul(lds[a], p == 1 ? tw1[i] : tw2[t4][i]);
                        }
                    }
                }
                if constexpr (SHARED) {  // the exchange is read: the buffer takes the next transform
                    ds_barrier();
                    if (tr + gridDim.x < batch)
                        dma_transform(reinterpret_cast<float2*>(stage),
                                      reinterpret_cast<const float2*>(in + (tr + gridDim.x) * N), kCount, tid,
                                      256);
                }
            }
            float2 y[4][4];
#pragma unroll
            for (int t4 = 0; t4 < 4; ++t4) {
                float xb[8];
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    xb[2 * i] = v[t4][i].x;
                    xb[2 * i + 1] = v[t4][i].y;
                }
                V8 b_hi, b_lo;
                split8<BF, SPLIT>(xb, b_hi, b_lo);
                const f4 yr = gemm_half<BF, SPLIT>(are_hi, are_lo, b_hi, b_lo);
                const f4 yi = gemm_half<BF, SPLIT>(aim_hi, aim_lo, b_hi, b_lo);
#pragma unroll
                for (int i = 0; i < 4; ++i) y[t4][i] = make_float2(yr[i], yi[i]);
            }
            if constexpr (R64) {
                // y[t4][i] = Z_t4[k16], k16 = 4 g + i: times W_64^{t4 k16}, then
                // the DFT-4 over t4 gives Y[k16 + 16 k4] in y[k4][i]
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const float2 w1 = tw2[1][i];
                    const float2 w2 = cmul(w1, w1);
                    const float2 a0 = y[0][i], a1 = cmul(y[1][i], w1), a2 = cmul(y[2][i], w2),
                                 a3 = cmul(y[3][i], cmul(w1, w2));
                    const float2 s02 = make_float2(a0.x + a2.x, a0.y + a2.y);
                    const float2 d02 = make_float2(a0.x - a2.x, a0.y - a2.y);
                    const float2 s13 = make_float2(a1.x + a3.x, a1.y + a3.y);
                    const float2 d13 = make_float2(a1.x - a3.x, a1.y - a3.y);
                    y[0][i] = make_float2(s02.x + s13.x, s02.y + s13.y);
                    y[2][i] = make_float2(s02.x - s13.x, s02.y - s13.y);
                    y[1][i] = make_float2(d02.x + d13.y, d02.y - d13.x);  // d02 - i d13
                    y[3][i] = make_float2(d02.x - d13.y, d02.y + d13.x);  // d02 + i d13
                }
            }
            if (p < PASSES - 1) {
                if (p > 0) ds_barrier();  // all reads of lds for this pass are done
#pragma unroll
                for (int t4 = 0; t4 < 4; ++t4) {
                    const int b = 16 * (4 * w + t4) + c;
                    const int base = (b / s) * (16 * s) + (b & (s - 1));
#pragma unroll
                    for (int i = 0; i < 4; ++i) {
                        int a;
                        if constexpr (R64)  // Y[4 g + i + 16 t4] of butterfly 16 w + c, stride 1
                            a = r64idx(64 * (16 * w + c) + 4 * g + i + 16 * t4);
                        else
                            a = !CLOSED ? midx(base + (4 * g + i) * s)
                                : p == 0 ? wr0 + 272 * t4 + i : wr1 + 272 * t4 + 17 * i;
                        lds[a] = y[t4][i];
                    }
                }
                ds_barrier();
            } else {
                // last pass: y_k of butterfly b goes to b + 256 k (R64: Y_K of
                // butterfly 16 w + c, K = 4 g + i + 16 t4, to 16 w + c + 64 K)
                float osc = ldexpf(out_scale, e);
                if constexpr (INV) osc *= 1.0f / (float)N;
#pragma unroll
                for (int t4 = 0; t4 < 4; ++t4) {
                    const int b = 16 * (4 * w + t4) + c;
#pragma unroll
                    for (int i = 0; i < 4; ++i) {
                        float2 r = make_float2(y[t4][i].x * osc, y[t4][i].y * osc);
                        if constexpr (INV) r.y = -r.y;
                        IO::st(dst + (R64 ? 16 * w + c + 64 * (4 * g + i + 16 * t4) : b + 256 * (4 * g + i)), r);
                    }
                }
            }
        }
    }
}

int device_cus() {
    static int cus = [] {
        int dev = 0, n = 0;
        if (hipGetDevice(&dev) != hipSuccess) return 256;
        if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
            return 256;
        return n > 0 ? n : 256;
    }();
    return cus;
}

template <bool BF, bool SPLIT, class IO, bool STAGE, bool R64>
hipError_t launch_mfma(const void* in, void* out, const float2* tw, int64_t batch, bool inverse,
                       float out_scale, hipStream_t st) {
    if (batch <= 0) return hipSuccess;
    int64_t blocks = batch;  // one-shot: one transform per workgroup
    if constexpr (STAGE) {
        // resident workgroups only (each loops over transforms, prefetching)
        static int per_cu = [] {
            int nb = 0;
            if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
                    &nb, reinterpret_cast<const void*>(&k_fft4096_mfma<BF, SPLIT, false, IO, true, R64>), 256,
                    0) != hipSuccess || nb <= 0)
                nb = 1;
            return nb;
        }();
        const int64_t cap = (int64_t)device_cus() * per_cu;
        if (blocks > cap) blocks = cap;
    }
    if (blocks > 0x7fffffffLL) return hipErrorInvalidValue;
    auto i = static_cast<const typename IO::T*>(in);
    auto o = static_cast<typename IO::T*>(out);
    if (inverse)
        hipLaunchKernelGGL((k_fft4096_mfma<BF, SPLIT, true, IO, STAGE, R64>), dim3((unsigned)blocks), dim3(256),
                           0, st, i, o, tw, batch, out_scale);
    else
        hipLaunchKernelGGL((k_fft4096_mfma<BF, SPLIT, false, IO, STAGE, R64>), dim3((unsigned)blocks),
                           dim3(256), 0, st, i, o, tw, batch, out_scale);
    return hipGetLastError();
}

template <bool BF, bool SPLIT, class IO>
hipError_t launch_mfma_form(int form, const void* in, void* out, const float2* tw, int64_t batch, bool inverse,
                            float out_scale, hipStream_t st) {
    switch (form) {
        case kMfmaStaged: return launch_mfma<BF, SPLIT, IO, true, false>(in, out, tw, batch, inverse, out_scale, st);
        case kMfmaOneShot: return launch_mfma<BF, SPLIT, IO, false, false>(in, out, tw, batch, inverse, out_scale, st);
        case kMfmaStaged64: return launch_mfma<BF, SPLIT, IO, true, true>(in, out, tw, batch, inverse, out_scale, st);
        case kMfmaOneShot64:
            return launch_mfma<BF, SPLIT, IO, false, true>(in, out, tw, batch, inverse, out_scale, st);
    }
    return hipErrorInvalidValue;
}

template <bool BF, bool SPLIT>
hipError_t launch_mfma_io(int io, int form, const void* in, void* out, const float2* tw, int64_t batch,
                          bool inverse, float out_scale, hipStream_t st) {
    if (io == 0) return launch_mfma_form<BF, SPLIT, IoC64>(form, in, out, tw, batch, inverse, out_scale, st);
    return launch_mfma_form<BF, SPLIT, IoC32h>(form, in, out, tw, batch, inverse, out_scale, st);
}

}  // namespace

hipError_t launch_fft_mfma(int n, int precision, int io, int form, const void* in, void* out,
                           const float2* tw, int64_t batch, bool inverse, float out_scale, hipStream_t st) {
    if (n != 4096 || (io != 0 && io != 1) || form < 0 || form >= kMfmaForms) return hipErrorInvalidValue;
    switch (precision) {
        case 1: return launch_mfma_io<false, false>(io, form, in, out, tw, batch, inverse, out_scale, st);
        case 2: return launch_mfma_io<true, false>(io, form, in, out, tw, batch, inverse, out_scale, st);
        case 3: return launch_mfma_io<false, true>(io, form, in, out, tw, batch, inverse, out_scale, st);
        case 4: return launch_mfma_io<true, true>(io, form, in, out, tw, batch, inverse, out_scale, st);
    }
    return hipErrorInvalidValue;
}

const char* mfma_form_name(int precision, int io, int form) {
    static const char* names[kMfmaForms][2][5] = {
        {{"", "mfma_f16_4096_c64", "mfma_bf16_4096_c64", "mfma_f16split_4096_c64", "mfma_bf16split_4096_c64"},
         {"", "mfma_f16_4096_c32h", "mfma_bf16_4096_c32h", "mfma_f16split_4096_c32h",
          "mfma_bf16split_4096_c32h"}},
        {{"", "mfma_f16_4096_c64_oneshot", "mfma_bf16_4096_c64_oneshot", "mfma_f16split_4096_c64_oneshot",
          "mfma_bf16split_4096_c64_oneshot"},
         {"", "mfma_f16_4096_c32h_oneshot", "mfma_bf16_4096_c32h_oneshot", "mfma_f16split_4096_c32h_oneshot",
          "mfma_bf16split_4096_c32h_oneshot"}},
        {{"", "mfma_f16_4096_c64_r64", "mfma_bf16_4096_c64_r64", "mfma_f16split_4096_c64_r64",
          "mfma_bf16split_4096_c64_r64"},
         {"", "mfma_f16_4096_c32h_r64", "mfma_bf16_4096_c32h_r64", "mfma_f16split_4096_c32h_r64",
          "mfma_bf16split_4096_c32h_r64"}},
        {{"", "mfma_f16_4096_c64_r64_oneshot", "mfma_bf16_4096_c64_r64_oneshot",
          "mfma_f16split_4096_c64_r64_oneshot", "mfma_bf16split_4096_c64_r64_oneshot"},
         {"", "mfma_f16_4096_c32h_r64_oneshot", "mfma_bf16_4096_c32h_r64_oneshot",
          "mfma_f16split_4096_c32h_r64_oneshot", "mfma_bf16split_4096_c32h_r64_oneshot"}}};
    if (precision < 1 || precision > 4 || io < 0 || io > 1 || form < 0 || form >= kMfmaForms) return "";
    return names[form][io][precision];
}

}  // namespace mifft
