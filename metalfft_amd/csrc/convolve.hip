// Fused FFT -> multiply -> IFFT FIR convolution of many real signals in one
// launch, and its hand-written adjoint.
//
// Reference: Convolver (Sources/MetalFFT/DSP/Convolve.swift:19-76) and
// FusedConvolver (Sources/MetalFFT/DSP/FusedConvolver.swift:55-133) compute the
// full linear convolution y = x * h (length S + K - 1) block by block with
// FFT size N and hop N - K + 1, one synchronous dispatch per block, adding
// overlapping block outputs on the CPU.
//
// Here every block of every signal runs in one launch, arranged so each
// output sample is written by exactly one workgroup (no atomics, no CPU
// overlap-add, deterministic):
//   * overlap-save: block b reads x[b*hop - (K-1) + n], n < N, and its
//     circular convolution with h holds y[b*hop + j] at n = j + K - 1 for
//     j < hop (the first K - 1 outputs are wrapped and dropped);
//   * two real blocks per complex FFT: z = w_{2q} + i w_{2q+1}; h is real, so
//     IFFT(Z * H) = (w_{2q} (*) h) + i (w_{2q+1} (*) h).
// Same linear convolution as the reference's overlap-add to FP32 rounding.
//
// Adjoint (no reference counterpart).  With g = dL/dy:
//   grad_x[n] = sum_k h[k] g[n + k]           (correlation of g with h)
//   grad_h[k] = sum_s sum_n x_s[n] g_s[n + k] (correlation of g with x)
// Block b of hop samples: window gw_b[n] = g[b*hop + n] (n < N) and
// xb_b[n] = x[b*hop + n] (n < hop, zero-padded).  For j < hop and k < K no
// index wraps, so
//   grad_x[b*hop + j] = IFFT(GW_b * conj(H))[j]
//   sum_b grad_h-part = Re IFFT( sum_q conj(U_q) * Z_q )[k],
// with Z_q = GW_{2q} + i GW_{2q+1}, U_q = XB_{2q} + i XB_{2q+1}: the real part
// of conj(U) Z is conj(XB_0) GW_0 + conj(XB_1) GW_1, the two blocks' cross
// spectra.  The signal spectra are recomputed from x, never stored.  Each
// workgroup accumulates its pairs' cross spectra in registers, runs one
// inverse FFT and writes K partial tap gradients; a second kernel sums the
// partials in a fixed order.
#include <hip/hip_runtime.h>

#include <string>
#include <type_traits>

#include "convolve.hpp"
#include "fp16_core.hpp"
#include "runtime.hpp"
#include "stockham_core.hpp"

namespace mifft {

namespace {

struct ConvGeom {
    int64_t S;      // samples per signal
    int64_t count;  // signals
    int K;          // taps
    int hop;        // N - K + 1
    int64_t nblk;   // blocks per signal
    int64_t npair;  // block pairs per signal
};

// A block's window into one signal: element n (0 <= n < N) is sig[w + n]
// when 0 <= w + n < len, else 0.  base + clo .. base + chi - 1 is always
// valid memory (len >= 1), even for an empty window.
struct Win {
    const float* base;
    int lo, hi;    // valid n
    int clo, chi;  // clamp range for the unconditional load
};

__device__ __forceinline__ Win window(const float* sig, int64_t len, int64_t w, int n_max) {
    int64_t lo = w < 0 ? -w : 0, hi = len - w;
    if (hi > n_max) hi = n_max;
    if (lo >= hi) return Win{sig, 0, 0, 0, 1};
    return Win{sig + w, (int)lo, (int)hi, (int)lo, (int)hi};
}

// Window element n, +0 outside.  The load is unconditional, of a clamped
// index, and masked with an opaque all-ones/zero word: a load the compiler
// may skip on some path gets sunk under a branch and waited for at the
// join, one HBM round trip per element, where unconditional loads all stay
// in flight together.  32-bit offsets from the window base keep the address
// arithmetic small.
__device__ __forceinline__ float ld_win(const Win& W, int n) {
    const int c = n < W.clo ? W.clo : (n >= W.chi ? W.chi - 1 : n);
    const float v = W.base[c];
    uint32_t keep = (n >= W.lo && n < W.hi) ? 0xffffffffu : 0u;
    __asm__("" : "+v"(keep));
    return __uint_as_float(__float_as_uint(v) & keep);
}

// Signal s = item / npair and pair q = item % npair, in 32-bit arithmetic
// when the whole grid fits (a 64-bit division expands to ~270 instructions,
// the 32-bit one to a few dozen).
__device__ __forceinline__ void pair_of(int64_t item, int64_t total, int64_t npair, int64_t& s,
                                        int64_t& q) {
    if (total <= 0x7fffffffLL) {
        const uint32_t i32 = (uint32_t)item, n32 = (uint32_t)npair;
        const uint32_t s32 = i32 / n32;
        s = s32;
        q = i32 - s32 * n32;
    } else {
        s = item / npair;
        q = item % npair;
    }
}

// Element n of a float array at a 32-bit byte offset from its uniform base
// (global_load / global_store with a scalar base and a 32-bit lane offset).
__device__ __forceinline__ float ld_at(const float* base, uint32_t n) {
    return *reinterpret_cast<const float*>(reinterpret_cast<const char*>(base) + n * 4u);
}
__device__ __forceinline__ void st_at(float* base, uint32_t n, float v) {
    *reinterpret_cast<float*>(reinterpret_cast<char*>(base) + n * 4u) = v;
}

// --------------------------------------------------------------- forward
template <int N, int E, int TPW>
__global__ void __launch_bounds__((N / E) * TPW)
k_conv_fwd(const float* __restrict__ x, float* __restrict__ y, const float2* __restrict__ H,
           const float2* __restrict__ tw, ConvGeom g) {
    constexpr int T = N / E;
    __shared__ float2 lds[TPW * LdsPad::len(N)];
    // one pair per workgroup (TPW == 1): item, signal and pair are uniform,
    // their 64-bit division scalar
    const int local = TPW == 1 ? 0 : (int)threadIdx.x / T;
    const int t = TPW == 1 ? (int)threadIdx.x : (int)threadIdx.x % T;
    float2* buf = lds + local * LdsPad::len(N);
    const int64_t item = (int64_t)blockIdx.x * TPW + local;
    const int64_t total = g.count * g.npair;
    const bool active = item < total;
    int64_t s = 0, q = 0;
    if (active) pair_of(item, total, g.npair, s, q);
    const int64_t b0 = 2 * q, b1 = 2 * q + 1;
    const float* xs = x + s * g.S;
    const int64_t ylen = g.S + g.K - 1;
    float* ys = y + s * ylen;
    const int64_t w0 = b0 * g.hop - (g.K - 1);
    const int64_t w1 = b1 * g.hop - (g.K - 1);
    const bool has1 = b1 < g.nblk;

    float2 v[E];
    // interior pair (both windows inside the signal, every output in range):
    // plain loads at 32-bit offsets from a uniform base, no clamp or mask
    const bool interior = active && has1 && w0 >= 0 && w1 + N <= g.S;
    if (interior) {
        const float* p0 = xs + w0;
        const float* p1 = xs + w1;
#pragma unroll
        for (int m = 0; m < E; ++m) {
            const uint32_t n = (uint32_t)(t + T * m);
            v[m] = make_float2(ld_at(p0, n), ld_at(p1, n));
        }
    } else {
        const Win W0 = window(xs, g.S, w0, N), W1 = window(xs, g.S, w1, N);  // S >= 1
#pragma unroll
        for (int m = 0; m < E; ++m) {
            const int n = t + T * m;
            const float re = ld_win(W0, n);
            const float im = ld_win(W1, n);
            v[m] = make_float2(active ? re : 0.f, active && has1 ? im : 0.f);
        }
    }
    RegTwiddles<N, E> ct;
    ct.load(tw, t);
    ct.template fft<false>(v, buf, t, tw);
    // multiply by the filter spectrum, then inverse via conj-FFT-conj
#pragma unroll
    for (int m = 0; m < E; ++m) v[m] = conjf2(cmul(v[m], H[t + T * m]));
    ct.template fft<true>(v, buf, t, tw);
    if (!active) return;
    constexpr float inv_n = 1.0f / (float)N;
    const int ts = fresh(t);
    if (interior) {  // outputs j = n - (K - 1) < hop of both blocks, all in range
        float* q0 = ys + b0 * g.hop;
        float* q1 = ys + b1 * g.hop;
#pragma unroll
        for (int m = 0; m < E; ++m) {
            const int n = ts + T * m;
            if (n >= g.K - 1) {
                const uint32_t j = (uint32_t)(n - (g.K - 1));
                st_at(q0, j, v[m].x * inv_n);   // Re of conj(.)/N
                st_at(q1, j, -v[m].y * inv_n);  // Im of conj(.)/N
            }
        }
        return;
    }
#pragma unroll
    for (int m = 0; m < E; ++m) {
        const int n = ts + T * m;
        const int j = n - (g.K - 1);
        if (j >= 0 && j < g.hop) {
            const int64_t o0 = b0 * g.hop + j;
            if (o0 < ylen) ys[o0] = v[m].x * inv_n;          // Re of conj(.)/N
            const int64_t o1 = b1 * g.hop + j;
            if (has1 && o1 < ylen) ys[o1] = -v[m].y * inv_n;  // Im of conj(.)/N
        }
    }
}

// --------------------------------------------------------------- adjoint
// Two kernels: grad_x (two transforms per block pair, like the forward) and
// grad_h (two forward transforms per pair, spectra accumulated per slot,
// one inverse per slot).  One fused kernel needs both pair spectra, the
// accumulator and a third transform live at once, which costs it most of
// its occupancy; the split spends one more transform per pair instead.

// Block pair q of signal s: z = two gradient windows gw_b[n] = gy[b*hop + n].
template <int N, int E>
__device__ __forceinline__ void load_grad_pair(float2 (&z)[E], const float* gs, int64_t ylen,
                                               int64_t p0, int64_t p1, bool active, bool a1, int t) {
    constexpr int T = N / E;
    const Win G0 = window(gs, ylen, p0, N), G1 = window(gs, ylen, p1, N);
    const int tz = fresh(t);
#pragma unroll
    for (int m = 0; m < E; ++m) {
        const int n = tz + T * m;
        const float zr = ld_win(G0, n);
        const float zi = ld_win(G1, n);
        z[m] = make_float2(active ? zr : 0.f, a1 ? zi : 0.f);
    }
}

// grad_x[b*hop + j] = IFFT(GW_b * conj(H))[j], j < hop: one block pair per
// transform slot.
template <int N, int E, int TPW>
__global__ void __launch_bounds__((N / E) * TPW)
k_conv_bwd_x(const float* __restrict__ gy, float* __restrict__ gx, const float2* __restrict__ H,
             const float2* __restrict__ tw, ConvGeom g, int64_t nbx, int64_t npx) {
    constexpr int T = N / E;
    __shared__ float2 lds[TPW * LdsPad::len(N)];
    const int local = threadIdx.x / T;
    const int t = threadIdx.x % T;
    float2* buf = lds + local * LdsPad::len(N);
    const int64_t item = (int64_t)blockIdx.x * TPW + local;
    const int64_t total = g.count * npx;
    const bool active = item < total;
    int64_t s = 0, q = 0;
    if (active) pair_of(item, total, npx, s, q);
    const int64_t ylen = g.S + g.K - 1;
    const int64_t p0 = 2 * q * g.hop, p1 = p0 + g.hop;
    const bool has1 = 2 * q + 1 < nbx;
    // interior pair: both gradient windows inside gy (p1 + N <= ylen, which
    // is p1 + hop <= S) and every grad_x output in range
    const bool interior = active && has1 && p1 + N <= ylen;
    float2 z[E];
    if (interior) {
        const float* g0 = gy + s * ylen + p0;
        const float* g1 = g0 + g.hop;
#pragma unroll
        for (int m = 0; m < E; ++m) {
            const uint32_t n = (uint32_t)(t + T * m);
            z[m] = make_float2(ld_at(g0, n), ld_at(g1, n));
        }
    } else {
        load_grad_pair<N, E>(z, gy + s * ylen, ylen, p0, p1, active, active && has1, t);
    }
    RegTwiddles<N, E> ct;
    ct.load(tw, t);
    ct.template fft<false>(z, buf, t, tw);
#pragma unroll
    for (int m = 0; m < E; ++m) z[m] = conjf2(cmulc(z[m], H[t + T * m]));  // conj(Z conj(H))
    ct.template fft<true>(z, buf, t, tw);
    if (!active) return;
    constexpr float inv_n = 1.0f / (float)N;
    float* gxs = gx + s * g.S;
    const int tg = fresh(t);
    if (interior) {
        float* q0 = gxs + p0;
        float* q1 = gxs + p1;
#pragma unroll
        for (int m = 0; m < E; ++m) {
            const int n = tg + T * m;
            if (n < g.hop) {
                st_at(q0, (uint32_t)n, z[m].x * inv_n);
                st_at(q1, (uint32_t)n, -z[m].y * inv_n);
            }
        }
        return;
    }
#pragma unroll
    for (int m = 0; m < E; ++m) {
        const int n = tg + T * m;
        if (n < g.hop) {
            if (p0 + n < g.S) gxs[p0 + n] = z[m].x * inv_n;
            if (has1 && p1 + n < g.S) gxs[p1 + n] = -z[m].y * inv_n;
        }
    }
}

// Partial tap gradient of one slot (ppw block pairs): Re IFFT(sum conj(U) Z)[k],
// k < K, with U the spectra of the pairs' signal blocks (hop samples each).
template <int N, int E, int TPW>
__global__ void __launch_bounds__((N / E) * TPW)
k_conv_bwd_h(const float* __restrict__ x, const float* __restrict__ gy,
             float* __restrict__ gh_part, const float2* __restrict__ tw, ConvGeom g, int64_t nbx,
             int64_t npx, int ppw) {
    constexpr int T = N / E;
    __shared__ float2 lds[TPW * LdsPad::len(N)];
    const int local = threadIdx.x / T;
    const int t = threadIdx.x % T;
    float2* buf = lds + local * LdsPad::len(N);
    const int64_t total = g.count * npx;
    const int64_t ylen = g.S + g.K - 1;

    float2 acc[E];
#pragma unroll
    for (int m = 0; m < E; ++m) acc[m] = make_float2(0.f, 0.f);
    RegTwiddles<N, E> ct;
    ct.load(tw, t);

    for (int p = 0; p < ppw; ++p) {  // uniform trip count across the workgroup
        const int64_t item = ((int64_t)blockIdx.x * ppw + p) * TPW + local;
        const bool active = item < total;
        int64_t s = 0, q = 0;
        if (active) pair_of(item, total, npx, s, q);
        const int64_t p0 = 2 * q * g.hop, p1 = p0 + g.hop;
        const bool a1 = active && 2 * q + 1 < nbx;
        // interior pair: both gradient windows inside gy and both signal
        // blocks inside x (p1 + N <= ylen, i.e. p1 + hop <= S)
        const bool interior = a1 && p1 + N <= ylen;
        // opaque copies of t per iteration: LDS and global index arithmetic
        // is recomputed each time instead of hoisted out of the loop and
        // held live through it
        const int tl = fresh(t);
        float2 z[E], u[E];
        if (interior) {
            const float* g0 = gy + s * ylen + p0;
            const float* g1 = g0 + g.hop;
#pragma unroll
            for (int m = 0; m < E; ++m) {
                const uint32_t n = (uint32_t)(tl + T * m);
                z[m] = make_float2(ld_at(g0, n), ld_at(g1, n));
            }
        } else {
            load_grad_pair<N, E>(z, gy + s * ylen, ylen, p0, p1, active, a1, tl);
        }
        ct.template fft<true>(z, buf, tl, tw);
        __builtin_amdgcn_sched_barrier(0);  // the signal loads stay below the transform
        {
            const float* xs = x + s * g.S;
            const int tu = fresh(t);
            if (interior) {  // blocks of hop samples, zero past hop
                const float* x0 = xs + p0;
                const float* x1 = xs + p1;
#pragma unroll
                for (int m = 0; m < E; ++m) {
                    const int n = tu + T * m;
                    const uint32_t c = (uint32_t)(n < g.hop ? n : g.hop - 1);
                    uint32_t keep = n < g.hop ? 0xffffffffu : 0u;
                    __asm__("" : "+v"(keep));
                    u[m] = make_float2(__uint_as_float(__float_as_uint(ld_at(x0, c)) & keep),
                                       __uint_as_float(__float_as_uint(ld_at(x1, c)) & keep));
                }
            } else {
                const Win X0 = window(xs, g.S, p0, g.hop), X1 = window(xs, g.S, p1, g.hop);
#pragma unroll
                for (int m = 0; m < E; ++m) {
                    const int n = tu + T * m;
                    const float ur = ld_win(X0, n);
                    const float ui = ld_win(X1, n);
                    u[m] = make_float2(active ? ur : 0.f, a1 ? ui : 0.f);
                }
            }
        }
        ct.template fft<true>(u, buf, fresh(t), tw);
#pragma unroll
        for (int m = 0; m < E; ++m) acc[m] = cadd(acc[m], cmulc(z[m], u[m]));  // conj(U) Z
    }
#pragma unroll
    for (int m = 0; m < E; ++m) acc[m] = conjf2(acc[m]);
    ct.template fft<true>(acc, buf, t, tw);
    constexpr float inv_n = 1.0f / (float)N;
    float* part = gh_part + ((int64_t)blockIdx.x * TPW + local) * g.K;
#pragma unroll
    for (int m = 0; m < E; ++m) {
        const int n = t + T * m;
        if (n < g.K) part[n] = acc[m].x * inv_n;
    }
}

// Column sums of a rows x cols matrix, 64 rows per output row:
// out[rb][c] = sum of in[r][c] over rb*64 <= r < min(rows, rb*64 + 64), as
// four 16-row sums (independent loads in flight together) added in a fixed
// order.  Applied until one row is left, it reduces the per-slot tap
// gradient partials in the same order on every run (deterministic).
__global__ void __launch_bounds__(256)
k_sum_rows64(const float* __restrict__ in, float* __restrict__ out, int64_t rows, int cols) {
    __shared__ float part[4][64];
    const int cl = threadIdx.x & 63, sub = threadIdx.x >> 6;
    const int c = blockIdx.x * 64 + cl;
    const int cc = c < cols ? c : cols - 1;  // in-bounds column for the loads
    const int64_t r0 = (int64_t)blockIdx.y * 64 + sub * 16;
    float v[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) {  // unconditional (clamped) loads, masked sums
        const int64_t r = r0 + i;
        const float x = in[(r < rows ? r : rows - 1) * cols + cc];
        uint32_t keep = r < rows ? 0xffffffffu : 0u;
        __asm__("" : "+v"(keep));
        v[i] = __uint_as_float(__float_as_uint(x) & keep);
    }
    float s = 0.f;
#pragma unroll
    for (int i = 0; i < 16; ++i) s += v[i];
    part[sub][cl] = s;
    __syncthreads();
    if (sub == 0 && c < cols)
        out[(int64_t)blockIdx.y * cols + c] = (part[0][cl] + part[1][cl]) + (part[2][cl] + part[3][cl]);
}

// --------------------------------------------------- FP16 precision modes
// FusedConvolver.Precision (Sources/MetalFFT/DSP/FusedConvolver.swift:24-35,
// kernels Sources/MetalFFT/Resources/fft_fused_convolve_fp16.metal:160-498),
// N = 4096 only as in the reference:
//   PURE     butterflies, twiddle powers, twiddle and filter products and the
//            1/N scale in FP16 (packed f16 VALU), FP16 LDS exchange;
//   STORAGE  every LDS exchange stored as FP16, all arithmetic in FP32;
//   MIXED    twiddle powers, twiddle and filter products in FP16,
//            butterflies in FP32, FP16 LDS exchange.
// Where the reference keeps data in threadgroup memory (between passes, and
// the spectrum around the filter product) the value is FP16 here too; its
// radix-4 passes are radix-16 here (3 roundings per transform instead of 6).
// Same overlap-save geometry as k_conv_fwd; the half2 exchange buffer is
// half the FP32 one.
using namespace fp16;

template <int PM>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(PM == kConvStorage ? 4 : 1, 8)))
k_conv_fwd_h(const float* __restrict__ x, float* __restrict__ y, const float2* __restrict__ H,
             const h2* __restrict__ H16, const float2* __restrict__ tw, ConvGeom g) {
    constexpr int N = 4096, T = 256;
    __shared__ h2 lds[LdsPad::len(N)];
    const int t = threadIdx.x;
    const int64_t item = blockIdx.x;  // one block pair per workgroup; grid == count * npair
    int64_t s, q;
    pair_of(item, g.count * g.npair, g.npair, s, q);
    const int64_t b0 = 2 * q, b1 = 2 * q + 1;
    const float* xs = x + s * g.S;
    const int64_t ylen = g.S + g.K - 1;
    float* ys = y + s * ylen;
    const int64_t w0 = b0 * g.hop - (g.K - 1);
    const int64_t w1 = b1 * g.hop - (g.K - 1);
    const bool has1 = b1 < g.nblk;

    ConvReg<PM> v[16];
    const bool interior = has1 && w0 >= 0 && w1 + N <= g.S;  // as in k_conv_fwd
    if (interior) {
        const float* p0 = xs + w0;
        const float* p1 = xs + w1;
#pragma unroll
        for (int m = 0; m < 16; ++m) {
            const uint32_t n = (uint32_t)(t + T * m);
            const float2 x2 = make_float2(ld_at(p0, n), ld_at(p1, n));
            if constexpr (PM == kConvPure) v[m] = to_h2(x2);
            else v[m] = x2;
        }
    } else {
        const Win W0 = window(xs, g.S, w0, N), W1 = window(xs, g.S, w1, N);  // S >= 1
#pragma unroll
        for (int m = 0; m < 16; ++m) {
            const int n = t + T * m;
            const float re = ld_win(W0, n);
            const float l1 = ld_win(W1, n);
            const float im = has1 ? l1 : 0.f;
            if constexpr (PM == kConvPure) v[m] = to_h2(make_float2(re, im));
            else v[m] = make_float2(re, im);
        }
    }
    PreBases<N, 16> pb;
    pb.load_packed(tw, t);
    conv_passes_h<PM, 0, false>(v, lds, t, pb.w);
    // spectrum (FP16, as it sits in threadgroup memory) times H, conjugated
    // for the inverse transform
#pragma unroll
    for (int m = 0; m < 16; ++m) {
        if constexpr (PM == kConvPure) v[m] = hconj(hmul(v[m], H16[t + T * m]));
        else if constexpr (PM == kConvMixed) v[m] = to_f2(hconj(hmul(to_h2(v[m]), H16[t + T * m])));
        else v[m] = conjf2(round_h(cmul(round_h(v[m]), H[t + T * m])));
    }
    pb.opaque();
    conv_passes_h<PM, 0, true>(v, lds, t, pb.w);
    const int ts = fresh(t);
#pragma unroll
    for (int m = 0; m < 16; ++m) {
        const int n = ts + T * m;
        const int j = n - (g.K - 1);
        if (j >= 0 && j < g.hop) {
            float2 r;  // conj(.) / N
            if constexpr (PM == kConvPure) {
                const _Float16 inv_n = (_Float16)(1.0f / (float)N);
                r = to_f2((h2){v[m].x * inv_n, -v[m].y * inv_n});
            } else {
                constexpr float inv_n = 1.0f / (float)N;
                r = make_float2(v[m].x * inv_n, -v[m].y * inv_n);
            }
            if (interior) {
                st_at(ys + b0 * g.hop, (uint32_t)j, r.x);
                st_at(ys + b1 * g.hop, (uint32_t)j, r.y);
            } else {
                const int64_t o0 = b0 * g.hop + j;
                if (o0 < ylen) ys[o0] = r.x;
                const int64_t o1 = b1 * g.hop + j;
                if (has1 && o1 < ylen) ys[o1] = r.y;
            }
        }
    }
}

hipError_t conv_fwd_fp16(int precision, const float* x, float* y, const float2* H, const float2* tw,
                         ConvGeom g, hipStream_t st) {
    const int64_t blocks = g.count * g.npair;
    if (blocks <= 0) return hipSuccess;
    if (blocks > 0x7fffffffLL) return hipErrorInvalidValue;
    const h2* H16 = reinterpret_cast<const h2*>(H + 4096);  // after H (convolver_spectrum)
    switch (precision) {
        case kConvPure:
            hipLaunchKernelGGL(k_conv_fwd_h<kConvPure>, dim3((unsigned)blocks), dim3(256), 0, st, x, y,
                               H, H16, tw, g);
            break;
        case kConvStorage:
            hipLaunchKernelGGL(k_conv_fwd_h<kConvStorage>, dim3((unsigned)blocks), dim3(256), 0, st, x,
                               y, H, H16, tw, g);
            break;
        case kConvMixed:
            hipLaunchKernelGGL(k_conv_fwd_h<kConvMixed>, dim3((unsigned)blocks), dim3(256), 0, st, x, y,
                               H, H16, tw, g);
            break;
        default:
            return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

__global__ void k_zero(float* __restrict__ p, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.f;
}

// H rounded to FP16 once (the FP16 modes' filter products read it: one
// 4-byte load per bin instead of an 8-byte load and two conversions)
__global__ void k_spectrum_h(const float2* __restrict__ H, h2* __restrict__ H16, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) H16[i] = to_h2(H[i]);
}

// taps (real, K) -> zero-padded complex N
__global__ void k_pack_taps(const float* __restrict__ taps, int K, float2* __restrict__ out, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = make_float2(i < K ? taps[i] : 0.f, 0.f);
}

template <int N>
struct ConvCfg {
    static constexpr int E = N >= 256 ? 16 : 8;
    static constexpr int TPW = tpw_for<N, E>();
    static constexpr int threads = (N / E) * TPW;
};

template <int N>
hipError_t conv_fwd_n(const float* x, float* y, const float2* H, const float2* tw, ConvGeom g,
                      hipStream_t st) {
    using C = ConvCfg<N>;
    const int64_t total = g.count * g.npair;
    const int64_t blocks = (total + C::TPW - 1) / C::TPW;
    if (blocks <= 0) return hipSuccess;
    if (blocks > 0x7fffffffLL) return hipErrorInvalidValue;
    hipLaunchKernelGGL((k_conv_fwd<N, C::E, C::TPW>), dim3((unsigned)blocks), dim3(C::threads), 0, st,
                       x, y, H, tw, g);
    return hipGetLastError();
}

template <int N>
hipError_t conv_bwd_n(const float* x, const float* gy, float* gx, float* part, int64_t slots,
                      const float2* H, const float2* tw, ConvGeom g, int64_t nbx, int64_t npx,
                      int ppw, hipStream_t st) {
    using C = ConvCfg<N>;
    const int64_t xblocks = (g.count * npx + C::TPW - 1) / C::TPW;
    const int64_t hblocks = slots / C::TPW;
    if (xblocks <= 0 || hblocks <= 0) return hipSuccess;
    if (xblocks > 0x7fffffffLL) return hipErrorInvalidValue;
    hipLaunchKernelGGL((k_conv_bwd_x<N, C::E, C::TPW>), dim3((unsigned)xblocks), dim3(C::threads), 0,
                       st, gy, gx, H, tw, g, nbx, npx);
    hipLaunchKernelGGL((k_conv_bwd_h<N, C::E, C::TPW>), dim3((unsigned)hblocks), dim3(C::threads), 0,
                       st, x, gy, part, tw, g, nbx, npx, ppw);
    return hipGetLastError();
}

template <int N>
int conv_tpw() { return ConvCfg<N>::TPW; }

#define MIFFT_CONV_SWITCH(n, F, ...)                     \
    switch (n) {                                         \
        case 64: return F<64>(__VA_ARGS__);              \
        case 128: return F<128>(__VA_ARGS__);            \
        case 256: return F<256>(__VA_ARGS__);            \
        case 512: return F<512>(__VA_ARGS__);            \
        case 1024: return F<1024>(__VA_ARGS__);          \
        case 2048: return F<2048>(__VA_ARGS__);          \
        case 4096: return F<4096>(__VA_ARGS__);          \
        case 8192: return F<8192>(__VA_ARGS__);          \
        case 16384: return F<16384>(__VA_ARGS__);        \
    }

hipError_t conv_fwd(int n, const float* x, float* y, const float2* H, const float2* tw, ConvGeom g,
                    hipStream_t st) {
    MIFFT_CONV_SWITCH(n, conv_fwd_n, x, y, H, tw, g, st)
    return hipErrorInvalidValue;
}

hipError_t conv_bwd(int n, const float* x, const float* gy, float* gx, float* part, int64_t slots,
                    const float2* H, const float2* tw, ConvGeom g, int64_t nbx, int64_t npx, int ppw,
                    hipStream_t st) {
    MIFFT_CONV_SWITCH(n, conv_bwd_n, x, gy, gx, part, slots, H, tw, g, nbx, npx, ppw, st)
    return hipErrorInvalidValue;
}

int conv_tpw_for(int n) {
    MIFFT_CONV_SWITCH(n, conv_tpw)
    return 1;
}


// --------------------------------------------------------- correlation
// r[m] = Re IFFT(A * conj(B))[m] for real a, b truncated or zero-padded to N
// (Sources/MetalFFT/DSP/Correlate.swift:33-61: two forward FFTs, the product
// A*conj(B) written as (:40-45), the conjugate-trick inverse, real part).
// Here a and b share one FFT, z = a + i b, and are split with
// A[k] = (Z[k] + conj Z[-k]) / 2, B[k] = -i (Z[k] - conj Z[-k]) / 2, the
// mirrored bin read back through LDS.  P = A conj(B) is Hermitian (r is
// real), so a transform slot takes two pairs and inverts Q = P1 + i P2 once:
// r1 is the real part of IFFT(Q) and r2 its imaginary part, three
// transforms per two correlations instead of four.
template <int N, int E, int TPW>
__global__ void __launch_bounds__((N / E) * TPW)
k_correlate(const float* __restrict__ a, int64_t na, const float* __restrict__ b, int64_t nb,
            float* __restrict__ r, const float2* __restrict__ tw, int64_t count) {
    constexpr int T = N / E;
    __shared__ float2 lds[TPW * LdsPad::len(N)];
    const int local = threadIdx.x / T;
    const int t = threadIdx.x % T;
    float2* buf = lds + local * LdsPad::len(N);
    const int64_t item = (int64_t)blockIdx.x * TPW + local;  // pairs 2 item, 2 item + 1
    RegTwiddles<N, E> ct;
    ct.load(tw, t);
    float2 q[E];
#pragma unroll 1
    for (int h = 0; h < 2; ++h) {
        const int64_t pi = 2 * item + h;
        const bool active = pi < count;
        const float* as = a + (active ? pi : 0) * na;
        const float* bs = b + (active ? pi : 0) * nb;
        const Win A = window(as, na > 0 ? na : 1, 0, N), Bw = window(bs, nb > 0 ? nb : 1, 0, N);
        const int tl = fresh(t);  // index arithmetic redone per pair, not held
        float2 v[E];
#pragma unroll
        for (int m = 0; m < E; ++m) {  // branch-free loads (ld_win); an empty input reads as 0
            const int n = tl + T * m;
            const float re = na > 0 ? ld_win(A, n) : 0.f;
            const float im = nb > 0 ? ld_win(Bw, n) : 0.f;
            v[m] = make_float2(active ? re : 0.f, active ? im : 0.f);
        }
        ct.template fft<true>(v, buf, tl, tw);  // SYNC0: the first pair's split read the LDS
        lds_barrier();
#pragma unroll
        for (int m = 0; m < E; ++m) buf[LdsPad::idx(tl + T * m)] = v[m];
        lds_barrier();
#pragma unroll
        for (int m = 0; m < E; ++m) {
            const int k = tl + T * m;
            const float2 zm = conjf2(buf[LdsPad::idx((N - k) & (N - 1))]);
            const float2 A = make_float2(0.5f * (v[m].x + zm.x), 0.5f * (v[m].y + zm.y));
            const float2 d = make_float2(0.5f * (v[m].x - zm.x), 0.5f * (v[m].y - zm.y));
            const float2 B = mul_mi(d);
            // A * conj(B) as (A.x B.x + A.y B.y, A.y B.x - A.x B.y)
            const float2 P = make_float2(A.x * B.x + A.y * B.y, A.y * B.x - A.x * B.y);
            q[m] = h == 0 ? P : make_float2(q[m].x - P.y, q[m].y + P.x);  // Q = P1 + i P2
        }
    }
    // IFFT(Q) = conj(FFT(conj Q)) / N = r1 + i r2
#pragma unroll
    for (int m = 0; m < E; ++m) q[m] = conjf2(q[m]);
    ct.template fft<true>(q, buf, t, tw);
    const int64_t p0 = 2 * item;
    if (p0 >= count) return;
    const bool has1 = p0 + 1 < count;
    constexpr float inv_n = 1.0f / (float)N;
    float* r0 = r + p0 * N;
    const int ts = fresh(t);
#pragma unroll
    for (int m = 0; m < E; ++m) {
        r0[ts + T * m] = q[m].x * inv_n;
        if (has1) r0[N + ts + T * m] = -q[m].y * inv_n;
    }
}

template <int N>
hipError_t corr_n(const float* a, int64_t na, const float* b, int64_t nb, float* r, const float2* tw,
                  int64_t count, hipStream_t st) {
    using C = ConvCfg<N>;
    const int64_t slots = (count + 1) / 2;  // two pairs per transform slot
    const int64_t blocks = (slots + C::TPW - 1) / C::TPW;
    if (blocks <= 0) return hipSuccess;
    if (blocks > 0x7fffffffLL) return hipErrorInvalidValue;
    hipLaunchKernelGGL((k_correlate<N, C::E, C::TPW>), dim3((unsigned)blocks), dim3(C::threads), 0, st,
                       a, na, b, nb, r, tw, count);
    return hipGetLastError();
}

hipError_t correlate(int n, const float* a, int64_t na, const float* b, int64_t nb, float* r,
                     const float2* tw, int64_t count, hipStream_t st) {
    MIFFT_CONV_SWITCH(n, corr_n, a, na, b, nb, r, tw, count, st)
    return hipErrorInvalidValue;
}
}  // namespace

// Geometry of a convolver call; pairs cover ceil((S + K - 1) / hop) output
// blocks (forward) or ceil(S / hop) gradient blocks (adjoint).
static ConvGeom geometry(const mifft_convolver* cv, int64_t S, int64_t count) {
    ConvGeom g;
    g.S = S;
    g.count = count;
    g.K = (int)cv->K;
    g.hop = (int)(cv->n - cv->K + 1);
    g.nblk = (S + cv->K - 1 + g.hop - 1) / g.hop;
    g.npair = (g.nblk + 1) / 2;
    return g;
}

mifft_status convolver_spectrum(mifft_convolver* cv, const float* taps_dev, hipStream_t st) {
    mifft_status s;
    const int n = (int)cv->n;
    hipLaunchKernelGGL(k_pack_taps, dim3((n + 255) / 256), dim3(256), 0, st, taps_dev, (int)cv->K,
                       static_cast<float2*>(cv->packed.ptr), n);
    if ((s = check_hip(hipGetLastError(), "pack taps")) != MIFFT_OK) return s;
    if ((s = check_hip(cv->plan->variant->launch(static_cast<const float2*>(cv->packed.ptr),
                                                 static_cast<float2*>(cv->spectrum.ptr),
                                                 cv->plan->twiddles, 1, false, st),
                       "taps FFT")) != MIFFT_OK)
        return s;
    const float2* H = static_cast<const float2*>(cv->spectrum.ptr);
    hipLaunchKernelGGL(k_spectrum_h, dim3((n + 255) / 256), dim3(256), 0, st, H,
                       reinterpret_cast<h2*>(const_cast<float2*>(H) + n), n);
    return check_hip(hipGetLastError(), "FP16 spectrum");
}

mifft_status convolver_forward(const mifft_convolver* cv, const float* x, int64_t S, int64_t count,
                               float* y, hipStream_t st) {
    if (S == 0) return zero_floats(y, count * (cv->K - 1), st);  // y = x * h is all zeros
    ConvGeom g = geometry(cv, S, count);
    if (cv->precision != MIFFT_CONV_FP32) {
        if (cv->n != 4096) return fail(MIFFT_ERR_KERNEL_NOT_FOUND, "FP16 fused convolve is n = 4096 only");
        return check_hip(conv_fwd_fp16(cv->precision, x, y, static_cast<const float2*>(cv->spectrum.ptr),
                                       cv->plan->twiddles, g, st),
                         "fused convolve (FP16)");
    }
    return check_hip(conv_fwd((int)cv->n, x, y, static_cast<const float2*>(cv->spectrum.ptr),
                              cv->plan->twiddles, g, st),
                     "fused convolve");
}

AdjointGeom adjoint_geometry(const mifft_convolver* cv, int64_t S, int64_t count, int64_t target_wg) {
    ConvGeom g = geometry(cv, S, count);
    AdjointGeom a;
    a.nbx = (S + g.hop - 1) / g.hop;
    a.npx = (a.nbx + 1) / 2;
    const int tpw = conv_tpw_for((int)cv->n);
    const int64_t items = count * a.npx;
    a.ppw = (items + target_wg * tpw - 1) / (target_wg * tpw);
    if (a.ppw < 1) a.ppw = 1;
    const int64_t blocks = (items + a.ppw * tpw - 1) / (a.ppw * tpw);
    a.slots = blocks * tpw;
    a.level1 = (a.slots + 63) / 64;
    return a;
}

int64_t adjoint_scratch_floats(const AdjointGeom& a, int64_t K) { return (a.slots + 2 * a.level1) * K; }

mifft_status convolver_backward_with(const mifft_convolver* cv, const float* x, int64_t S, int64_t count,
                                     const float* gy, float* gx, float* gh, float* part, const AdjointGeom& a,
                                     hipStream_t st) {
    ConvGeom g = geometry(cv, S, count);
    mifft_status s;
    if ((s = check_hip(conv_bwd((int)cv->n, x, gy, gx, part, a.slots,
                                static_cast<const float2*>(cv->spectrum.ptr), cv->plan->twiddles, g,
                                a.nbx, a.npx, (int)a.ppw, st),
                       "fused convolve adjoint")) != MIFFT_OK)
        return s;
    // grad_h = column sums of the slot partials, ping-pong between two
    // scratch areas after `part`
    const int K = (int)cv->K;
    return check_hip(sum_rows(part, a.slots, K, part + a.slots * (int64_t)K,
                              part + (a.slots + a.level1) * (int64_t)K, gh, st),
                     "tap gradient reduction");
}

mifft_status convolver_backward(mifft_convolver* cv, const float* x, int64_t S, int64_t count,
                                 const float* gy, float* gx, float* gh, hipStream_t st) {
    // ~2048 workgroups: enough to fill 256 CUs, few partial rows to reduce
    const AdjointGeom a = adjoint_geometry(cv, S, count, kAdjointWorkgroups);
    // partials, then the 64-row sums of them (tree reduction scratch)
    mifft_status s = cv->partial.reserve(sizeof(float) * (size_t)adjoint_scratch_floats(a, cv->K));
    if (s != MIFFT_OK) return s;
    return convolver_backward_with(cv, x, S, count, gy, gx, gh, static_cast<float*>(cv->partial.ptr), a, st);
}

hipError_t sum_rows(const float* in, int64_t rows, int cols, float* s0, float* s1, float* out,
                    hipStream_t st) {
    const float* src = in;
    float* scratch[2] = {s0, s1};
    for (int lvl = 0; rows > 1; ++lvl) {
        const int64_t next = (rows + 63) / 64;
        float* dst = next == 1 ? out : scratch[lvl & 1];
        hipLaunchKernelGGL(k_sum_rows64, dim3((unsigned)((cols + 63) / 64), (unsigned)next), dim3(256), 0, st,
                           src, dst, rows, cols);
        src = dst;
        rows = next;
    }
    if (src == in)  // one row: copied as it is
        hipLaunchKernelGGL(k_sum_rows64, dim3((unsigned)((cols + 63) / 64), 1), dim3(256), 0, st, in, out,
                           (int64_t)1, cols);
    return hipGetLastError();
}

// Fill kernel rather than hipMemsetAsync, so callers may capture it in a graph.
mifft_status zero_floats(float* p, int64_t n, hipStream_t st) {
    if (n <= 0) return MIFFT_OK;
    hipLaunchKernelGGL(k_zero, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, p, n);
    return check_hip(hipGetLastError(), "zero fill");
}

mifft_status correlate_device(const mifft_plan* plan, const float* a, int64_t na, const float* b,
                              int64_t nb, float* r, int64_t count, hipStream_t st) {
    return check_hip(correlate((int)plan->n, a, na, b, nb, r, plan->twiddles, count, st), "correlate");
}

}  // namespace mifft
