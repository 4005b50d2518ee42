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

#include "convolve.hpp"
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

__device__ __forceinline__ float ld_or0(const float* p, int64_t i, int64_t lo, int64_t hi) {
    return (i >= lo && i < hi) ? p[i] : 0.f;
}

// --------------------------------------------------------------- forward
template <int N, int E, int TPW>
__global__ void __launch_bounds__((N / E) * TPW)
k_conv_fwd(const float* __restrict__ x, float* __restrict__ y, const float2* __restrict__ H,
           const float2* __restrict__ tw, ConvGeom g) {
    constexpr int T = N / E;
    __shared__ float2 lds[TPW * LdsPad::len(N)];
    const int local = threadIdx.x / T;
    const int t = threadIdx.x % T;
    float2* buf = lds + local * LdsPad::len(N);
    const int64_t item = (int64_t)blockIdx.x * TPW + local;
    const int64_t total = g.count * g.npair;
    const bool active = item < total;
    const int64_t s = active ? item / g.npair : 0;
    const int64_t q = active ? item % g.npair : 0;
    const int64_t b0 = 2 * q, b1 = 2 * q + 1;
    const float* xs = x + s * g.S;
    const int64_t ylen = g.S + g.K - 1;
    float* ys = y + s * ylen;
    const int64_t w0 = b0 * g.hop - (g.K - 1);
    const int64_t w1 = b1 * g.hop - (g.K - 1);
    const bool has1 = b1 < g.nblk;

    float2 v[E];
#pragma unroll
    for (int m = 0; m < E; ++m) {
        const int n = t + T * m;
        float re = 0.f, im = 0.f;
        if (active) {
            re = ld_or0(xs, w0 + n, 0, g.S);
            if (has1) im = ld_or0(xs, w1 + n, 0, g.S);
        }
        v[m] = make_float2(re, im);
    }
    stockham_regs<N, E, LdsPad, 0, false>(v, buf, t, tw);
    // multiply by the filter spectrum, then inverse via conj-FFT-conj
#pragma unroll
    for (int m = 0; m < E; ++m) v[m] = conjf2(cmul(v[m], H[t + T * m]));
    stockham_regs<N, E, LdsPad, 0, true>(v, buf, t, tw);
    if (!active) return;
    constexpr float inv_n = 1.0f / (float)N;
#pragma unroll
    for (int m = 0; m < E; ++m) {
        const int n = t + T * m;
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
template <int N, int E, int TPW>
__global__ void __launch_bounds__((N / E) * TPW)
k_conv_bwd(const float* __restrict__ x, const float* __restrict__ gy,
           float* __restrict__ gx, float* __restrict__ gh_part, const float2* __restrict__ H,
           const float2* __restrict__ tw, ConvGeom g, int64_t nbx, int64_t npx, int ppw) {
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

    for (int p = 0; p < ppw; ++p) {  // uniform trip count across the workgroup
        const int64_t item = ((int64_t)blockIdx.x * ppw + p) * TPW + local;
        const bool active = item < total;
        const int64_t s = active ? item / npx : 0;
        const int64_t q = active ? item % npx : 0;
        const int64_t b0 = 2 * q, b1 = 2 * q + 1;
        const bool has1 = b1 < nbx;
        const float* xs = x + s * g.S;
        const float* gs = gy + s * ylen;
        const int64_t p0 = b0 * g.hop, p1 = b1 * g.hop;

        float2 z[E], u[E];
#pragma unroll
        for (int m = 0; m < E; ++m) {
            const int n = t + T * m;
            float zr = 0.f, zi = 0.f, ur = 0.f, ui = 0.f;
            if (active) {
                zr = ld_or0(gs, p0 + n, 0, ylen);
                if (n < g.hop) ur = ld_or0(xs, p0 + n, 0, g.S);
                if (has1) {
                    zi = ld_or0(gs, p1 + n, 0, ylen);
                    if (n < g.hop) ui = ld_or0(xs, p1 + n, 0, g.S);
                }
            }
            z[m] = make_float2(zr, zi);
            u[m] = make_float2(ur, ui);
        }
        stockham_regs<N, E, LdsPad, 0, true>(z, buf, t, tw);
        stockham_regs<N, E, LdsPad, 0, true>(u, buf, t, tw);
#pragma unroll
        for (int m = 0; m < E; ++m) {
            acc[m] = cadd(acc[m], cmulc(z[m], u[m]));  // conj(U) Z
            z[m] = conjf2(cmulc(z[m], H[t + T * m]));  // conj(Z conj(H)) for the inverse
        }
        stockham_regs<N, E, LdsPad, 0, true>(z, buf, t, tw);
        if (active) {
            constexpr float inv_n = 1.0f / (float)N;
            float* gxs = gx + s * g.S;
#pragma unroll
            for (int m = 0; m < E; ++m) {
                const int n = t + T * m;
                if (n < g.hop) {
                    if (p0 + n < g.S) gxs[p0 + n] = z[m].x * inv_n;
                    if (has1 && p1 + n < g.S) gxs[p1 + n] = -z[m].y * inv_n;
                }
            }
        }
    }
    // partial tap gradient of this slot: Re IFFT(acc)[k], k < K
#pragma unroll
    for (int m = 0; m < E; ++m) acc[m] = conjf2(acc[m]);
    stockham_regs<N, E, LdsPad, 0, true>(acc, buf, t, tw);
    constexpr float inv_n = 1.0f / (float)N;
    float* part = gh_part + ((int64_t)blockIdx.x * TPW + local) * g.K;
#pragma unroll
    for (int m = 0; m < E; ++m) {
        const int n = t + T * m;
        if (n < g.K) part[n] = acc[m].x * inv_n;
    }
}

// grad_h[k] = sum over slots w of part[w][k], in slot order.
__global__ void k_sum_columns(const float* __restrict__ part, float* __restrict__ out, int64_t rows,
                              int cols) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= cols) return;
    float s = 0.f;
    for (int64_t r = 0; r < rows; ++r) s += part[r * cols + k];
    out[k] = s;
}

__global__ void k_zero(float* __restrict__ p, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.f;
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
    const int64_t blocks = slots / C::TPW;
    if (blocks <= 0) return hipSuccess;
    hipLaunchKernelGGL((k_conv_bwd<N, C::E, C::TPW>), dim3((unsigned)blocks), dim3(C::threads), 0, st,
                       x, gy, gx, part, H, tw, g, nbx, npx, ppw);
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
// mirrored bin read back through LDS.
template <int N, int E, int TPW>
__global__ void __launch_bounds__((N / E) * TPW)
k_correlate(const float* __restrict__ a, int64_t na, const float* __restrict__ b, int64_t nb,
            float* __restrict__ r, const float2* __restrict__ tw, int64_t count) {
    constexpr int T = N / E;
    __shared__ float2 lds[TPW * LdsPad::len(N)];
    const int local = threadIdx.x / T;
    const int t = threadIdx.x % T;
    float2* buf = lds + local * LdsPad::len(N);
    const int64_t item = (int64_t)blockIdx.x * TPW + local;
    const bool active = item < count;
    const float* as = a + (active ? item : 0) * na;
    const float* bs = b + (active ? item : 0) * nb;
    float2 v[E];
#pragma unroll
    for (int m = 0; m < E; ++m) {
        const int n = t + T * m;
        float re = 0.f, im = 0.f;
        if (active) {
            if (n < na) re = as[n];
            if (n < nb) im = bs[n];
        }
        v[m] = make_float2(re, im);
    }
    stockham_regs<N, E, LdsPad, 0, false>(v, buf, t, tw);
    lds_barrier();
#pragma unroll
    for (int m = 0; m < E; ++m) buf[LdsPad::idx(t + T * m)] = v[m];
    lds_barrier();
#pragma unroll
    for (int m = 0; m < E; ++m) {
        const int k = t + T * m;
        const float2 zm = conjf2(buf[LdsPad::idx((N - k) & (N - 1))]);
        const float2 A = make_float2(0.5f * (v[m].x + zm.x), 0.5f * (v[m].y + zm.y));
        const float2 d = make_float2(0.5f * (v[m].x - zm.x), 0.5f * (v[m].y - zm.y));
        const float2 B = mul_mi(d);
        // A * conj(B) as (A.x B.x + A.y B.y, A.y B.x - A.x B.y); conj for the inverse
        const float2 P = make_float2(A.x * B.x + A.y * B.y, A.y * B.x - A.x * B.y);
        v[m] = conjf2(P);
    }
    stockham_regs<N, E, LdsPad, 0, true>(v, buf, t, tw);
    if (!active) return;
    constexpr float inv_n = 1.0f / (float)N;
    float* rs = r + item * N;
#pragma unroll
    for (int m = 0; m < E; ++m) rs[t + T * m] = v[m].x * inv_n;
}

template <int N>
hipError_t corr_n(const float* a, int64_t na, const float* b, int64_t nb, float* r, const float2* tw,
                  int64_t count, hipStream_t st) {
    using C = ConvCfg<N>;
    const int64_t blocks = (count + C::TPW - 1) / C::TPW;
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
    return check_hip(cv->plan->variant->launch(static_cast<const float2*>(cv->packed.ptr),
                                               static_cast<float2*>(cv->spectrum.ptr),
                                               cv->plan->twiddles, 1, false, st),
                     "taps FFT");
}

mifft_status convolver_forward(const mifft_convolver* cv, const float* x, int64_t S, int64_t count,
                               float* y, hipStream_t st) {
    ConvGeom g = geometry(cv, S, count);
    return check_hip(conv_fwd((int)cv->n, x, y, static_cast<const float2*>(cv->spectrum.ptr),
                              cv->plan->twiddles, g, st),
                     "fused convolve");
}

mifft_status convolver_backward(mifft_convolver* cv, const float* x, int64_t S, int64_t count,
                                 const float* gy, float* gx, float* gh, hipStream_t st) {
    ConvGeom g = geometry(cv, S, count);
    const int64_t nbx = (S + g.hop - 1) / g.hop;
    const int64_t npx = (nbx + 1) / 2;
    const int tpw = conv_tpw_for((int)cv->n);
    const int64_t items = count * npx;
    // ~2048 workgroups: enough to fill 256 CUs, few partial rows to reduce
    const int64_t target_wg = 2048;
    int64_t ppw = (items + target_wg * tpw - 1) / (target_wg * tpw);
    if (ppw < 1) ppw = 1;
    const int64_t blocks = (items + ppw * tpw - 1) / (ppw * tpw);
    const int64_t slots = blocks * tpw;
    mifft_status s = cv->partial.reserve(sizeof(float) * (size_t)slots * (size_t)cv->K);
    if (s != MIFFT_OK) return s;
    float* part = static_cast<float*>(cv->partial.ptr);
    if ((s = check_hip(conv_bwd((int)cv->n, x, gy, gx, part, slots,
                                static_cast<const float2*>(cv->spectrum.ptr), cv->plan->twiddles, g,
                                nbx, npx, (int)ppw, st),
                       "fused convolve adjoint")) != MIFFT_OK)
        return s;
    hipLaunchKernelGGL(k_sum_columns, dim3((unsigned)((cv->K + 255) / 256)), dim3(256), 0, st, part, gh,
                       slots, (int)cv->K);
    return check_hip(hipGetLastError(), "tap gradient reduction");
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
