// Multi-channel cross-spectral analysis (Welch averaging over frames).
//
// Reference: PSD.crossSpectral(channels:fftSize:hopSize:sampleRate:window:)
// (Sources/MetalFFT/DSP/CrossSpectral.swift:55-173) with the per-bin kernel
// fft_cross_spectral (Sources/MetalFFT/Resources/fft_cross_spectral.metal:
// 29-64).  The reference windows and FFTs one frame of all channels at a
// time (MultiChannelFFT.forward), runs the kernel on that frame, and sums the
// results on the CPU; frames start at 0, hop, ... while pos + N <= samples.
//
// Here the whole signal set is three kinds of launch:
//   1. every windowed frame of every channel transformed at once, two real
//      frames per complex FFT, the two spectra split with the mirrored bin
//      (X_a[k] = (Z[k] + conj Z[-k]) / 2, X_b[k] = (Z[k] - conj Z[-k]) / 2i),
//      written to a spectra scratch [channel][frame][N];
//   2. partial sums over chunks of 64 frames, one thread per (series, bin,
//      chunk), a series being a channel's power |X_c|^2 or a pair's cross
//      spectrum X_i conj(X_j) and per-frame coherence
//      |S_ij|^2 / max(P_i P_j, 1e-30) (the kernel's definition, :56-61);
//   3. the chunk partials added in a fixed order (deterministic) and divided
//      by the frame count; optionally the magnitude-squared coherence of the
//      averaged spectra, |<S_ij>|^2 / max(<P_i><P_j>, 1e-30).
// Pairs are the upper triangle (0,1), (0,2), ..., (C-2, C-1) as in
// CrossSpectralResult.pairIndex (CrossSpectral.swift:31-38).
#include <hip/hip_runtime.h>

#include <string>

#include "runtime.hpp"
#include "stockham_core.hpp"

struct mifft_xspec {
    int64_t n = 0;
    int64_t channels = 0;
    mifft_plan* plan = nullptr;           // length-n twiddles and pass rows
    mifft::DeviceBuffer spectra, partial;  // scratch, grown on demand
};

namespace mifft {
namespace {

constexpr int kChunk = 64;  // frames per partial sum

template <int N>
struct XCfg {
    static constexpr int E = N >= 256 ? 16 : 8;
    static constexpr int TPW = tpw_for<N, E>();
    static constexpr int threads = (N / E) * TPW;
};

// Frame g of the C x F list: channel g / F, frame g % F, samples
// sig[c][f*hop + n] times window[n] (window null: rectangular).
template <int N>
__global__ void __launch_bounds__(XCfg<N>::threads)
k_xspec_fft(const float* __restrict__ sig, int64_t samples, int64_t hop, int64_t frames,
            int64_t total, const float* __restrict__ window, float2* __restrict__ spec,
            const float2* __restrict__ tw) {
    using C = XCfg<N>;
    constexpr int E = C::E, T = N / E;
    __shared__ float2 lds[C::TPW * LdsPad::len(N)];
    const int local = threadIdx.x / T;
    const int t = threadIdx.x % T;
    float2* buf = lds + local * LdsPad::len(N);
    const int64_t q = (int64_t)blockIdx.x * C::TPW + local;  // frame pair
    const int64_t ga = 2 * q, gb = 2 * q + 1;
    const bool act_a = ga < total, act_b = gb < total;
    const int64_t ca = act_a ? ga : 0, cb = act_b ? gb : 0;  // in-range frames for the loads
    const float* pa = sig + (ca / frames) * samples + (ca % frames) * hop;
    const float* pb = sig + (cb / frames) * samples + (cb % frames) * hop;
    float2 v[E];
#pragma unroll
    for (int m = 0; m < E; ++m) {
        const int n = t + T * m;
        const float w = window ? window[n] : 1.f;
        const float a = pa[n] * w, b = pb[n] * w;  // whole frames: always in bounds
        v[m] = make_float2(act_a ? a : 0.f, act_b ? b : 0.f);
    }
    RegTwiddles<N, E> rt;
    rt.load(tw, t);
    rt.template fft<false>(v, buf, t, tw);
    // mirrored bins through LDS
    lds_barrier();
#pragma unroll
    for (int m = 0; m < E; ++m) buf[LdsPad::idx(t + T * m)] = v[m];
    lds_barrier();
    float2* sa = spec + ca * N;
    float2* sb = spec + cb * N;
#pragma unroll
    for (int m = 0; m < E; ++m) {
        const int k = t + T * m;
        const float2 zm = conjf2(buf[LdsPad::idx((N - k) & (N - 1))]);
        const float2 xa = make_float2(0.5f * (v[m].x + zm.x), 0.5f * (v[m].y + zm.y));
        const float2 d = make_float2(0.5f * (v[m].x - zm.x), 0.5f * (v[m].y - zm.y));
        if (act_a) sa[k] = xa;
        if (act_b) sb[k] = mul_mi(d);  // d / i
    }
}

// Pair p of the upper triangle -> (i, j).
__device__ __forceinline__ void pair_of(int p, int c, int& i, int& j) {
    i = 0;
    int row = c - 1;  // pairs starting at channel i
    while (p >= row) {
        p -= row;
        ++i;
        --row;
    }
    j = i + 1 + p;
}

// series s < C: power of channel s; s >= C: pair s - C.  Partial record per
// (chunk h, series s, bin k): {sum re S, sum im S, sum coherence} for pairs,
// {sum power, 0, 0} for channels, at partial[((h * nser + s) * N + k) * 3].
__global__ void __launch_bounds__(256)
k_xspec_partial(const float2* __restrict__ spec, float* __restrict__ partial, int64_t n, int c,
                int64_t frames, int nser) {
    const int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int s = blockIdx.y;
    const int64_t h = blockIdx.z;
    if (k >= n) return;
    const int64_t f0 = h * kChunk;
    const int64_t f1 = f0 + kChunk < frames ? f0 + kChunk : frames;
    float a0 = 0.f, a1 = 0.f, a2 = 0.f;
    if (s < c) {
        const float2* x = spec + (int64_t)s * frames * n + k;
        for (int64_t f = f0; f < f1; ++f) {
            const float2 u = x[f * n];
            a0 += u.x * u.x + u.y * u.y;
        }
    } else {
        int i, j;
        pair_of(s - c, c, i, j);
        const float2* xi = spec + (int64_t)i * frames * n + k;
        const float2* xj = spec + (int64_t)j * frames * n + k;
        for (int64_t f = f0; f < f1; ++f) {
            const float2 u = xi[f * n], w = xj[f * n];
            const float2 sx = make_float2(u.x * w.x + u.y * w.y, u.y * w.x - u.x * w.y);  // u conj(w)
            const float pi = u.x * u.x + u.y * u.y, pj = w.x * w.x + w.y * w.y;
            a0 += sx.x;
            a1 += sx.y;
            a2 += (sx.x * sx.x + sx.y * sx.y) / fmaxf(pi * pj, 1e-30f);
        }
    }
    float* o = partial + ((h * nser + s) * n + k) * 3;
    o[0] = a0;
    o[1] = a1;
    o[2] = a2;
}

// Outputs: power[c][k], cross[p][k] (complex), coherence[p][k], all / frames.
__global__ void __launch_bounds__(256)
k_xspec_reduce(const float* __restrict__ partial, int64_t n, int c, int64_t chunks, int nser,
               float inv_frames, float* __restrict__ power, float2* __restrict__ cross,
               float* __restrict__ coherence) {
    const int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int s = blockIdx.y;
    if (k >= n) return;
    float a0 = 0.f, a1 = 0.f, a2 = 0.f;
    for (int64_t h = 0; h < chunks; ++h) {
        const float* o = partial + ((h * nser + s) * n + k) * 3;
        a0 += o[0];
        a1 += o[1];
        a2 += o[2];
    }
    if (s < c) {
        power[(int64_t)s * n + k] = a0 * inv_frames;
    } else {
        const int p = s - c;
        cross[(int64_t)p * n + k] = make_float2(a0 * inv_frames, a1 * inv_frames);
        coherence[(int64_t)p * n + k] = a2 * inv_frames;
    }
}

__global__ void k_xspec_msc(const float* __restrict__ power, const float2* __restrict__ cross,
                            float* __restrict__ msc, int64_t n, int c) {
    const int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int p = blockIdx.y;
    if (k >= n) return;
    int i, j;
    pair_of(p, c, i, j);
    const float2 s = cross[(int64_t)p * n + k];
    msc[(int64_t)p * n + k] =
        (s.x * s.x + s.y * s.y) / fmaxf(power[(int64_t)i * n + k] * power[(int64_t)j * n + k], 1e-30f);
}

__global__ void k_fill_zero(float* __restrict__ p, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.f;
}

template <int N>
hipError_t xspec_fft_n(const float* sig, int64_t samples, int64_t hop, int64_t frames, int64_t total,
                       const float* window, float2* spec, const float2* tw, hipStream_t st) {
    using C = XCfg<N>;
    const int64_t pairs = (total + 1) / 2;
    const int64_t blocks = (pairs + C::TPW - 1) / C::TPW;
    if (blocks > 0x7fffffffLL) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_xspec_fft<N>, dim3((unsigned)blocks), dim3(C::threads), 0, st, sig, samples,
                       hop, frames, total, window, spec, tw);
    return hipGetLastError();
}

hipError_t xspec_fft(int n, const float* sig, int64_t samples, int64_t hop, int64_t frames,
                     int64_t total, const float* window, float2* spec, const float2* tw, hipStream_t st) {
    switch (n) {
        case 64: return xspec_fft_n<64>(sig, samples, hop, frames, total, window, spec, tw, st);
        case 128: return xspec_fft_n<128>(sig, samples, hop, frames, total, window, spec, tw, st);
        case 256: return xspec_fft_n<256>(sig, samples, hop, frames, total, window, spec, tw, st);
        case 512: return xspec_fft_n<512>(sig, samples, hop, frames, total, window, spec, tw, st);
        case 1024: return xspec_fft_n<1024>(sig, samples, hop, frames, total, window, spec, tw, st);
        case 2048: return xspec_fft_n<2048>(sig, samples, hop, frames, total, window, spec, tw, st);
        case 4096: return xspec_fft_n<4096>(sig, samples, hop, frames, total, window, spec, tw, st);
        case 8192: return xspec_fft_n<8192>(sig, samples, hop, frames, total, window, spec, tw, st);
        case 16384: return xspec_fft_n<16384>(sig, samples, hop, frames, total, window, spec, tw, st);
    }
    return hipErrorInvalidValue;
}

mifft_status zero(float* p, int64_t n, hipStream_t st) {
    if (!p || n <= 0) return MIFFT_OK;
    hipLaunchKernelGGL(k_fill_zero, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, p, n);
    return check_hip(hipGetLastError(), "zero fill");
}

}  // namespace
}  // namespace mifft

using namespace mifft;

extern "C" {

mifft_status mifft_xspec_create(int64_t fft_size, int64_t channels, mifft_xspec** out) {
    if (!out) return fail(MIFFT_ERR_INVALID_ARGUMENT, "null pointer");
    *out = nullptr;
    if (channels < 1 || channels > 16)
        return fail(MIFFT_ERR_PRECONDITION, "crossSpectral requires 1-16 channels");
    mifft_plan* plan = nullptr;
    mifft_status s = mifft_plan_create(fft_size, &plan);  // unsupportedFFTSize surfaces here
    if (s != MIFFT_OK) return s;
    mifft_xspec* x = new mifft_xspec();
    x->n = fft_size;
    x->channels = channels;
    x->plan = plan;
    *out = x;
    return MIFFT_OK;
}

mifft_status mifft_xspec_destroy(mifft_xspec* x) {
    if (!x) return MIFFT_OK;
    if (x->plan && x->plan->stream) (void)hipStreamSynchronize(x->plan->stream);
    (void)hipDeviceSynchronize();  // scratch may be in use on a caller stream
    x->spectra.release();
    x->partial.release();
    mifft_plan_destroy(x->plan);
    delete x;
    return MIFFT_OK;
}

int64_t mifft_xspec_frames(const mifft_xspec* x, int64_t samples, int64_t hop) {
    if (!x || hop < 1 || samples < x->n) return 0;
    return (samples - x->n) / hop + 1;
}

mifft_status mifft_xspec_device(mifft_xspec* x, const float* signals, int64_t samples, int64_t hop,
                                const float* window, float* power, float* cross, float* coherence,
                                float* welch_msc, void* stream) {
    if (!x) return fail(MIFFT_ERR_INVALID_ARGUMENT, "null analyser");
    if (samples < 0) return fail(MIFFT_ERR_INVALID_ARGUMENT, "negative size");
    if (hop < 1) return fail(MIFFT_ERR_PRECONDITION, "hopSize must be >= 1");
    const int64_t n = x->n, c = x->channels, pairs = c * (c - 1) / 2;
    if (!power || (pairs > 0 && (!cross || !coherence)))
        return fail(MIFFT_ERR_INVALID_ARGUMENT, "null output");
    hipStream_t st = static_cast<hipStream_t>(stream);
    const int64_t frames = mifft_xspec_frames(x, samples, hop);
    mifft_status s;
    if (frames == 0) {  // the reference returns its zeroed accumulators
        if ((s = zero(power, c * n, st)) != MIFFT_OK || (s = zero(cross, 2 * pairs * n, st)) != MIFFT_OK ||
            (s = zero(coherence, pairs * n, st)) != MIFFT_OK || (s = zero(welch_msc, pairs * n, st)) != MIFFT_OK)
            return s;
        return MIFFT_OK;
    }
    if (!signals) return fail(MIFFT_ERR_INVALID_ARGUMENT, "null signals");
    const int64_t total = c * frames;
    const int64_t chunks = (frames + kChunk - 1) / kChunk;
    const int nser = (int)(c + pairs);
    if ((s = x->spectra.reserve(sizeof(float2) * (size_t)(total * n))) != MIFFT_OK ||
        (s = x->partial.reserve(sizeof(float) * 3 * (size_t)(chunks * nser * n))) != MIFFT_OK)
        return s;
    float2* spec = static_cast<float2*>(x->spectra.ptr);
    float* part = static_cast<float*>(x->partial.ptr);
    if ((s = check_hip(xspec_fft((int)n, signals, samples, hop, frames, total, window, spec,
                                 x->plan->twiddles, st),
                       "cross-spectral FFT")) != MIFFT_OK)
        return s;
    const unsigned kb = (unsigned)((n + 255) / 256);
    if (chunks > 65535) return fail(MIFFT_ERR_INVALID_ARGUMENT, "too many frames");
    hipLaunchKernelGGL(k_xspec_partial, dim3(kb, (unsigned)nser, (unsigned)chunks), dim3(256), 0, st, spec,
                       part, n, (int)c, frames, nser);
    hipLaunchKernelGGL(k_xspec_reduce, dim3(kb, (unsigned)nser), dim3(256), 0, st, part, n, (int)c, chunks,
                       nser, 1.0f / (float)frames, power, reinterpret_cast<float2*>(cross), coherence);
    if (welch_msc && pairs > 0)
        hipLaunchKernelGGL(k_xspec_msc, dim3(kb, (unsigned)pairs), dim3(256), 0, st, power,
                           reinterpret_cast<const float2*>(cross), welch_msc, n, (int)c);
    return check_hip(hipGetLastError(), "cross-spectral reduction");
}

}  // extern "C"
