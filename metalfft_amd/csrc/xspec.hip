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
//   1. every windowed frame of every channel transformed at once by the
//      real-input kernel (realfft.hip: an N/2-point complex FFT per frame and
//      a split), bins 0..N/2 written to a spectra scratch [channel][frame]
//      [N/2 + 1] (a real frame's other bins are their conjugates);
//   2. partial sums over chunks of frames (16, or more for long signals),
//      one workgroup per (chunk, 64 bins), of every series: a channel's power |X_c|^2, a pair's cross
//      spectrum X_i conj(X_j) and per-frame coherence
//      |S_ij|^2 / max(P_i P_j, 1e-30) (the kernel's definition, :56-61);
//   3. the chunk partials added by a 64-row tree in a fixed order
//      (deterministic), divided by the frame count and written for all N bins (bin N - k from bin k,
//      conjugated for the cross spectra); optionally the magnitude-squared
//      coherence of the averaged spectra, |<S_ij>|^2 / (<P_i><P_j>), 0 where
//      the denominator is not > 0 (PSD.coherence, Sources/MetalFFT/DSP/
//      PSD.swift:127-133).
// The same frame transform writing all N bins, with |X|^2 and dB beside the
// spectra, serves STFT.analyze; power alone, scaled by 1 / (sum w^2 * fs), is
// PSD.welch.
// Pairs are the upper triangle (0,1), (0,2), ..., (C-2, C-1) as in
// CrossSpectralResult.pairIndex (CrossSpectral.swift:31-38).
#include <hip/hip_runtime.h>

#include <string>

#include "convolve.hpp"
#include "realfft.hpp"
#include "runtime.hpp"
#include "stockham_core.hpp"

struct mifft_xspec {
    int64_t n = 0;
    int64_t channels = 0;
    mifft_plan* plan = nullptr;                 // real-input table of length n
    mifft::DeviceBuffer spectra, partial, aux;  // scratch, grown on demand
    mifft::DeviceBuffer hin, hout;              // staging of the host entry points
};

namespace mifft {
namespace {

// Frames per partial sum: at least 16, more when the partials of 16-frame
// chunks would exceed 256 MiB or 4096 chunks.
int64_t chunk_frames(int64_t frames, int64_t nser, int64_t nb) {
    int64_t cap = ((int64_t)256 << 20) / (12 * nser * nb);
    cap = cap < 1 ? 1 : (cap > 4096 ? 4096 : cap);
    const int64_t len = (frames + cap - 1) / cap;
    return len < 16 ? 16 : len;
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
// (chunk h, series s, bin k <= N/2): {sum re S, sum im S, sum coherence} for
// pairs, {sum power, 0, 0} for channels, at partial[((h * nser + s) * n + k)
// * 3], n = N/2 + 1 bins per spectrum.
//
// A workgroup is 64 bins x 4 groups over one chunk of frames: per frame the
// spectra of all channels at its 64 bins are read once into LDS (each group
// loads channels g, g + 4, ...; the next frame's loads are issued before the
// current frame is used), then group g accumulates channel powers g, g + 4,
// ... and pairs g, g + 4, ... from LDS.  Each spectrum value is read from
// memory once, whatever the number of pairs.
// KC: channel capacity (4, 8 or 16), sizing the accumulators.
constexpr int kGroups = 4;

template <int KC>
__global__ void __launch_bounds__(256)
k_xspec_partial(const float2* __restrict__ spec, float* __restrict__ partial, int64_t n, int c,
                int64_t frames, int nser, int64_t chunk) {
    constexpr int kMaxC = KC, kMaxP = KC * (KC - 1) / 2;
    constexpr int kPairsPerGroup = (kMaxP + kGroups - 1) / kGroups;
    constexpr int kChanPerGroup = KC / kGroups;
    constexpr int D = KC == 16 ? 2 : 4;
    __shared__ float2 xs[KC == 16 ? 2 : 4][kMaxC][64];
    __shared__ int pi_[kMaxP], pj_[kMaxP];
    const int lane = threadIdx.x & 63, g = threadIdx.x >> 6;
    const int64_t k = (int64_t)blockIdx.y * 64 + lane;
    const int64_t kc = k < n ? k : n - 1;  // in-range bin for the loads
    const int64_t h = blockIdx.x;
    const int np = c * (c - 1) / 2;
    for (int p = threadIdx.x; p < np; p += 256) pair_of(p, c, pi_[p], pj_[p]);
    const int64_t f0 = h * chunk;
    const int64_t f1 = f0 + chunk < frames ? f0 + chunk : frames;
    float acc[kPairsPerGroup][3], pw[kChanPerGroup];
#pragma unroll
    for (int q = 0; q < kPairsPerGroup; ++q) acc[q][0] = acc[q][1] = acc[q][2] = 0.f;
#pragma unroll
    for (int q = 0; q < kChanPerGroup; ++q) pw[q] = 0.f;
    // loads of channel min(ch, c - 1) at every frame: unconditional, rows of
    // absent channels are never read back
    // D frames per step: their loads are issued one step ahead
    float2 cur[D][kChanPerGroup], nxt[D][kChanPerGroup];
    auto load = [&](float2 (&dst)[D][kChanPerGroup], int64_t fb) {
#pragma unroll
        for (int d = 0; d < D; ++d) {
            const int64_t fr = fb + d < f1 ? fb + d : f1 - 1;
#pragma unroll
            for (int q = 0; q < kChanPerGroup; ++q) {
                const int ch = g + kGroups * q < c ? g + kGroups * q : c - 1;
                dst[d][q] = spec[((int64_t)ch * frames + fr) * n + kc];
            }
        }
    };
    load(cur, f0);
    for (int64_t fb = f0; fb < f1; fb += D) {
        load(nxt, fb + D);
        __syncthreads();  // the previous step's LDS reads (and the pair table) are done
#pragma unroll
        for (int d = 0; d < D; ++d)
#pragma unroll
            for (int q = 0; q < kChanPerGroup; ++q) xs[d][g + kGroups * q][lane] = cur[d][q];
        __syncthreads();
#pragma unroll
        for (int d = 0; d < D; ++d) {
            if (fb + d >= f1) break;  // uniform across the workgroup
#pragma unroll
            for (int q = 0; q < kChanPerGroup; ++q) {
                const float2 u = cur[d][q];
                pw[q] += u.x * u.x + u.y * u.y;
            }
#pragma unroll
            for (int q = 0; q < kPairsPerGroup; ++q) {
                const int p = g + kGroups * q;
                if (p < np) {
                    const float2 u = xs[d][pi_[p]][lane], w = xs[d][pj_[p]][lane];
                    const float2 sx = make_float2(u.x * w.x + u.y * w.y, u.y * w.x - u.x * w.y);  // u conj(w)
                    const float pi = u.x * u.x + u.y * u.y, pj = w.x * w.x + w.y * w.y;
                    acc[q][0] += sx.x;
                    acc[q][1] += sx.y;
                    // hardware reciprocal (1 ulp): the quotient is a per-frame coherence in [0, 1]
                    acc[q][2] += (sx.x * sx.x + sx.y * sx.y) * __builtin_amdgcn_rcpf(fmaxf(pi * pj, 1e-30f));
                }
            }
        }
#pragma unroll
        for (int d = 0; d < D; ++d)
#pragma unroll
            for (int q = 0; q < kChanPerGroup; ++q) cur[d][q] = nxt[d][q];
    }
    if (k >= n) return;
#pragma unroll
    for (int q = 0; q < kChanPerGroup; ++q) {
        const int ch = g + kGroups * q;
        if (ch < c) {
            float* o = partial + ((h * nser + ch) * n + k) * 3;
            o[0] = pw[q];
            o[1] = 0.f;
            o[2] = 0.f;
        }
    }
#pragma unroll
    for (int q = 0; q < kPairsPerGroup; ++q) {
        const int p = g + kGroups * q;
        if (p < np && c + p < nser) {
            float* o = partial + ((h * nser + c + p) * n + k) * 3;
            o[0] = acc[q][0];
            o[1] = acc[q][1];
            o[2] = acc[q][2];
        }
    }
}

// Outputs for all N bins: power[c][k], cross[p][k] (complex), coherence[p][k],
// all / frames; bin k > N/2 is bin N - k (cross conjugated).
__global__ void __launch_bounds__(256)
k_xspec_reduce(const float* __restrict__ partial, int64_t n, int c, int64_t chunks, int nser,
               float inv_frames, float* __restrict__ power, float2* __restrict__ cross,
               float* __restrict__ coherence) {
    const int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int s = blockIdx.y;
    if (k >= n) return;
    const int64_t nb = n / 2 + 1;
    const bool mirror = k >= nb;
    const int64_t kk = mirror ? n - k : k;
    float a0 = 0.f, a1 = 0.f, a2 = 0.f;
    for (int64_t h = 0; h < chunks; ++h) {
        const float* o = partial + ((h * nser + s) * nb + kk) * 3;
        a0 += o[0];
        a1 += o[1];
        a2 += o[2];
    }
    if (s < c) {
        power[(int64_t)s * n + k] = a0 * inv_frames;
    } else {
        const int p = s - c;
        cross[(int64_t)p * n + k] = make_float2(a0 * inv_frames, (mirror ? -a1 : a1) * inv_frames);
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
    const float den = power[(int64_t)i * n + k] * power[(int64_t)j * n + k];
    msc[(int64_t)p * n + k] = den > 0.f ? (s.x * s.x + s.y * s.y) / den : 0.f;
}

__global__ void k_fill_zero(float* __restrict__ p, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.f;
}

// psd[c][k] = sum of |X[k]|^2 over frames / (frames * window sum of squares
// * sample rate) (PSD.swift:55-58), k < N, from sums[c][k'] over bins
// k' <= N/2 (bin k > N/2 is bin N - k).  Every 256-thread block forms the
// window's sum of squares itself (a strided sum and a fixed tree, the same
// in every block, N <= 16384 values), so no separate launch precedes this
// one.
__global__ void __launch_bounds__(256) k_psd_finish(const float* __restrict__ sums,
                                                    const float* __restrict__ w, int64_t n,
                                                    int64_t total, float inv_frames_rate,
                                                    float* __restrict__ psd) {
    __shared__ float part[256];
    float s = 0.f;
    for (int64_t i = threadIdx.x; i < n; i += 256) s += w ? w[i] * w[i] : 1.f;
    part[threadIdx.x] = s;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) part[threadIdx.x] += part[threadIdx.x + off];
        __syncthreads();
    }
    const float sumsq = part[0];
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= total) return;
    const int64_t c = i / n, k = i % n, nb = n / 2 + 1;
    psd[i] = sums[c * nb + (k < nb ? k : n - k)] * (inv_frames_rate / sumsq);
}

mifft_status zero(float* p, int64_t n, hipStream_t st) {
    if (!p || n <= 0) return MIFFT_OK;
    hipLaunchKernelGGL(k_fill_zero, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, p, n);
    return check_hip(hipGetLastError(), "zero fill");
}

// Frame transforms, chunk partials and their reduction for `frames` >= 1
// frames of every channel; pair series only when `pairs` > 0 (power only
// otherwise).
mifft_status xspec_run(mifft_xspec* x, const float* signals, int64_t samples, int64_t hop,
                       int64_t frames, const float* window, float* power, float* cross,
                       float* coherence, int64_t pairs, hipStream_t st) {
    const int64_t n = x->n, c = x->channels, nb = n / 2 + 1;
    const int64_t total = c * frames;
    const int nser = (int)(c + pairs);
    const int64_t chunk = chunk_frames(frames, nser, nb);
    const int64_t chunks = (frames + chunk - 1) / chunk;
    mifft_status s;
    const int64_t cols = 3 * nser * nb, lvl1 = (chunks + 63) / 64;
    if ((s = x->spectra.reserve(sizeof(float2) * (size_t)(total * nb))) != MIFFT_OK ||
        (s = x->partial.reserve(sizeof(float) * (size_t)((chunks + 2 * lvl1 + 1) * cols))) != MIFFT_OK)
        return s;
    float2* spec = static_cast<float2*>(x->spectra.ptr);
    float* part = static_cast<float*>(x->partial.ptr);
    const RealFrames a{signals, samples, hop, frames, total, window, spec, nullptr, nullptr, 0.f, nb,
                       x->plan->real_tw};
    if ((s = check_hip(real_frames_forward((int)n, a, st), "cross-spectral FFT")) != MIFFT_OK) return s;
    const dim3 pgrid((unsigned)chunks, (unsigned)((nb + 63) / 64));
    if (c <= 4)
        hipLaunchKernelGGL(k_xspec_partial<4>, pgrid, dim3(256), 0, st, spec, part, nb, (int)c, frames, nser, chunk);
    else if (c <= 8)
        hipLaunchKernelGGL(k_xspec_partial<8>, pgrid, dim3(256), 0, st, spec, part, nb, (int)c, frames, nser, chunk);
    else
        hipLaunchKernelGGL(k_xspec_partial<16>, pgrid, dim3(256), 0, st, spec, part, nb, (int)c, frames, nser,
                           chunk);
    // chunk partials summed by the 64-row tree (fixed order), then scaled and
    // mirrored to all N bins
    float* sums = part + (chunks + 2 * lvl1) * cols;
    if ((s = check_hip(sum_rows(part, chunks, (int)cols, part + chunks * cols, part + (chunks + lvl1) * cols, sums,
                                st),
                       "cross-spectral sums")) != MIFFT_OK)
        return s;
    hipLaunchKernelGGL(k_xspec_reduce, dim3((unsigned)((n + 255) / 256), (unsigned)nser), dim3(256), 0, st,
                       sums, n, (int)c, (int64_t)1, nser, 1.0f / (float)frames, power,
                       reinterpret_cast<float2*>(cross), coherence);
    return check_hip(hipGetLastError(), "cross-spectral reduction");
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
    mifft_status s = require_supported_size(fft_size);
    if (s != MIFFT_OK) return s;
    s = mifft_plan_create(fft_size, &plan);
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
    x->aux.release();
    x->hin.release();
    x->hout.release();
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
    if ((s = xspec_run(x, signals, samples, hop, frames, window, power, cross, coherence, pairs, st)) !=
        MIFFT_OK)
        return s;
    if (welch_msc && pairs > 0)
        hipLaunchKernelGGL(k_xspec_msc, dim3((unsigned)((n + 255) / 256), (unsigned)pairs), dim3(256), 0, st,
                           power, reinterpret_cast<const float2*>(cross), welch_msc, n, (int)c);
    return check_hip(hipGetLastError(), "magnitude-squared coherence");
}

mifft_status mifft_stft_device(mifft_xspec* x, const float* signals, int64_t samples, int64_t hop,
                               const float* window, float* spectra, float* power, float* power_db,
                               float db_alpha, void* stream) {
    if (!x) return fail(MIFFT_ERR_INVALID_ARGUMENT, "null analyser");
    if (samples < 0) return fail(MIFFT_ERR_INVALID_ARGUMENT, "negative size");
    if (hop < 1) return fail(MIFFT_ERR_PRECONDITION, "hopSize must be >= 1");
    const int64_t frames = mifft_xspec_frames(x, samples, hop);
    if (frames == 0) return MIFFT_OK;
    if (!signals || !spectra) return fail(MIFFT_ERR_INVALID_ARGUMENT, "null buffer");
    const RealFrames a{signals, samples,  hop,      frames, x->channels * frames, window,
                       reinterpret_cast<float2*>(spectra), power, power_db, db_alpha, x->n,
                       x->plan->real_tw};
    return check_hip(real_frames_forward((int)x->n, a, static_cast<hipStream_t>(stream)), "STFT");
}

mifft_status mifft_welch_device(mifft_xspec* x, const float* signals, int64_t samples, int64_t hop,
                                const float* window, double sample_rate, float* psd, void* stream) {
    if (!x) return fail(MIFFT_ERR_INVALID_ARGUMENT, "null analyser");
    if (!(sample_rate > 0)) return fail(MIFFT_ERR_PRECONDITION, "sampleRate must be > 0");
    if (!psd) return fail(MIFFT_ERR_INVALID_ARGUMENT, "null output");
    if (samples < 0) return fail(MIFFT_ERR_INVALID_ARGUMENT, "negative size");
    if (hop < 1) return fail(MIFFT_ERR_PRECONDITION, "hopSize must be >= 1");
    hipStream_t st = static_cast<hipStream_t>(stream);
    const int64_t n = x->n, c = x->channels;
    const int64_t frames = mifft_xspec_frames(x, samples, hop);
    if (frames == 0) return zero(psd, c * n, st);  // the reference returns its zeroed accumulator
    if (!signals) return fail(MIFFT_ERR_INVALID_ARGUMENT, "null signals");
    // workgroups: about 2048 over all channels, each summing `iters` groups
    const int64_t nb = n / 2 + 1, cols = c * nb;
    const int64_t grp = real_frames_per_group((int)n);
    const int64_t groups = (frames + grp - 1) / grp;
    int64_t wpc = (2048 + c - 1) / c;
    if (wpc > groups) wpc = groups;
    const int64_t iters = (groups + wpc - 1) / wpc;
    const int64_t lvl1 = (wpc + 63) / 64;
    mifft_status s;
    if ((s = x->aux.reserve(sizeof(float) * (size_t)((wpc + 2 * lvl1 + 1) * cols))) != MIFFT_OK) return s;
    float* part = static_cast<float*>(x->aux.ptr);
    float* sums = part + (wpc + 2 * lvl1) * cols;
    const RealPower a{signals, samples, hop, frames, c, wpc, iters, window, part, x->plan->real_tw};
    if ((s = check_hip(real_power((int)n, a, st), "Welch power")) != MIFFT_OK ||
        (s = check_hip(sum_rows(part, wpc, (int)cols, part + wpc * cols, part + (wpc + lvl1) * cols, sums, st),
                       "Welch reduction")) != MIFFT_OK)
        return s;
    hipLaunchKernelGGL(k_psd_finish, dim3((unsigned)((c * n + 255) / 256)), dim3(256), 0, st, sums, window, n,
                       c * n, (float)(1.0 / ((double)frames * sample_rate)), psd);
    return check_hip(hipGetLastError(), "Welch PSD");
}

// ---- host entry points: stage through device buffers on the analyser's
// stream, run the device form, copy back, block (the reference's calls are
// synchronous).

mifft_status mifft_xspec_host(mifft_xspec* x, const float* signals, int64_t samples, int64_t hop,
                              const float* window, float* power, float* cross, float* coherence,
                              float* welch_msc) {
    if (!x) return fail(MIFFT_ERR_INVALID_ARGUMENT, "null analyser");
    if (samples < 0) return fail(MIFFT_ERR_INVALID_ARGUMENT, "negative size");
    if (hop < 1) return fail(MIFFT_ERR_PRECONDITION, "hopSize must be >= 1");
    const int64_t n = x->n, c = x->channels, pairs = c * (c - 1) / 2;
    if (!power || (pairs > 0 && (!cross || !coherence)) || (samples > 0 && !signals))
        return fail(MIFFT_ERR_INVALID_ARGUMENT, "null buffer");
    hipStream_t st = x->plan->stream;
    const size_t in_f = (size_t)(c * samples), w_f = window ? (size_t)n : 0;
    const size_t o_pw = (size_t)(c * n), o_cr = (size_t)(2 * pairs * n), o_co = (size_t)(pairs * n);
    const size_t o_ms = welch_msc ? (size_t)(pairs * n) : 0;
    mifft_status s;
    if ((s = x->hin.reserve(4 * (in_f + w_f + 1))) != MIFFT_OK ||
        (s = x->hout.reserve(4 * (o_pw + o_cr + o_co + o_ms + 1))) != MIFFT_OK)
        return s;
    float* di = static_cast<float*>(x->hin.ptr);
    float* dout = static_cast<float*>(x->hout.ptr);
    if (in_f && (s = check_hip(hipMemcpyAsync(di, signals, 4 * in_f, hipMemcpyHostToDevice, st), "hipMemcpyAsync")) !=
                    MIFFT_OK)
        return s;
    if (w_f && (s = check_hip(hipMemcpyAsync(di + in_f, window, 4 * w_f, hipMemcpyHostToDevice, st),
                              "hipMemcpyAsync")) != MIFFT_OK)
        return s;
    float* dpw = dout;
    float* dcr = dpw + o_pw;
    float* dco = dcr + o_cr;
    float* dms = dco + o_co;
    if ((s = mifft_xspec_device(x, in_f ? di : nullptr, samples, hop, w_f ? di + in_f : nullptr, dpw,
                                pairs ? dcr : nullptr, pairs ? dco : nullptr, o_ms ? dms : nullptr, st)) != MIFFT_OK)
        return s;
    struct Out {
        float* h;
        const float* d;
        size_t n;
    } outs[] = {{power, dpw, o_pw}, {cross, dcr, o_cr}, {coherence, dco, o_co}, {welch_msc, dms, o_ms}};
    for (const Out& o : outs)
        if (o.h && o.n && (s = check_hip(hipMemcpyAsync(o.h, o.d, 4 * o.n, hipMemcpyDeviceToHost, st),
                                         "hipMemcpyAsync")) != MIFFT_OK)
            return s;
    return check_hip(hipStreamSynchronize(st), "hipStreamSynchronize");
}

mifft_status mifft_stft_host(mifft_xspec* x, const float* signals, int64_t samples, int64_t hop,
                             const float* window, float* spectra, float* power, float* power_db,
                             float db_alpha) {
    if (!x) return fail(MIFFT_ERR_INVALID_ARGUMENT, "null analyser");
    if (samples < 0) return fail(MIFFT_ERR_INVALID_ARGUMENT, "negative size");
    if (hop < 1) return fail(MIFFT_ERR_PRECONDITION, "hopSize must be >= 1");
    const int64_t frames = mifft_xspec_frames(x, samples, hop);
    if (frames == 0) return MIFFT_OK;
    if (!signals || !spectra) return fail(MIFFT_ERR_INVALID_ARGUMENT, "null buffer");
    hipStream_t st = x->plan->stream;
    const int64_t n = x->n, c = x->channels;
    const size_t in_f = (size_t)(c * samples), w_f = window ? (size_t)n : 0;
    const size_t bins = (size_t)(c * frames * n);
    const size_t o_pw = power ? bins : 0, o_db = power_db ? bins : 0;
    mifft_status s;
    if ((s = x->hin.reserve(4 * (in_f + w_f + 1))) != MIFFT_OK ||
        (s = x->hout.reserve(4 * (2 * bins + o_pw + o_db))) != MIFFT_OK)
        return s;
    float* di = static_cast<float*>(x->hin.ptr);
    float* dsp = static_cast<float*>(x->hout.ptr);
    float* dpw = dsp + 2 * bins;
    float* ddb = dpw + o_pw;
    if ((s = check_hip(hipMemcpyAsync(di, signals, 4 * in_f, hipMemcpyHostToDevice, st), "hipMemcpyAsync")) !=
            MIFFT_OK ||
        (w_f && (s = check_hip(hipMemcpyAsync(di + in_f, window, 4 * w_f, hipMemcpyHostToDevice, st),
                               "hipMemcpyAsync")) != MIFFT_OK))
        return s;
    if ((s = mifft_stft_device(x, di, samples, hop, w_f ? di + in_f : nullptr, dsp, o_pw ? dpw : nullptr,
                               o_db ? ddb : nullptr, db_alpha, st)) != MIFFT_OK)
        return s;
    if ((s = check_hip(hipMemcpyAsync(spectra, dsp, 8 * bins, hipMemcpyDeviceToHost, st), "hipMemcpyAsync")) !=
            MIFFT_OK ||
        (o_pw && (s = check_hip(hipMemcpyAsync(power, dpw, 4 * bins, hipMemcpyDeviceToHost, st),
                                "hipMemcpyAsync")) != MIFFT_OK) ||
        (o_db && (s = check_hip(hipMemcpyAsync(power_db, ddb, 4 * bins, hipMemcpyDeviceToHost, st),
                                "hipMemcpyAsync")) != MIFFT_OK))
        return s;
    return check_hip(hipStreamSynchronize(st), "hipStreamSynchronize");
}

mifft_status mifft_welch_host(mifft_xspec* x, const float* signals, int64_t samples, int64_t hop,
                              const float* window, double sample_rate, float* psd) {
    if (!x) return fail(MIFFT_ERR_INVALID_ARGUMENT, "null analyser");
    if (samples < 0) return fail(MIFFT_ERR_INVALID_ARGUMENT, "negative size");
    if (!psd || (samples > 0 && !signals)) return fail(MIFFT_ERR_INVALID_ARGUMENT, "null buffer");
    hipStream_t st = x->plan->stream;
    const int64_t n = x->n, c = x->channels;
    const size_t in_f = (size_t)(c * samples), w_f = window ? (size_t)n : 0;
    mifft_status s;
    if ((s = x->hin.reserve(4 * (in_f + w_f + 1))) != MIFFT_OK ||
        (s = x->hout.reserve(4 * (size_t)(c * n))) != MIFFT_OK)
        return s;
    float* di = static_cast<float*>(x->hin.ptr);
    if (in_f && (s = check_hip(hipMemcpyAsync(di, signals, 4 * in_f, hipMemcpyHostToDevice, st), "hipMemcpyAsync")) !=
                    MIFFT_OK)
        return s;
    if (w_f && (s = check_hip(hipMemcpyAsync(di + in_f, window, 4 * w_f, hipMemcpyHostToDevice, st),
                              "hipMemcpyAsync")) != MIFFT_OK)
        return s;
    float* dpsd = static_cast<float*>(x->hout.ptr);
    if ((s = mifft_welch_device(x, in_f ? di : nullptr, samples, hop, w_f ? di + in_f : nullptr, sample_rate,
                                dpsd, st)) != MIFFT_OK)
        return s;
    if ((s = check_hip(hipMemcpyAsync(psd, dpsd, 4 * (size_t)(c * n), hipMemcpyDeviceToHost, st),
                       "hipMemcpyAsync")) != MIFFT_OK)
        return s;
    return check_hip(hipStreamSynchronize(st), "hipStreamSynchronize");
}

}  // extern "C"
