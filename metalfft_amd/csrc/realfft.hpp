// Real-input transforms: an N-point real FFT as one N/2-point complex FFT of
// the even/odd sample pairs and a split.  Shared by the real-input plan entry
// points and the framed analyses (STFT, Welch, cross-spectral).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mifft {

// Real-FFT table of a length-n plan: the twiddle table of the n/2-point
// transform (twiddle_table_len(n/2) entries, n/2 = 32 included), then
// W_n^k for k < n/2.
int real_table_len(int n);
int real_roots_offset(int n);
hipError_t launch_real_table(float2* tab, int n, hipStream_t st);

// Frame g < total of a C x F frame list: channel g / frames, frame g % frames,
// samples sig[c * samples + f * hop + j] times window[j] (window null:
// rectangular), j < N.  Writes `bins` values per frame (N: the whole
// spectrum; N/2 + 1: bins 0..N/2) to spec[g * bins], and when not null
// |X|^2 to pw[g * bins] and db_alpha * log10 |X|^2 floored at -120 to
// db[g * bins].
struct RealFrames {
    const float* sig;
    int64_t samples, hop, frames, total;
    const float* window;
    float2* spec;
    float* pw;
    float* db;
    float db_alpha;
    int64_t bins;
    const float2* tab;  // real_table of N
};
hipError_t real_frames_forward(int n, const RealFrames& a, hipStream_t st);

// Welch power sums without a spectra pass: frames of `channels` signals as
// in RealFrames (per channel, `frames` frames), workgroup (c, j), j < wpc,
// sums |X[k]|^2, k <= N/2, over frames [j * iters * G, (j + 1) * iters * G)
// (G = real_frames_per_group(n) frames per transform group) and writes
// partial[j * channels * (N/2 + 1) + c * (N/2 + 1) + k].  Needs
// wpc * iters * G >= frames.
struct RealPower {
    const float* sig;
    int64_t samples, hop, frames, channels, wpc, iters;
    const float* window;
    float* partial;
    const float2* tab;
};
int real_frames_per_group(int n);
hipError_t real_power(int n, const RealPower& a, hipStream_t st);

// Inverse of half spectra: rows y[b * in_stride + k], k <= N/2, of Hermitian
// spectra -> x[b * N + j] = (1/N) sum_k X[k] W^{-jk} (real).
hipError_t real_inverse(int n, const float2* y, int64_t in_stride, float* x, int64_t batch,
                        const float2* tab, hipStream_t st);

}  // namespace mifft
