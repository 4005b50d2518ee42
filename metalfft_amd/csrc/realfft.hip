// Real-input transforms.
//
// The reference's callers turn real signals into complex ones with a zero
// imaginary part and run the complex FFT (Convolve.swift:48-53,
// FusedConvolver.swift:101-107, Correlate.swift:53-61, STFT.swift:116-122,
// PSD.swift:37-45, FFTAnalyzer.swift:93-99).  Here a length-N real frame is
// read as M = N/2 complex values z[n] = x[2n] + i x[2n+1], transformed with
// the M-point register/LDS Stockham of the complex path, and split:
//   E[k] = (Z[k] + conj Z[M-k]) / 2,  O[k] = (Z[k] - conj Z[M-k]) / 2i,
//   X[k] = E[k] + W_N^k O[k],  X[k+M] = E[k] - W_N^k O[k],  k < M,
// half the transform work of the (x, 0) form, one frame per transform (no
// frame shares a transform with another, so a zero frame gives exact zeros).
// The inverse runs the same identities backwards on bins 0..M of a Hermitian
// spectrum: Z[k] = E[k] + i O[k], O[k] = (X[k] - conj X[M-k]) conj(W_N^k) / 2,
// z = IFFT_M(Z), x[2n] = Re z[n], x[2n+1] = Im z[n].
#include <hip/hip_runtime.h>

#include "kernels.hpp"
#include "realfft.hpp"
#include "runtime.hpp"
#include "stockham_core.hpp"

namespace mifft {
namespace {

// Exchange layout of the half-length transforms: XOR-swizzled, no padding,
// so two 2048-point transforms take 32 KiB (five workgroups per CU; the
// padded layout's 34 KiB allowed four).
using RLds = LdsSwz;

template <int M>
struct RCfg {
    static constexpr int E = M >= 256 ? 16 : 8;
    static constexpr int TPW = tpw_for<M, E>();
    static constexpr int threads = (M / E) * TPW;
};

// W_n^k for k < n/2, FP64 sincospi rounded once to FP32.
__global__ void k_half_roots(float2* __restrict__ r, int n) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n / 2) {
        double s, c;
        sincospi(-2.0 * (double)k / (double)n, &s, &c);
        r[k] = make_float2((float)c, (float)s);
    }
}

// Samples j, j + 1 of a frame (times the window): one 8-byte load when the
// frame start and the window are 8-byte aligned (PAIR), else two.
template <bool WIN, bool PAIR>
__device__ __forceinline__ float2 ld_pair(const float* p, const float* w, int j) {
    float2 x;
    if constexpr (PAIR) {
        x = *reinterpret_cast<const float2*>(p + j);
    } else {
        x = make_float2(p[j], p[j + 1]);
    }
    if constexpr (WIN) {
        const float2 g = PAIR ? *reinterpret_cast<const float2*>(w + j) : make_float2(w[j], w[j + 1]);
        x = make_float2(x.x * g.x, x.y * g.y);
    }
    return x;
}

template <int N, bool WIN, bool PAIR>
__global__ void __launch_bounds__(RCfg<N / 2>::threads) k_real_frames(RealFrames a, const float2* roots) {
    constexpr int M = N / 2;
    using C = RCfg<M>;
    constexpr int E = C::E, T = M / E;
    __shared__ float2 lds[C::TPW * RLds::len(M)];
    const int local = threadIdx.x / T;
    const int t = threadIdx.x % T;
    float2* buf = lds + local * RLds::len(M);
    const int64_t g = (int64_t)blockIdx.x * C::TPW + local;
    const bool act = g < a.total;
    const int64_t gc = act ? g : 0;  // an in-range frame for the loads
    const float* p = a.sig + (gc / a.frames) * a.samples + (gc % a.frames) * a.hop;
    float2 v[E];
#pragma unroll
    for (int m = 0; m < E; ++m) v[m] = ld_pair<WIN, PAIR>(p, a.window, 2 * (t + T * m));  // whole frames
    RegTwiddles<M, E> rt;
    rt.load(a.tab, t);
    rt.template fft<false, RLds>(v, buf, t, a.tab);
    lds_barrier();
#pragma unroll
    for (int m = 0; m < E; ++m) buf[RLds::idx(t + T * m)] = v[m];
    lds_barrier();
    const bool full = a.bins == N;
    float2* out = a.spec + gc * a.bins;
#pragma unroll
    for (int m = 0; m < E; ++m) {
        const int k = t + T * m;
        const float2 z = v[m];
        const float2 zc = conjf2(buf[RLds::idx((M - k) & (M - 1))]);
        const float2 e = make_float2(0.5f * (z.x + zc.x), 0.5f * (z.y + zc.y));
        const float2 o = mul_mi(make_float2(0.5f * (z.x - zc.x), 0.5f * (z.y - zc.y)));
        const float2 wo = cmul(roots[k], o);
        const float2 lo = cadd(e, wo), hi = csub(e, wo);
        const bool put_hi = full || k == 0;
        const int khi = full ? k + M : M;
        if (act) {  // streamed outputs: non-temporal stores
            gstore<true>(out + k, lo);
            if (put_hi) gstore<true>(out + khi, hi);
        }
        if (a.pw || a.db) {
            const float ql = lo.x * lo.x + lo.y * lo.y, qh = hi.x * hi.x + hi.y * hi.y;
            if (a.pw && act) {
                __builtin_nontemporal_store(ql, a.pw + gc * a.bins + k);
                if (put_hi) __builtin_nontemporal_store(qh, a.pw + gc * a.bins + khi);
            }
            if (a.db && act) {
                __builtin_nontemporal_store(fmaxf(a.db_alpha * log10f(ql), -120.f), a.db + gc * a.bins + k);
                if (put_hi)
                    __builtin_nontemporal_store(fmaxf(a.db_alpha * log10f(qh), -120.f), a.db + gc * a.bins + khi);
            }
        }
    }
}

// Welch power without a spectra pass: workgroup (c, j) of the wpc per
// channel transforms frame groups j*iters .. j*iters + iters - 1 (TPW frames
// each) of channel c and sums |X[k]|^2, k <= N/2, over them in registers;
// the TPW slots are then added through LDS in a fixed order and written as
// partial[j * cols + c * (N/2 + 1) + k].
template <int N, bool WIN, bool PAIR>
__global__ void __launch_bounds__(RCfg<N / 2>::threads) __attribute__((amdgpu_waves_per_eu(4, 8)))
k_real_power(const float* __restrict__ sig, int64_t samples, int64_t hop, int64_t frames, int64_t wpc,
             int64_t iters, const float* __restrict__ window, float* __restrict__ partial, int64_t cols,
             const float2* __restrict__ tab, const float2* __restrict__ roots) {
    constexpr int M = N / 2;
    using C = RCfg<M>;
    constexpr int E = C::E, T = M / E;
    __shared__ float2 lds[C::TPW * RLds::len(M)];
    const int local = threadIdx.x / T;
    const int t = threadIdx.x % T;
    float2* buf = lds + local * RLds::len(M);
    const int64_t c = blockIdx.x / wpc, j = blockIdx.x % wpc;
    const float* base = sig + c * samples;
    float acc[E], acc_m = 0.f;
#pragma unroll
    for (int m = 0; m < E; ++m) acc[m] = 0.f;
    RegTwiddles<M, E> rt;
    rt.load(tab, t);
    for (int64_t it = 0; it < iters; ++it) {
        const int tl = fresh(t);  // per-iteration index arithmetic (not hoisted)
        const int64_t f = (j * iters + it) * C::TPW + local;
        const bool act = f < frames;
        const float* p = base + (act ? f : 0) * hop;
        float2 v[E];
#pragma unroll
        for (int m = 0; m < E; ++m) v[m] = ld_pair<WIN, PAIR>(p, window, 2 * (tl + T * m));
        rt.template fft<true, RLds>(v, buf, tl, tab);
        lds_barrier();
#pragma unroll
        for (int m = 0; m < E; ++m) buf[RLds::idx(tl + T * m)] = v[m];
        lds_barrier();
        const float keep = act ? 1.f : 0.f;
#pragma unroll
        for (int m = 0; m < E; ++m) {
            const int k = tl + T * m;
            const float2 z = v[m];
            const float2 zc = conjf2(buf[RLds::idx((M - k) & (M - 1))]);
            const float2 e = make_float2(0.5f * (z.x + zc.x), 0.5f * (z.y + zc.y));
            const float2 o = mul_mi(make_float2(0.5f * (z.x - zc.x), 0.5f * (z.y - zc.y)));
            const float2 lo = cadd(e, cmul(roots[k], o));
            acc[m] += keep * (lo.x * lo.x + lo.y * lo.y);
            if (m == 0) {  // bin M from k = 0: E[0] - O[0]
                const float hm = e.x - o.x;
                acc_m += keep * (hm * hm);
            }
        }
    }
    lds_barrier();
    float* red = reinterpret_cast<float*>(lds);  // TPW x (M + 1) floats
#pragma unroll
    for (int m = 0; m < E; ++m) red[local * (M + 1) + t + T * m] = acc[m];
    if (t == 0) red[local * (M + 1) + M] = acc_m;
    lds_barrier();
    for (int i = threadIdx.x; i <= M; i += C::threads) {
        float s = 0.f;
#pragma unroll
        for (int l = 0; l < C::TPW; ++l) s += red[l * (M + 1) + i];
        partial[j * cols + c * (M + 1) + i] = s;
    }
}

template <int N, bool PAIR>
__global__ void __launch_bounds__(RCfg<N / 2>::threads)
k_real_inverse(const float2* __restrict__ y, int64_t in_stride, float* __restrict__ x, int64_t batch,
               const float2* __restrict__ tab, const float2* __restrict__ roots) {
    constexpr int M = N / 2;
    using C = RCfg<M>;
    constexpr int E = C::E, T = M / E;
    __shared__ float2 lds[C::TPW * RLds::len(M)];
    const int local = threadIdx.x / T;
    const int t = threadIdx.x % T;
    float2* buf = lds + local * RLds::len(M);
    const int64_t b = (int64_t)blockIdx.x * C::TPW + local;
    const bool act = b < batch;
    const float2* row = y + (act ? b : 0) * in_stride;
    float2 v[E];
#pragma unroll
    for (int m = 0; m < E; ++m) {
        const int k = t + T * m;
        const float2 xa = row[k];
        const float2 xc = conjf2(row[M - k]);  // k = 0: bin M
        const float2 e = make_float2(0.5f * (xa.x + xc.x), 0.5f * (xa.y + xc.y));
        const float2 o = cmulc(make_float2(0.5f * (xa.x - xc.x), 0.5f * (xa.y - xc.y)), roots[k]);
        // conj(E + i O): a forward transform of it is M conj(z)
        v[m] = make_float2(e.x - o.y, -(e.y + o.x));
    }
    RegTwiddles<M, E> rt;
    rt.load(tab, t);
    rt.template fft<false, RLds>(v, buf, t, tab);
    if (act) {
        constexpr float inv_m = 1.0f / (float)M;
        float* out = x + b * N;
#pragma unroll
        for (int m = 0; m < E; ++m) {
            const int j = 2 * (t + T * m);
            const float2 y = make_float2(v[m].x * inv_m, -v[m].y * inv_m);
            if constexpr (PAIR) {
                gstore<true>(reinterpret_cast<float2*>(out + j), y);
            } else {
                __builtin_nontemporal_store(y.x, out + j);
                __builtin_nontemporal_store(y.y, out + j + 1);
            }
        }
    }
}

inline bool aligned8(const void* p) { return (reinterpret_cast<uintptr_t>(p) & 7) == 0; }

template <int N>
hipError_t frames_n(const RealFrames& a, hipStream_t st) {
    using C = RCfg<N / 2>;
    const int64_t blocks = (a.total + C::TPW - 1) / C::TPW;
    if (blocks > 0x7fffffffLL) return hipErrorInvalidValue;
    const float2* roots = a.tab + real_roots_offset(N);
    const bool pair = aligned8(a.sig) && aligned8(a.window) && a.samples % 2 == 0 && a.hop % 2 == 0;
    const dim3 grid((unsigned)blocks), block(C::threads);
    if (a.window) {
        if (pair) hipLaunchKernelGGL((k_real_frames<N, true, true>), grid, block, 0, st, a, roots);
        else hipLaunchKernelGGL((k_real_frames<N, true, false>), grid, block, 0, st, a, roots);
    } else {
        if (pair) hipLaunchKernelGGL((k_real_frames<N, false, true>), grid, block, 0, st, a, roots);
        else hipLaunchKernelGGL((k_real_frames<N, false, false>), grid, block, 0, st, a, roots);
    }
    return hipGetLastError();
}

template <int N>
hipError_t power_n(const RealPower& a, hipStream_t st) {
    constexpr int TPW = RCfg<N / 2>::TPW;
    const int64_t blocks = a.channels * a.wpc;
    if (blocks > 0x7fffffffLL || a.iters * a.wpc * TPW < a.frames) return hipErrorInvalidValue;
    const float2* roots = a.tab + real_roots_offset(N);
    const int64_t cols = a.channels * (N / 2 + 1);
    const bool pair = aligned8(a.sig) && aligned8(a.window) && a.samples % 2 == 0 && a.hop % 2 == 0;
    const dim3 grid((unsigned)blocks), block(RCfg<N / 2>::threads);
#define MIFFT_POW(W, P)                                                                                   \
    hipLaunchKernelGGL((k_real_power<N, W, P>), grid, block, 0, st, a.sig, a.samples, a.hop, a.frames, a.wpc, \
                       a.iters, a.window, a.partial, cols, a.tab, roots)
    if (a.window) {
        if (pair) MIFFT_POW(true, true);
        else MIFFT_POW(true, false);
    } else {
        if (pair) MIFFT_POW(false, true);
        else MIFFT_POW(false, false);
    }
#undef MIFFT_POW
    return hipGetLastError();
}

template <int N>
hipError_t inverse_n(const float2* y, int64_t in_stride, float* x, int64_t batch, const float2* tab,
                     hipStream_t st) {
    using C = RCfg<N / 2>;
    const int64_t blocks = (batch + C::TPW - 1) / C::TPW;
    if (blocks > 0x7fffffffLL) return hipErrorInvalidValue;
    if (aligned8(x))
        hipLaunchKernelGGL((k_real_inverse<N, true>), dim3((unsigned)blocks), dim3(C::threads), 0, st, y, in_stride,
                           x, batch, tab, tab + real_roots_offset(N));
    else
        hipLaunchKernelGGL((k_real_inverse<N, false>), dim3((unsigned)blocks), dim3(C::threads), 0, st, y,
                           in_stride, x, batch, tab, tab + real_roots_offset(N));
    return hipGetLastError();
}

}  // namespace

int real_roots_offset(int n) { return twiddle_table_len(n / 2); }  // n/2 = 32: roots only
int real_table_len(int n) { return real_roots_offset(n) + n / 2; }

hipError_t launch_real_table(float2* tab, int n, hipStream_t st) {
    hipError_t e = launch_twiddles(tab, n / 2, st);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_half_roots, dim3((n / 2 + 255) / 256), dim3(256), 0, st, tab + real_roots_offset(n), n);
    return hipGetLastError();
}

#define MIFFT_REAL_SIZES(X) X(64) X(128) X(256) X(512) X(1024) X(2048) X(4096) X(8192) X(16384)

hipError_t real_frames_forward(int n, const RealFrames& a, hipStream_t st) {
    if (a.total <= 0) return hipSuccess;
#define MIFFT_RF(NN) \
    case NN: return frames_n<NN>(a, st);
    switch (n) { MIFFT_REAL_SIZES(MIFFT_RF) }
#undef MIFFT_RF
    return hipErrorInvalidValue;
}

int real_frames_per_group(int n) {
#define MIFFT_RT(NN) \
    case NN: return RCfg<NN / 2>::TPW;
    switch (n) { MIFFT_REAL_SIZES(MIFFT_RT) }
#undef MIFFT_RT
    return 0;
}

hipError_t real_power(int n, const RealPower& a, hipStream_t st) {
    if (a.frames <= 0) return hipSuccess;
#define MIFFT_RP(NN) \
    case NN: return power_n<NN>(a, st);
    switch (n) { MIFFT_REAL_SIZES(MIFFT_RP) }
#undef MIFFT_RP
    return hipErrorInvalidValue;
}

hipError_t real_inverse(int n, const float2* y, int64_t in_stride, float* x, int64_t batch,
                        const float2* tab, hipStream_t st) {
    if (batch <= 0) return hipSuccess;
#define MIFFT_RI(NN) \
    case NN: return inverse_n<NN>(y, in_stride, x, batch, tab, st);
    switch (n) { MIFFT_REAL_SIZES(MIFFT_RI) }
#undef MIFFT_RI
    return hipErrorInvalidValue;
}

}  // namespace mifft

using namespace mifft;

extern "C" {

mifft_status mifft_forward_real_device(const mifft_plan* plan, const float* x, float* y, int64_t batch,
                                       int half, void* stream) {
    if (!plan) return fail(MIFFT_ERR_INVALID_ARGUMENT, "null plan");
    if (batch < 0) return fail(MIFFT_ERR_INVALID_ARGUMENT, "negative batch");
    if (half != 0 && half != 1) return fail(MIFFT_ERR_INVALID_ARGUMENT, "half must be 0 or 1");
    if (batch == 0) return MIFFT_OK;
    if (!x || !y) return fail(MIFFT_ERR_INVALID_ARGUMENT, "null buffer");
    if (!plan->real_tw) return fail(MIFFT_ERR_KERNEL_NOT_FOUND, "HIP kernel not found: no real-input table");
    const int64_t n = plan->n;
    RealFrames a{x, batch * n, n, batch, batch, nullptr, reinterpret_cast<float2*>(y), nullptr, nullptr,
                 0.f, half ? n / 2 + 1 : n, plan->real_tw};
    return check_hip(real_frames_forward((int)n, a, static_cast<hipStream_t>(stream)), "real FFT");
}

mifft_status mifft_inverse_real_device(const mifft_plan* plan, const float* y, int half, float* x,
                                       int64_t batch, void* stream) {
    if (!plan) return fail(MIFFT_ERR_INVALID_ARGUMENT, "null plan");
    if (batch < 0) return fail(MIFFT_ERR_INVALID_ARGUMENT, "negative batch");
    if (half != 0 && half != 1) return fail(MIFFT_ERR_INVALID_ARGUMENT, "half must be 0 or 1");
    if (batch == 0) return MIFFT_OK;
    if (!x || !y) return fail(MIFFT_ERR_INVALID_ARGUMENT, "null buffer");
    if (!plan->real_tw) return fail(MIFFT_ERR_KERNEL_NOT_FOUND, "HIP kernel not found: no real-input table");
    const int64_t n = plan->n;
    return check_hip(real_inverse((int)n, reinterpret_cast<const float2*>(y), half ? n / 2 + 1 : n, x, batch,
                                  plan->real_tw, static_cast<hipStream_t>(stream)),
                     "real inverse FFT");
}

}  // extern "C"
