// LDS-resident Stockham autosort FFT for N = 64 ... 16384 on gfx950.
//
// One workgroup holds TPW whole transforms in LDS (16384 complex64 plus
// padding is 136 KiB, inside gfx950's 160 KiB per workgroup).  Each transform
// is worked by T = N/E threads that keep E complex values in registers:
//
//   pass 0   : coalesced global load in[t + T*m]  -> DFT-R      -> LDS
//   pass p   : LDS read  buf[t + T*m] -> twiddle -> DFT-R       -> LDS
//   last     : LDS read               -> twiddle -> DFT-R -> coalesced
//              global store out[b + j*N/R]
//
// Butterfly b of a radix-R pass at stride s reads x_j = a[b + j*N/R],
// multiplies by W_N^{j*pos*N/(R*s)} (pos = b mod s), runs a DFT-R and writes
// y_j to a'[(b/s)*R*s + pos + j*s]: the self-sorting Stockham DIT pass used by
// the reference's radix-4 and radix-8 kernels
// (Sources/MetalFFT/Resources/fft_multisize.metal:92-111,
//  Sources/MetalFFT/Resources/fft_4096_batched.metal:108-203), here with
// E in {4, 8, 16} so the same template yields the radix-4 family, the radix-8
// batched form and a radix-16 form, all with twiddles read from one FP32 table
// generated in FP64 instead of per-butterfly sincos.
//
// The inverse follows MetalFFT.inverse (Sources/MetalFFT/FFT/MetalFFT.swift:56-65):
// conjugate on load, forward transform, conjugate and multiply by 1/N on store.
#include "fft_device.hpp"
#include "kernels.hpp"

namespace mifft {

template <int N, int E, int P, bool INV>
__device__ __forceinline__ void stockham_pass(float2 (&v)[E], float2* buf, int t,
                                              const float2* __restrict__ tw,
                                              float2* __restrict__ dst, bool active) {
    using S = Schedule<N, E>;
    constexpr int R = S::radix(P);
    constexpr int SD = S::stride(P);
    constexpr int B = E / R;  // butterflies per thread in this pass
    constexpr int T = S::T;
    constexpr int TWS = N / (R * SD);
    constexpr bool LAST = (P == S::passes - 1);

#pragma unroll
    for (int i = 0; i < B; ++i) {
        float2 x[R];
#pragma unroll
        for (int j = 0; j < R; ++j) x[j] = v[i + B * j];
        if constexpr (SD > 1) {
            const int pos = (t + i * T) & (SD - 1);
#pragma unroll
            for (int j = 1; j < R; ++j) x[j] = cmul(x[j], tw[j * pos * TWS]);
        }
        dft<R>(x);
#pragma unroll
        for (int j = 0; j < R; ++j) v[i + B * j] = x[j];
    }

    if constexpr (LAST) {
        // R * SD == N, so grp == 0 and y_j goes to b + j*SD: coalesced.
        if (active) {
#pragma unroll
            for (int i = 0; i < B; ++i) {
#pragma unroll
                for (int j = 0; j < R; ++j) {
                    float2 y = v[i + B * j];
                    if constexpr (INV) {
                        constexpr float inv_n = 1.0f / (float)N;
                        y = make_float2(y.x * inv_n, -y.y * inv_n);
                    }
                    dst[t + i * T + j * SD] = y;
                }
            }
        }
    } else {
        if constexpr (P > 0) __syncthreads();  // all reads of this pass done
#pragma unroll
        for (int i = 0; i < B; ++i) {
            const int b = t + i * T;
            const int pos = b & (SD - 1);
            const int base = (b / SD) * (R * SD) + pos;
#pragma unroll
            for (int j = 0; j < R; ++j) buf[lds_idx(base + j * SD)] = v[i + B * j];
        }
        __syncthreads();
#pragma unroll
        for (int m = 0; m < E; ++m) v[m] = buf[lds_idx(t + T * m)];
        stockham_pass<N, E, P + 1, INV>(v, buf, t, tw, dst, active);
    }
}

template <int N, int E, int TPW, bool INV>
__global__ void __launch_bounds__((N / E) * TPW)
k_stockham(const float2* __restrict__ in, float2* __restrict__ out,
           const float2* __restrict__ tw, int64_t batch) {
    constexpr int T = Schedule<N, E>::T;
    __shared__ float2 lds[TPW * lds_len(N)];
    const int local = threadIdx.x / T;
    const int t = threadIdx.x % T;
    const int64_t tr = (int64_t)blockIdx.x * TPW + local;
    const bool active = tr < batch;
    float2* buf = lds + local * lds_len(N);
    const float2* src = in + tr * N;
    float2* dst = out + tr * N;

    float2 v[E];
    if (active) {
#pragma unroll
        for (int m = 0; m < E; ++m) {
            float2 x = src[t + T * m];
            v[m] = INV ? conjf2(x) : x;
        }
    } else {
#pragma unroll
        for (int m = 0; m < E; ++m) v[m] = make_float2(0.f, 0.f);
    }
    stockham_pass<N, E, 0, INV>(v, buf, t, tw, dst, active);
}

// Transforms per workgroup: pack small transforms up to 256 threads.
template <int N, int E>
constexpr int tpw_for() {
    constexpr int T = N / E;
    return T >= 256 ? 1 : 256 / T;
}

template <int N, int E>
static hipError_t launch_stockham(const float2* in, float2* out, const float2* tw, int64_t batch,
                                  bool inverse, hipStream_t stream) {
    constexpr int TPW = tpw_for<N, E>();
    constexpr int threads = (N / E) * TPW;
    static_assert(threads <= 1024, "workgroup too large");
    static_assert(TPW * lds_len(N) * 8 <= 160 * 1024, "LDS over 160 KiB");
    if (batch <= 0) return hipSuccess;
    const int64_t blocks = (batch + TPW - 1) / TPW;
    if (blocks > 0x7fffffffLL) return hipErrorInvalidValue;
    if (inverse)
        hipLaunchKernelGGL((k_stockham<N, E, TPW, true>), dim3((unsigned)blocks), dim3(threads), 0,
                           stream, in, out, tw, batch);
    else
        hipLaunchKernelGGL((k_stockham<N, E, TPW, false>), dim3((unsigned)blocks), dim3(threads), 0,
                           stream, in, out, tw, batch);
    return hipGetLastError();
}

#define MIFFT_V(N, E) {N, E, tpw_for<N, E>(), &launch_stockham<N, E>, "stockham_r" #E "_" #N}

static const StockhamVariant kVariants[] = {
    MIFFT_V(64, 4),    MIFFT_V(64, 8),    MIFFT_V(64, 16),
    MIFFT_V(128, 4),   MIFFT_V(128, 8),   MIFFT_V(128, 16),
    MIFFT_V(256, 4),   MIFFT_V(256, 8),   MIFFT_V(256, 16),
    MIFFT_V(512, 4),   MIFFT_V(512, 8),   MIFFT_V(512, 16),
    MIFFT_V(1024, 4),  MIFFT_V(1024, 8),  MIFFT_V(1024, 16),
    MIFFT_V(2048, 4),  MIFFT_V(2048, 8),  MIFFT_V(2048, 16),
    MIFFT_V(4096, 4),  MIFFT_V(4096, 8),  MIFFT_V(4096, 16),
    MIFFT_V(8192, 8),  MIFFT_V(8192, 16),
    MIFFT_V(16384, 16),
};
#undef MIFFT_V

const StockhamVariant* stockham_variants(int* count) {
    *count = (int)(sizeof(kVariants) / sizeof(kVariants[0]));
    return kVariants;
}

const StockhamVariant* find_stockham(int n, int e) {
    int c = 0;
    const StockhamVariant* v = stockham_variants(&c);
    for (int i = 0; i < c; ++i)
        if (v[i].n == n && v[i].e == e) return &v[i];
    return nullptr;
}

// W_N^k = exp(-2 pi i k / N), k in [0, N), evaluated in FP64 with sincospi
// and rounded once to FP32.
__global__ void k_twiddles(float2* __restrict__ tw, int n) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) {
        double s, c;
        sincospi(-2.0 * (double)k / (double)n, &s, &c);
        tw[k] = make_float2((float)c, (float)s);
    }
}

hipError_t launch_twiddles(float2* tw, int n, hipStream_t stream) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_twiddles, dim3((n + 255) / 256), dim3(256), 0, stream, tw, n);
    return hipGetLastError();
}

}  // namespace mifft
