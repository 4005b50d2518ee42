// Compiled LDS Stockham kernel variants and their launchers.
//
// Variants per length N: E in {4, 8, 16} complex values per thread (radix-4
// family as in Sources/MetalFFT/Resources/fft_multisize.metal, radix-8 as in
// Sources/MetalFFT/Resources/fft_4096_batched.metal, and radix-16), in the
// forms:
//   one-shot      : ceil(batch / TPW) workgroups, padded LDS layout;
//   _swz_nt       : one-shot, XOR-swizzled LDS (exactly 8*N bytes per
//                   transform), non-temporal global I/O;
//   _stream       : persistent grid prefetching the next transform into
//                   registers during the LDS passes;
//   _dma          : persistent grid, next transforms copied HBM -> LDS by
//                   global_load_lds DMA into the other half of a double
//                   buffer, twiddles held in registers (N >= 128);
//   _swz_nt_pt    : _swz_nt whose threads load their twiddle table entries
//                   for every pass right after the transform's data, as
//                   16-byte reads of packed per-pass rows (PassTable), so no
//                   pass waits on an L2 round trip for them;
//   _stream_pt    : _stream with those entries held in registers for the
//                   whole loop.
// All read twiddles as W^k, W^2k, W^4k, W^8k and derive the other powers.
// The autotuner picks one per size (autotune.hip).
#include "kernels.hpp"
#include "stockham_core.hpp"

namespace mifft {

namespace {
int device_cu_count() {
    static int cus = [] {
        int dev = 0, n = 0;
        if (hipGetDevice(&dev) != hipSuccess) return 256;
        if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
            return 256;
        return n > 0 ? n : 256;
    }();
    return cus;
}
}  // namespace

template <int N, int E>
static hipError_t launch_dma(const float2* in, float2* out, const float2* tw, int64_t batch,
                             bool inverse, hipStream_t stream) {
    constexpr int TPW = tpw_for<N, E>();
    constexpr int threads = (N / E) * TPW;
    static_assert(2 * TPW * N * 8 <= 160 * 1024, "LDS over 160 KiB");
    if (batch <= 0) return hipSuccess;
    static int per_cu = [] {
        int nb = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
                &nb, reinterpret_cast<const void*>(&k_stockham_dma<N, E, TPW, false>), threads, 0) !=
                hipSuccess || nb <= 0)
            nb = 1;
        return nb;
    }();
    int64_t blocks = (batch + TPW - 1) / TPW;
    const int64_t resident = (int64_t)per_cu * device_cu_count();
    if (blocks > resident) blocks = resident;
    if (inverse)
        hipLaunchKernelGGL((k_stockham_dma<N, E, TPW, true>), dim3((unsigned)blocks), dim3(threads), 0,
                           stream, in, out, tw, batch);
    else
        hipLaunchKernelGGL((k_stockham_dma<N, E, TPW, false>), dim3((unsigned)blocks), dim3(threads), 0,
                           stream, in, out, tw, batch);
    return hipGetLastError();
}

template <int N, int E, class L, bool NT, bool PERS, int TWM = 0>
static hipError_t launch_stockham(const float2* in, float2* out, const float2* tw, int64_t batch,
                                  bool inverse, hipStream_t stream) {
    constexpr int TPW = tpw_for<N, E>();
    constexpr int threads = (N / E) * TPW;
    static_assert(threads <= 1024, "workgroup too large");
    static_assert(TPW * L::len(N) * 8 <= 160 * 1024, "LDS over 160 KiB");
    if (batch <= 0) return hipSuccess;
    int64_t blocks = (batch + TPW - 1) / TPW;
    if constexpr (PERS) {
        // resident workgroups per CU, from the occupancy calculator (once)
        static int per_cu = [] {
            int nb = 0;
            if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
                    &nb,
                    reinterpret_cast<const void*>(
                        &k_stockham_stream<N, E, TPW, false, L, NT, true, 0, TWM>),
                    threads, 0) != hipSuccess || nb <= 0)
                nb = 1;
            return nb;
        }();
        const int64_t resident = (int64_t)per_cu * device_cu_count();
        if (blocks > resident) blocks = resident;
        if (inverse)
            hipLaunchKernelGGL((k_stockham_stream<N, E, TPW, true, L, NT, true, 0, TWM>),
                               dim3((unsigned)blocks), dim3(threads), 0, stream, in, out, tw, batch);
        else
            hipLaunchKernelGGL((k_stockham_stream<N, E, TPW, false, L, NT, true, 0, TWM>),
                               dim3((unsigned)blocks), dim3(threads), 0, stream, in, out, tw, batch);
    } else {
        if (blocks > 0x7fffffffLL) return hipErrorInvalidValue;
        if (inverse)
            hipLaunchKernelGGL((k_stockham<N, E, TPW, true, L, NT, true, TWM>), dim3((unsigned)blocks),
                               dim3(threads), 0, stream, in, out, tw, batch);
        else
            hipLaunchKernelGGL((k_stockham<N, E, TPW, false, L, NT, true, TWM>),
                               dim3((unsigned)blocks), dim3(threads), 0, stream, in, out, tw, batch);
    }
    return hipGetLastError();
}

// Half-LDS exchange form (k_stockham_half): one transform per workgroup.
template <int N, int WPE>
static hipError_t launch_half(const float2* in, float2* out, const float2* tw, int64_t batch,
                              bool inverse, hipStream_t stream) {
    if (batch <= 0) return hipSuccess;
    if (batch > 0x7fffffffLL) return hipErrorInvalidValue;
    if (inverse)
        hipLaunchKernelGGL((k_stockham_half<N, true, true, WPE>), dim3((unsigned)batch),
                           dim3(N / 32), 0, stream, in, out, tw, batch);
    else
        hipLaunchKernelGGL((k_stockham_half<N, false, true, WPE>), dim3((unsigned)batch),
                           dim3(N / 32), 0, stream, in, out, tw, batch);
    return hipGetLastError();
}

// Half-LDS exchange with 16 values per thread (k_stockham_half16).
template <int N, int WPE>
static hipError_t launch_half16(const float2* in, float2* out, const float2* tw, int64_t batch,
                                bool inverse, hipStream_t stream) {
    if (batch <= 0) return hipSuccess;
    if (batch > 0x7fffffffLL) return hipErrorInvalidValue;
    if (inverse)
        hipLaunchKernelGGL((k_stockham_half16<N, true, true, WPE>), dim3((unsigned)batch),
                           dim3(N / 16), 0, stream, in, out, tw, batch);
    else
        hipLaunchKernelGGL((k_stockham_half16<N, false, true, WPE>), dim3((unsigned)batch),
                           dim3(N / 16), 0, stream, in, out, tw, batch);
    return hipGetLastError();
}

#define MIFFT_V(N, E) \
    {N, E, tpw_for<N, E>(), &launch_stockham<N, E, LdsPad, false, false>, "stockham_r" #E "_" #N}
#define MIFFT_VS(N, E)                                                                          \
    {N, E, tpw_for<N, E>(), &launch_stockham<N, E, LdsSwz, true, false>,                        \
     "stockham_r" #E "_" #N "_swz_nt"}
#define MIFFT_VP(N, E)                                                                          \
    {N, E, tpw_for<N, E>(), &launch_stockham<N, E, LdsSwz, true, true>,                         \
     "stockham_r" #E "_" #N "_stream"}
// packed bases rows (PassTable<N, 16, 0>, after the N-entry table) loaded
// with the transform's data
#define MIFFT_VT(N)                                                                             \
    {N, 16, tpw_for<N, 16>(), &launch_stockham<N, 16, LdsSwz, true, false, 2>,                  \
     "stockham_r16_" #N "_swz_nt_pt"}
// _stream with the packed rows in registers for the whole loop
#define MIFFT_VPT(N)                                                                            \
    {N, 16, tpw_for<N, 16>(), &launch_stockham<N, 16, LdsSwz, true, true, 2>,                   \
     "stockham_r16_" #N "_stream_pt"}
// _stream with the packed bases rows read at each pass (nothing held
// across the loop, which leaves room for the prefetched transform at 16384:
// 61.0 us against 62.7 us for _swz_nt_pt per 1024 transforms on one box)
#define MIFFT_VP3(N)                                                                            \
    {N, 16, tpw_for<N, 16>(), &launch_stockham<N, 16, LdsSwz, true, true, 3>,                   \
     "stockham_r16_" #N "_stream_p3"}
// half-LDS exchange, E = 32, at least WPE waves per SIMD
#define MIFFT_VH(N, WPE)                                                                        \
    {N, 32, 1, &launch_half<N, WPE>, "stockham_r16_" #N "_half" #WPE}
// half-LDS exchange, E = 16, at least WPE waves per SIMD
#define MIFFT_VH16(N, WPE)                                                                      \
    {N, 16, 1, &launch_half16<N, WPE>, "stockham_r16_" #N "_half16w" #WPE}
#define MIFFT_VD(N, E) \
    {N, E, tpw_for<N, E>(), &launch_dma<N, E>, "stockham_r" #E "_" #N "_dma"}

static const StockhamVariant kVariants[] = {
    MIFFT_V(64, 4),    MIFFT_V(64, 8),    MIFFT_V(64, 16),    MIFFT_VS(64, 8),    MIFFT_VP(64, 8),
    MIFFT_VP(64, 16),
    MIFFT_V(128, 4),   MIFFT_V(128, 8),   MIFFT_V(128, 16),   MIFFT_VS(128, 16),  MIFFT_VP(128, 8),
    MIFFT_VP(128, 16), MIFFT_VD(128, 8),  MIFFT_VD(128, 16),
    MIFFT_V(256, 4),   MIFFT_V(256, 8),   MIFFT_V(256, 16),   MIFFT_VS(256, 16),  MIFFT_VP(256, 8),
    MIFFT_VP(256, 16), MIFFT_VD(256, 8),  MIFFT_VD(256, 16),
    MIFFT_V(512, 4),   MIFFT_V(512, 8),   MIFFT_V(512, 16),   MIFFT_VS(512, 16),  MIFFT_VP(512, 8),
    MIFFT_VP(512, 16), MIFFT_VD(512, 8),  MIFFT_VD(512, 16),
    MIFFT_V(1024, 4),  MIFFT_V(1024, 8),  MIFFT_V(1024, 16),  MIFFT_VS(1024, 16), MIFFT_VP(1024, 8),
    MIFFT_VP(1024, 16), MIFFT_VD(1024, 8), MIFFT_VD(1024, 16),
    MIFFT_V(2048, 4),  MIFFT_V(2048, 8),  MIFFT_V(2048, 16),  MIFFT_VS(2048, 16), MIFFT_VP(2048, 8),
    MIFFT_VP(2048, 16), MIFFT_VD(2048, 8), MIFFT_VD(2048, 16),
    MIFFT_V(4096, 4),  MIFFT_V(4096, 8),  MIFFT_V(4096, 16),  MIFFT_VS(4096, 16), MIFFT_VP(4096, 8),
    MIFFT_VP(4096, 16), MIFFT_VD(4096, 8), MIFFT_VD(4096, 16),
    MIFFT_V(8192, 8),  MIFFT_V(8192, 16), MIFFT_VS(8192, 16), MIFFT_VP(8192, 8),  MIFFT_VP(8192, 16),
    MIFFT_VD(8192, 8), MIFFT_VD(8192, 16),
    MIFFT_V(16384, 16), MIFFT_VS(16384, 16), MIFFT_VP(16384, 16),
    MIFFT_VT(64),      MIFFT_VT(128),     MIFFT_VT(256),      MIFFT_VT(512),      MIFFT_VT(1024),
    // not at 8192: there it ties with the half-LDS form over the tuner's
    // short bursts of launches and runs 6 % slower over a sustained stream
    // (52.7-53.0 against 49.4-50.0 us per 128 MiB), so it was sometimes chosen
    MIFFT_VT(2048),    MIFFT_VT(4096),    MIFFT_VT(16384),
    MIFFT_VPT(1024),   MIFFT_VPT(2048),   MIFFT_VPT(4096),    MIFFT_VPT(8192),
    MIFFT_VP3(16384),
    MIFFT_VH(2048, 4),  MIFFT_VH(4096, 4), MIFFT_VH(8192, 4),
    MIFFT_VH(16384, 4),
    // 32768 in one workgroup: 1024 threads x 32 values, 128 KiB of LDS
    // exchanged in halves (one transform per CU); the four-step kernels are
    // the other candidates for this length
    MIFFT_VH(32768, 4),
    MIFFT_VH16(8192, 8), MIFFT_VH16(16384, 8),
};
#undef MIFFT_V
#undef MIFFT_VS
#undef MIFFT_VP
#undef MIFFT_VD
#undef MIFFT_VT
#undef MIFFT_VPT
#undef MIFFT_VP3
#undef MIFFT_VH
#undef MIFFT_VH16

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

namespace {
template <int N>
constexpr int pass_rows() { return PassTable<N, 16, 0>::size; }

template <int N>
hipError_t launch_pass_rows(float2* tab, hipStream_t stream) {
    constexpr int size = pass_rows<N>();
    if constexpr (size > 0)
        hipLaunchKernelGGL((k_pass_table<N, 16, 0>), dim3((size + 255) / 256), dim3(256), 0, stream,
                           tab);
    return hipGetLastError();
}
}  // namespace

int twiddle_table_len(int n) {
    switch (n) {
        case 64: return 64 + pass_rows<64>();
        case 128: return 128 + pass_rows<128>();
        case 256: return 256 + pass_rows<256>();
        case 512: return 512 + pass_rows<512>();
        case 1024: return 1024 + pass_rows<1024>();
        case 2048: return 2048 + pass_rows<2048>();
        case 4096: return 4096 + pass_rows<4096>();
        case 8192: return 8192 + pass_rows<8192>();
        case 16384: return 16384 + pass_rows<16384>();
        case 32768: return 32768 + pass_rows<32768>();
        default: return n;
    }
}

hipError_t launch_twiddles(float2* tw, int n, hipStream_t stream) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_twiddles, dim3((n + 255) / 256), dim3(256), 0, stream, tw, n);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    switch (n) {
        case 64: return launch_pass_rows<64>(tw + n, stream);
        case 128: return launch_pass_rows<128>(tw + n, stream);
        case 256: return launch_pass_rows<256>(tw + n, stream);
        case 512: return launch_pass_rows<512>(tw + n, stream);
        case 1024: return launch_pass_rows<1024>(tw + n, stream);
        case 2048: return launch_pass_rows<2048>(tw + n, stream);
        case 4096: return launch_pass_rows<4096>(tw + n, stream);
        case 8192: return launch_pass_rows<8192>(tw + n, stream);
        case 16384: return launch_pass_rows<16384>(tw + n, stream);
        case 32768: return launch_pass_rows<32768>(tw + n, stream);
        default: return hipSuccess;
    }
}

}  // namespace mifft
