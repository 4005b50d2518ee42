// SAR range-Doppler focusing (range compression -> corner turn -> azimuth
// FFT / matched filter / IFFT) on complex-FP16 or complex64 data.
//
// Reference: the fused RDA pipeline (src/radar/rda_fused_pipeline.swift:131-369)
// with kernels fused_range_compression / fused_multiply_ifft
// (src/fft_sar_fused.metal:246-293, :364-410), transpose_2d and
// generate_azimuth_matched_filter (src/radar/rda_kernels.metal:76-91, :195-238)
// and the FP16 variants (src/fft_sar_fused_fp16.metal), run as one
// synchronous command buffer per step with device data in FP32.
//
// Here, three launches over data that stays in HBM (complex FP16 halves the
// bytes of every sweep; arithmetic is FP32 in registers and LDS):
//   1. range compression, one workgroup per pulse: FFT_Nr -> x conj(FFT(chirp))
//      -> IFFT_Nr, times 2^-sr;
//   2. corner turn (Na x Nr -> Nr x Na) through 64 x 64 LDS tiles, coalesced
//      on both sides;
//   3. azimuth compression, one workgroup per range bin: FFT_Na -> x H_a ->
//      IFFT_Na, times 2^-sa, with
//        H_a(f_a, R0) = exp(+i 4 pi R0 / lambda sqrt(1 - (lambda f_a / 2V)^2))
//      (zero where the ratio reaches 0.95, as the reference), R0 the slant
//      range of the bin: the conjugate of the echo's Doppler phase.  The
//      reference's kernel writes exp(-i ...) (rda_kernels.metal:231-232),
//      which equals the echo spectrum itself and leaves targets unfocused;
//      that sign is available as MIFFT_RDA_REFERENCE_FORMULA.  The phase, ~8e6 rad at 20 km in X band, is formed in
//      FP64 as cycles A(k) + m B D(k) (D = the square root, m = bin - Nr/2)
//      and reduced mod 1 before the FP32 sincos; the reference evaluates it in
//      FP32, where 8e6 rad carries ~0.5 rad of rounding.
// The image comes out range-bin-major (Nr x Na), the layout after the corner
// turn (the reference transposes back with a fourth naive transpose).
#include <hip/hip_runtime.h>

#include <cmath>
#include <string>
#include <vector>

#include "convolve.hpp"
#include "fp16_core.hpp"
#include "runtime.hpp"
#include "sar.hpp"
#include "stockham_core.hpp"

namespace mifft {

namespace {

using fp16::h2;

template <int IO>
struct SarIo;
template <>
struct SarIo<MIFFT_IO_C64> {
    using T = float2;
    static __device__ __forceinline__ float2 cvt(T v) { return v; }
    static __device__ __forceinline__ float2 ld(const T* p) { return *p; }
    static __device__ __forceinline__ void st(T* p, float2 v) { *p = v; }
};
template <>
struct SarIo<MIFFT_IO_C32H> {
    using T = h2;
    static __device__ __forceinline__ float2 cvt(T v) { return make_float2((float)v.x, (float)v.y); }
    static __device__ __forceinline__ float2 ld(const T* p) { return cvt(*p); }
    static __device__ __forceinline__ void st(T* p, float2 v) {
        h2 o = {(_Float16)v.x, (_Float16)v.y};
        *p = o;
    }
};

struct RangeFilter {
    const float2* H;  // conj(FFT(chirp reference)), n entries
    __device__ __forceinline__ float2 at(int64_t, int k) const { return H[k]; }
};

struct AzimuthFilter {
    const double* A;  // frac(2 R0c D(k) / lambda), cycles
    const double* D;  // sqrt(1 - ratio^2), or -1 where the filter is zero
    double B;         // 2 * range pixel / lambda, cycles per bin per unit D
    int64_t half_nr;  // Nr / 2
    float sign;       // +1: exp(+i 2 pi ph) (matched), -1: exp(-i 2 pi ph)
    __device__ __forceinline__ float2 at(int64_t row, int k) const {
        const double d = D[k];
        if (d < 0.0) return make_float2(0.f, 0.f);
        double ph = A[k] + (double)(row - half_nr) * B * d;
        ph -= rint(ph);
        float s, c;
        sincospif(2.0f * (float)ph, &s, &c);
        return make_float2(c, sign * s);
    }
};

// The same filter from two tables built once per processor: the phase is
// linear in the row, so with row = 64 rh + rl
//   H[row][k] = E1[rh][k] E2[rl][k],
//   E1[rh][k] = exp(i s 2 pi (A[k] + (64 rh - Nr/2) B D[k])),
//   E2[rl][k] = exp(i s 2 pi rl B D[k]),
// each entry's phase in FP64 with an FP64 sincospi, rounded once to FP32
// (k_az_factors), zero where the filter is.  Two reads that stay in L2 (64
// rows of E2 and the current rows of E1: 4 MiB at 4096 Doppler bins) and one
// complex product per element, instead of an FP64 phase and an FP32 sincos,
// or of a read of an Nr x Na table from HBM (128 MiB at 4096 x 4096, as much
// as the frame itself).
constexpr int kAzRowBits = 6;
struct AzimuthTable {
    const float2* E1;
    const float2* E2;
    int64_t n;  // Doppler bins per row
    __device__ __forceinline__ float2 at(int64_t row, int k) const {
        return cmul(E1[(row >> kAzRowBits) * n + k], E2[(row & ((1 << kAzRowBits) - 1)) * n + k]);
    }
};

// rows_hi rows of E1, then 64 rows of E2.
__global__ void k_az_factors(const double* __restrict__ A, const double* __restrict__ D, double B, int64_t half_nr,
                             double sign, int64_t rows_hi, int64_t n, float2* __restrict__ E) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (rows_hi + (1 << kAzRowBits)) * n) return;
    const int64_t r = i / n, k = i % n;
    const double d = D[k];
    float2 h = make_float2(0.f, 0.f);
    if (d >= 0.0) {
        double ph = r < rows_hi ? A[k] + (double)((r << kAzRowBits) - half_nr) * B * d
                                : (double)(r - rows_hi) * B * d;
        ph -= rint(ph);
        double sn, cs;
        sincospi(2.0 * ph, &sn, &cs);
        h = make_float2((float)cs, (float)(sign * sn));
    }
    E[i] = h;
}

int sar_cu_count() {
    static int cus = [] {
        int dev = 0, n = 0;
        if (hipGetDevice(&dev) != hipSuccess) return 256;
        if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 256;
        return n > 0 ? n : 256;
    }();
    return cus;
}

// Row kernel form at N = 4096: persistent with the next row prefetched
// (default), or one workgroup per row (MIFFT_SAR_ROWS=oneshot).
bool sar_rows_stream() {
    static bool on = [] {
        const char* e = std::getenv("MIFFT_SAR_ROWS");
        return !(e && std::string(e) == "oneshot");
    }();
    return on;
}

template <int N>
constexpr int rows_e() { return N >= 256 ? 16 : 8; }
template <int N>
constexpr int rows_threads() { return (N / rows_e<N>()) * tpw_for<N, rows_e<N>()>(); }

// RCMC on the range-Doppler frame laid out rows = range bins, columns =
// Doppler bins (the layout after the corner turn and azimuth FFT):
//   out[r][f] = sum_{k=0..7} w[f][k] in[r + si[f] + k - 3][f],
// taps outside the frame dropped (rda_kernels.metal:150-180).  The shift is
// the same for every range bin, so si and the 8 windowed-sinc weights per
// Doppler bin are formed once on the host in FP64.  A thread computes
// kRcmcRows consecutive range bins of one column from a window of
// kRcmcRows + 7 input rows held in registers (every input read once per
// thread instead of once per tap); a workgroup is 64 Doppler columns x 4
// such runs, so every load is one coalesced row segment.
constexpr int kRcmcRows = 16;

template <int IO>
__global__ void __launch_bounds__(256)
k_rcmc(const typename SarIo<IO>::T* __restrict__ in, typename SarIo<IO>::T* __restrict__ out,
       const int* __restrict__ si, const float4* __restrict__ w, int64_t nr, int64_t na) {
    const int64_t f = (int64_t)blockIdx.x * 64 + (threadIdx.x & 63);
    const int64_t r0 = ((int64_t)blockIdx.y * 4 + (threadIdx.x >> 6)) * kRcmcRows;
    if (f >= na || r0 >= nr) return;
    const int s = si[f];
    const float4 wa = w[2 * f], wb = w[2 * f + 1];
    const float wk[8] = {wa.x, wa.y, wa.z, wa.w, wb.x, wb.y, wb.z, wb.w};
    float2 x[kRcmcRows + 7];
#pragma unroll
    for (int q = 0; q < kRcmcRows + 7; ++q) {
        const int64_t rr = r0 + s + q - 3;
        const bool ok = rr >= 0 && rr < nr;
        const float2 v = SarIo<IO>::ld(in + (rr < 0 ? 0 : (rr >= nr ? nr - 1 : rr)) * na + f);
        uint32_t m = ok ? 0xffffffffu : 0u;  // unconditional load, opaque mask
        __asm__("" : "+v"(m));
        x[q] = make_float2(__uint_as_float(__float_as_uint(v.x) & m), __uint_as_float(__float_as_uint(v.y) & m));
    }
#pragma unroll
    for (int j = 0; j < kRcmcRows; ++j) {
        float2 acc = make_float2(0.f, 0.f);
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            acc.x += wk[k] * x[j + k].x;
            acc.y += wk[k] * x[j + k].y;
        }
        if (r0 + j < nr) SarIo<IO>::st(out + (r0 + j) * na + f, acc);
    }
}

enum { kRowsCompress = 0, kRowsForward = 1, kRowsFilterInverse = 2 };

// One workgroup per row (TPW rows for small N):
//   kRowsCompress       FFT -> x filter -> IFFT, times scale;
//   kRowsForward        FFT, times scale (the azimuth FFT before RCMC);
//   kRowsFilterInverse  x filter -> IFFT, times scale (after RCMC).
// A row is read whole before it is written: in == out is allowed.
template <int N, int IO, class F, int MODE>
__global__ void __launch_bounds__(rows_threads<N>())
k_rows(const typename SarIo<IO>::T* in, typename SarIo<IO>::T* out, const float2* __restrict__ tw,
       F filt, int64_t rows, float scale) {
    constexpr int E = N >= 256 ? 16 : 8;
    constexpr int TPW = tpw_for<N, E>();
    constexpr int T = N / E;
    __shared__ float2 lds[TPW * LdsPad::len(N)];
    const int local = threadIdx.x / T;
    const int t = threadIdx.x % T;
    const int64_t row = (int64_t)blockIdx.x * TPW + local;
    const bool active = row < rows;
    float2* buf = lds + local * LdsPad::len(N);
    float2 v[E];
    RegTwiddles<N, E> rt;  // table entries in registers for both transforms
    rt.load(tw, t);
    {
        const typename SarIo<IO>::T* src = in + (active ? row : 0) * N;
#pragma unroll
        for (int m = 0; m < E; ++m)
            v[m] = active ? SarIo<IO>::ld(src + t + T * m) : make_float2(0.f, 0.f);
        if constexpr (MODE != kRowsFilterInverse) rt.template fft<false>(v, buf, t, tw);
    }
    if constexpr (MODE == kRowsForward) {
        if (!active) return;
        typename SarIo<IO>::T* dst = out + row * N;
        const int ts = fresh(t);
#pragma unroll
        for (int m = 0; m < E; ++m) SarIo<IO>::st(dst + ts + T * m, make_float2(v[m].x * scale, v[m].y * scale));
        return;
    } else {
        {
            const int tf = fresh(t);  // filter index arithmetic not shared with the loads
#pragma unroll
            for (int m = 0; m < E; ++m) v[m] = conjf2(cmul(v[m], filt.at(row, tf + T * m)));
        }
        rt.template fft<MODE == kRowsCompress>(v, buf, t, tw);  // SYNC0: LDS used by the forward
        if (!active) return;
        const float s = scale / (float)N;
        typename SarIo<IO>::T* dst = out + row * N;
        const int ts = fresh(t);
#pragma unroll
        for (int m = 0; m < E; ++m) SarIo<IO>::st(dst + ts + T * m, make_float2(v[m].x * s, -v[m].y * s));
    }
}

// Persistent form of k_rows for one row per 256-thread workgroup (E = 16,
// N = 4096): the grid is the resident workgroups, each loops over rows
// blockIdx.x, + gridDim.x, ... and issues the loads of its next row before
// the transforms of the current one, so a row's HBM read overlaps the
// previous row's arithmetic instead of every resident workgroup loading at
// once and then computing at once.  The prefetched row is held as loaded
// (a complex-FP16 row is 16 VGPRs per thread).  The next row is always
// loaded (the current one again past the end: in bounds, never used), so no
// load sits under a branch.  Rows are read whole before they are written
// and a workgroup's prefetch is of another row than any being written, so
// in == out is allowed as in k_rows.
template <int N, int IO, class F, int MODE>
__global__ void __launch_bounds__(N / 16)
k_rows_stream(const typename SarIo<IO>::T* in, typename SarIo<IO>::T* out, const float2* __restrict__ tw,
              F filt, int64_t rows, float scale) {
    constexpr int E = 16;
    constexpr int T = N / E;
    static_assert(tpw_for<N, E>() == 1, "one row per workgroup");
    using Tp = typename SarIo<IO>::T;
    __shared__ float2 lds[LdsPad::len(N)];
    const int t = threadIdx.x;
    const int64_t step = gridDim.x;  // the launch keeps gridDim.x <= rows
    RegTwiddles<N, E> rt;
    rt.load(tw, t);
    Tp nxt[E];
    {
        const Tp* src = in + (int64_t)blockIdx.x * N;
#pragma unroll
        for (int m = 0; m < E; ++m) nxt[m] = src[t + T * m];
    }
    // the trip count depends on blockIdx only: barriers stay uniform
    for (int64_t row = blockIdx.x; row < rows; row += step) {
        const int tl = fresh(t);  // per-iteration index arithmetic (not hoisted)
        float2 v[E];
#pragma unroll
        for (int m = 0; m < E; ++m) v[m] = SarIo<IO>::cvt(nxt[m]);
        {
            const int64_t rn = row + step < rows ? row + step : row;
            const Tp* src = in + rn * N;
#pragma unroll
            for (int m = 0; m < E; ++m) nxt[m] = src[tl + T * m];
        }
        // SYNC0 on every transform: the previous row's passes used the LDS
        if constexpr (MODE != kRowsFilterInverse) rt.template fft<true>(v, lds, tl, tw);
        Tp* dst = out + row * N;
        if constexpr (MODE == kRowsForward) {
            const int ts = fresh(t);
#pragma unroll
            for (int m = 0; m < E; ++m)
                SarIo<IO>::st(dst + ts + T * m, make_float2(v[m].x * scale, v[m].y * scale));
        } else {
            {
                const int tf = fresh(t);
#pragma unroll
                for (int m = 0; m < E; ++m) v[m] = conjf2(cmul(v[m], filt.at(row, tf + T * m)));
            }
            rt.template fft<true>(v, lds, fresh(t), tw);
            const float s = scale / (float)N;
            const int ts = fresh(t);
#pragma unroll
            for (int m = 0; m < E; ++m) SarIo<IO>::st(dst + ts + T * m, make_float2(v[m].x * s, -v[m].y * s));
        }
    }
}

// FP16 precision modes of the rows (src/fft_sar_fused_fp16.metal: fused range
// compression :250-806, fused azimuth compression :807-941, multiply + IFFT
// :942-1059, each in modes pure / storage / mixed; N = 4096 as there), one
// row per 256-thread workgroup, the mode arithmetic of the fused convolver:
//   kRowsCompress       FFT (every pass / 16, so the spectrum is X / N and
//                       FP16 does not overflow) -> x filter -> IFFT;
//   kRowsFilterInverse  x filter -> IFFT (input: a stored spectrum).
// The inverse is not normalised; the output is conj(.) times scale
// (kRowsCompress) or scale / N (kRowsFilterInverse), powers of two, so
// the result equals the FP32 rows' up to FP16 rounding.  The reference
// divides by N only after the inverse and overflows FP16 for radar data
// (its own "risk: overflow" note, fft_sar_fused_fp16.metal:18).
template <int PM, int IO, class F, int MODE>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4, 8)))
k_rows_h(const typename SarIo<IO>::T* in, typename SarIo<IO>::T* out, const float2* __restrict__ tw,
         F filt, float scale) {
    using namespace fp16;
    constexpr int N = 4096, T = 256;
    __shared__ h2 lds[LdsPad::len(N)];
    const int t = threadIdx.x;
    const int64_t row = blockIdx.x;  // grid == rows
    ConvReg<PM> v[16];
    const typename SarIo<IO>::T* src = in + row * N;
#pragma unroll
    for (int m = 0; m < 16; ++m) {
        const float2 x = SarIo<IO>::ld(src + t + T * m);
        if constexpr (PM == kConvPure) v[m] = to_h2(x);
        else v[m] = x;
    }
    PreBases<N, 16> pb;
    pb.load_packed(tw, t);
    if constexpr (MODE == kRowsCompress) conv_passes_h<PM, 0, false, true>(v, lds, t, pb.w);
    {
        const int tf = fresh(t);  // filter index arithmetic not shared with the loads
#pragma unroll
        for (int m = 0; m < 16; ++m) {
            const float2 hm = filt.at(row, tf + T * m);
            if constexpr (PM == kConvPure) v[m] = hconj(hmul(v[m], to_h2(hm)));
            else if constexpr (PM == kConvMixed) v[m] = to_f2(hconj(hmul(to_h2(v[m]), to_h2(hm))));
            else v[m] = conjf2(round_h(cmul(round_h(v[m]), hm)));
        }
    }
    pb.opaque();
    conv_passes_h<PM, 0, MODE == kRowsCompress, false>(v, lds, t, pb.w);
    const float s = MODE == kRowsCompress ? scale : scale / (float)N;
    typename SarIo<IO>::T* dst = out + row * N;
    const int ts = fresh(t);
#pragma unroll
    for (int m = 0; m < 16; ++m) {
        float2 r;  // conj(.) * s
        if constexpr (PM == kConvPure) {
            const _Float16 sh = (_Float16)s;
            r = to_f2((h2){v[m].x * sh, -v[m].y * sh});
        } else {
            r = make_float2(v[m].x * s, -v[m].y * s);
        }
        SarIo<IO>::st(dst + ts + T * m, r);
    }
}

template <int IO, class F, int MODE>
hipError_t rows_h_io(int precision, const void* in, void* out, const float2* tw, F f, int64_t rows, float scale,
                     hipStream_t st) {
    using namespace fp16;
    if (rows <= 0) return hipSuccess;
    if (rows > 0x7fffffffLL) return hipErrorInvalidValue;
    using Tp = typename SarIo<IO>::T;
    const Tp* i = static_cast<const Tp*>(in);
    Tp* o = static_cast<Tp*>(out);
    switch (precision) {
        case kConvPure:
            hipLaunchKernelGGL((k_rows_h<kConvPure, IO, F, MODE>), dim3((unsigned)rows), dim3(256), 0, st, i, o, tw,
                               f, scale);
            break;
        case kConvStorage:
            hipLaunchKernelGGL((k_rows_h<kConvStorage, IO, F, MODE>), dim3((unsigned)rows), dim3(256), 0, st, i, o,
                               tw, f, scale);
            break;
        case kConvMixed:
            hipLaunchKernelGGL((k_rows_h<kConvMixed, IO, F, MODE>), dim3((unsigned)rows), dim3(256), 0, st, i, o, tw,
                               f, scale);
            break;
        default:
            return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// Rows of length 4096 in an FP16 mode (precision 1-3).
template <class F, int MODE = kRowsCompress>
hipError_t rows_h(int precision, int io, const void* in, void* out, const float2* tw, F f, int64_t rows,
                  float scale, hipStream_t st) {
    return io == MIFFT_IO_C32H ? rows_h_io<MIFFT_IO_C32H, F, MODE>(precision, in, out, tw, f, rows, scale, st)
                               : rows_h_io<MIFFT_IO_C64, F, MODE>(precision, in, out, tw, f, rows, scale, st);
}

// out[c][r] = in[r][c] for a rows x cols matrix, 64 x 64 tiles via LDS.
template <int IO>
__global__ void __launch_bounds__(256)
k_corner_turn(const typename SarIo<IO>::T* __restrict__ in, typename SarIo<IO>::T* __restrict__ out,
              int64_t rows, int64_t cols) {
    using T = typename SarIo<IO>::T;
    __shared__ T tile[64][65];
    const int64_t r0 = (int64_t)blockIdx.y * 64, c0 = (int64_t)blockIdx.x * 64;
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;  // 64 x 4
#pragma unroll
    for (int i = 0; i < 64; i += 4) {
        const int64_t r = r0 + ty + i, c = c0 + tx;
        if (r < rows && c < cols) tile[ty + i][tx] = in[r * cols + c];
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < 64; i += 4) {
        const int64_t c = c0 + ty + i, r = r0 + tx;  // output row c, column r
        if (r < rows && c < cols) out[c * rows + r] = tile[tx][ty + i];
    }
}

// Complex-FP16 corner turn (4-byte elements) for rows and cols multiples of
// 64 and 16-byte aligned buffers: 16-byte loads and stores (a 16-lane group
// moves one 256-byte tile row), the 64 x 64 tile in LDS with element
// (r, col) at word r*64 + (col ^ 4*((r >> 2) & 15)): the 16-byte chunks a
// row is written in stay whole, and the column reads of the store phase
// (element (4q + i, c) for lanes q < 16, four consecutive c per wave) land
// on 64 distinct banks.
__global__ void __launch_bounds__(256)
k_corner_turn_h16(const uint4* __restrict__ in, uint4* __restrict__ out, int64_t rows, int64_t cols) {
    __shared__ uint32_t tile[64 * 64];
    const int tid = threadIdx.x;
    const int q = tid & 15, rr = tid >> 4;
    const int64_t r0 = (int64_t)blockIdx.y * 64, c0 = (int64_t)blockIdx.x * 64;
    const int64_t cols4 = cols >> 2, rows4 = rows >> 2;
#pragma unroll
    for (int p = 0; p < 4; ++p) {
        const int r = rr + 16 * p;
        const uint4 v = in[(r0 + r) * cols4 + (c0 >> 2) + q];
        reinterpret_cast<uint4*>(tile)[r * 16 + (q ^ ((r >> 2) & 15))] = v;
    }
    __syncthreads();
#pragma unroll
    for (int p = 0; p < 4; ++p) {
        const int c = rr + 16 * p;  // output row c0 + c, columns r0 + 4q .. + 3
        uint4 o;
        o.x = tile[(4 * q + 0) * 64 + (c ^ (4 * q))];
        o.y = tile[(4 * q + 1) * 64 + (c ^ (4 * q))];
        o.z = tile[(4 * q + 2) * 64 + (c ^ (4 * q))];
        o.w = tile[(4 * q + 3) * 64 + (c ^ (4 * q))];
        out[(c0 + c) * rows4 + (r0 >> 2) + q] = o;
    }
}

template <int N, int IO, class F, int MODE>
hipError_t rows_n(const void* in, void* out, const float2* tw, F f, int64_t rows, float scale,
                  hipStream_t st) {
    constexpr int E = N >= 256 ? 16 : 8;
    constexpr int TPW = tpw_for<N, E>();
    const int64_t blocks = (rows + TPW - 1) / TPW;
    if (blocks <= 0) return hipSuccess;
    if constexpr (N == 4096) {
        // persistent form when the rows outnumber the resident workgroups
        static int per_cu = [] {
            int nb = 0;
            if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
                    &nb, reinterpret_cast<const void*>(&k_rows_stream<N, IO, F, MODE>), N / 16, 0) !=
                    hipSuccess || nb <= 0)
                nb = 1;
            return nb;
        }();
        const int64_t resident = (int64_t)per_cu * sar_cu_count();
        if (sar_rows_stream() && rows > resident) {
            hipLaunchKernelGGL((k_rows_stream<N, IO, F, MODE>), dim3((unsigned)resident), dim3(N / 16), 0, st,
                               static_cast<const typename SarIo<IO>::T*>(in),
                               static_cast<typename SarIo<IO>::T*>(out), tw, f, rows, scale);
            return hipGetLastError();
        }
    }
    if (blocks > 0x7fffffffLL) return hipErrorInvalidValue;
    hipLaunchKernelGGL((k_rows<N, IO, F, MODE>), dim3((unsigned)blocks), dim3((N / E) * TPW), 0, st,
                       static_cast<const typename SarIo<IO>::T*>(in),
                       static_cast<typename SarIo<IO>::T*>(out), tw, f, rows, scale);
    return hipGetLastError();
}

template <int IO, class F, int MODE>
hipError_t rows_dispatch(int n, const void* in, void* out, const float2* tw, F f, int64_t rows,
                         float scale, hipStream_t st) {
    switch (n) {
        case 64: return rows_n<64, IO, F, MODE>(in, out, tw, f, rows, scale, st);
        case 128: return rows_n<128, IO, F, MODE>(in, out, tw, f, rows, scale, st);
        case 256: return rows_n<256, IO, F, MODE>(in, out, tw, f, rows, scale, st);
        case 512: return rows_n<512, IO, F, MODE>(in, out, tw, f, rows, scale, st);
        case 1024: return rows_n<1024, IO, F, MODE>(in, out, tw, f, rows, scale, st);
        case 2048: return rows_n<2048, IO, F, MODE>(in, out, tw, f, rows, scale, st);
        case 4096: return rows_n<4096, IO, F, MODE>(in, out, tw, f, rows, scale, st);
        case 8192: return rows_n<8192, IO, F, MODE>(in, out, tw, f, rows, scale, st);
        case 16384: return rows_n<16384, IO, F, MODE>(in, out, tw, f, rows, scale, st);
    }
    return hipErrorInvalidValue;
}

template <class F, int MODE = kRowsCompress>
hipError_t rows_io(int io, int n, const void* in, void* out, const float2* tw, F f, int64_t rows,
                   float scale, hipStream_t st) {
    return io == MIFFT_IO_C32H ? rows_dispatch<MIFFT_IO_C32H, F, MODE>(n, in, out, tw, f, rows, scale, st)
                               : rows_dispatch<MIFFT_IO_C64, F, MODE>(n, in, out, tw, f, rows, scale, st);
}

hipError_t rcmc(int io, const void* in, void* out, const int* si, const float4* w, int64_t nr,
                int64_t na, hipStream_t st) {
    dim3 grid((unsigned)((na + 63) / 64), (unsigned)((nr + 4 * kRcmcRows - 1) / (4 * kRcmcRows)));
    if (io == MIFFT_IO_C32H)
        hipLaunchKernelGGL(k_rcmc<MIFFT_IO_C32H>, grid, dim3(256), 0, st, static_cast<const h2*>(in),
                           static_cast<h2*>(out), si, w, nr, na);
    else
        hipLaunchKernelGGL(k_rcmc<MIFFT_IO_C64>, grid, dim3(256), 0, st, static_cast<const float2*>(in),
                           static_cast<float2*>(out), si, w, nr, na);
    return hipGetLastError();
}

hipError_t corner_turn(int io, const void* in, void* out, int64_t rows, int64_t cols, hipStream_t st) {
    dim3 grid((unsigned)((cols + 63) / 64), (unsigned)((rows + 63) / 64));
    if (io == MIFFT_IO_C32H && rows % 64 == 0 && cols % 64 == 0 &&
        ((reinterpret_cast<uintptr_t>(in) | reinterpret_cast<uintptr_t>(out)) & 15) == 0) {
        hipLaunchKernelGGL(k_corner_turn_h16, grid, dim3(256), 0, st, static_cast<const uint4*>(in),
                           static_cast<uint4*>(out), rows, cols);
        return hipGetLastError();
    }
    if (io == MIFFT_IO_C32H)
        hipLaunchKernelGGL(k_corner_turn<MIFFT_IO_C32H>, grid, dim3(256), 0, st,
                           static_cast<const h2*>(in), static_cast<h2*>(out), rows, cols);
    else
        hipLaunchKernelGGL(k_corner_turn<MIFFT_IO_C64>, grid, dim3(256), 0, st,
                           static_cast<const float2*>(in), static_cast<float2*>(out), rows, cols);
    return hipGetLastError();
}

int ceil_log2(double x) {
    int e = 0;
    while (std::ldexp(1.0, e) < x) ++e;
    return e;
}

}  // namespace

mifft_status rda_range(const mifft_rda* r, const void* raw, void* out, hipStream_t st) {
    RangeFilter f{static_cast<const float2*>(r->range_filter.ptr)};
    if (r->p.precision != MIFFT_CONV_FP32)
        return check_hip(rows_h(r->p.precision, r->p.io, raw, out, r->range_plan->twiddles, f, r->p.n_azimuth,
                                (float)r->range_scale, st),
                         "range compression (FP16)");
    return check_hip(rows_io(r->p.io, (int)r->p.n_range, raw, out, r->range_plan->twiddles, f,
                             r->p.n_azimuth, (float)r->range_scale, st),
                     "range compression");
}

mifft_status rda_turn(const mifft_rda* r, const void* in, void* out, hipStream_t st) {
    return check_hip(corner_turn(r->p.io, in, out, r->p.n_azimuth, r->p.n_range, st), "corner turn");
}

template <class F>
mifft_status rda_azimuth_with(const mifft_rda* r, const F& f, const void* in, void* out, hipStream_t st);

mifft_status rda_azimuth(const mifft_rda* r, const void* in, void* out, hipStream_t st) {
    if (r->az_table.ptr) {
        const float2* e1 = static_cast<const float2*>(r->az_table.ptr);
        const int64_t rows_hi = (r->p.n_range + (1 << kAzRowBits) - 1) >> kAzRowBits;
        return rda_azimuth_with(r, AzimuthTable{e1, e1 + rows_hi * r->p.n_azimuth, r->p.n_azimuth}, in, out,
                                st);
    }
    AzimuthFilter f{static_cast<const double*>(r->az_a.ptr), static_cast<const double*>(r->az_d.ptr),
                    r->az_b, r->p.n_range / 2,
                    r->p.filter_convention == MIFFT_RDA_REFERENCE_FORMULA ? -1.0f : 1.0f};
    return rda_azimuth_with(r, f, in, out, st);
}

template <class F>
mifft_status rda_azimuth_with(const mifft_rda* r, const F& f, const void* in, void* out, hipStream_t st) {
    const bool half = r->p.precision != MIFFT_CONV_FP32;
    if (!r->p.rcmc) {
        if (half)
            return check_hip(rows_h(r->p.precision, r->p.io, in, out, r->az_plan->twiddles, f, r->p.n_range,
                                    (float)r->az_scale, st),
                             "azimuth compression (FP16)");
        return check_hip(rows_io(r->p.io, (int)r->p.n_azimuth, in, out, r->az_plan->twiddles, f,
                                 r->p.n_range, (float)r->az_scale, st),
                         "azimuth compression");
    }
    // azimuth FFT (scaled: the range-Doppler frame is stored, FP16 must not
    // overflow) into the work frame, RCMC from it into out, then filter +
    // IFFT in place
    mifft_status s = check_hip(rows_io<F, kRowsForward>(
                                   r->p.io, (int)r->p.n_azimuth, in, r->work.ptr,
                                   r->az_plan->twiddles, f, r->p.n_range, (float)r->az_scale, st),
                               "azimuth FFT");
    if (s != MIFFT_OK) return s;
    if ((s = check_hip(rcmc(r->p.io, r->work.ptr, out, static_cast<const int*>(r->rcmc_si.ptr),
                            static_cast<const float4*>(r->rcmc_w.ptr), r->p.n_range, r->p.n_azimuth,
                            st),
                       "RCMC")) != MIFFT_OK)
        return s;
    if (half)
        return check_hip(rows_h<F, kRowsFilterInverse>(r->p.precision, r->p.io, out, out,
                                                                   r->az_plan->twiddles, f, r->p.n_range, 1.0f, st),
                         "azimuth filter + IFFT (FP16)");
    return check_hip(rows_io<F, kRowsFilterInverse>(r->p.io, (int)r->p.n_azimuth, out, out,
                                                                 r->az_plan->twiddles, f, r->p.n_range,
                                                                 1.0f, st),
                     "azimuth filter + IFFT");
}

// Chirp reference of src/radar/sar_simulator.swift:207-224: Tp*fs samples of
// exp(i pi Kr t^2), t centred, stored wrapped around index 0.
std::vector<float2> chirp_reference(const mifft_rda_params& p, int64_t* samples) {
    std::vector<float2> c((size_t)p.n_range, make_float2(0.f, 0.f));
    const double kr = p.bandwidth / p.pulse_duration;
    const int64_t ns = (int64_t)(p.pulse_duration * p.sampling_rate);
    const int64_t half = ns / 2;
    for (int64_t i = 0; i < ns; ++i) {
        const double t = (double)(i - half) / p.sampling_rate;
        const double ph = M_PI * kr * t * t;
        const int64_t idx = i < half ? (p.n_range - half + i) : (i - half);
        if (idx >= 0 && idx < p.n_range) c[(size_t)idx] = make_float2((float)std::cos(ph), (float)std::sin(ph));
    }
    *samples = ns;
    return c;
}

mifft_status rda_setup(mifft_rda* r) {
    const mifft_rda_params& p = r->p;
    mifft_status s;
    // range matched filter conj(FFT(chirp)) from the FP32 FFT of the chirp
    int64_t ns = 0;
    std::vector<float2> chirp = chirp_reference(p, &ns);
    const size_t nb = sizeof(float2) * (size_t)p.n_range;
    if ((s = r->range_filter.reserve(nb)) != MIFFT_OK) return s;
    DeviceBuffer tmp;
    if ((s = tmp.reserve(nb)) != MIFFT_OK) return s;
    if ((s = check_hip(hipMemcpy(tmp.ptr, chirp.data(), nb, hipMemcpyHostToDevice), "hipMemcpy")) !=
        MIFFT_OK ||
        (s = check_hip(r->range_plan->variant->launch(static_cast<const float2*>(tmp.ptr),
                                                      static_cast<float2*>(r->range_filter.ptr),
                                                      r->range_plan->twiddles, 1, false, r->stream),
                       "chirp FFT")) != MIFFT_OK ||
        (s = check_hip(hipStreamSynchronize(r->stream), "hipStreamSynchronize")) != MIFFT_OK) {
        tmp.release();
        return s;
    }
    std::vector<float2> H((size_t)p.n_range);
    s = check_hip(hipMemcpy(H.data(), r->range_filter.ptr, nb, hipMemcpyDeviceToHost), "hipMemcpy");
    tmp.release();
    if (s != MIFFT_OK) return s;
    for (auto& h : H) h.y = -h.y;
    if ((s = check_hip(hipMemcpy(r->range_filter.ptr, H.data(), nb, hipMemcpyHostToDevice),
                       "hipMemcpy")) != MIFFT_OK)
        return s;

    // azimuth phase model in FP64 (rda_kernels.metal:195-238 in cycles)
    const double c = 299792458.0;
    const double lambda = c / p.carrier_freq;
    const double range_pixel = c / (2.0 * p.sampling_rate);
    std::vector<double> A((size_t)p.n_azimuth), D((size_t)p.n_azimuth);
    for (int64_t k = 0; k < p.n_azimuth; ++k) {
        const double fa_norm = k < p.n_azimuth / 2 ? (double)k / (double)p.n_azimuth
                                                   : (double)(k - p.n_azimuth) / (double)p.n_azimuth;
        const double ratio = lambda * fa_norm * p.prf / (2.0 * p.velocity);
        if (std::fabs(ratio) < 0.95) {
            const double d = std::sqrt(1.0 - ratio * ratio);
            const double cyc = 2.0 * p.center_range * d / lambda;
            D[(size_t)k] = d;
            A[(size_t)k] = cyc - std::floor(cyc);
        } else {
            D[(size_t)k] = -1.0;
            A[(size_t)k] = 0.0;
        }
    }
    r->az_b = 2.0 * range_pixel / lambda;
    const size_t ab = sizeof(double) * (size_t)p.n_azimuth;
    if ((s = r->az_a.reserve(ab)) != MIFFT_OK || (s = r->az_d.reserve(ab)) != MIFFT_OK) return s;
    if ((s = check_hip(hipMemcpy(r->az_a.ptr, A.data(), ab, hipMemcpyHostToDevice), "hipMemcpy")) !=
            MIFFT_OK ||
        (s = check_hip(hipMemcpy(r->az_d.ptr, D.data(), ab, hipMemcpyHostToDevice), "hipMemcpy")) !=
            MIFFT_OK)
        return s;
    // the filter's two factor tables, (Nr / 64 + 64) x Na entries (4 MiB at
    // 4096 x 4096), when they hold at most 2^26 entries
    const int64_t rows_hi = (p.n_range + (1 << kAzRowBits) - 1) >> kAzRowBits;
    const int64_t cells = (rows_hi + (1 << kAzRowBits)) * p.n_azimuth;
    if (cells <= ((int64_t)1 << 26)) {
        if ((s = r->az_table.reserve(sizeof(float2) * (size_t)cells)) != MIFFT_OK) return s;
        hipLaunchKernelGGL(k_az_factors, dim3((unsigned)((cells + 255) / 256)), dim3(256), 0, r->stream,
                           static_cast<const double*>(r->az_a.ptr), static_cast<const double*>(r->az_d.ptr), r->az_b,
                           p.n_range / 2, p.filter_convention == MIFFT_RDA_REFERENCE_FORMULA ? -1.0 : 1.0, rows_hi,
                           p.n_azimuth, static_cast<float2*>(r->az_table.ptr));
        if ((s = check_hip(hipGetLastError(), "azimuth filter tables")) != MIFFT_OK ||
            (s = check_hip(hipStreamSynchronize(r->stream), "hipStreamSynchronize")) != MIFFT_OK)
            return s;
    }
    // RCMC (rda_kernels.metal:103-185): shift per Doppler bin in range bins,
    // dR = R0 (1/sqrt(1 - ratio^2) - 1) at the centre range, 0 where
    // |ratio| >= 0.95; output bin r reads r + shift (matched: migrated energy
    // pulled back) or r - shift (the kernel's literal srcIdx = rgIdx - shift).
    // Integer part and the 8 windowed-sinc weights of the fraction, in FP64.
    if (p.rcmc) {
        std::vector<int> si((size_t)p.n_azimuth);
        std::vector<float> w((size_t)p.n_azimuth * 8);
        const double sgn = p.filter_convention == MIFFT_RDA_REFERENCE_FORMULA ? -1.0 : 1.0;
        for (int64_t k = 0; k < p.n_azimuth; ++k) {
            const double fa_norm = k < p.n_azimuth / 2 ? (double)k / (double)p.n_azimuth
                                                       : (double)(k - p.n_azimuth) / (double)p.n_azimuth;
            const double ratio = lambda * fa_norm * p.prf / (2.0 * p.velocity);
            const double dr =
                std::fabs(ratio) < 0.95 ? p.center_range * (1.0 / std::sqrt(1.0 - ratio * ratio) - 1.0) : 0.0;
            const double shift = sgn * dr / range_pixel;
            const double fl = std::floor(shift);
            const double frac = shift - fl;
            si[(size_t)k] = (int)fl;
            for (int j = 0; j < 8; ++j) {
                const double x = frac - (double)(j - 3);
                const double sinc = std::fabs(x) < 1e-6 ? 1.0 : std::sin(M_PI * x) / (M_PI * x);
                const double win = 0.54 - 0.46 * std::cos(2.0 * M_PI * (double)j / 7.0);
                w[(size_t)k * 8 + j] = (float)(sinc * win);
            }
        }
        if ((s = r->rcmc_si.reserve(sizeof(int) * si.size())) != MIFFT_OK ||
            (s = r->rcmc_w.reserve(sizeof(float) * w.size())) != MIFFT_OK)
            return s;
        if ((s = check_hip(hipMemcpy(r->rcmc_si.ptr, si.data(), sizeof(int) * si.size(),
                                     hipMemcpyHostToDevice), "hipMemcpy")) != MIFFT_OK ||
            (s = check_hip(hipMemcpy(r->rcmc_w.ptr, w.data(), sizeof(float) * w.size(),
                                     hipMemcpyHostToDevice), "hipMemcpy")) != MIFFT_OK)
            return s;
    }
    // one frame of work memory: range-compressed data in process_device, the
    // range-Doppler frame before RCMC
    const size_t esz = p.io == MIFFT_IO_C32H ? 4 : 8;
    if ((s = r->work.reserve(esz * (size_t)p.n_range * (size_t)p.n_azimuth)) != MIFFT_OK) return s;
    // power-of-two stage gains keep complex FP16 in range: a point target's
    // range peak is ~(chirp samples), its azimuth peak ~(pulses) times that
    r->chirp_samples = ns;
    r->range_scale = std::ldexp(1.0, -ceil_log2((double)(ns > 0 ? ns : 1)));
    r->az_scale = std::ldexp(1.0, -ceil_log2((double)p.n_azimuth));
    return MIFFT_OK;
}

}  // namespace mifft

using namespace mifft;

extern "C" {

mifft_status mifft_rda_default_params(mifft_rda_params* p) {
    if (!p) return fail(MIFFT_ERR_INVALID_ARGUMENT, "null params");
    // src/radar/sar_simulator.swift:22-67
    p->n_range = 4096;
    p->n_azimuth = 4096;
    p->bandwidth = 100e6;
    p->pulse_duration = 10e-6;
    p->sampling_rate = 1.2 * 100e6;
    p->prf = 1000.0;
    p->velocity = 100.0;
    p->carrier_freq = 10e9;
    p->center_range = 20000.0;
    p->io = MIFFT_IO_C32H;
    p->filter_convention = MIFFT_RDA_MATCHED;
    p->rcmc = 1;
    p->precision = MIFFT_CONV_FP32;
    return MIFFT_OK;
}

mifft_status mifft_rda_create(const mifft_rda_params* params, mifft_rda** out) {
    if (!out || !params) return fail(MIFFT_ERR_INVALID_ARGUMENT, "null argument");
    *out = nullptr;
    const mifft_rda_params& p = *params;
    if (!supported_size(p.n_range))
        return fail(MIFFT_ERR_UNSUPPORTED_SIZE, "Unsupported FFT size: " + std::to_string(p.n_range),
                    -1, p.n_range);
    if (!supported_size(p.n_azimuth))
        return fail(MIFFT_ERR_UNSUPPORTED_SIZE, "Unsupported FFT size: " + std::to_string(p.n_azimuth),
                    -1, p.n_azimuth);
    if (p.io != MIFFT_IO_C32H && p.io != MIFFT_IO_C64)
        return fail(MIFFT_ERR_INVALID_ARGUMENT, "unknown io type");
    if (p.filter_convention != MIFFT_RDA_MATCHED && p.filter_convention != MIFFT_RDA_REFERENCE_FORMULA)
        return fail(MIFFT_ERR_INVALID_ARGUMENT, "unknown filter convention");
    if (p.rcmc != 0 && p.rcmc != 1) return fail(MIFFT_ERR_INVALID_ARGUMENT, "rcmc must be 0 or 1");
    if (p.precision < MIFFT_CONV_FP32 || p.precision > MIFFT_CONV_FP16_MIXED)
        return fail(MIFFT_ERR_INVALID_ARGUMENT, "unknown precision");
    if (p.precision != MIFFT_CONV_FP32 && (p.n_range != 4096 || p.n_azimuth != 4096))
        return fail(MIFFT_ERR_KERNEL_NOT_FOUND,
                    "HIP kernel not found: the FP16 SAR modes are built for 4096 x 4096 frames");
    if (!(p.bandwidth > 0 && p.pulse_duration > 0 && p.sampling_rate > 0 && p.prf > 0 &&
          p.velocity > 0 && p.carrier_freq > 0 && p.center_range > 0))
        return fail(MIFFT_ERR_PRECONDITION, "SAR parameters must be positive");
    mifft_rda* r = new mifft_rda();
    r->p = p;
    mifft_status s;
    if ((s = mifft_plan_create(p.n_range, &r->range_plan)) != MIFFT_OK ||
        (s = mifft_plan_create(p.n_azimuth, &r->az_plan)) != MIFFT_OK) {
        mifft_rda_destroy(r);
        return s;
    }
    if (hipStreamCreateWithFlags(&r->stream, hipStreamNonBlocking) != hipSuccess) {
        mifft_rda_destroy(r);
        return fail(MIFFT_ERR_NO_QUEUE, "Could not create HIP stream");
    }
    if ((s = rda_setup(r)) != MIFFT_OK) {
        mifft_rda_destroy(r);
        return s;
    }
    *out = r;
    return MIFFT_OK;
}

mifft_status mifft_rda_destroy(mifft_rda* r) {
    if (!r) return MIFFT_OK;
    if (r->stream) (void)hipStreamSynchronize(r->stream);
    r->range_filter.release();
    r->az_a.release();
    r->az_d.release();
    r->az_table.release();
    r->rcmc_si.release();
    r->rcmc_w.release();
    r->work.release();
    mifft_plan_destroy(r->range_plan);
    mifft_plan_destroy(r->az_plan);
    if (r->stream) (void)hipStreamDestroy(r->stream);
    delete r;
    return MIFFT_OK;
}

mifft_status mifft_rda_scales(const mifft_rda* r, double* range_scale, double* azimuth_scale,
                              int64_t* chirp_samples) {
    if (!r) return fail(MIFFT_ERR_INVALID_ARGUMENT, "null rda");
    if (range_scale) *range_scale = r->range_scale;
    if (azimuth_scale) *azimuth_scale = r->az_scale;
    if (chirp_samples) *chirp_samples = r->chirp_samples;
    return MIFFT_OK;
}

mifft_status mifft_rda_range_compress_device(const mifft_rda* r, const void* raw, void* out,
                                             void* stream) {
    if (!r || !raw || !out) return fail(MIFFT_ERR_INVALID_ARGUMENT, "null argument");
    return rda_range(r, raw, out, static_cast<hipStream_t>(stream));
}

mifft_status mifft_rda_corner_turn_device(const mifft_rda* r, const void* in, void* out,
                                          void* stream) {
    if (!r || !in || !out) return fail(MIFFT_ERR_INVALID_ARGUMENT, "null argument");
    if (in == out) return fail(MIFFT_ERR_INVALID_ARGUMENT, "corner turn is out of place");
    return rda_turn(r, in, out, static_cast<hipStream_t>(stream));
}

mifft_status mifft_rda_azimuth_compress_device(const mifft_rda* r, const void* in, void* out,
                                               void* stream) {
    if (!r || !in || !out) return fail(MIFFT_ERR_INVALID_ARGUMENT, "null argument");
    return rda_azimuth(r, in, out, static_cast<hipStream_t>(stream));
}

mifft_status mifft_rda_process_device(mifft_rda* r, const void* raw, void* image, void* stream) {
    if (!r || !raw || !image) return fail(MIFFT_ERR_INVALID_ARGUMENT, "null argument");
    if (raw == image) return fail(MIFFT_ERR_INVALID_ARGUMENT, "raw and image must differ");
    mifft_status s;
    hipStream_t st = static_cast<hipStream_t>(stream);
    if ((s = rda_range(r, raw, r->work.ptr, st)) != MIFFT_OK) return s;
    if ((s = rda_turn(r, r->work.ptr, image, st)) != MIFFT_OK) return s;
    return rda_azimuth(r, image, image, st);
}

}  // extern "C"
