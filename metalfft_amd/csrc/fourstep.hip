// Four-step complex FFT for lengths above one workgroup's reach
// (n = 32768 ... 2^20), two launches per batch.
//
// Reference counterpart: MetalFFT's four-step for 8192/16384
// (Sources/MetalFFT/FFT/MetalFFT.swift:206-277: transpose, n1 row FFTs of
// size n2, twiddle W_N^{row*col} + transpose, n2 row FFTs of size n1,
// transpose, with the kernels of Sources/MetalFFT/Resources/fft_multisize.metal:700-749),
// five dispatches and four device round trips per transform.  Here the
// transposes and the twiddle are folded into the two transform passes
// through LDS:
//
//   n = n1 + N1 n2 (n1 < N1, n2 < N2),  X[k2 + N2 k1] =
//       sum_n1 W_N1^{n1 k1} [ W_N^{n1 k2} sum_n2 x[n1 + N1 n2] W_N2^{n2 k2} ]
//
//   pass A (columns): a workgroup takes CA adjacent columns n1 of the
//     N2 x N1 input, loads them as rows of CA contiguous complex values
//     (8 CA-byte segments), stages the tile in LDS column-major, runs the
//     N2-point Stockham transform of every column, multiplies by
//     W_N^{n1 k2} and stores each column as a contiguous row of the
//     N1 x N2 intermediate y[n1][k2].
//   pass B (rows): a workgroup takes CB adjacent columns k2 of y, loads
//     rows of CB contiguous values, runs the N1-point transforms and writes
//     X[k2 + N2 k1] back through LDS as rows of CB contiguous values.
//
// Every element is read and written once per pass (32 n bytes per
// transform against 16 n for the single-pass lengths).  When in != out the
// intermediate is `out` itself: pass B reads and writes the same column set
// of its transform, and loads all of it before its first store.  The
// inverse is conj -> forward -> conj / n (MetalFFT.swift:56-65), conjugated
// on pass A's load and scaled on pass B's store.
#include "fourstep.hpp"
#include "kernels.hpp"
#include "stockham_core.hpp"

namespace mifft {

namespace {

// --------------------------------------------------------- twiddle layout
template <int L>
constexpr int sub_len() { return L + PassTable<L, 16, 0>::size; }
// offset of the length-L sub table (L = 128 ... 1024)
constexpr int sub_off(int L) {
    return L <= 128 ? 0
         : L == 256 ? sub_len<128>()
         : L == 512 ? sub_len<128>() + sub_len<256>()
                    : sub_len<128>() + sub_len<256>() + sub_len<512>();
}
constexpr int kSubTotal = sub_off(1024) + sub_len<1024>();

constexpr int ilog2i(int n) { return n <= 1 ? 0 : 1 + ilog2i(n / 2); }
// bits of the low factor of W_n^e
constexpr int lo_bits(int n) { return (ilog2i(n) + 1) / 2; }

// lo[i] = W_n^i (i < 2^S), hi[i] = W_n^{i 2^S} (i < n / 2^S): FP64 sincospi,
// rounded once to FP32.
__global__ void k_fs_factors(float2* __restrict__ tab, int n, int s) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int nlo = 1 << s, nhi = n >> s;
    if (i >= nlo + nhi) return;
    const int64_t e = i < nlo ? i : (int64_t)(i - nlo) << s;
    double sn, cs;
    sincospi(-2.0 * (double)e / (double)n, &sn, &cs);
    tab[i] = make_float2((float)cs, (float)sn);
}

// ------------------------------------------------------------ LDS tiles
// C columns of L complex values; column c occupies [c * LEN, c * LEN + L)
// with the XOR-swizzled Stockham layout inside (LdsSwz).  The pad makes the
// staging accesses (C consecutive columns at one row per group of C lanes)
// fall on distinct 8-byte slots: 32 / C extra entries per column for C < 32,
// one for C >= 32.
template <int L, int C>
struct FsTile {
    static constexpr int T = L / 16;  // threads per column
    static constexpr int THREADS = C * T;
    static constexpr int PAD = C >= 32 ? 1 : 32 / C;
    static constexpr int LEN = L + PAD;
    static_assert(THREADS <= 1024, "workgroup too large");
    static_assert(C * LEN * 8 <= 160 * 1024, "LDS over 160 KiB");
};

// Element (c, r) of the C x L tile at src[c + r * ld] -> column c, row r.
// Thread tid moves elements tid + THREADS q, q < 16: consecutive lanes take
// consecutive columns of one row, so each group of C lanes reads 8 C
// contiguous bytes.
template <int L, int C, bool CONJ, int LP = 0>
__device__ __forceinline__ void stage_in(float2* lds, const float2* src, int ld, int tid) {
    using Tl = FsTile<L, C>;
    float2 x[16];
#pragma unroll
    for (int q = 0; q < 16; ++q) {
        const int f = tid + Tl::THREADS * q;
        x[q] = gload_p<LP>(src + (f % C) + (int64_t)(f / C) * ld);
    }
#pragma unroll
    for (int q = 0; q < 16; ++q) {
        const int f = tid + Tl::THREADS * q;
        lds[(f % C) * Tl::LEN + LdsSwz::idx(f / C)] = CONJ ? conjf2(x[q]) : x[q];
    }
}

// Column c, row r of the tile -> dst[c + r * ld] (the transpose of
// stage_in), conj and scale by 1/n for the inverse.
template <int L, int C, bool INV, int NTOT, int SP = 0>
__device__ __forceinline__ void stage_out(const float2* lds, float2* dst, int ld, int tid) {
    using Tl = FsTile<L, C>;
#pragma unroll
    for (int q = 0; q < 16; ++q) {
        const int f = tid + Tl::THREADS * q;
        float2 y = lds[(f % C) * Tl::LEN + LdsSwz::idx(f / C)];
        if constexpr (INV) {
            constexpr float inv_n = 1.0f / (float)NTOT;
            y = make_float2(y.x * inv_n, -y.y * inv_n);
        }
        gstore_p<SP>(dst + (f % C) + (int64_t)(f / C) * ld, y);
    }
}

// Length-L transform of the column this thread works on (column `local`,
// lanes t < L/16 of it), staged in LDS: v[m] = X[t + T m] on return.
template <int L>
__device__ __forceinline__ void column_fft(float2 (&v)[16], float2* buf, int t,
                                           const float2* __restrict__ twL) {
    PreBases<L, 16> pb;
    pb.load_packed(twL, t);  // in flight through the staging barrier
    lds_barrier();           // the whole tile is staged
    lds_read_column<16, L / 16, LdsSwz>(v, buf, t);
    stockham_regs<L, 16, LdsSwz, 0, true, true, false, 2>(v, buf, t, twL, pb.w);
}

// Pass A: columns n1 in [c0, c0 + CA) of transform tr; y row n1 = the
// N2-point transform of column n1 times W_N^{n1 k2}, stored under cache
// policy YP (gstore_p).
template <int N1, int N2, int CA, bool INV, int YP>
__global__ void __launch_bounds__((N2 / 16) * CA)
k_fourstep_cols(const float2* in, float2* y, const float2* __restrict__ tw, int64_t batch) {
    using Tl = FsTile<N2, CA>;
    constexpr int N = N1 * N2;
    constexpr int TILES = N1 / CA;
    constexpr int S = lo_bits(N);
    constexpr int T = Tl::T;
    __shared__ float2 lds[CA * Tl::LEN];
    const int64_t tr = blockIdx.x / TILES;
    const int c0 = (int)(blockIdx.x % TILES) * CA;
    const int tid = threadIdx.x;
    stage_in<N2, CA, INV>(lds, in + tr * N + c0, N1, tid);
    const int local = tid / T, t = tid % T;
    float2 v[16];
    column_fft<N2>(v, lds + local * Tl::LEN, t, tw + sub_off(N2));
    // n1 is the same for a whole wave when a column spans whole waves
    const int n1 = T >= 64 ? __builtin_amdgcn_readfirstlane(c0 + local) : c0 + local;
    const float2* lo = tw + kSubTotal;
    const float2* hi = lo + (1 << S);
    float2* dst = y + tr * N + (int64_t)n1 * N2;
    // W_N^{n1 k2}, k2 = t + T m, as W_N^{n1 t} W_N^{n1 T m}: one gathered
    // factor pair per thread, the other per m the same for the column (one
    // cache line per wave load instead of up to 64); exponents < N, no
    // reduction mod N
    const int ea = n1 * t;
    const float2 wa = cmul(hi[ea >> S], lo[ea & ((1 << S) - 1)]);
    gstore_p<YP>(dst + t, cmul(v[0], wa));
#pragma unroll
    for (int m = 1; m < 16; ++m) {
        const int eb = n1 * (T * m);
        const float2 w = cmul(wa, cmul(hi[eb >> S], lo[eb & ((1 << S) - 1)]));
        gstore_p<YP>(dst + t + T * m, cmul(v[m], w));
    }
}

// Pass B: columns k2 in [k0, k0 + CB) of y (N1 rows of N2) -> X[k2 + N2 k1];
// y read under policy YP, X stored under OP.
template <int N1, int N2, int CB, bool INV, int YP, int OP>
__global__ void __launch_bounds__((N1 / 16) * CB)
k_fourstep_rows(const float2* y, float2* out, const float2* __restrict__ tw, int64_t batch) {
    using Tl = FsTile<N1, CB>;
    constexpr int N = N1 * N2;
    constexpr int TILES = N2 / CB;
    constexpr int T = Tl::T;
    __shared__ float2 lds[CB * Tl::LEN];
    const int64_t tr = blockIdx.x / TILES;
    const int k0 = (int)(blockIdx.x % TILES) * CB;
    const int tid = threadIdx.x;
    stage_in<N1, CB, false, YP>(lds, y + tr * N + k0, N2, tid);
    const int local = tid / T, t = tid % T;
    float2* buf = lds + local * Tl::LEN;
    float2 v[16];
    column_fft<N1>(v, buf, t, tw + sub_off(N1));
    lds_barrier();  // every column's last LDS reads are done
#pragma unroll
    for (int m = 0; m < 16; ++m) buf[LdsSwz::idx(t + T * m)] = v[m];
    lds_barrier();
    stage_out<N1, CB, INV, N, OP>(lds, out + tr * N + k0, N2, tid);
}

template <int N1, int N2, int CA, int CB, int YP, int OP>
hipError_t launch_fs(const float2* in, float2* out, float2* y, const float2* tw, int64_t batch,
                     bool inverse, hipStream_t st) {
    if (batch <= 0) return hipSuccess;
    const int64_t ga = batch * (N1 / CA), gb = batch * (N2 / CB);
    if (ga > 0x7fffffffLL || gb > 0x7fffffffLL) return hipErrorInvalidValue;
    constexpr int ta = FsTile<N2, CA>::THREADS, tb = FsTile<N1, CB>::THREADS;
    if (inverse) {
        hipLaunchKernelGGL((k_fourstep_cols<N1, N2, CA, true, YP>), dim3((unsigned)ga), dim3(ta), 0, st, in, y,
                           tw, batch);
        hipLaunchKernelGGL((k_fourstep_rows<N1, N2, CB, true, YP, OP>), dim3((unsigned)gb), dim3(tb), 0, st, y,
                           out, tw, batch);
    } else {
        hipLaunchKernelGGL((k_fourstep_cols<N1, N2, CA, false, YP>), dim3((unsigned)ga), dim3(ta), 0, st, in,
                           y, tw, batch);
        hipLaunchKernelGGL((k_fourstep_rows<N1, N2, CB, false, YP, OP>), dim3((unsigned)gb), dim3(tb), 0, st, y,
                           out, tw, batch);
    }
    return hipGetLastError();
}

#define MIFFT_FS(N1, N2, CA, CB)                                                      \
    {N1 * N2, N1, N2, CA, CB, &launch_fs<N1, N2, CA, CB, 0, 0>,                       \
     "fourstep_" #N1 "x" #N2 "_c" #CA "_c" #CB}
#define MIFFT_FSP(N1, N2, CA, CB, YP, OP)                                             \
    {N1 * N2, N1, N2, CA, CB, &launch_fs<N1, N2, CA, CB, YP, OP>,                     \
     "fourstep_" #N1 "x" #N2 "_c" #CA "_c" #CB "_y" #YP "o" #OP}
// Every shape again with the intermediate stored (and read back) under the
// default or the system-coherent policy instead of the streaming one, and
// the output streaming or system-coherent: pass B then finds part of y in
// the caches (at 2^20, y2o2: 133 -> 121 us per 128 MiB)
#define MIFFT_FSPS(N1, N2, CA, CB)                                                    \
    MIFFT_FSP(N1, N2, CA, CB, 1, 0), MIFFT_FSP(N1, N2, CA, CB, 2, 0),                 \
    MIFFT_FSP(N1, N2, CA, CB, 1, 2), MIFFT_FSP(N1, N2, CA, CB, 2, 2)

// Default first for each length (find_fourstep); the autotuner times all.
const FourStepVariant kFourStep[] = {
    MIFFT_FS(256, 128, 32, 16),   MIFFT_FS(128, 256, 16, 32),   MIFFT_FS(256, 128, 16, 16),
    MIFFT_FS(256, 256, 16, 16),   MIFFT_FS(256, 256, 32, 32),
    MIFFT_FS(512, 256, 16, 16),   MIFFT_FS(256, 512, 16, 16),   MIFFT_FS(512, 256, 32, 16),
    MIFFT_FS(512, 512, 16, 16),   MIFFT_FS(512, 512, 8, 8),
    MIFFT_FS(1024, 512, 16, 16),  MIFFT_FS(512, 1024, 16, 16),  MIFFT_FS(1024, 512, 16, 8),
    MIFFT_FS(1024, 1024, 16, 16), MIFFT_FS(1024, 1024, 8, 8),
    MIFFT_FSPS(256, 128, 32, 16), MIFFT_FSPS(128, 256, 16, 32), MIFFT_FSPS(256, 128, 16, 16),
    MIFFT_FSPS(256, 256, 16, 16), MIFFT_FSPS(256, 256, 32, 32),
    MIFFT_FSPS(512, 256, 16, 16), MIFFT_FSPS(256, 512, 16, 16), MIFFT_FSPS(512, 256, 32, 16),
    MIFFT_FSPS(512, 512, 16, 16), MIFFT_FSPS(512, 512, 8, 8),
    MIFFT_FSPS(1024, 512, 16, 16), MIFFT_FSPS(512, 1024, 16, 16), MIFFT_FSPS(1024, 512, 16, 8),
    MIFFT_FSPS(1024, 1024, 16, 16), MIFFT_FSPS(1024, 1024, 8, 8),
};
#undef MIFFT_FSPS
#undef MIFFT_FSP
#undef MIFFT_FS

}  // namespace

const FourStepVariant* fourstep_variants(int* count) {
    *count = (int)(sizeof(kFourStep) / sizeof(kFourStep[0]));
    return kFourStep;
}

const FourStepVariant* find_fourstep(int n) {
    int c = 0;
    const FourStepVariant* v = fourstep_variants(&c);
    for (int i = 0; i < c; ++i)
        if (v[i].n == n) return &v[i];
    return nullptr;
}

bool fourstep_size(int64_t n) { return n >= kFourStepMin && n <= kFourStepMax && (n & (n - 1)) == 0; }

int fourstep_table_len(int n) {
    const int s = lo_bits(n);
    return kSubTotal + (1 << s) + (n >> s);
}

hipError_t launch_fourstep_twiddles(float2* tw, int n, hipStream_t st) {
    for (int L = 128; L <= 1024; L *= 2) {
        const hipError_t e = launch_twiddles(tw + sub_off(L), L, st);
        if (e != hipSuccess) return e;
    }
    const int s = lo_bits(n);
    const int cnt = (1 << s) + (n >> s);
    hipLaunchKernelGGL(k_fs_factors, dim3((cnt + 255) / 256), dim3(256), 0, st, tw + kSubTotal, n, s);
    return hipGetLastError();
}

}  // namespace mifft
