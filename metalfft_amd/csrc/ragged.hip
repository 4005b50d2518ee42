// Ragged batches: thousands of complex transforms of different power-of-two
// lengths (64 ... 16384) in one launch.
//
// No reference counterpart: MetalFFT.forward(batch:) rejects uneven lengths
// (Sources/MetalFFT/FFT/MetalFFT.swift:108-114).
//
// The transforms are sorted by length, longest first, and cut into tickets:
// the transforms of one length that one 512-thread workgroup pass holds (one
// of 16384, two of 8192, 8192 or 4096 elements of a shorter length; the tile
// shape per length is measured at plan creation).  Launch structures over
// the same ticket list, all capturable in a hipGraph:
//   default    one launch, workgroup b runs ticket b (k_ragged_tickets);
//   per size   the same kernel once per length;
//   work queue one persistent launch drawing tickets from a device counter
//              that resets itself (k_ragged_queue);
//   split      one launch per length of 4096 and longer (one transform per
//              workgroup, k_ragged_one / k_ragged_half) and one persistent
//              launch over 4096-element tiles of the shorter lengths
//              (k_ragged).
// The three ticket forms run the same code per ticket and give the same
// bits.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "kernels.hpp"
#include "ragged.hpp"
#include "stockham_core.hpp"

namespace mifft {

namespace {

int ragged_cus() {
    static int cus = [] {
        int dev = 0, n = 0;
        if (hipGetDevice(&dev) != hipSuccess) return 256;
        if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
            return 256;
        return n > 0 ? n : 256;
    }();
    return cus;
}

// Tile elements of length 2^l: one transform for 4096 and longer, else 4096.
constexpr int tile_elems(int l) { return l >= 12 ? (1 << l) : 4096; }

// Twiddle tables of every length back to back: the plan table of N = 2^l
// (W_N^k, k < N, then the packed bases rows of PassTable<N, 16, 0>) starts at
// table_off(l).
constexpr int table_len(int l) { return l == 6 ? 64 + PassTable<64, 16, 0>::size
                                      : l == 7 ? 128 + PassTable<128, 16, 0>::size
                                      : l == 8 ? 256 + PassTable<256, 16, 0>::size
                                      : l == 9 ? 512 + PassTable<512, 16, 0>::size
                                      : l == 10 ? 1024 + PassTable<1024, 16, 0>::size
                                      : l == 11 ? 2048 + PassTable<2048, 16, 0>::size
                                      : l == 12 ? 4096 + PassTable<4096, 16, 0>::size
                                      : l == 13 ? 8192 + PassTable<8192, 16, 0>::size
                                      : 16384 + PassTable<16384, 16, 0>::size; }
constexpr int table_off(int l) { return l <= 6 ? 0 : table_off(l - 1) + table_len(l - 1); }
constexpr int kTablesLen = table_off(15);

// One transform of a tile: thread `tid` of the workgroup works on transform
// tid / (N/16) at offset `off` (`active`: it exists).  The data loads are
// unconditional (an inactive slot reads the tile's last transform and never
// stores), non-temporal, and issued before the packed twiddle bases rows
// (one 32-byte row per butterfly and pass), so no pass waits on a table read.
template <int N, bool INV>
__device__ __forceinline__ void tile_body(const float2* in, float2* out, const float2* tw, int64_t off,
                                          bool active, float2* lds) {
    constexpr int E = 16;
    constexpr int T = N / E;
    const int tid = fresh((int)threadIdx.x);  // index arithmetic redone per tile, not hoisted
    const int local = tid / T;
    const int t = tid % T;
    float2* buf = lds + local * LdsSwz::len(N);
    float2 v[E];
    load_column<N, E, INV, true>(v, in + off, t, true);
    PreBases<N, E> pb;
    pb.load_packed(tw, t);
    stockham_regs<N, E, LdsSwz, 0, true, true, false, 2>(v, buf, t, tw, pb.w);
    if (active) store_column<N, E, INV, true>(v, out + off, t);
}

// Workgroup b takes tiles tile_begin + b, + gridDim.x, ... of the sorted
// list; the tiles of one length form a contiguous range, so that sequence is
// walked as one loop per length (largest first), each with its own
// transform code.  The next tile's descriptor and this thread's offset in it
// are loaded while the current tile runs (clamped indices, no branch).
template <int L, bool INV>
__device__ __forceinline__ void length_loop(const float2* in, float2* out, const float2* tw_all,
                                            const RaggedTile* tiles, const int64_t* offs, int lo, int hi,
                                            int tile_begin, int tile_end, float2* lds) {
    constexpr int T = (1 << L) / 16;
    // a constant expression: as a plain call the recursive table_off became
    // a device function call per tile (with a full s_waitcnt on entry)
    constexpr int kOff = table_off(L);
    const int b = lo > tile_begin ? lo : tile_begin;
    const int e = hi < tile_end ? hi : tile_end;
    const int stride = (int)gridDim.x;
    const int r = tile_begin + (int)blockIdx.x;
    int ti = b + (((r - b) % stride) + stride) % stride;
    if (ti >= e) return;
    const int local = (int)threadIdx.x / T;
    RaggedTile d = tiles[ti];
    int64_t off = offs[d.first + (local < d.count ? local : d.count - 1)];
    for (; ti < e; ti += stride) {
        const int tn = ti + stride < e ? ti + stride : ti;
        const RaggedTile dn = tiles[tn];
        const int64_t offn = offs[dn.first + (local < dn.count ? local : dn.count - 1)];
        tile_body<(1 << L), INV>(in, out, tw_all + kOff, off, local < d.count, lds);
        d = dn;
        off = offn;
    }
}

// One tile class: TILE elements per tile, TILE/16 threads, lengths
// LMIN..LMAX.  LDS: TILE/N transforms of LdsSwz::len(N) = N.
template <bool INV, int TILE, int LMIN, int LMAX>
__global__ void __launch_bounds__(TILE / 16)
k_ragged(const float2* in, float2* out, const float2* __restrict__ tw_all,
         const RaggedTile* __restrict__ tiles, const int64_t* __restrict__ offs, RaggedRanges rr,
         int tile_begin, int tile_end) {
    __shared__ float2 lds[TILE];
    if constexpr (LMAX >= 14 && LMIN <= 14) length_loop<14, INV>(in, out, tw_all, tiles, offs, rr.lo[8], rr.hi[8], tile_begin, tile_end, lds);
    if constexpr (LMAX >= 13 && LMIN <= 13) length_loop<13, INV>(in, out, tw_all, tiles, offs, rr.lo[7], rr.hi[7], tile_begin, tile_end, lds);
    if constexpr (LMAX >= 12 && LMIN <= 12) length_loop<12, INV>(in, out, tw_all, tiles, offs, rr.lo[6], rr.hi[6], tile_begin, tile_end, lds);
    if constexpr (LMAX >= 11 && LMIN <= 11) length_loop<11, INV>(in, out, tw_all, tiles, offs, rr.lo[5], rr.hi[5], tile_begin, tile_end, lds);
    if constexpr (LMAX >= 10 && LMIN <= 10) length_loop<10, INV>(in, out, tw_all, tiles, offs, rr.lo[4], rr.hi[4], tile_begin, tile_end, lds);
    if constexpr (LMAX >= 9 && LMIN <= 9) length_loop<9, INV>(in, out, tw_all, tiles, offs, rr.lo[3], rr.hi[3], tile_begin, tile_end, lds);
    if constexpr (LMAX >= 8 && LMIN <= 8) length_loop<8, INV>(in, out, tw_all, tiles, offs, rr.lo[2], rr.hi[2], tile_begin, tile_end, lds);
    if constexpr (LMAX >= 7 && LMIN <= 7) length_loop<7, INV>(in, out, tw_all, tiles, offs, rr.lo[1], rr.hi[1], tile_begin, tile_end, lds);
    if constexpr (LMAX >= 6 && LMIN <= 6) length_loop<6, INV>(in, out, tw_all, tiles, offs, rr.lo[0], rr.hi[0], tile_begin, tile_end, lds);
}

// Length 4096: one workgroup per transform (the uniform kernel's body,
// k_stockham with swizzled LDS, non-temporal I/O and packed twiddle rows) at
// offs[blockIdx.x], a uniform (scalar) load: one launch per length,
// scheduled by the hardware; 8192 and 16384 likewise with k_ragged_half.  A persistent loop over these tiles ran 25-35 %
// slower than the same transforms as one-shot workgroups.
template <int N, bool INV>
__global__ void __launch_bounds__(N / 16)
k_ragged_one(const float2* in, float2* out, const float2* __restrict__ tw, const int64_t* __restrict__ offs) {
    constexpr int E = 16;
    __shared__ float2 lds[N];
    const int t = threadIdx.x;
    const int64_t off = offs[blockIdx.x];
    float2 v[E];
    load_column<N, E, INV, true>(v, in + off, t, true);
    PreBases<N, E> pb;
    pb.load_packed(tw, t);
    stockham_regs<N, E, LdsSwz, 0, false, true, false, 2>(v, lds, t, tw, pb.w);
    store_column<N, E, INV, true>(v, out + off, t);
}

// 8192 and 16384: the half-LDS exchange form (32 values per thread, N/2
// entries of LDS), four and two transforms per CU instead of two and one.
template <int N, bool INV>
__global__ void __launch_bounds__(N / 32)
__attribute__((amdgpu_waves_per_eu(4, 8)))
k_ragged_half(const float2* in, float2* out, const float2* __restrict__ tw, const int64_t* __restrict__ offs) {
    __shared__ float2 lds[N / 2];
    const int64_t off = offs[blockIdx.x];
    half_transform<N, INV, true>(in + off, out + off, tw, lds, threadIdx.x, true);
}

template <int L>
hipError_t launch_one(const float2* in, float2* out, const float2* tw_all, const int64_t* offs, int count,
                      bool inverse, hipStream_t st) {
    constexpr int N = 1 << L;
    if (count <= 0) return hipSuccess;
    if constexpr (L >= 13) {
        if (inverse)
            hipLaunchKernelGGL((k_ragged_half<N, true>), dim3((unsigned)count), dim3(N / 32), 0, st, in, out,
                               tw_all + table_off(L), offs);
        else
            hipLaunchKernelGGL((k_ragged_half<N, false>), dim3((unsigned)count), dim3(N / 32), 0, st, in, out,
                               tw_all + table_off(L), offs);
    } else {
        if (inverse)
            hipLaunchKernelGGL((k_ragged_one<N, true>), dim3((unsigned)count), dim3(N / 16), 0, st, in, out,
                               tw_all + table_off(L), offs);
        else
            hipLaunchKernelGGL((k_ragged_one<N, false>), dim3((unsigned)count), dim3(N / 16), 0, st, in, out,
                               tw_all + table_off(L), offs);
    }
    return hipGetLastError();
}

// Launches the class kernel over tiles [tile_begin, tile_end) (all of one
// class) with as many workgroups as are resident on the device at once.
template <int TILE, int LMIN, int LMAX>
hipError_t launch_class(const float2* in, float2* out, const float2* tw_all, const RaggedTile* tiles,
                        const int64_t* offs, const RaggedRanges& rr, int tile_begin, int tile_end,
                        bool inverse, hipStream_t st) {
    const int ntiles = tile_end - tile_begin;
    if (ntiles <= 0) return hipSuccess;
    static const int per_cu = [] {
        int b = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, k_ragged<false, TILE, LMIN, LMAX>, TILE / 16, 0) !=
                hipSuccess ||
            b < 1)
            b = 1;
        return b;
    }();
    const int grid = std::min(ntiles, ragged_cus() * per_cu);
    if (inverse)
        hipLaunchKernelGGL((k_ragged<true, TILE, LMIN, LMAX>), dim3(grid), dim3(TILE / 16), 0, st, in, out,
                           tw_all, tiles, offs, rr, tile_begin, tile_end);
    else
        hipLaunchKernelGGL((k_ragged<false, TILE, LMIN, LMAX>), dim3(grid), dim3(TILE / 16), 0, st, in, out,
                           tw_all, tiles, offs, rr, tile_begin, tile_end);
    return hipGetLastError();
}

// ---------------------------------------------------------------- tickets
// The one-launch forms run 512-thread workgroups over a list of tickets.  A
// ticket is the transforms of one length 2^L that one 512-thread pass holds,
// in one of two tile shapes (the shape per length is measured at plan
// creation):
//   16384         : one transform, half-LDS form (32 values per thread);
//   8192          : two transforms, half form on 256 threads each;
//   <= 4096, E=16 : 8192 elements, 16 values per thread (radix-16 passes,
//                   packed twiddle rows);
//   <= 4096, E=8  : 4096 elements, 8 values per thread (radix-8 passes,
//                   table twiddles), more threads per transform.
// All within 64 KiB of LDS and 128 VGPRs: two workgroups per CU.  A ticket
// code is log2 length + 16 * shape; the list is sorted by length, longest
// first.
constexpr int ticket_code(int l, int shape) { return l + 16 * shape; }
constexpr int ticket_count(int l, int shape) {
    return l == 14 ? 1 : l == 13 ? 2 : (shape ? 4096 : 8192) >> l;
}

// The next-ticket draw of lane 0 in the work-queue form: issued by draw()
// before a ticket's data loads, written to its LDS slot by publish() right
// after them (its latency hides behind the loads and its register is not
// held through the passes).  A no-op in the one-shot form.
struct TicketDraw {
    int* counter;
    int* slot;
    int value;
    __device__ __forceinline__ void draw() {
        if (threadIdx.x == 0) value = atomicAdd(counter, 1);
    }
    __device__ __forceinline__ void publish() {
        if (threadIdx.x == 0) *slot = value;
    }
};
// The one-shot form draws nothing (no object on the stack: a kernel with a
// private segment, even an unused one, is dispatched with scratch set up).
struct NoDraw {
    __device__ __forceinline__ void publish() {}
};

// With T >= 64 threads per transform every wave works on one transform:
// its offset moves to scalar registers, so the loads and stores address
// from a scalar base (no 64-bit pointer per lane).
template <bool UNIFORM>
__device__ __forceinline__ int64_t wave_uniform(int64_t v) {
    if constexpr (UNIFORM) {
        const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
        const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)((uint64_t)v >> 32));
        return (int64_t)(((uint64_t)hi << 32) | lo);
    } else {
        return v;
    }
}

// One ticket: `count` transforms of length 2^L at sorted offsets
// offs[first ...]; a slot past `count` transforms the ticket's last
// transform again and stores nothing.
template <int L, int SHAPE, bool INV, class Draw>
__device__ __forceinline__ void ticket_body(const float2* in, float2* out, const float2* tw,
                                            const int64_t* offs, int64_t first, int count, float2* lds,
                                            Draw& td) {
    constexpr int N = 1 << L;
    const int tid = fresh((int)threadIdx.x);  // index arithmetic redone per ticket, not hoisted
    if constexpr (L >= 13) {
        // half-LDS form (k_stockham_half's body): 16384 on all 512 threads,
        // 8192 on each 256-thread half
        constexpr int E = 32;
        const int h = L == 14 ? 0 : tid >> 8;
        const int t = L == 14 ? tid : tid & 255;
        const bool active = h < count;
        const int64_t off = wave_uniform<true>(offs[first + (active ? h : count - 1)]);
        float2 v[E];
        load_column<N, E, INV, true>(v, in + off, t, true);
        td.publish();
        float2 w[E / 16 * 4];
        half_bases<N, E, 1>(w, tw, t);
        stockham_half<N, E, LdsSwz, 0>(v, lds + h * (N / 2), t, tw, w);
        if (active) store_column<N, E, INV, true>(v, out + off, t);
    } else {
        constexpr int E = SHAPE ? 8 : 16;
        constexpr int T = N / E;
        const int local = tid / T, t = tid % T;
        const int64_t off = wave_uniform<T >= 64>(offs[first + (local < count ? local : count - 1)]);
        float2 v[E];
        load_column<N, E, INV, true>(v, in + off, t, true);
        td.publish();
        if constexpr (E == 16) {
            PreBases<N, 16> pb;
            pb.load_packed(tw, t);
            stockham_regs<N, 16, LdsSwz, 0, false, true, false, 2>(v, lds + local * N, t, tw, pb.w);
        } else {
            stockham_regs<N, 8, LdsSwz, 0, false, true, false, 0>(v, lds + local * N, t, tw);
        }
        if (local < count) store_column<N, E, INV, true>(v, out + off, t);
    }
}

template <bool INV, class Draw>
__device__ __forceinline__ void run_ticket(const RaggedTile& d, const float2* in, float2* out,
                                           const float2* tw_all, const int64_t* offs, float2* lds, Draw& td) {
    const int64_t first = d.first;
    const int count = d.count;
    // the table offsets as constants: a plain call of the recursive
    // table_off here is emitted as a device call (stack, full wait on entry)
#define MIFFT_TK(L, S)                                                                                   \
    case ticket_code(L, S): {                                                                            \
        constexpr int off = table_off(L);                                                                \
        ticket_body<L, S, INV, Draw>(in, out, tw_all + off, offs, first, count, lds, td);               \
        break;                                                                                           \
    }
    switch (d.log2n) {
        MIFFT_TK(14, 0) MIFFT_TK(13, 0) MIFFT_TK(12, 0) MIFFT_TK(12, 1) MIFFT_TK(11, 0) MIFFT_TK(11, 1)
        MIFFT_TK(10, 0) MIFFT_TK(10, 1) MIFFT_TK(9, 0) MIFFT_TK(9, 1) MIFFT_TK(8, 0) MIFFT_TK(8, 1)
        MIFFT_TK(7, 0) MIFFT_TK(7, 1) MIFFT_TK(6, 0) MIFFT_TK(6, 1)
        default: break;
    }
#undef MIFFT_TK
}

// Default form: workgroup b runs ticket b of [begin, end) and exits.  The
// hardware dispatcher hands workgroups to CUs in order as slots free up, so
// the longest lengths start first and the short tickets fill the end.
template <bool INV>
__global__ void __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(4, 8)))
k_ragged_tickets(const float2* in, float2* out, const float2* __restrict__ tw_all,
                 const RaggedTile* __restrict__ tickets, const int64_t* __restrict__ offs, int begin) {
    __shared__ float2 lds[8192];
    NoDraw td;
    const RaggedTile d = tickets[begin + (int)blockIdx.x];
    run_ticket<INV>(d, in, out, tw_all, offs, lds, td);
}

// Work-queue form: a persistent grid (the resident workgroups) takes the
// tickets in list order from one device counter, each workgroup drawing its
// next ticket while it runs the current one.  Between tickets it meets at an
// LDS-only barrier.  A workgroup that draws past the end leaves; the last one
// to leave (a second counter) sets both counters back to zero, so the queue
// is ready for the next launch or graph replay without a memset.  (Measured
// slower than the one-shot form: on this GPU one in-order counter, vmcnt,
// covers loads and stores, so a persistent workgroup's next data wait also
// waits for the previous ticket's 64-128 KiB of stores, while one-shot
// workgroups overlap them with their successors' loads.)
template <bool INV>
__global__ void __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(4, 8)))
k_ragged_queue(const float2* in, float2* out, const float2* __restrict__ tw_all,
               const RaggedTile* __restrict__ tickets, const int64_t* __restrict__ offs, int ntickets,
               int* queue) {
    __shared__ float2 lds[8192];
    __shared__ int s_next[2];
    if (threadIdx.x == 0) s_next[0] = atomicAdd(&queue[0], 1);
    lds_barrier();
    int cur = __builtin_amdgcn_readfirstlane(s_next[0]);
    int par = 1;
    while (cur < ntickets) {  // uniform
        TicketDraw td{&queue[0], &s_next[par], 0};
        td.draw();
        run_ticket<INV>(tickets[cur], in, out, tw_all, offs, lds, td);
        // LDS only: this ticket's stores stay in flight.  The slot written in
        // this ticket is read after the barrier; the other slot, read after
        // the previous barrier, is written again only in the next ticket.
        lds_barrier();
        cur = __builtin_amdgcn_readfirstlane(s_next[par]);
        par ^= 1;
    }
    if (threadIdx.x == 0) {
        __threadfence();  // this workgroup's draws precede its exit count
        if (atomicAdd(&queue[1], 1) == (int)gridDim.x - 1) {
            atomicExch(&queue[0], 0);
            atomicExch(&queue[1], 0);
        }
    }
}

hipError_t launch_tickets(const float2* in, float2* out, const float2* tw_all, const RaggedTile* tickets,
                          const int64_t* offs, int begin, int end, bool inverse, hipStream_t st) {
    if (end <= begin) return hipSuccess;
    if (inverse)
        hipLaunchKernelGGL(k_ragged_tickets<true>, dim3(end - begin), dim3(512), 0, st, in, out, tw_all, tickets,
                           offs, begin);
    else
        hipLaunchKernelGGL(k_ragged_tickets<false>, dim3(end - begin), dim3(512), 0, st, in, out, tw_all, tickets,
                           offs, begin);
    return hipGetLastError();
}

hipError_t launch_queue(const float2* in, float2* out, const float2* tw_all, const RaggedTile* tickets,
                        const int64_t* offs, int ntickets, int* queue, bool inverse, hipStream_t st) {
    if (ntickets <= 0) return hipSuccess;
    static const int per_cu = [] {
        int b = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, k_ragged_queue<false>, 512, 0) != hipSuccess || b < 1)
            b = 1;
        return b;
    }();
    const int grid = std::min(ntickets, ragged_cus() * per_cu);
    if (inverse)
        hipLaunchKernelGGL(k_ragged_queue<true>, dim3(grid), dim3(512), 0, st, in, out, tw_all, tickets, offs,
                           ntickets, queue);
    else
        hipLaunchKernelGGL(k_ragged_queue<false>, dim3(grid), dim3(512), 0, st, in, out, tw_all, tickets, offs,
                           ntickets, queue);
    return hipGetLastError();
}

}  // namespace

int ragged_twiddles_len() { return kTablesLen; }

hipError_t launch_ragged_twiddles(float2* tw_all, hipStream_t st) {
    for (int l = 6; l <= 14; ++l) {
        const hipError_t e = launch_twiddles(tw_all + table_off(l), 1 << l, st);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

// Lengths 2^l >= 4096: the `count` transforms whose sorted offsets start at
// offs (one launch).
hipError_t launch_ragged_one(int l, const float2* in, float2* out, const float2* tw_all, const int64_t* offs,
                             int count, bool inverse, hipStream_t st) {
    switch (l) {
        case 12: return launch_one<12>(in, out, tw_all, offs, count, inverse, st);
        case 13: return launch_one<13>(in, out, tw_all, offs, count, inverse, st);
        case 14: return launch_one<14>(in, out, tw_all, offs, count, inverse, st);
    }
    return hipErrorInvalidValue;
}

// Lengths 64 ... 2048: tiles [tile_begin, tile_end) of 4096 elements, one
// persistent launch.
hipError_t launch_ragged_tiles(const float2* in, float2* out, const float2* tw_all, const RaggedTile* tiles,
                               const int64_t* offs, const RaggedRanges& rr, int tile_begin, int tile_end,
                               bool inverse, hipStream_t st) {
    return launch_class<4096, 6, 11>(in, out, tw_all, tiles, offs, rr, tile_begin, tile_end, inverse, st);
}

}  // namespace mifft

using namespace mifft;

namespace {

mifft_status enqueue(mifft_ragged_plan* rp, const float2* in, float2* out, bool inverse, int mode,
                     hipStream_t st) {
    const RaggedTile* tiles = static_cast<const RaggedTile*>(rp->d_tiles.ptr);
    const int64_t* offs = static_cast<const int64_t*>(rp->d_offsets.ptr);
    const float2* tw = static_cast<const float2*>(rp->d_twiddles.ptr);
    const RaggedTile* tickets = static_cast<const RaggedTile*>(rp->d_tickets.ptr);
    if (mode & MIFFT_RAGGED_QUEUE)
        return check_hip(launch_queue(in, out, tw, tickets, offs, rp->ntickets, static_cast<int*>(rp->d_queue.ptr),
                                      inverse, st),
                         "ragged fft (work queue)");
    if (!(mode & MIFFT_RAGGED_SPLIT)) {
        if (!(mode & MIFFT_RAGGED_PER_SIZE))
            return check_hip(launch_tickets(in, out, tw, tickets, offs, 0, rp->ntickets, inverse, st), "ragged fft");
        for (size_t q = 0; q + 1 < rp->size_ticket_begin.size(); ++q) {
            const mifft_status r = check_hip(launch_tickets(in, out, tw, tickets, offs, rp->size_ticket_begin[q],
                                                            rp->size_ticket_begin[q + 1], inverse, st),
                                             "ragged fft (per size)");
            if (r != MIFFT_OK) return r;
        }
        return MIFFT_OK;
    }
    // lengths >= 4096: one launch each (their tiles hold one transform, so
    // the tile range is the sorted transform range)
    int small_begin = rp->ntiles;
    for (size_t s = 0; s + 1 < rp->size_tile_begin.size(); ++s) {
        const int l = rp->size_log2[s], b = rp->size_tile_begin[s], e = rp->size_tile_begin[s + 1];
        mifft_status r = MIFFT_OK;
        if (l >= 12)
            r = check_hip(launch_ragged_one(l, in, out, tw, offs + rp->size_first[s], e - b, inverse, st),
                          "ragged fft");
        else if (b < small_begin)
            small_begin = b;
        if (r != MIFFT_OK) return r;
    }
    if (small_begin < rp->ntiles)
        return check_hip(launch_ragged_tiles(in, out, tw, tiles, offs, rp->ranges, small_begin, rp->ntiles, inverse,
                                             st),
                         "ragged fft");
    return MIFFT_OK;
}

// Tickets of the one-launch forms: the transforms of each length (sorted
// order, longest length first) in runs of ticket_count(l, shape), and the
// first ticket of each length (plus the end).
std::vector<RaggedTile> build_tickets(const int64_t* lengths, const std::vector<int64_t>& order, int64_t count,
                                      const int* shape, std::vector<int>* begins) {
    std::vector<RaggedTile> t;
    int64_t i = 0;
    while (i < count) {
        const int64_t n = lengths[order[i]];
        int l = 0;
        while ((int64_t(1) << l) < n) ++l;
        int64_t j = i;
        while (j < count && lengths[order[j]] == n) ++j;
        const int sh = l <= 12 ? shape[l - 6] : 0;
        const int64_t per = ticket_count(l, sh);
        begins->push_back((int)t.size());
        for (int64_t a = i; a < j; a += per) {
            RaggedTile d;
            d.log2n = ticket_code(l, sh);
            d.count = (int)std::min<int64_t>(per, j - a);
            d.first = a;
            t.push_back(d);
        }
        i = j;
    }
    begins->push_back((int)t.size());
    return t;
}

// Tile shape of the one-launch forms for length 2^l (l <= 12): 16 or 8
// values per thread, timed as ticket launches over 2^24 / n transforms and
// recorded under "ragged fp32 c64 <n>".  Built-in choice: 8 values per
// thread for 64 and 128 (where the uniform kernels' radix-8 form is the
// faster one), 16 otherwise.
int choose_shape(int l, const float2* tw_all, hipStream_t st) {
    const int n = 1 << l;
    const int fallback = l <= 7 ? 1 : 0;
    if (!tuning_enabled()) return fallback;
    const int64_t batch = (int64_t(1) << 24) / n;
    std::vector<int64_t> offs(batch);
    for (int64_t i = 0; i < batch; ++i) offs[i] = i * n;
    std::vector<RaggedTile> tk[2];
    for (int sh = 0; sh < 2; ++sh)
        for (int64_t a = 0; a < batch; a += ticket_count(l, sh)) {
            RaggedTile d;
            d.log2n = ticket_code(l, sh);
            d.count = (int)std::min<int64_t>(ticket_count(l, sh), batch - a);
            d.first = a;
            tk[sh].push_back(d);
        }
    DeviceBuffer d_offs, d_tk0, d_tk1;
    if (d_offs.reserve(sizeof(int64_t) * batch) != MIFFT_OK ||
        d_tk0.reserve(sizeof(RaggedTile) * tk[0].size()) != MIFFT_OK ||
        d_tk1.reserve(sizeof(RaggedTile) * tk[1].size()) != MIFFT_OK ||
        hipMemcpy(d_offs.ptr, offs.data(), sizeof(int64_t) * batch, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(d_tk0.ptr, tk[0].data(), sizeof(RaggedTile) * tk[0].size(), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(d_tk1.ptr, tk[1].data(), sizeof(RaggedTile) * tk[1].size(), hipMemcpyHostToDevice) != hipSuccess) {
        (void)hipGetLastError();
        d_offs.release();
        d_tk0.release();
        d_tk1.release();
        return fallback;
    }
    const int64_t* o = static_cast<const int64_t*>(d_offs.ptr);
    const RaggedTile* t0 = static_cast<const RaggedTile*>(d_tk0.ptr);
    const RaggedTile* t1 = static_cast<const RaggedTile*>(d_tk1.ptr);
    const int n0 = (int)tk[0].size(), n1 = (int)tk[1].size();
    std::vector<Candidate> cand = {
        {"ragged_tile_e16",
         [=](const void* in, void* out, int64_t) {
             return launch_tickets(static_cast<const float2*>(in), static_cast<float2*>(out), tw_all, t0, o, 0, n0,
                                   false, st);
         },
         false},
        {"ragged_tile_e8",
         [=](const void* in, void* out, int64_t) {
             return launch_tickets(static_cast<const float2*>(in), static_cast<float2*>(out), tw_all, t1, o, 0, n1,
                                   false, st);
         },
         false}};
    const int q = tuned_choice("ragged fp32 c64 " + std::to_string(n), n, 8, cand, st, "ragged");
    (void)hipStreamSynchronize(st);
    d_offs.release();
    d_tk0.release();
    d_tk1.release();
    return q < 0 ? fallback : q;
}

void drop_slot(mifft_ragged_plan::GraphSlot& g) {
    if (g.exec) (void)hipGraphExecDestroy(g.exec);
    if (g.graph) (void)hipGraphDestroy(g.graph);
    g = mifft_ragged_plan::GraphSlot{};
}

void drop_graphs(mifft_ragged_plan* rp) {
    for (auto& g : rp->graphs) drop_slot(g);
}

}  // namespace

extern "C" {

mifft_status mifft_ragged_plan_create(const int64_t* lengths, const int64_t* offsets, int64_t count,
                                      mifft_ragged_plan** out) {
    if (!out) return fail(MIFFT_ERR_INVALID_ARGUMENT, "null plan pointer");
    *out = nullptr;
    if (count < 0) return fail(MIFFT_ERR_INVALID_ARGUMENT, "negative count");
    if (count > 0 && !lengths) return fail(MIFFT_ERR_INVALID_ARGUMENT, "null lengths");
    std::vector<int64_t> offs(count);
    int64_t pos = 0, extent = 0;
    for (int64_t i = 0; i < count; ++i) {
        const int64_t n = lengths[i];
        if (!supported_size(n))
            return fail(MIFFT_ERR_UNSUPPORTED_SIZE,
                        "Batch element " + std::to_string(i) + ": unsupported FFT size " +
                            std::to_string(n),
                        -1, n, i);
        offs[i] = offsets ? offsets[i] : pos;
        if (offs[i] < 0) return fail(MIFFT_ERR_INVALID_ARGUMENT, "negative offset");
        pos += n;
        extent = std::max(extent, offs[i] + n);
    }
    mifft_status s = context_init();
    if (s != MIFFT_OK) return s;

    // order: length descending, then input order
    std::vector<int64_t> order(count);
    for (int64_t i = 0; i < count; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(),
                     [&](int64_t a, int64_t b) { return lengths[a] > lengths[b]; });
    std::vector<int64_t> sorted_offs(count);
    std::vector<RaggedTile> tiles;
    std::vector<int> begins, log2s;
    int64_t i = 0;
    while (i < count) {
        const int64_t n = lengths[order[i]];
        int l = 0;
        while ((int64_t(1) << l) < n) ++l;
        begins.push_back((int)tiles.size());
        log2s.push_back(l);
        const int per_tile = (int)(tile_elems(l) / n);
        int64_t j = i;
        while (j < count && lengths[order[j]] == n) ++j;
        for (int64_t a = i; a < j; a += per_tile) {
            RaggedTile t;
            t.log2n = l;
            t.count = (int)std::min<int64_t>(per_tile, j - a);
            t.first = a;
            tiles.push_back(t);
        }
        for (int64_t a = i; a < j; ++a) sorted_offs[a] = offs[order[a]];
        i = j;
    }
    begins.push_back((int)tiles.size());

    mifft_ragged_plan* rp = new mifft_ragged_plan();
    rp->count = count;
    rp->extent = extent;
    rp->ntiles = (int)tiles.size();
    rp->size_tile_begin = begins;
    rp->size_log2 = log2s;
    for (size_t q = 0; q + 1 < begins.size(); ++q) {
        const int l = log2s[q];
        rp->ranges.lo[l - 6] = begins[q];
        rp->ranges.hi[l - 6] = begins[q + 1];
        rp->size_first.push_back(tiles[(size_t)begins[q]].first);
    }
    for (int64_t a = 0; a < count; ++a) rp->elements += lengths[a];
    if (hipStreamCreateWithFlags(&rp->stream, hipStreamNonBlocking) != hipSuccess) {
        mifft_ragged_plan_destroy(rp);
        return fail(MIFFT_ERR_NO_QUEUE, "Could not create HIP stream");
    }
    if ((s = rp->d_queue.reserve(2 * sizeof(int))) != MIFFT_OK ||
        (s = check_hip(hipMemset(rp->d_queue.ptr, 0, 2 * sizeof(int)), "hipMemset")) != MIFFT_OK ||
        (s = rp->d_tiles.reserve(sizeof(RaggedTile) * std::max<size_t>(1, tiles.size()))) != MIFFT_OK ||
        (s = rp->d_offsets.reserve(sizeof(int64_t) * std::max<int64_t>(1, count))) != MIFFT_OK ||
        (s = rp->d_twiddles.reserve(sizeof(float2) * ragged_twiddles_len())) != MIFFT_OK) {
        mifft_ragged_plan_destroy(rp);
        return s;
    }
    if ((!tiles.empty() &&
         (s = check_hip(hipMemcpy(rp->d_tiles.ptr, tiles.data(), sizeof(RaggedTile) * tiles.size(),
                                  hipMemcpyHostToDevice),
                        "hipMemcpy")) != MIFFT_OK) ||
        (count > 0 && (s = check_hip(hipMemcpy(rp->d_offsets.ptr, sorted_offs.data(),
                                               sizeof(int64_t) * count, hipMemcpyHostToDevice),
                                     "hipMemcpy")) != MIFFT_OK) ||
        (s = check_hip(launch_ragged_twiddles(static_cast<float2*>(rp->d_twiddles.ptr), rp->stream),
                       "twiddles")) != MIFFT_OK ||
        (s = check_hip(hipStreamSynchronize(rp->stream), "hipStreamSynchronize")) != MIFFT_OK) {
        mifft_ragged_plan_destroy(rp);
        return s;
    }
    // tile shape per length present (measured once per length and process)
    for (int l : log2s)
        if (l <= 12) rp->shape[l - 6] = choose_shape(l, static_cast<const float2*>(rp->d_twiddles.ptr), rp->stream);
    std::vector<RaggedTile> tickets = build_tickets(lengths, order, count, rp->shape, &rp->size_ticket_begin);
    rp->ntickets = (int)tickets.size();
    if ((s = rp->d_tickets.reserve(sizeof(RaggedTile) * std::max<size_t>(1, tickets.size()))) != MIFFT_OK ||
        (!tickets.empty() &&
         (s = check_hip(hipMemcpy(rp->d_tickets.ptr, tickets.data(), sizeof(RaggedTile) * tickets.size(),
                                  hipMemcpyHostToDevice),
                        "hipMemcpy")) != MIFFT_OK)) {
        mifft_ragged_plan_destroy(rp);
        return s;
    }
    *out = rp;
    return MIFFT_OK;
}

mifft_status mifft_ragged_plan_destroy(mifft_ragged_plan* rp) {
    if (!rp) return MIFFT_OK;
    if (rp->stream) (void)hipStreamSynchronize(rp->stream);
    drop_graphs(rp);
    rp->d_tiles.release();
    rp->d_tickets.release();
    rp->d_queue.release();
    rp->d_offsets.release();
    rp->d_twiddles.release();
    if (rp->stream) (void)hipStreamDestroy(rp->stream);
    delete rp;
    return MIFFT_OK;
}

int64_t mifft_ragged_extent(const mifft_ragged_plan* rp) { return rp ? rp->extent : -1; }
int64_t mifft_ragged_elements(const mifft_ragged_plan* rp) { return rp ? rp->elements : -1; }
int64_t mifft_ragged_launches(const mifft_ragged_plan* rp, int mode) {
    if (!rp) return -1;
    if (rp->ntiles == 0) return 0;
    if (mode & MIFFT_RAGGED_PER_SIZE) return (int64_t)rp->size_tile_begin.size() - 1;
    if (!(mode & MIFFT_RAGGED_SPLIT)) return 1;  // tickets or work queue
    int64_t n = 0;
    bool small = false;
    for (int l : rp->size_log2) {
        if (l >= 12) ++n;
        else small = true;
    }
    return n + (small ? 1 : 0);
}

mifft_status mifft_ragged_execute_device(mifft_ragged_plan* rp, const float* in, float* out,
                                         mifft_direction dir, int mode, void* stream) {
    if (!rp) return fail(MIFFT_ERR_INVALID_ARGUMENT, "null plan");
    if (dir != MIFFT_FORWARD && dir != MIFFT_INVERSE)
        return fail(MIFFT_ERR_INVALID_ARGUMENT, "direction must be MIFFT_FORWARD or MIFFT_INVERSE");
    if (mode & ~(MIFFT_RAGGED_PER_SIZE | MIFFT_RAGGED_GRAPH | MIFFT_RAGGED_QUEUE | MIFFT_RAGGED_SPLIT))
        return fail(MIFFT_ERR_INVALID_ARGUMENT, "unknown mode bits");
    if (((mode & MIFFT_RAGGED_PER_SIZE) != 0) + ((mode & MIFFT_RAGGED_QUEUE) != 0) +
            ((mode & MIFFT_RAGGED_SPLIT) != 0) > 1)
        return fail(MIFFT_ERR_INVALID_ARGUMENT,
                    "MIFFT_RAGGED_PER_SIZE, MIFFT_RAGGED_QUEUE and MIFFT_RAGGED_SPLIT exclude each other");
    if (rp->count == 0) return MIFFT_OK;
    if (!in || !out) return fail(MIFFT_ERR_INVALID_ARGUMENT, "null buffer");
    hipStream_t st = static_cast<hipStream_t>(stream);
    const float2* i2 = reinterpret_cast<const float2*>(in);
    float2* o2 = reinterpret_cast<float2*>(out);
    const bool inverse = dir == MIFFT_INVERSE;
    if (!(mode & MIFFT_RAGGED_GRAPH)) return enqueue(rp, i2, o2, inverse, mode, st);

    // Graph form: captured once per (buffers, direction, mode), replayed after.
    mifft_ragged_plan::GraphSlot* slot = nullptr;
    for (auto& g : rp->graphs)
        if (g.exec && g.in == in && g.out == out && g.dir == dir && g.mode == mode) slot = &g;
    if (!slot) {
        slot = &rp->graphs[0];  // an empty slot, else the least recently used
        for (auto& g : rp->graphs)
            if (!g.exec || (slot->exec && g.used < slot->used)) slot = &g;
        drop_slot(*slot);
        hipStream_t cap = rp->stream;
        mifft_status s =
            check_hip(hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
        if (s != MIFFT_OK) return s;
        mifft_status se = enqueue(rp, i2, o2, inverse, mode & ~MIFFT_RAGGED_GRAPH, cap);
        hipGraph_t g = nullptr;
        s = check_hip(hipStreamEndCapture(cap, &g), "hipStreamEndCapture");
        if (se != MIFFT_OK || s != MIFFT_OK) {
            if (g) (void)hipGraphDestroy(g);
            return se != MIFFT_OK ? se : s;
        }
        slot->graph = g;
        if ((s = check_hip(hipGraphInstantiate(&slot->exec, g, nullptr, nullptr, 0), "hipGraphInstantiate")) !=
            MIFFT_OK) {
            drop_slot(*slot);
            return s;
        }
        slot->in = in;
        slot->out = out;
        slot->dir = dir;
        slot->mode = mode;
    }
    slot->used = ++rp->graph_clock;
    return check_hip(hipGraphLaunch(slot->exec, st), "hipGraphLaunch");
}

}  // extern "C"
