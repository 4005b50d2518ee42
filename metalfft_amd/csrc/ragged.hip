// Ragged batches: thousands of complex transforms of different power-of-two
// lengths (64 ... 16384) in one persistent launch.
//
// No reference counterpart: MetalFFT.forward(batch:) rejects uneven lengths
// (Sources/MetalFFT/FFT/MetalFFT.swift:108-114).
//
// Lengths of 4096 and longer run one transform per workgroup, one launch per
// length (k_ragged_one, and k_ragged_half for 8192 and 16384: the uniform
// kernels' bodies at a per-transform offset).  Shorter lengths are cut into tiles of 4096 elements (256 threads,
// 32 KiB of LDS): 2 transforms of 2048, ..., 64 of 64 (the last tile of a
// length may hold fewer); length N uses N/16 threads per transform
// (radix-16 passes, 16 values per thread), so every tile keeps all its
// threads and LDS busy whatever its length.  Tiles are sorted by length,
// largest first, and one persistent launch (as many workgroups as are
// resident) strides over them.  The same kernels launched once per length
// are the one-launch-per-size baseline.  Both forms can be captured in a
// hipGraph and replayed.
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

// ------------------------------------------------------------- work queue
// One persistent launch of 512-thread workgroups for every length.  A
// ticket is `count` transforms of one length 2^L, taken in units of what one
// 512-thread pass holds (queue_unit: 16384: one, half-LDS form with every
// thread; 8192: two, half form on 256 threads each; 4096 and shorter: 8192
// elements, 16 values per thread), all within 64 KiB of LDS.  Tickets are
// sorted by length, largest first; each workgroup takes the next ticket
// from one device counter (an atomic increment by one lane, broadcast
// through LDS), drawing its following ticket while it works on the current
// one.  A workgroup that draws past the end leaves; the last one to leave
// (a second counter) sets both counters back to zero, so the queue is ready
// for the next launch or graph replay without a memset.
constexpr int queue_unit(int l) { return l == 14 ? 1 : l == 13 ? 2 : 8192 >> l; }

template <int L, bool INV>
__device__ __forceinline__ void queue_unit_body(const float2* in, float2* out, const float2* tw,
                                                const int64_t* offs, int64_t first, int count, float2* lds) {
    constexpr int N = 1 << L;
    const int tid = fresh((int)threadIdx.x);  // index arithmetic redone per unit, not hoisted
    if constexpr (L == 14) {
        const int64_t off = offs[first];
        half_transform<N, INV, true>(in + off, out + off, tw, lds, tid, true);
    } else if constexpr (L == 13) {
        const int h = tid >> 8, t = tid & 255;
        const int64_t off = offs[first + (h < count ? h : count - 1)];
        half_transform<N, INV, true>(in + off, out + off, tw, lds + h * (N / 2), t, h < count);
    } else {
        constexpr int T = N / 16;
        const int local = tid / T, t = tid % T;
        // a slot past `count` transforms the unit's last transform again and
        // stores nothing
        const int64_t off = offs[first + (local < count ? local : count - 1)];
        float2 v[16];
        load_column<N, 16, INV, true>(v, in + off, t, true);
        PreBases<N, 16> pb;
        pb.load_packed(tw, t);
        stockham_regs<N, 16, LdsSwz, 0, true, true, false, 2>(v, lds + local * N, t, tw, pb.w);
        if (local < count) store_column<N, 16, INV, true>(v, out + off, t);
    }
}

template <int L, bool INV>
__device__ __forceinline__ void queue_ticket(const float2* in, float2* out, const float2* tw_all,
                                             const int64_t* offs, int64_t first, int count, float2* lds) {
    constexpr int U = queue_unit(L);
    constexpr int kOff = table_off(L);
    for (int b = 0; b < count; b += U) {  // count is uniform
        if (b > 0) lds_barrier();        // the previous unit's LDS reads are done
        queue_unit_body<L, INV>(in, out, tw_all + kOff, offs, first + b, count - b < U ? count - b : U, lds);
    }
}

template <bool INV>
__global__ void __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(4, 8)))
k_ragged_queue(const float2* in, float2* out, const float2* __restrict__ tw_all,
               const RaggedTile* __restrict__ tickets, const int64_t* __restrict__ offs, int ntickets,
               int* queue) {
    __shared__ float2 lds[8192];
    __shared__ int s_next;
    if (threadIdx.x == 0) s_next = atomicAdd(&queue[0], 1);
    __syncthreads();
    int cur = __builtin_amdgcn_readfirstlane(s_next);
    while (cur < ntickets) {  // uniform
        int nxt = 0;
        if (threadIdx.x == 0) nxt = atomicAdd(&queue[0], 1);  // in flight during this ticket
        const RaggedTile d = tickets[cur];
        const int64_t first = d.first;
        const int count = d.count;
        switch (d.log2n) {
            case 14: queue_ticket<14, INV>(in, out, tw_all, offs, first, count, lds); break;
            case 13: queue_ticket<13, INV>(in, out, tw_all, offs, first, count, lds); break;
            case 12: queue_ticket<12, INV>(in, out, tw_all, offs, first, count, lds); break;
            case 11: queue_ticket<11, INV>(in, out, tw_all, offs, first, count, lds); break;
            case 10: queue_ticket<10, INV>(in, out, tw_all, offs, first, count, lds); break;
            case 9: queue_ticket<9, INV>(in, out, tw_all, offs, first, count, lds); break;
            case 8: queue_ticket<8, INV>(in, out, tw_all, offs, first, count, lds); break;
            case 7: queue_ticket<7, INV>(in, out, tw_all, offs, first, count, lds); break;
            default: queue_ticket<6, INV>(in, out, tw_all, offs, first, count, lds); break;
        }
        __syncthreads();  // every read of s_next and of this ticket's LDS is done
        if (threadIdx.x == 0) s_next = nxt;
        __syncthreads();
        cur = __builtin_amdgcn_readfirstlane(s_next);
    }
    if (threadIdx.x == 0) {
        __threadfence();  // this workgroup's draws precede its exit count
        if (atomicAdd(&queue[1], 1) == (int)gridDim.x - 1) {
            atomicExch(&queue[0], 0);
            atomicExch(&queue[1], 0);
        }
    }
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
    if (mode & MIFFT_RAGGED_QUEUE)
        return check_hip(launch_queue(in, out, tw, static_cast<const RaggedTile*>(rp->d_tickets.ptr), offs,
                                      rp->ntickets, static_cast<int*>(rp->d_queue.ptr), inverse, st),
                         "ragged fft (work queue)");
    // lengths >= 4096: one launch each (their tiles hold one transform, so
    // the tile range is the sorted transform range)
    int small_begin = rp->ntiles;
    for (size_t s = 0; s + 1 < rp->size_tile_begin.size(); ++s) {
        const int l = rp->size_log2[s], b = rp->size_tile_begin[s], e = rp->size_tile_begin[s + 1];
        mifft_status r = MIFFT_OK;
        if (l >= 12)
            r = check_hip(launch_ragged_one(l, in, out, tw, offs + rp->size_first[s], e - b, inverse, st),
                          "ragged fft");
        else if (mode & MIFFT_RAGGED_PER_SIZE)
            r = check_hip(launch_ragged_tiles(in, out, tw, tiles, offs, rp->ranges, b, e, inverse, st),
                          "ragged fft (per size)");
        else if (b < small_begin)
            small_begin = b;
        if (r != MIFFT_OK) return r;
    }
    if ((mode & MIFFT_RAGGED_PER_SIZE) == 0 && small_begin < rp->ntiles)
        return check_hip(launch_ragged_tiles(in, out, tw, tiles, offs, rp->ranges, small_begin, rp->ntiles, inverse,
                                             st),
                         "ragged fft");
    return MIFFT_OK;
}

// Work-queue tickets: the transforms of each length (sorted order, largest
// length first) in runs of units_per_ticket[l - 6] * queue_unit(l).
std::vector<RaggedTile> build_tickets(const int64_t* lengths, const std::vector<int64_t>& order, int64_t count,
                                      const int* units_per_ticket) {
    std::vector<RaggedTile> t;
    int64_t i = 0;
    while (i < count) {
        const int64_t n = lengths[order[i]];
        int l = 0;
        while ((int64_t(1) << l) < n) ++l;
        int64_t j = i;
        while (j < count && lengths[order[j]] == n) ++j;
        const int64_t per = (int64_t)queue_unit(l) * units_per_ticket[l - 6];
        for (int64_t a = i; a < j; a += per) {
            RaggedTile d;
            d.log2n = l;
            d.count = (int)std::min<int64_t>(per, j - a);
            d.first = a;
            t.push_back(d);
        }
        i = j;
    }
    return t;
}

void drop_graph(mifft_ragged_plan* rp) {
    if (rp->graph_exec) (void)hipGraphExecDestroy(rp->graph_exec);
    if (rp->graph) (void)hipGraphDestroy(rp->graph);
    rp->graph_exec = nullptr;
    rp->graph = nullptr;
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
    std::vector<RaggedTile> tickets = build_tickets(lengths, order, count, rp->units_per_ticket);
    rp->ntickets = (int)tickets.size();
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
    if ((s = rp->d_tickets.reserve(sizeof(RaggedTile) * std::max<size_t>(1, tickets.size()))) != MIFFT_OK ||
        (s = rp->d_queue.reserve(2 * sizeof(int))) != MIFFT_OK ||
        (s = check_hip(hipMemset(rp->d_queue.ptr, 0, 2 * sizeof(int)), "hipMemset")) != MIFFT_OK ||
        (!tickets.empty() &&
         (s = check_hip(hipMemcpy(rp->d_tickets.ptr, tickets.data(), sizeof(RaggedTile) * tickets.size(),
                                  hipMemcpyHostToDevice),
                        "hipMemcpy")) != MIFFT_OK) ||
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
    *out = rp;
    return MIFFT_OK;
}

mifft_status mifft_ragged_plan_destroy(mifft_ragged_plan* rp) {
    if (!rp) return MIFFT_OK;
    if (rp->stream) (void)hipStreamSynchronize(rp->stream);
    drop_graph(rp);
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
    if (mode & MIFFT_RAGGED_QUEUE) return 1;
    if (mode & MIFFT_RAGGED_PER_SIZE) return (int64_t)rp->size_tile_begin.size() - 1;
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
    if (mode & ~(MIFFT_RAGGED_PER_SIZE | MIFFT_RAGGED_GRAPH | MIFFT_RAGGED_QUEUE))
        return fail(MIFFT_ERR_INVALID_ARGUMENT, "unknown mode bits");
    if ((mode & MIFFT_RAGGED_PER_SIZE) && (mode & MIFFT_RAGGED_QUEUE))
        return fail(MIFFT_ERR_INVALID_ARGUMENT, "MIFFT_RAGGED_PER_SIZE and MIFFT_RAGGED_QUEUE exclude each other");
    if (rp->count == 0) return MIFFT_OK;
    if (!in || !out) return fail(MIFFT_ERR_INVALID_ARGUMENT, "null buffer");
    hipStream_t st = static_cast<hipStream_t>(stream);
    const float2* i2 = reinterpret_cast<const float2*>(in);
    float2* o2 = reinterpret_cast<float2*>(out);
    const bool inverse = dir == MIFFT_INVERSE;
    if (!(mode & MIFFT_RAGGED_GRAPH)) return enqueue(rp, i2, o2, inverse, mode, st);

    // Graph form: captured once per (buffers, direction, mode), replayed after.
    if (!rp->graph_exec || rp->g_in != in || rp->g_out != out || rp->g_dir != dir ||
        rp->g_mode != mode) {
        drop_graph(rp);
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
        rp->graph = g;
        if ((s = check_hip(hipGraphInstantiate(&rp->graph_exec, g, nullptr, nullptr, 0),
                           "hipGraphInstantiate")) != MIFFT_OK) {
            drop_graph(rp);
            return s;
        }
        rp->g_in = in;
        rp->g_out = out;
        rp->g_dir = dir;
        rp->g_mode = mode;
    }
    return check_hip(hipGraphLaunch(rp->graph_exec, st), "hipGraphLaunch");
}

}  // extern "C"
