// Ragged-batch plan shared by ragged.hip and the C-ABI.
#pragma once
#include <cstdint>
#include <vector>

#include "runtime.hpp"

namespace mifft {
// One workgroup-iteration: `count` transforms of length 2^log2n whose
// offsets are offs[first .. first + count).  In a one-launch ticket, log2n
// holds the ticket code (log2 length + 16 * tile shape).
struct RaggedTile {
    int32_t log2n;
    int32_t count;
    int64_t first;
};
// Tiles of length 2^(6 + i) are [lo[i], hi[i]) of the sorted tile list
// (empty: lo == hi == 0).
struct RaggedRanges {
    int32_t lo[9];
    int32_t hi[9];
};
}  // namespace mifft

struct mifft_ragged_plan {
    int64_t count = 0;     // transforms
    int64_t elements = 0;  // sum of lengths
    int64_t extent = 0;    // max(offset + length): buffer length needed
    int ntiles = 0;
    std::vector<int> size_tile_begin;  // tile range per length, largest length first
    std::vector<int> size_log2;        // log2 length of each of those ranges
    mifft::RaggedRanges ranges = {};   // the same ranges indexed by log2 length
    std::vector<int64_t> size_first;   // sorted index of each length's first transform
    mifft::DeviceBuffer d_tiles, d_offsets, d_twiddles;
    // one-launch forms: tickets (code, count, first) sorted longest length
    // first, the ticket range of each length (aligned with size_log2), the
    // tile shape per length (0: 16 values per thread, 1: 8; by log2 - 6) and
    // the work-queue form's two device counters (next ticket, workgroups done)
    int ntickets = 0;
    std::vector<int> size_ticket_begin;
    int shape[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    mifft::DeviceBuffer d_tickets, d_queue;
    hipStream_t stream = nullptr;  // capture stream
    // graph form: one captured graph per (in, out, direction, mode), the
    // least recently used one replaced when all slots are taken (a forward
    // and an inverse on two buffer pairs alternate without re-capturing)
    struct GraphSlot {
        hipGraph_t graph = nullptr;
        hipGraphExec_t exec = nullptr;
        const void* in = nullptr;
        const void* out = nullptr;
        int dir = 0, mode = -1;
        uint64_t used = 0;
    };
    static constexpr int kGraphSlots = 4;
    GraphSlot graphs[kGraphSlots];
    uint64_t graph_clock = 0;
};
