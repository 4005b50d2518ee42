// Library-internal runtime: error state, device context, plan object.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "../../include/mifft.h"
#include "kernels.hpp"

namespace mifft {

// Records the calling thread's last error and returns `s`.
mifft_status fail(mifft_status s, const std::string& msg, int64_t expected = -1, int64_t got = -1,
                  int64_t index = -1);
mifft_status ok();
// Maps a HIP error to MIFFT_ERR_COMMAND_FAILED (allocation errors to
// MIFFT_ERR_ALLOCATION) with the HIP error text; hipSuccess -> MIFFT_OK.
mifft_status check_hip(hipError_t e, const char* what);

bool supported_size(int64_t n);

// Process-wide device context (thread-safe, idempotent).
mifft_status context_init();

// Device buffer that grows on demand (host-pointer entry points).
struct DeviceBuffer {
    void* ptr = nullptr;
    size_t bytes = 0;
    mifft_status reserve(size_t want);
    void release();
};

// Kernel variant per size: picked by the autotuner (timed on first use and
// cached on disk) or the built-in default.
const StockhamVariant* choose_variant(int n, hipStream_t stream, const float2* tw);

}  // namespace mifft

struct mifft_plan {
    int64_t n = 0;
    int precision = 0;  // mifft_precision
    int io = 0;         // mifft_io
    const mifft::StockhamVariant* variant = nullptr;
    float2* twiddles = nullptr;  // W_N^k, k < N, then pass rows (twiddle_table_len)
    float2* real_tw = nullptr;   // real-input transforms (real_table_len)
    hipStream_t stream = nullptr;
    mifft::DeviceBuffer in, out;
};
