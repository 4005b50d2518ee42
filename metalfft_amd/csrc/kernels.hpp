// Host-visible launchers of the gfx950 kernels (one translation unit each).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mifft {

using StockhamLaunch = hipError_t (*)(const float2* in, float2* out, const float2* tw,
                                      int64_t batch, bool inverse, hipStream_t stream);

// One compiled LDS Stockham kernel: length n, e complex values per thread
// (largest radix), tpw transforms per workgroup.
struct StockhamVariant {
    int n;
    int e;
    int tpw;
    StockhamLaunch launch;
    const char* name;
};

const StockhamVariant* stockham_variants(int* count);
const StockhamVariant* find_stockham(int n, int e);
// Plan twiddle memory: W_N^k for k < n, then the packed per-pass bases rows
// of the radix-16 schedule (PassTable<n, 16, 0>) the "_pt" variants read.
// twiddle_table_len(n) entries; launch_twiddles fills all of them.
int twiddle_table_len(int n);
hipError_t launch_twiddles(float2* tw, int n, hipStream_t stream);

}  // namespace mifft
