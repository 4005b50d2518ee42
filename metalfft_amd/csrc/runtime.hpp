// Library-internal runtime: error state, device context, plan object.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <functional>
#include <string>
#include <vector>

#include "../../include/mifft.h"
#include "fourstep.hpp"
#include "kernels.hpp"

namespace mifft {

// Records the calling thread's last error and returns `s`.
mifft_status fail(mifft_status s, const std::string& msg, int64_t expected = -1, int64_t got = -1,
                  int64_t index = -1);
mifft_status ok();
// Maps a HIP error to MIFFT_ERR_COMMAND_FAILED (allocation errors to
// MIFFT_ERR_ALLOCATION) with the HIP error text; hipSuccess -> MIFFT_OK.
mifft_status check_hip(hipError_t e, const char* what);

// Lengths of the single-workgroup kernels (64 ... 16384): every object
// built on the transform (convolver, correlator, analysers, ragged, SAR).
bool supported_size(int64_t n);
// Lengths a plan accepts: those, and the four-step lengths 32768 ... 2^20.
bool plan_size(int64_t n);
// MIFFT_OK for supported_size(n), else the unsupportedFFTSize error.
mifft_status require_supported_size(int64_t n);

// Process-wide device context (thread-safe, idempotent).
mifft_status context_init();

// Device buffer that grows on demand.  A grown buffer does not free the
// allocation it replaces: that one moves to `retired` and is freed by
// release() (the owning object's destroy), because a hipGraph captured
// before the growth still holds its address and may be replayed after it.
// Growth is geometric (at least 1.5x), so the retired allocations add up to
// at most about twice the final size.
struct DeviceBuffer {
    void* ptr = nullptr;
    size_t bytes = 0;
    std::vector<void*> retired;
    mifft_status reserve(size_t want);
    void release();
};

// Kernel variant per size: picked by the autotuner (timed on first use and
// cached on disk) or the built-in default.
const StockhamVariant* choose_variant(int n, hipStream_t stream, const float2* tw);
// Autotuner core.  One timing candidate: a launch of `batch` transforms of
// the tuned length from in to out (device buffers of 2^24 / n transforms).
struct Candidate {
    const char* name;
    std::function<hipError_t(const void* in, void* out, int64_t batch)> run;
    bool prefer;  // stateless one-shot form, kept when within 3 % of the fastest
};
// The recorded choice for `key` ("<kind> <precision> <io> <n>"), or the
// fastest of `cand` timed now on `elem_bytes`-byte elements (recorded);
// the candidate index, or -1 if none ran.
int tuned_choice(const std::string& key, int n, size_t elem_bytes, const std::vector<Candidate>& cand,
                 hipStream_t stream, const char* what);
// False under MIFFT_AUTOTUNE=0 (built-in defaults, nothing timed).
bool tuning_enabled();

// MFMA plans (n = 4096): the faster kernel form (kMfmaStaged or
// kMfmaOneShot) for the precision and device element type.
int choose_mfma(int precision, int io, hipStream_t stream, const float2* tw);

// Lengths above 16384: the fastest of the single-workgroup variants of n
// (tw: their table, or null) and the four-step variants (fs_tw, or null);
// exactly one of *variant / *fourstep is set.
void choose_long(int n, hipStream_t stream, const float2* tw, const float2* fs_tw,
                 const StockhamVariant** variant, const FourStepVariant** fourstep);

}  // namespace mifft

struct mifft_plan {
    int64_t n = 0;
    int precision = 0;  // mifft_precision
    int io = 0;         // mifft_io
    const mifft::StockhamVariant* variant = nullptr;   // n <= 16384
    const mifft::FourStepVariant* fourstep = nullptr;  // n > 16384
    int mfma_form = 0;                                 // MFMA plans: kMfmaStaged / kMfmaOneShot
    float2* twiddles = nullptr;     // single-workgroup kernels: W_N^k, k < N, then the
                                    // pass rows (twiddle_table_len); null if none exists
    float2* fs_twiddles = nullptr;  // four-step kernels (fourstep_table_len), in the same
                                    // allocation after `twiddles`; null below 32768
    float2* tw_alloc = nullptr;     // the allocation holding both
    float2* real_tw = nullptr;   // real-input transforms (real_table_len)
    hipStream_t stream = nullptr;
    mifft::DeviceBuffer in, out;
    // four-step intermediate of an in-place call (an out-of-place call uses
    // the output buffer); grown by eager calls, see DeviceBuffer
    mutable mifft::DeviceBuffer fs_tmp;
};
