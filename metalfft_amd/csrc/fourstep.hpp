// Four-step transforms for lengths above 16384 (32768 ... 2^20): declarations
// shared by fourstep.hip, the plan runtime and the autotuner.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mifft {

// in -> out through y (N1 x N2 intermediate, batch * n complex64): pass A
// writes y, pass B reads it.  y may be `out` when in != out.
using FourStepLaunch = hipError_t (*)(const float2* in, float2* out, float2* y, const float2* tw,
                                      int64_t batch, bool inverse, hipStream_t stream);

// One compiled four-step form: n = n1 * n2, pass A runs the n2-point
// column transforms over ca columns per workgroup, pass B the n1-point
// transforms over cb columns per workgroup.
struct FourStepVariant {
    int n, n1, n2, ca, cb;
    FourStepLaunch launch;
    const char* name;
};

constexpr int64_t kFourStepMin = 32768;
constexpr int64_t kFourStepMax = int64_t(1) << 20;

const FourStepVariant* fourstep_variants(int* count);
const FourStepVariant* find_fourstep(int n);  // built-in default for length n
bool fourstep_size(int64_t n);                // n a power of two in [32768, 2^20]
// Plan twiddle memory of a four-step length n: the sub-length tables of
// 128 ... 1024 (each W_L^k then its packed pass rows) and the two factors
// of W_n^e, e = eh * 2^S + el: W_n^el (el < 2^S) and W_n^(eh 2^S).
int fourstep_table_len(int n);
hipError_t launch_fourstep_twiddles(float2* tw, int n, hipStream_t stream);

}  // namespace mifft
