// Diagnostic copy (no counterpart in the reference): the achievable HBM
// streaming rate on this device, against which the FFT kernels' roofline
// fractions are read.  Same access style as the transforms: one-shot
// workgroups, 16-byte non-temporal loads and stores, every load of a thread
// issued before its first store.
#include <hip/hip_runtime.h>

#include "runtime.hpp"

namespace mifft {
namespace {

typedef float f4v __attribute__((ext_vector_type(4)));
constexpr int kU = 8;                  // 16-byte vectors per thread
constexpr int kBlock = 256 * kU;       // vectors per workgroup

__global__ void __launch_bounds__(256) k_copy_nt(const f4v* __restrict__ src, f4v* __restrict__ dst,
                                                 int64_t n4) {
    const int64_t base = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if ((int64_t)(blockIdx.x + 1) * kBlock <= n4) {  // full block (uniform)
        f4v v[kU];
#pragma unroll
        for (int j = 0; j < kU; ++j) v[j] = __builtin_nontemporal_load(src + base + 256 * j);
#pragma unroll
        for (int j = 0; j < kU; ++j) __builtin_nontemporal_store(v[j], dst + base + 256 * j);
    } else {
        for (int j = 0; j < kU; ++j)
            if (base + 256 * j < n4) dst[base + 256 * j] = src[base + 256 * j];
    }
}

__global__ void k_copy_tail(const char* __restrict__ src, char* __restrict__ dst, int64_t from,
                            int64_t bytes) {
    const int64_t i = from + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < bytes) dst[i] = src[i];
}

}  // namespace
}  // namespace mifft

using namespace mifft;

extern "C" {

mifft_status mifft_copy_device(const void* src, void* dst, int64_t bytes, void* stream) {
    if (bytes < 0) return fail(MIFFT_ERR_INVALID_ARGUMENT, "negative size");
    if (bytes == 0) return MIFFT_OK;
    if (!src || !dst) return fail(MIFFT_ERR_INVALID_ARGUMENT, "null buffer");
    hipStream_t st = static_cast<hipStream_t>(stream);
    const bool aligned = ((reinterpret_cast<uintptr_t>(src) | reinterpret_cast<uintptr_t>(dst)) & 15) == 0;
    const int64_t n4 = aligned ? bytes / 16 : 0;
    if (n4 > 0) {
        const int64_t blocks = (n4 + kBlock - 1) / kBlock;
        if (blocks > 0x7fffffffLL) return fail(MIFFT_ERR_INVALID_ARGUMENT, "copy too large");
        hipLaunchKernelGGL(k_copy_nt, dim3((unsigned)blocks), dim3(256), 0, st,
                           static_cast<const f4v*>(src), static_cast<f4v*>(dst), n4);
    }
    const int64_t from = n4 * 16;
    if (from < bytes) {
        const int64_t rest = bytes - from;
        hipLaunchKernelGGL(k_copy_tail, dim3((unsigned)((rest + 255) / 256)), dim3(256), 0, st,
                           static_cast<const char*>(src), static_cast<char*>(dst), from, bytes);
    }
    return check_hip(hipGetLastError(), "copy");
}

}  // extern "C"
