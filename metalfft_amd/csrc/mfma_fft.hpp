// Matrix-core (MFMA) DFT-as-GEMM FFT launcher.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mifft {
// Kernel forms: persistent workgroups with the next transform staged in LDS
// by DMA, or one transform per workgroup loaded straight into registers;
// three radix-16 passes, or two radix-64 passes (each four DFT-16s on the
// matrix cores and a DFT-4 in registers).
enum { kMfmaStaged = 0, kMfmaOneShot = 1, kMfmaStaged64 = 2, kMfmaOneShot64 = 3, kMfmaForms = 4 };
// precision: 1 FP16, 2 BF16, 3 FP16 hi/lo split, 4 BF16 hi/lo split
// (mifft_precision); io: 0 complex64, 1 complex FP16 (mifft_io).  n = 4096.
// out = out_scale * FFT(in) (inverse: conj(FFT(conj in)) * out_scale / n).
hipError_t launch_fft_mfma(int n, int precision, int io, int form, const void* in, void* out,
                           const float2* tw, int64_t batch, bool inverse, float out_scale,
                           hipStream_t st);
// Variant name of a (precision, io, form), as mifft_plan_kernel_name reports it.
const char* mfma_form_name(int precision, int io, int form);
}  // namespace mifft
