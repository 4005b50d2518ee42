// Matrix-core (MFMA) DFT-as-GEMM FFT launcher.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mifft {
// precision: 1 FP16, 2 BF16, 3 FP16 hi/lo split, 4 BF16 hi/lo split
// (mifft_precision); io: 0 complex64, 1 complex FP16 (mifft_io).  n = 4096.
// out = out_scale * FFT(in) (inverse: conj(FFT(conj in)) * out_scale / n).
hipError_t launch_fft_mfma(int n, int precision, int io, const void* in, void* out,
                           const float2* tw, int64_t batch, bool inverse, float out_scale,
                           hipStream_t st);
}  // namespace mifft
