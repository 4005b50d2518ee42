// Convolver object shared by the fused convolution kernels and the C-ABI.
#pragma once
#include "runtime.hpp"

struct mifft_convolver {
    int64_t n = 0;       // FFT size (blockSize)
    int64_t K = 0;       // taps (kernelLen)
    int fused = 0;       // created as FusedConvolver
    int precision = 0;   // mifft_conv_precision
    mifft_plan* plan = nullptr;       // length-n FFT: twiddles + kernel variant
    mifft::DeviceBuffer spectrum;     // H = FFT(taps zero-padded to n), complex64
    mifft::DeviceBuffer packed;       // taps as complex, n entries
    mifft::DeviceBuffer partial;      // adjoint: per-workgroup tap partials
    mifft::DeviceBuffer hx, hy;       // host-call staging
    hipStream_t stream = nullptr;     // host-call stream
};

namespace mifft {
// H <- FFT(taps) from K device floats.
mifft_status convolver_spectrum(mifft_convolver* cv, const float* taps_dev, hipStream_t st);
// y[count][S + K - 1] <- x[count][S] (*) taps.
mifft_status convolver_forward(const mifft_convolver* cv, const float* x, int64_t S, int64_t count,
                               float* y, hipStream_t st);
// gx[count][S], gh[K] from x[count][S] and gy[count][S + K - 1].
mifft_status convolver_backward(mifft_convolver* cv, const float* x, int64_t S, int64_t count,
                                const float* gy, float* gx, float* gh, hipStream_t st);
// Work split of the adjoint's tap-gradient kernel: block pairs per signal,
// pairs per workgroup, one partial row of K floats per slot (transform
// slot of a workgroup) and the first level of the 64-row sum tree.
struct AdjointGeom {
    int64_t nbx, npx, ppw, slots, level1;
};
constexpr int64_t kAdjointWorkgroups = 2048;  // enough to fill 256 CUs
AdjointGeom adjoint_geometry(const mifft_convolver* cv, int64_t S, int64_t count, int64_t target_wg);
// Scratch floats the split needs: the partial rows and two tree levels.
int64_t adjoint_scratch_floats(const AdjointGeom& a, int64_t K);
// The adjoint with its partials in `part` (adjoint_scratch_floats(a, K)
// floats, every one written before it is read).
mifft_status convolver_backward_with(const mifft_convolver* cv, const float* x, int64_t S, int64_t count,
                                     const float* gy, float* gx, float* gh, float* part, const AdjointGeom& a,
                                     hipStream_t st);
// r[count][N] <- circular cross-correlation of a[count][na], b[count][nb]
// (each truncated or zero-padded to N = plan size).
mifft_status correlate_device(const mifft_plan* plan, const float* a, int64_t na, const float* b,
                              int64_t nb, float* r, int64_t count, hipStream_t st);
// out[c] = sum over r < rows of in[r * cols + c], in 64-row levels added in a
// fixed order (deterministic); s0, s1: scratch of ceil(rows / 64) * cols
// floats each (only used when rows > 64).
hipError_t sum_rows(const float* in, int64_t rows, int cols, float* s0, float* s1, float* out,
                    hipStream_t st);
// p[0:n] <- 0 with a fill kernel.
mifft_status zero_floats(float* p, int64_t n, hipStream_t st);
}  // namespace mifft
