// C-ABI of the convolver (Convolver / FusedConvolver) and correlator objects.
//
// Reference: Sources/MetalFFT/DSP/Convolve.swift:8-87,
// Sources/MetalFFT/DSP/FusedConvolver.swift:17-134,
// Sources/MetalFFT/DSP/Correlate.swift:7-62.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "convolve.hpp"
#include "runtime.hpp"

using namespace mifft;

struct mifft_correlator {
    int64_t n = 0;
    mifft_plan* plan = nullptr;
    DeviceBuffer a, b, r;
    hipStream_t stream = nullptr;
};

namespace {
mifft_status stream_err() { return fail(MIFFT_ERR_NO_QUEUE, "Could not create HIP stream"); }
}  // namespace

extern "C" {

int64_t mifft_convolver_recommended_fft_size(int64_t kernel_count) {
    for (int64_t n = 64; n <= 16384; n *= 2)
        if (n > kernel_count) return n;
    return -1;
}

mifft_status mifft_convolver_create(const float* taps, int64_t k, int64_t fft_size, int fused,
                                    int precision, mifft_convolver** out) {
    if (!out) return fail(MIFFT_ERR_INVALID_ARGUMENT, "null convolver pointer");
    *out = nullptr;
    if (!taps && k > 0) return fail(MIFFT_ERR_INVALID_ARGUMENT, "null taps");
    if (fused && fft_size != 4096)
        return fail(MIFFT_ERR_PRECONDITION, "FusedConvolver only supports fftSize=4096");
    if (!fused && precision != MIFFT_CONV_FP32)
        return fail(MIFFT_ERR_PRECONDITION, "Convolver computes in FP32 only");
    if (precision < MIFFT_CONV_FP32 || precision > MIFFT_CONV_FP16_MIXED)
        return fail(MIFFT_ERR_INVALID_ARGUMENT, "unknown precision");
    if (!(fft_size > k))
        return fail(MIFFT_ERR_PRECONDITION,
                    "fftSize must be > kernel.count; got fftSize=" + std::to_string(fft_size) +
                        ", kernel.count=" + std::to_string(k));
    if (k < 1) return fail(MIFFT_ERR_PRECONDITION, "kernel must have at least one tap");
    mifft_plan* plan = nullptr;
    mifft_status s = require_supported_size(fft_size);
    if (s != MIFFT_OK) return s;
    s = mifft_plan_create(fft_size, &plan);
    if (s != MIFFT_OK) return s;
    mifft_convolver* cv = new mifft_convolver();
    cv->n = fft_size;
    cv->K = k;
    cv->fused = fused;
    cv->precision = precision;
    cv->plan = plan;
    if (hipStreamCreateWithFlags(&cv->stream, hipStreamNonBlocking) != hipSuccess) {
        mifft_convolver_destroy(cv);
        return stream_err();
    }
    if ((s = cv->spectrum.reserve(12 * (size_t)fft_size)) != MIFFT_OK ||  // H, then H in FP16
        (s = cv->packed.reserve(8 * (size_t)fft_size)) != MIFFT_OK ||
        (s = cv->hx.reserve(4 * (size_t)k)) != MIFFT_OK) {
        mifft_convolver_destroy(cv);
        return s;
    }
    if ((s = check_hip(hipMemcpyAsync(cv->hx.ptr, taps, 4 * (size_t)k, hipMemcpyHostToDevice,
                                      cv->stream),
                       "hipMemcpyAsync")) != MIFFT_OK ||
        (s = convolver_spectrum(cv, static_cast<const float*>(cv->hx.ptr), cv->stream)) != MIFFT_OK ||
        (s = check_hip(hipStreamSynchronize(cv->stream), "hipStreamSynchronize")) != MIFFT_OK) {
        mifft_convolver_destroy(cv);
        return s;
    }
    *out = cv;
    return MIFFT_OK;
}

mifft_status mifft_convolver_destroy(mifft_convolver* cv) {
    if (!cv) return MIFFT_OK;
    if (cv->stream) (void)hipStreamSynchronize(cv->stream);
    cv->spectrum.release();
    cv->packed.release();
    cv->partial.release();
    cv->hx.release();
    cv->hy.release();
    if (cv->stream) (void)hipStreamDestroy(cv->stream);
    mifft_plan_destroy(cv->plan);
    delete cv;
    return MIFFT_OK;
}

int64_t mifft_convolver_block_size(const mifft_convolver* cv) { return cv ? cv->n : -1; }
int64_t mifft_convolver_hop_size(const mifft_convolver* cv) { return cv ? cv->n - cv->K + 1 : -1; }
int64_t mifft_convolver_kernel_len(const mifft_convolver* cv) { return cv ? cv->K : -1; }

mifft_status mifft_convolver_set_taps_device(mifft_convolver* cv, const float* taps, void* stream) {
    if (!cv || !taps) return fail(MIFFT_ERR_INVALID_ARGUMENT, "null argument");
    return convolver_spectrum(cv, taps, static_cast<hipStream_t>(stream));
}

mifft_status mifft_convolver_apply_device(const mifft_convolver* cv, const float* signals,
                                          int64_t samples, int64_t count, float* out, void* stream) {
    if (!cv) return fail(MIFFT_ERR_INVALID_ARGUMENT, "null convolver");
    if (samples < 0 || count < 0) return fail(MIFFT_ERR_INVALID_ARGUMENT, "negative size");
    if (count == 0 || samples + cv->K - 1 <= 0) return MIFFT_OK;
    if (!signals || !out) return fail(MIFFT_ERR_INVALID_ARGUMENT, "null buffer");
    return convolver_forward(cv, signals, samples, count, out, static_cast<hipStream_t>(stream));
}

mifft_status mifft_convolver_apply_host(mifft_convolver* cv, const float* signal, int64_t samples,
                                        float* out) {
    if (!cv) return fail(MIFFT_ERR_INVALID_ARGUMENT, "null convolver");
    if (samples < 0) return fail(MIFFT_ERR_INVALID_ARGUMENT, "negative size");
    const int64_t out_len = samples + cv->K - 1;
    if (samples == 0) {  // the reference returns K - 1 zeros (Convolve.swift:37-44)
        for (int64_t i = 0; i < out_len; ++i) out[i] = 0.f;
        return MIFFT_OK;
    }
    if (!signal || !out) return fail(MIFFT_ERR_INVALID_ARGUMENT, "null buffer");
    mifft_status s;
    if ((s = cv->hx.reserve(4 * (size_t)samples)) != MIFFT_OK ||
        (s = cv->hy.reserve(4 * (size_t)out_len)) != MIFFT_OK)
        return s;
    if ((s = check_hip(hipMemcpyAsync(cv->hx.ptr, signal, 4 * (size_t)samples,
                                      hipMemcpyHostToDevice, cv->stream),
                       "hipMemcpyAsync")) != MIFFT_OK)
        return s;
    if ((s = convolver_forward(cv, static_cast<const float*>(cv->hx.ptr), samples, 1,
                               static_cast<float*>(cv->hy.ptr), cv->stream)) != MIFFT_OK)
        return s;
    if ((s = check_hip(hipMemcpyAsync(out, cv->hy.ptr, 4 * (size_t)out_len, hipMemcpyDeviceToHost,
                                      cv->stream),
                       "hipMemcpyAsync")) != MIFFT_OK)
        return s;
    return check_hip(hipStreamSynchronize(cv->stream), "hipStreamSynchronize");
}

mifft_status mifft_convolver_backward_device(mifft_convolver* cv, const float* signals,
                                             int64_t samples, int64_t count, const float* grad_out,
                                             float* grad_signal, float* grad_taps, void* stream) {
    if (!cv) return fail(MIFFT_ERR_INVALID_ARGUMENT, "null convolver");
    if (cv->precision != MIFFT_CONV_FP32)
        return fail(MIFFT_ERR_KERNEL_NOT_FOUND, "HIP kernel not found: adjoint is FP32 only");
    if (samples < 0 || count < 0) return fail(MIFFT_ERR_INVALID_ARGUMENT, "negative size");
    if (!grad_taps) return fail(MIFFT_ERR_INVALID_ARGUMENT, "null grad_taps");
    if (count > 0 && samples > 0 && (!signals || !grad_out || !grad_signal))
        return fail(MIFFT_ERR_INVALID_ARGUMENT, "null buffer");
    hipStream_t st = static_cast<hipStream_t>(stream);
    if (count == 0 || samples == 0) {
        return zero_floats(grad_taps, cv->K, st);
    }
    return convolver_backward(cv, signals, samples, count, grad_out, grad_signal, grad_taps, st);
}

// --------------------------------------------- adjoint without allocation
// Callers that cannot allocate (code under hipGraph capture, tight serving
// loops) get the adjoint with its tap partials in this file-scope array
// instead of convolver-owned memory that grows on demand.  Each GPU holds
// its own copy for the life of the process; host code uses the address
// hipGetSymbolAddress returns for the current device, never the array's
// name.  Every call on one GPU shares it: calls that overlap in time on one
// GPU (two streams, two threads) race.
constexpr size_t kNoAllocScratchFloats = (size_t(16) << 20) / sizeof(float);  // 16 MiB
static __device__ __align__(256) float scratch[kNoAllocScratchFloats];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
float* g_scratch_dev[kMaxDevices] = {};

// Address of the current device's scratch array, looked up once per device.
mifft_status noalloc_scratch(float** out) {
    int dev = 0;
    mifft_status s = check_hip(hipGetDevice(&dev), "hipGetDevice");
    if (s != MIFFT_OK) return s;
    if (dev < 0 || dev >= kMaxDevices) return fail(MIFFT_ERR_NO_DEVICE, "device index out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_dev[dev]) {
        void* p = nullptr;
        if ((s = check_hip(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress")) != MIFFT_OK)
            return s;
        g_scratch_dev[dev] = static_cast<float*>(p);
    }
    *out = g_scratch_dev[dev];
    return MIFFT_OK;
}
}  // namespace

mifft_status mifft_convolver_backward_device_noalloc(const mifft_convolver* cv, const float* signals,
                                                     int64_t samples, int64_t count, const float* grad_out,
                                                     float* grad_signal, float* grad_taps, void* stream) {
    if (!cv) return fail(MIFFT_ERR_INVALID_ARGUMENT, "null convolver");
    if (cv->precision != MIFFT_CONV_FP32)
        return fail(MIFFT_ERR_KERNEL_NOT_FOUND, "HIP kernel not found: adjoint is FP32 only");
    if (samples < 0 || count < 0) return fail(MIFFT_ERR_INVALID_ARGUMENT, "negative size");
    if (!grad_taps) return fail(MIFFT_ERR_INVALID_ARGUMENT, "null grad_taps");
    if (count > 0 && samples > 0 && (!signals || !grad_out || !grad_signal))
        return fail(MIFFT_ERR_INVALID_ARGUMENT, "null buffer");
    hipStream_t st = static_cast<hipStream_t>(stream);
    if (count == 0 || samples == 0) return zero_floats(grad_taps, cv->K, st);
    float* part = nullptr;
    mifft_status s = noalloc_scratch(&part);
    if (s != MIFFT_OK) return s;
    // the allocating entry point's work split when its partials fit (same
    // bits), else fewer, longer-running workgroups
    const int64_t cap = (int64_t)kNoAllocScratchFloats;
    int64_t target = kAdjointWorkgroups;
    AdjointGeom a = adjoint_geometry(cv, samples, count, target);
    while (adjoint_scratch_floats(a, cv->K) > cap && target > 1) {
        target = target * 3 / 4;
        a = adjoint_geometry(cv, samples, count, target);
    }
    const int64_t need = adjoint_scratch_floats(a, cv->K);
    if (need > cap)
        return fail(MIFFT_ERR_INVALID_ARGUMENT,
                    "no-allocation adjoint: tap partials need " + std::to_string(need * 4) +
                        " bytes, the fixed scratch holds " + std::to_string(cap * 4),
                    cap * 4, need * 4);
    return convolver_backward_with(cv, signals, samples, count, grad_out, grad_signal, grad_taps, part, a, st);
}

// ------------------------------------------------------------ correlator
mifft_status mifft_correlator_create(int64_t fft_size, mifft_correlator** out) {
    if (!out) return fail(MIFFT_ERR_INVALID_ARGUMENT, "null correlator pointer");
    *out = nullptr;
    mifft_plan* plan = nullptr;
    mifft_status s = require_supported_size(fft_size);
    if (s != MIFFT_OK) return s;
    s = mifft_plan_create(fft_size, &plan);
    if (s != MIFFT_OK) return s;
    mifft_correlator* c = new mifft_correlator();
    c->n = fft_size;
    c->plan = plan;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        mifft_correlator_destroy(c);
        return stream_err();
    }
    *out = c;
    return MIFFT_OK;
}

mifft_status mifft_correlator_destroy(mifft_correlator* c) {
    if (!c) return MIFFT_OK;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    c->a.release();
    c->b.release();
    c->r.release();
    if (c->stream) (void)hipStreamDestroy(c->stream);
    mifft_plan_destroy(c->plan);
    delete c;
    return MIFFT_OK;
}

int64_t mifft_correlator_fft_size(const mifft_correlator* c) { return c ? c->n : -1; }

mifft_status mifft_correlator_cross_device(const mifft_correlator* c, const float* a, int64_t na,
                                           const float* b, int64_t nb, int64_t count, float* out,
                                           void* stream) {
    if (!c) return fail(MIFFT_ERR_INVALID_ARGUMENT, "null correlator");
    if (na < 0 || nb < 0 || count < 0) return fail(MIFFT_ERR_INVALID_ARGUMENT, "negative size");
    if (count == 0) return MIFFT_OK;
    if ((na > 0 && !a) || (nb > 0 && !b) || !out) return fail(MIFFT_ERR_INVALID_ARGUMENT, "null buffer");
    return correlate_device(c->plan, a, na, b, nb, out, count, static_cast<hipStream_t>(stream));
}

mifft_status mifft_correlator_cross_host(mifft_correlator* c, const float* a, int64_t na,
                                         const float* b, int64_t nb, float* out) {
    if (!c) return fail(MIFFT_ERR_INVALID_ARGUMENT, "null correlator");
    if (na < 0 || nb < 0) return fail(MIFFT_ERR_INVALID_ARGUMENT, "negative size");
    if ((na > 0 && !a) || (nb > 0 && !b) || !out) return fail(MIFFT_ERR_INVALID_ARGUMENT, "null buffer");
    // Inputs longer than fftSize are truncated (Correlate.swift:54).
    const int64_t ta = na < c->n ? na : c->n, tb = nb < c->n ? nb : c->n;
    mifft_status s;
    if ((s = c->a.reserve(4 * (size_t)(ta > 0 ? ta : 1))) != MIFFT_OK ||
        (s = c->b.reserve(4 * (size_t)(tb > 0 ? tb : 1))) != MIFFT_OK ||
        (s = c->r.reserve(4 * (size_t)c->n)) != MIFFT_OK)
        return s;
    if (ta > 0 && (s = check_hip(hipMemcpyAsync(c->a.ptr, a, 4 * (size_t)ta, hipMemcpyHostToDevice,
                                                c->stream),
                                 "hipMemcpyAsync")) != MIFFT_OK)
        return s;
    if (tb > 0 && (s = check_hip(hipMemcpyAsync(c->b.ptr, b, 4 * (size_t)tb, hipMemcpyHostToDevice,
                                                c->stream),
                                 "hipMemcpyAsync")) != MIFFT_OK)
        return s;
    if ((s = correlate_device(c->plan, static_cast<const float*>(c->a.ptr), ta,
                              static_cast<const float*>(c->b.ptr), tb, static_cast<float*>(c->r.ptr),
                              1, c->stream)) != MIFFT_OK)
        return s;
    if ((s = check_hip(hipMemcpyAsync(out, c->r.ptr, 4 * (size_t)c->n, hipMemcpyDeviceToHost,
                                      c->stream),
                       "hipMemcpyAsync")) != MIFFT_OK)
        return s;
    return check_hip(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
}

mifft_status mifft_correlator_auto_host(mifft_correlator* c, const float* x, int64_t n, float* out) {
    return mifft_correlator_cross_host(c, x, n, x, n, out);
}

}  // extern "C"
