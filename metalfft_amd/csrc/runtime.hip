// C-ABI of libmifft: context, errors and the MetalFFT-equivalent plan.
//
// Reference counterparts: MetalContext / FFTError / FFTDescriptor
// (Sources/MetalFFT/Internal/MetalContext.swift:6-142), commitAndWait /
// makeBuffer (Sources/MetalFFT/Internal/Dispatch.swift:3-24) and the MetalFFT
// class (Sources/MetalFFT/FFT/MetalFFT.swift:10-289).
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "mfma_fft.hpp"
#include "realfft.hpp"
#include "runtime.hpp"

namespace mifft {

namespace {
struct ErrorState {
    mifft_status status = MIFFT_OK;
    int64_t expected = -1, got = -1, index = -1;
    std::string message;
};
thread_local ErrorState g_err;

std::mutex g_ctx_mutex;
bool g_ctx_ready = false;
int g_device = -1;
std::string g_device_name;
}  // namespace

mifft_status fail(mifft_status s, const std::string& msg, int64_t expected, int64_t got,
                  int64_t index) {
    g_err.status = s;
    g_err.expected = expected;
    g_err.got = got;
    g_err.index = index;
    g_err.message = msg;
    return s;
}

mifft_status ok() { return MIFFT_OK; }

mifft_status check_hip(hipError_t e, const char* what) {
    if (e == hipSuccess) return MIFFT_OK;
    std::string msg = std::string(what) + ": " + hipGetErrorString(e);
    if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation)
        return fail(MIFFT_ERR_ALLOCATION, "Buffer allocation failed (" + msg + ")");
    return fail(MIFFT_ERR_COMMAND_FAILED, "Command buffer error: " + msg);
}

bool supported_size(int64_t n) { return n >= 64 && n <= 16384 && (n & (n - 1)) == 0; }
bool plan_size(int64_t n) { return supported_size(n) || fourstep_size(n); }

mifft_status require_supported_size(int64_t n) {
    if (supported_size(n)) return MIFFT_OK;
    return fail(MIFFT_ERR_UNSUPPORTED_SIZE,
                "Unsupported FFT size: " + std::to_string(n) +
                    ". Supported: 64,128,256,512,1024,2048,4096,8192,16384",
                -1, n);
}

mifft_status DeviceBuffer::reserve(size_t want) {
    if (want <= bytes) return MIFFT_OK;
    const size_t grow = bytes + bytes / 2;
    const size_t size = want > grow ? want : grow;
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, size);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return fail(MIFFT_ERR_ALLOCATION, "Buffer allocation failed");
    }
    if (ptr) retired.push_back(ptr);
    ptr = p;
    bytes = size;
    return MIFFT_OK;
}

void DeviceBuffer::release() {
    if (ptr) (void)hipFree(ptr);
    for (void* r : retired) (void)hipFree(r);
    retired.clear();
    ptr = nullptr;
    bytes = 0;
}

mifft_status context_init() {
    std::lock_guard<std::mutex> lock(g_ctx_mutex);
    if (g_ctx_ready) return MIFFT_OK;
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess || count <= 0) {
        (void)hipGetLastError();
        return fail(MIFFT_ERR_NO_DEVICE, "No HIP device available");
    }
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return fail(MIFFT_ERR_NO_DEVICE, "No HIP device available");
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, dev) != hipSuccess)
        return fail(MIFFT_ERR_NO_DEVICE, "No HIP device available");
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        return fail(MIFFT_ERR_KERNEL_NOT_FOUND,
                    std::string("HIP kernel not found: library is built for gfx950, device is ") +
                        prop.gcnArchName);
    g_device = dev;
    g_device_name = std::string(prop.name) + " (" + prop.gcnArchName + ")";
    g_ctx_ready = true;
    return MIFFT_OK;
}

void context_shutdown() {
    std::lock_guard<std::mutex> lock(g_ctx_mutex);
    g_ctx_ready = false;
}

const std::string& device_name() { return g_device_name; }

}  // namespace mifft

using namespace mifft;

namespace {
mifft_status make_stream(hipStream_t* s) {
    return check_hip(hipStreamCreateWithFlags(s, hipStreamNonBlocking), "hipStreamCreate");
}

mifft_status validate_exec(const mifft_plan* plan, const void* in, const void* out, int64_t batch,
                           int dir) {
    if (!plan) return fail(MIFFT_ERR_INVALID_ARGUMENT, "null plan");
    if (batch < 0) return fail(MIFFT_ERR_INVALID_ARGUMENT, "negative batch");
    if (batch > 0 && (!in || !out)) return fail(MIFFT_ERR_INVALID_ARGUMENT, "null buffer");
    if (dir != MIFFT_FORWARD && dir != MIFFT_INVERSE)
        return fail(MIFFT_ERR_INVALID_ARGUMENT, "direction must be MIFFT_FORWARD or MIFFT_INVERSE");
    return MIFFT_OK;
}

mifft_status run(const mifft_plan* plan, const float* in, float* out, int64_t batch, int dir,
                 hipStream_t stream) {
    if (plan->precision != MIFFT_PREC_FP32)
        return check_hip(launch_fft_mfma((int)plan->n, plan->precision, plan->io, plan->mfma_form, in, out,
                                         plan->twiddles, batch, dir == MIFFT_INVERSE, 1.0f, stream),
                         "mfma fft");
    if (plan->fourstep) {
        // the intermediate goes through the output unless the call is in place
        float2* y = reinterpret_cast<float2*>(out);
        if (in == out) {
            mifft_status s = plan->fs_tmp.reserve(sizeof(float2) * (size_t)plan->n * (size_t)batch);
            if (s != MIFFT_OK) return s;
            y = static_cast<float2*>(plan->fs_tmp.ptr);
        }
        return check_hip(plan->fourstep->launch(reinterpret_cast<const float2*>(in),
                                                reinterpret_cast<float2*>(out), y, plan->fs_twiddles,
                                                batch, dir == MIFFT_INVERSE, stream),
                         plan->fourstep->name);
    }
    hipError_t e = plan->variant->launch(reinterpret_cast<const float2*>(in),
                                         reinterpret_cast<float2*>(out), plan->twiddles, batch,
                                         dir == MIFFT_INVERSE, stream);
    return check_hip(e, plan->variant->name);
}

// Copy-in, transform, copy-out, wait: the reference's synchronous host call.
mifft_status host_batch(mifft_plan* plan, const float* const* inputs, const int64_t* counts,
                        int64_t batch, float* const* outputs, int dir) {
    if (!plan) return fail(MIFFT_ERR_INVALID_ARGUMENT, "null plan");
    if (plan->io != MIFFT_IO_C64)
        return fail(MIFFT_ERR_INVALID_ARGUMENT, "host entry points take complex64 data");
    if (batch < 0) return fail(MIFFT_ERR_INVALID_ARGUMENT, "negative batch");
    if (batch == 0) return MIFFT_OK;
    if (!inputs || !counts || !outputs) return fail(MIFFT_ERR_INVALID_ARGUMENT, "null array");
    const int64_t n = plan->n;
    for (int64_t i = 0; i < batch; ++i) {
        if (counts[i] != n)
            return fail(MIFFT_ERR_BATCH_INPUT_SIZE,
                        "Batch element " + std::to_string(i) + ": expected " + std::to_string(n) +
                            ", got " + std::to_string(counts[i]),
                        n, counts[i], i);
        if (!inputs[i] || !outputs[i]) return fail(MIFFT_ERR_INVALID_ARGUMENT, "null element");
    }
    const size_t bytes = (size_t)n * 8;
    mifft_status s;
    if ((s = plan->in.reserve(bytes * batch)) != MIFFT_OK) return s;
    if ((s = plan->out.reserve(bytes * batch)) != MIFFT_OK) return s;
    char* din = static_cast<char*>(plan->in.ptr);
    char* dout = static_cast<char*>(plan->out.ptr);
    for (int64_t i = 0; i < batch; ++i)
        if ((s = check_hip(hipMemcpyAsync(din + i * bytes, inputs[i], bytes, hipMemcpyHostToDevice,
                                          plan->stream),
                           "hipMemcpyAsync")) != MIFFT_OK)
            return s;
    if ((s = run(plan, (const float*)din, (float*)dout, batch, dir, plan->stream)) != MIFFT_OK) return s;
    for (int64_t i = 0; i < batch; ++i)
        if ((s = check_hip(hipMemcpyAsync(outputs[i], dout + i * bytes, bytes, hipMemcpyDeviceToHost,
                                          plan->stream),
                           "hipMemcpyAsync")) != MIFFT_OK)
            return s;
    return check_hip(hipStreamSynchronize(plan->stream), "hipStreamSynchronize");
}

mifft_status host_single(mifft_plan* plan, const float* in, int64_t in_count, float* out, int dir) {
    if (!plan) return fail(MIFFT_ERR_INVALID_ARGUMENT, "null plan");
    if (in_count != plan->n)
        return fail(MIFFT_ERR_INVALID_INPUT_SIZE,
                    "Input size mismatch: expected " + std::to_string(plan->n) + ", got " +
                        std::to_string(in_count),
                    plan->n, in_count);
    if (!in || !out) return fail(MIFFT_ERR_INVALID_ARGUMENT, "null buffer");
    const float* ins[1] = {in};
    float* outs[1] = {out};
    int64_t counts[1] = {in_count};
    return host_batch(plan, ins, counts, 1, outs, dir);
}
}  // namespace

extern "C" {

const char* mifft_status_string(mifft_status s) {
    switch (s) {
        case MIFFT_OK: return "OK";
        case MIFFT_ERR_NO_DEVICE: return "No HIP device available";
        case MIFFT_ERR_NO_QUEUE: return "Could not create HIP stream";
        case MIFFT_ERR_LIBRARY_BUILD: return "HIP library build failed";
        case MIFFT_ERR_KERNEL_NOT_FOUND: return "HIP kernel not found";
        case MIFFT_ERR_COMMAND_FAILED: return "Command buffer error";
        case MIFFT_ERR_INVALID_INPUT_SIZE: return "Input size mismatch";
        case MIFFT_ERR_BATCH_INPUT_SIZE: return "Batch element size mismatch";
        case MIFFT_ERR_UNSUPPORTED_SIZE:
            return "Unsupported FFT size. Supported: 64,128,256,512,1024,2048,4096,8192,16384 "
                   "(four-step: 32768 ... 1048576)";
        case MIFFT_ERR_ALLOCATION: return "Buffer allocation failed";
        case MIFFT_ERR_PRECONDITION: return "Precondition failed";
        case MIFFT_ERR_INVALID_ARGUMENT: return "Invalid argument";
    }
    return "Unknown status";
}

mifft_status mifft_last_error(int64_t* expected, int64_t* got, int64_t* batch_index) {
    if (expected) *expected = g_err.expected;
    if (got) *got = g_err.got;
    if (batch_index) *batch_index = g_err.index;
    return g_err.status;
}

const char* mifft_last_error_message(void) { return g_err.message.c_str(); }

mifft_status mifft_init(void) { return context_init(); }

mifft_status mifft_shutdown(void) {
    context_shutdown();
    return MIFFT_OK;
}

mifft_status mifft_device_name(char* buf, size_t len) {
    mifft_status s = context_init();
    if (s != MIFFT_OK) return s;
    if (!buf || len == 0) return fail(MIFFT_ERR_INVALID_ARGUMENT, "null buffer");
    std::snprintf(buf, len, "%s", device_name().c_str());
    return MIFFT_OK;
}

mifft_status mifft_plan_create(int64_t n, mifft_plan** out) {
    return mifft_plan_create_ex(n, MIFFT_PREC_FP32, MIFFT_IO_C64, out);
}

mifft_status mifft_plan_create_ex(int64_t n, mifft_precision precision, mifft_io io,
                                  mifft_plan** out) {
    if (!out) return fail(MIFFT_ERR_INVALID_ARGUMENT, "null plan pointer");
    if (precision < MIFFT_PREC_FP32 || precision > MIFFT_PREC_BF16_MFMA_SPLIT)
        return fail(MIFFT_ERR_INVALID_ARGUMENT, "unknown precision");
    if (io != MIFFT_IO_C64 && io != MIFFT_IO_C32H)
        return fail(MIFFT_ERR_INVALID_ARGUMENT, "unknown io type");
    *out = nullptr;
    if (!plan_size(n))
        return fail(MIFFT_ERR_UNSUPPORTED_SIZE,
                    "Unsupported FFT size: " + std::to_string(n) +
                        ". Supported: 64,128,256,512,1024,2048,4096,8192,16384 "
                        "(four-step: 32768 ... 1048576)",
                    -1, n);
    mifft_status s = context_init();
    if (s != MIFFT_OK) return s;
    if ((precision != MIFFT_PREC_FP32 || io != MIFFT_IO_C64) && n != 4096)
        return fail(MIFFT_ERR_KERNEL_NOT_FOUND,
                    "HIP kernel not found: MFMA / complex-FP16 transforms are built for n = 4096");
    if (precision == MIFFT_PREC_FP32 && io != MIFFT_IO_C64)
        return fail(MIFFT_ERR_KERNEL_NOT_FOUND,
                    "HIP kernel not found: complex-FP16 I/O needs an MFMA precision");
    mifft_plan* p = new mifft_plan();
    p->n = n;
    p->precision = precision;
    p->io = io;
    if ((s = make_stream(&p->stream)) != MIFFT_OK) {
        delete p;
        return fail(MIFFT_ERR_NO_QUEUE, "Could not create HIP stream");
    }
    const bool four = fourstep_size(n);
    const bool single = find_stockham((int)n, 16) || find_stockham((int)n, 32);
    const size_t st_len = single ? (size_t)twiddle_table_len((int)n) : 0;  // even: 16-byte rows
    const size_t fs_len = four ? (size_t)fourstep_table_len((int)n) : 0;
    float2* tw_all = nullptr;
    if ((s = check_hip(hipMalloc(&tw_all, sizeof(float2) * (st_len + fs_len)), "hipMalloc")) != MIFFT_OK) {
        mifft_plan_destroy(p);
        return fail(MIFFT_ERR_ALLOCATION, "Buffer allocation failed");
    }
    p->twiddles = single ? tw_all : nullptr;
    p->fs_twiddles = four ? tw_all + st_len : nullptr;
    p->tw_alloc = tw_all;
    // real-input tables for the single-pass lengths only (the real entry
    // points report kernelNotFound for four-step plans)
    if (!four && (s = check_hip(hipMalloc(&p->real_tw, sizeof(float2) * real_table_len((int)n)),
                                "hipMalloc")) != MIFFT_OK) {
        mifft_plan_destroy(p);
        return fail(MIFFT_ERR_ALLOCATION, "Buffer allocation failed");
    }
    if ((single && (s = check_hip(launch_twiddles(p->twiddles, (int)n, p->stream), "twiddles")) !=
                       MIFFT_OK) ||
        (four && (s = check_hip(launch_fourstep_twiddles(p->fs_twiddles, (int)n, p->stream),
                                "four-step twiddles")) != MIFFT_OK) ||
        (!four && (s = check_hip(launch_real_table(p->real_tw, (int)n, p->stream), "real-input table")) !=
                      MIFFT_OK) ||
        (s = check_hip(hipStreamSynchronize(p->stream), "hipStreamSynchronize")) != MIFFT_OK) {
        mifft_plan_destroy(p);
        return s;
    }
    if (four) {
        choose_long((int)n, p->stream, p->twiddles, p->fs_twiddles, &p->variant, &p->fourstep);
        if (!p->fourstep && !p->variant) {
            mifft_plan_destroy(p);
            return fail(MIFFT_ERR_KERNEL_NOT_FOUND, "HIP kernel not found for size " + std::to_string(n));
        }
        *out = p;
        return MIFFT_OK;
    }
    // the Stockham variant serves FP32 plans (and their real-input entry
    // points); MFMA plans time their own kernel forms instead
    p->variant = precision == MIFFT_PREC_FP32 ? choose_variant((int)n, p->stream, p->twiddles)
                                              : find_stockham((int)n, 16);
    if (precision != MIFFT_PREC_FP32) p->mfma_form = choose_mfma(precision, io, p->stream, p->twiddles);
    if (!p->variant) {
        mifft_plan_destroy(p);
        return fail(MIFFT_ERR_KERNEL_NOT_FOUND, "HIP kernel not found for size " + std::to_string(n));
    }
    *out = p;
    return MIFFT_OK;
}

mifft_status mifft_plan_destroy(mifft_plan* plan) {
    if (!plan) return MIFFT_OK;
    if (plan->stream) (void)hipStreamSynchronize(plan->stream);
    plan->in.release();
    plan->out.release();
    plan->fs_tmp.release();
    if (plan->tw_alloc) (void)hipFree(plan->tw_alloc);
    if (plan->real_tw) (void)hipFree(plan->real_tw);
    if (plan->stream) (void)hipStreamDestroy(plan->stream);
    delete plan;
    return MIFFT_OK;
}

int64_t mifft_plan_size(const mifft_plan* plan) { return plan ? plan->n : -1; }

const char* mifft_plan_kernel_name(const mifft_plan* plan) {
    if (!plan) return "";
    if (plan->precision != MIFFT_PREC_FP32) return mfma_form_name(plan->precision, plan->io, plan->mfma_form);
    if (plan->fourstep) return plan->fourstep->name;
    return plan->variant ? plan->variant->name : "";
}

mifft_status mifft_execute_device(const mifft_plan* plan, const float* in, float* out, int64_t batch,
                                  mifft_direction dir, void* stream) {
    mifft_status s = validate_exec(plan, in, out, batch, dir);
    if (s != MIFFT_OK) return s;
    if (batch == 0) return MIFFT_OK;
    return run(plan, in, out, batch, dir, static_cast<hipStream_t>(stream));
}

mifft_status mifft_forward_host(mifft_plan* plan, const float* in, int64_t in_count, float* out) {
    return host_single(plan, in, in_count, out, MIFFT_FORWARD);
}

mifft_status mifft_inverse_host(mifft_plan* plan, const float* in, int64_t in_count, float* out) {
    return host_single(plan, in, in_count, out, MIFFT_INVERSE);
}

mifft_status mifft_forward_batch_host(mifft_plan* plan, const float* const* inputs,
                                      const int64_t* counts, int64_t batch, float* const* outputs) {
    return host_batch(plan, inputs, counts, batch, outputs, MIFFT_FORWARD);
}

mifft_status mifft_inverse_batch_host(mifft_plan* plan, const float* const* inputs,
                                      const int64_t* counts, int64_t batch, float* const* outputs) {
    return host_batch(plan, inputs, counts, batch, outputs, MIFFT_INVERSE);
}

}  // extern "C"
