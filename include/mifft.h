/*
 * mifft.h -- C-ABI of libmifft, the MI355X (gfx950) batched complex FFT
 * library that stands in for MetalFFT's Metal back end.
 *
 * Data: complex64 is interleaved (re, im) float pairs, the layout of
 * SIMD2<Float> in the reference (Sources/MetalFFT/FFT/MetalFFT.swift:7-8).
 * Forward is the unscaled DFT X[k] = sum_n x[n] exp(-2 pi i n k / N); inverse
 * is conj(FFT(conj X)) / N (MetalFFT.swift:56-65).
 *
 * Two calling conventions:
 *   *_device  : device pointers, asynchronous on the caller's HIP stream
 *               (`stream` is a hipStream_t, NULL = the null stream).
 *   *_host    : host pointers, blocking, copying through plan-owned device
 *               buffers -- the semantics of the reference's Swift methods,
 *               which copy into shared buffers and wait
 *               (MetalFFT.swift:74-83, Sources/MetalFFT/Internal/Dispatch.swift:3-10).
 *
 * Threading: library initialisation is thread-safe (the reference locks its
 * context singleton, Sources/MetalFFT/Internal/MetalContext.swift:79-88);
 * a plan, convolver or correlator object must be used by one thread at a
 * time (MetalFFT.swift:9).
 *
 * Every entry point returns mifft_status.  On failure, mifft_last_error()
 * gives the case's payload (expected / got / batch index) and
 * mifft_last_error_message() the text of the matching FFTError description.
 */
#ifndef MIFFT_H
#define MIFFT_H

#include <stddef.h>
#include <stdint.h>

#if defined(__GNUC__) || defined(__clang__)
#define MIFFT_API __attribute__((visibility("default")))
#else
#define MIFFT_API
#endif

#ifdef __cplusplus
extern "C" {
#endif

/* One value per FFTError case (Sources/MetalFFT/Internal/MetalContext.swift:6-30)
 * plus the Swift precondition traps, which a C library reports instead. */
typedef enum mifft_status {
    MIFFT_OK = 0,
    MIFFT_ERR_NO_DEVICE = 1,          /* FFTError.noMetalDevice          MetalContext.swift:7  */
    MIFFT_ERR_NO_QUEUE = 2,           /* FFTError.noCommandQueue         MetalContext.swift:8  */
    MIFFT_ERR_LIBRARY_BUILD = 3,      /* FFTError.libraryBuildFailed     MetalContext.swift:9  */
    MIFFT_ERR_KERNEL_NOT_FOUND = 4,   /* FFTError.kernelNotFound         MetalContext.swift:10 */
    MIFFT_ERR_COMMAND_FAILED = 5,     /* FFTError.commandBufferFailed    MetalContext.swift:11 */
    MIFFT_ERR_INVALID_INPUT_SIZE = 6, /* FFTError.invalidInputSize       MetalContext.swift:12 */
    MIFFT_ERR_BATCH_INPUT_SIZE = 7,   /* FFTError.batchInputSize         MetalContext.swift:13 */
    MIFFT_ERR_UNSUPPORTED_SIZE = 8,   /* FFTError.unsupportedFFTSize     MetalContext.swift:14 */
    MIFFT_ERR_ALLOCATION = 9,         /* FFTError.bufferAllocationFailed MetalContext.swift:15 */
    MIFFT_ERR_PRECONDITION = 10,      /* precondition traps: Convolve.swift:20-21,
                                         FusedConvolver.swift:56-58, MultiChannelFFT.swift:40-41 */
    MIFFT_ERR_INVALID_ARGUMENT = 11   /* null pointer / negative count (no Swift counterpart:
                                         Swift arrays cannot be null) */
} mifft_status;

typedef enum mifft_direction {
    MIFFT_FORWARD = -1, /* exp(-2 pi i n k / N), unscaled */
    MIFFT_INVERSE = 1   /* conj(FFT(conj x)) / N */
} mifft_direction;

/* ------------------------------------------------------------- library */
/* Text of the FFTError description for a status (MetalContext.swift:17-29). */
MIFFT_API const char *mifft_status_string(mifft_status s);
/* Payload of the calling thread's last error: expected size, received size,
 * batch index (-1 where the case has none). Any pointer may be NULL. */
MIFFT_API mifft_status mifft_last_error(int64_t *expected, int64_t *got, int64_t *batch_index);
MIFFT_API const char *mifft_last_error_message(void);

/* Picks the current HIP device and checks it is usable.  Replaces
 * MetalContext.shared() (MetalContext.swift:79-95); called implicitly by
 * every create function, safe to call from several threads. */
MIFFT_API mifft_status mifft_init(void);
/* Releases process-wide resources (autotune table, device info). */
MIFFT_API mifft_status mifft_shutdown(void);
/* Device marketing name and gfx arch, e.g. "AMD Instinct MI355X (gfx950)". */
MIFFT_API mifft_status mifft_device_name(char *buf, size_t len);

/* ------------------------------------------------------ plan (MetalFFT) */
typedef struct mifft_plan mifft_plan;

/* MetalFFT.init(size:) (MetalFFT.swift:27-51): n in {64, ..., 16384} powers of
 * two, else MIFFT_ERR_UNSUPPORTED_SIZE (FFTDescriptor, MetalContext.swift:48-67). */
MIFFT_API mifft_status mifft_plan_create(int64_t n, mifft_plan **plan);
MIFFT_API mifft_status mifft_plan_destroy(mifft_plan *plan);
/* MetalFFT.size */
MIFFT_API int64_t mifft_plan_size(const mifft_plan *plan);
/* Name of the kernel variant the plan runs (chosen by the autotuner). */
MIFFT_API const char *mifft_plan_kernel_name(const mifft_plan *plan);

/* Kernel variants compiled for length n: writes up to `cap` names into
 * `names` (may be NULL) and returns how many exist.  The reference's fixed
 * size -> kernel table (MetalContext.swift:48-67, MetalFFT.swift:163-165) is
 * replaced by an autotuner over these; see mifft_plan_select_kernel. */
MIFFT_API int mifft_kernel_variants(int64_t n, const char **names, int cap);
/* Pins the plan to one named variant (overrides the autotuner's choice);
 * MIFFT_ERR_KERNEL_NOT_FOUND for an unknown name or another length. */
MIFFT_API mifft_status mifft_plan_select_kernel(mifft_plan *plan, const char *name);

/* Batched transform of `batch` contiguous length-n complex64 signals on the
 * device, asynchronous on `stream`.  in == out is allowed (each workgroup
 * loads its whole transform before storing).  Replaces dispatchBatch
 * (MetalFFT.swift:159-204) and, with batch = 1, dispatchSingle (:137-157). */
MIFFT_API mifft_status mifft_execute_device(const mifft_plan *plan, const float *in, float *out,
                                  int64_t batch, mifft_direction dir, void *stream);

/* MetalFFT.forward(_:) (MetalFFT.swift:69-85) and forward(input:output:)
 * (:88-104): in_count must equal n (else MIFFT_ERR_INVALID_INPUT_SIZE with
 * expected/got); `out` holds n complex values.  Blocking. */
MIFFT_API mifft_status mifft_forward_host(mifft_plan *plan, const float *in, int64_t in_count, float *out);
/* MetalFFT.inverse(_:) (MetalFFT.swift:56-65). Blocking. */
MIFFT_API mifft_status mifft_inverse_host(mifft_plan *plan, const float *in, int64_t in_count, float *out);
/* MetalFFT.forward(batch:) (MetalFFT.swift:108-133): inputs[i] holds
 * counts[i] complex values, each must equal n (else
 * MIFFT_ERR_BATCH_INPUT_SIZE with the first failing index); outputs[i]
 * receives n values.  batch == 0 is a no-op (the reference returns []). */
MIFFT_API mifft_status mifft_forward_batch_host(mifft_plan *plan, const float *const *inputs,
                                      const int64_t *counts, int64_t batch, float *const *outputs);
/* Batched inverse: one launch for all channels (the reference loops
 * inverse per channel, Sources/MetalFFT/FFT/MultiChannelFFT.swift:49-53). */
MIFFT_API mifft_status mifft_inverse_batch_host(mifft_plan *plan, const float *const *inputs,
                                      const int64_t *counts, int64_t batch, float *const *outputs);

#ifdef __cplusplus
}
#endif

#endif /* MIFFT_H */
