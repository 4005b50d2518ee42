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
 *
 * Scratch: objects (plans, convolvers, analysers, SAR processors) own their
 * device scratch and grow it on demand in an eager call.  A grown buffer
 * never frees the one it replaces before the object is destroyed, so a
 * hipGraph captured at one size stays valid after eager calls at larger
 * sizes.  *_device entry points issue no memsets (fill kernels instead) and
 * can be captured once their scratch has grown to the captured size in an
 * eager call.
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

/* Arithmetic of a plan.  The MFMA forms run every radix-16 stage of the
 * N = 4096 transform as a DFT-16 GEMM on the matrix cores with FP32
 * accumulation (the reference's CT-MMA variant, src/fft_4096_ct_mma.metal:88-206,
 * does this with 8x8 simdgroup matrices); SPLIT adds the hi/lo product
 * F_lo x_hi + F_hi x_lo + F_hi x_hi.  No reference counterpart for FP16/BF16. */
typedef enum mifft_precision {
    MIFFT_PREC_FP32 = 0,            /* FP32 Stockham on the VALU (all sizes) */
    MIFFT_PREC_FP16_MFMA = 1,       /* n = 4096 */
    MIFFT_PREC_BF16_MFMA = 2,       /* n = 4096 */
    MIFFT_PREC_FP16_MFMA_SPLIT = 3, /* n = 4096 */
    MIFFT_PREC_BF16_MFMA_SPLIT = 4  /* n = 4096 */
} mifft_precision;

/* Element type in device memory. */
typedef enum mifft_io {
    MIFFT_IO_C64 = 0, /* complex64: interleaved FP32 (re, im), 8 bytes */
    MIFFT_IO_C32H = 1 /* complex FP16: interleaved FP16 (re, im), 4 bytes (MFMA plans only) */
} mifft_io;

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

/* MetalFFT.init(size:) (MetalFFT.swift:27-51): n a power of two from 64 to
 * 16384 as FFTDescriptor (MetalContext.swift:48-67), or 32768 ... 2^20 (the
 * four-step lengths), else MIFFT_ERR_UNSUPPORTED_SIZE.
 * Creation allocates the plan's twiddle tables and stream and, for FP32
 * plans, the first creation of a length in a process times the compiled
 * kernel variants of that length (about 1 GiB of temporary device memory
 * for some tens of milliseconds; MIFFT_AUTOTUNE=0 skips it).  The choice is
 * kept in memory; it is written to disk only when MIFFT_AUTOTUNE_CACHE
 * names a file. */
MIFFT_API mifft_status mifft_plan_create(int64_t n, mifft_plan **plan);
/* Plan with an explicit arithmetic and device element type; the host entry
 * points below accept complex64 plans only.  MFMA precisions and complex-FP16
 * I/O exist for n = 4096 (else MIFFT_ERR_KERNEL_NOT_FOUND). */
MIFFT_API mifft_status mifft_plan_create_ex(int64_t n, mifft_precision precision, mifft_io io,
                                            mifft_plan **plan);
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
 * MIFFT_ERR_KERNEL_NOT_FOUND for an unknown name or another length.  Plans
 * of another precision (mifft_plan_create_ex) take the names of their four
 * DFT-as-GEMM forms, "mfma_<f16|bf16|f16split|bf16split>_4096_<c64|c32h>"
 * followed by nothing (DMA-staged, three radix-16 passes), "_oneshot",
 * "_r64" (two radix-64 passes) or "_r64_oneshot". */
MIFFT_API mifft_status mifft_plan_select_kernel(mifft_plan *plan, const char *name);

/* Batched transform of `batch` contiguous length-n complex64 signals on the
 * device, asynchronous on `stream`.  in == out is allowed (each workgroup
 * loads its whole transform before storing).  Replaces dispatchBatch
 * (MetalFFT.swift:159-204) and, with batch = 1, dispatchSingle (:137-157). */
MIFFT_API mifft_status mifft_execute_device(const mifft_plan *plan, const float *in, float *out,
                                  int64_t batch, mifft_direction dir, void *stream);

/* Real-input transforms of `batch` contiguous length-n float signals,
 * asynchronous on `stream`.  The reference's callers pack each real sample
 * as (x, 0) and run the complex transform (Convolve.swift:48-53,
 * FusedConvolver.swift:101-107, Correlate.swift:53-61, STFT.swift:116-122,
 * PSD.swift:37-45, FFTAnalyzer.swift:93-99); this computes the same spectra
 * with an n/2-point complex FFT of (x[2j], x[2j+1]) and a split, half the
 * work.  half = 0: y holds all n bins per signal (complex64, Hermitian: what
 * forward(input:output:) returns for the packed input); half = 1: bins
 * 0..n/2 (n/2 + 1 complex64 per signal).  x and y must not overlap.  FP32
 * plans only. */
MIFFT_API mifft_status mifft_forward_real_device(const mifft_plan *plan, const float *x, float *y,
                                                 int64_t batch, int half, void *stream);
/* Inverse of Hermitian spectra to real signals: y rows of n (half = 0) or
 * n/2 + 1 (half = 1) complex64, of which bins 0..n/2 are read;
 * x[b * n + j] = (1/n) sum_k X[k] exp(+2 pi i j k / n), the real part of
 * MetalFFT.inverse (MetalFFT.swift:56-65) that Convolve.swift:70-75 and
 * FusedConvolver keep.  x and y must not overlap. */
MIFFT_API mifft_status mifft_inverse_real_device(const mifft_plan *plan, const float *y, int half,
                                                 float *x, int64_t batch, void *stream);

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

/* ------------------------------------ convolver (Convolver / FusedConvolver) */
typedef struct mifft_convolver mifft_convolver;

/* FusedConvolver.Precision (Sources/MetalFFT/DSP/FusedConvolver.swift:27-36). */
typedef enum mifft_conv_precision {
    MIFFT_CONV_FP32 = 0,         /* .float32 */
    MIFFT_CONV_FP16_PURE = 1,    /* .float16Pure: FP16 arithmetic and storage */
    MIFFT_CONV_FP16_STORAGE = 2, /* .float16Storage: FP16 LDS storage, FP32 arithmetic */
    MIFFT_CONV_FP16_MIXED = 3    /* .float16Mixed: FP16 twiddle/filter products, FP32 butterflies */
} mifft_conv_precision;

/* Convolver.init(kernel:fftSize:) (Convolve.swift:19-32) with fused = 0, or
 * FusedConvolver.init(kernel:fftSize:precision:) (FusedConvolver.swift:55-86)
 * with fused = 1.  fft_size must be a supported FFT size greater than k
 * (MIFFT_ERR_PRECONDITION otherwise; the reference traps), and 4096 when
 * fused.  The taps' spectrum H = FFT(taps zero-padded) is computed on the
 * device once.  Both forms run the same one-launch fused kernel. */
MIFFT_API mifft_status mifft_convolver_create(const float *taps, int64_t k, int64_t fft_size,
                                              int fused, int precision, mifft_convolver **cv);
MIFFT_API mifft_status mifft_convolver_destroy(mifft_convolver *cv);
/* blockSize, hopSize (= fftSize - K + 1), kernelLen */
MIFFT_API int64_t mifft_convolver_block_size(const mifft_convolver *cv);
MIFFT_API int64_t mifft_convolver_hop_size(const mifft_convolver *cv);
MIFFT_API int64_t mifft_convolver_kernel_len(const mifft_convolver *cv);
/* Convolver.recommendedFFTSize(forKernelCount:) (Convolve.swift:83-86):
 * smallest supported size > kernel_count, or -1 (Swift nil). */
MIFFT_API int64_t mifft_convolver_recommended_fft_size(int64_t kernel_count);
/* apply(to:) (Convolve.swift:36-76, FusedConvolver.swift:95-133): full linear
 * convolution of one host signal, out has samples + K - 1 floats. Blocking. */
MIFFT_API mifft_status mifft_convolver_apply_host(mifft_convolver *cv, const float *signal,
                                                  int64_t samples, float *out);
/* `count` signals of `samples` floats each (contiguous) on the device ->
 * out[count][samples + K - 1], one launch, asynchronous on `stream`. */
MIFFT_API mifft_status mifft_convolver_apply_device(const mifft_convolver *cv, const float *signals,
                                                    int64_t samples, int64_t count, float *out,
                                                    void *stream);
/* Replaces the taps (K device floats) and recomputes H on `stream`. */
MIFFT_API mifft_status mifft_convolver_set_taps_device(mifft_convolver *cv, const float *taps,
                                                       void *stream);
/* Adjoint of apply_device (no reference counterpart): from grad_out
 * [count][samples + K - 1] writes grad_signal[count][samples] (correlation of
 * grad_out with the taps) and grad_taps[K] (sum over signals of the
 * correlation of grad_out with the signal).  Signal spectra are recomputed,
 * not stored; the tap sum is reduced in a fixed order (deterministic).  The
 * per-workgroup tap partials live in convolver-owned scratch: backward calls
 * on one convolver must not overlap in time (on one stream they are
 * ordered; on two streams the caller orders them with events). */
MIFFT_API mifft_status mifft_convolver_backward_device(mifft_convolver *cv, const float *signals,
                                                       int64_t samples, int64_t count,
                                                       const float *grad_out, float *grad_signal,
                                                       float *grad_taps, void *stream);

/* Adjoint that allocates nothing.  No counterpart in the reference; it is
 * for callers that cannot allocate device memory (code under hipGraph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * arguments and results as mifft_convolver_backward_device, but the tap
 * partials live in one fixed 16 MiB array per GPU (a static device
 * variable) instead of convolver-owned memory grown on demand.  When the
 * allocating form's partials fit there -- (slots + 2 ceil(slots / 64)) * K
 * floats, slots <= 2048 workgroups x transforms per workgroup, e.g. 8.6 MB
 * for 512 x 65536 samples, K = 1023, fft size 4096 -- the results are bit
 * for bit the same; otherwise the work is split over fewer workgroups (the
 * same arithmetic per block pair, the tap sum in another order, within
 * FP32 rounding).  A split that cannot fit even with one workgroup returns
 * MIFFT_ERR_INVALID_ARGUMENT with the bytes needed (expected = the array's
 * size, got = the need).  Calls on one GPU share the array: calls that
 * overlap in time on one GPU race. */
MIFFT_API mifft_status mifft_convolver_backward_device_noalloc(const mifft_convolver *cv,
                                                               const float *signals, int64_t samples,
                                                               int64_t count, const float *grad_out,
                                                               float *grad_signal, float *grad_taps,
                                                               void *stream);

/* ------------------------------------------------------------ correlator */
typedef struct mifft_correlator mifft_correlator;

/* Correlator.init(fftSize:) (Sources/MetalFFT/DSP/Correlate.swift:16-23). */
MIFFT_API mifft_status mifft_correlator_create(int64_t fft_size, mifft_correlator **c);
MIFFT_API mifft_status mifft_correlator_destroy(mifft_correlator *c);
MIFFT_API int64_t mifft_correlator_fft_size(const mifft_correlator *c);
/* cross(_:_:) (Correlate.swift:33-49): out[m] = sum_n a[(n+m) mod N] b[n]
 * with a, b truncated or zero-padded to N (:53-61); out has N floats. */
MIFFT_API mifft_status mifft_correlator_cross_host(mifft_correlator *c, const float *a, int64_t na,
                                                   const float *b, int64_t nb, float *out);
/* auto(_:) (Correlate.swift:27-29) = cross(x, x). */
MIFFT_API mifft_status mifft_correlator_auto_host(mifft_correlator *c, const float *x, int64_t n,
                                                  float *out);
/* `count` correlations on the device: a[count][na], b[count][nb] ->
 * out[count][N], one launch, asynchronous on `stream`. */
MIFFT_API mifft_status mifft_correlator_cross_device(const mifft_correlator *c, const float *a,
                                                     int64_t na, const float *b, int64_t nb,
                                                     int64_t count, float *out, void *stream);

/* ---------------------------------------------------------- ragged batches */
/* Transforms of different lengths in a few launches (no reference counterpart:
 * MetalFFT.forward(batch:) rejects uneven lengths, MetalFFT.swift:108-114). */
typedef struct mifft_ragged_plan mifft_ragged_plan;

/* Execution mode bits for mifft_ragged_execute_device.  Tickets: the
 * transforms of one length that one 512-thread workgroup pass holds (one
 * 16384, two 8192, or 8192 / 4096 elements of a shorter length in the tile
 * shape measured at plan creation), sorted longest length first. */
#define MIFFT_RAGGED_DEFAULT 0    /* one launch, one workgroup per ticket: the hardware
                                     dispatcher starts them in ticket order as CU slots
                                     free up */
#define MIFFT_RAGGED_PER_SIZE 1   /* one launch per distinct length (baseline) */
#define MIFFT_RAGGED_GRAPH 2      /* capture the launch(es) in a hipGraph once per
                                     (in, out, direction, mode) and replay it; a plan
                                     keeps up to four such graphs, replacing the least
                                     recently used */
#define MIFFT_RAGGED_QUEUE 4      /* one persistent launch: the resident workgroups take
                                     tickets from a device counter (the plan's counter
                                     makes calls on one plan exclusive) */
#define MIFFT_RAGGED_SPLIT 8      /* one launch per length of 4096 and longer, one
                                     persistent launch of 4096-element tiles for the
                                     shorter lengths */

/* lengths[i] in {64, ..., 16384} (else MIFFT_ERR_UNSUPPORTED_SIZE with
 * batch_index = i); offsets[i] = first complex element of transform i, or
 * NULL for back-to-back packing in the given order. */
MIFFT_API mifft_status mifft_ragged_plan_create(const int64_t *lengths, const int64_t *offsets,
                                                int64_t count, mifft_ragged_plan **plan);
MIFFT_API mifft_status mifft_ragged_plan_destroy(mifft_ragged_plan *plan);
/* Complex elements the buffers must hold: max(offset + length). */
MIFFT_API int64_t mifft_ragged_extent(const mifft_ragged_plan *plan);
/* Sum of lengths (for byte / flop accounting). */
MIFFT_API int64_t mifft_ragged_elements(const mifft_ragged_plan *plan);
/* Kernel launches one execution issues in `mode` (lengths >= 4096 present
 * plus one if any is shorter, or the distinct lengths). */
MIFFT_API int64_t mifft_ragged_launches(const mifft_ragged_plan *plan, int mode);
/* Forward or inverse of every transform, complex64 device buffers (in == out
 * allowed), asynchronous on `stream`.  Not thread-safe per plan. */
MIFFT_API mifft_status mifft_ragged_execute_device(mifft_ragged_plan *plan, const float *in,
                                                   float *out, mifft_direction dir, int mode,
                                                   void *stream);

/* -------------------------------------------- SAR range-Doppler processor */
/* Point-target SAR focusing by the range-Doppler algorithm without RCMC:
 * range compression (FFT, x conj(FFT(chirp)), IFFT per pulse), corner turn,
 * azimuth compression (FFT, x H_a, IFFT per range bin).  Reference:
 * src/radar/rda_fused_pipeline.swift:131-369, src/fft_sar_fused.metal:246-410,
 * src/radar/rda_kernels.metal:76-91,195-238; parameters and chirp reference
 * of src/radar/sar_simulator.swift:22-67,207-224. */
typedef struct mifft_rda mifft_rda;

typedef struct mifft_rda_params {
    int64_t n_range;       /* range samples per pulse (FFT length 64..16384) */
    int64_t n_azimuth;     /* pulses (FFT length 64..16384) */
    double bandwidth;      /* chirp bandwidth, Hz */
    double pulse_duration; /* s */
    double sampling_rate;  /* range sampling rate, Hz */
    double prf;            /* pulse repetition frequency, Hz */
    double velocity;       /* platform velocity, m/s */
    double carrier_freq;   /* Hz */
    double center_range;   /* scene centre slant range, m */
    int io;                /* mifft_io: MIFFT_IO_C32H (complex FP16) or MIFFT_IO_C64 */
    int filter_convention; /* MIFFT_RDA_MATCHED or MIFFT_RDA_REFERENCE_FORMULA */
    int rcmc;              /* 1: range cell migration correction between the azimuth FFT and
                              the azimuth filter (rda_kernels.metal:103-185), 0: none */
    int precision;         /* mifft_conv_precision of range and azimuth compression: FP32, or
                              the FP16 modes pure / storage / mixed of
                              src/fft_sar_fused_fp16.metal:10-34 (4096 x 4096 frames only,
                              else MIFFT_ERR_KERNEL_NOT_FOUND); the azimuth FFT before RCMC
                              and RCMC stay FP32, as in precision_comparison.swift:226-230 */
} mifft_rda_params;

/* Azimuth filter phase sign and RCMC direction.  MATCHED: H_a = exp(+i phi),
 * the conjugate of the echo's Doppler phase exp(-i phi) (echo phase
 * -4 pi R / lambda, sar_simulator.swift:144), which focuses the targets, and
 * RCMC reading range bin r + shift (migrated energy pulled back).
 * REFERENCE_FORMULA: the literal formulas of rda_kernels.metal:231-232
 * (H_a = exp(-i phi)) and :150 (srcIdx = rgIdx - shift), kept for comparison
 * with the reference (they do not focus: the filter does not compress in
 * azimuth and the RCMC doubles the migration). */
#define MIFFT_RDA_MATCHED 0
#define MIFFT_RDA_REFERENCE_FORMULA 1

/* The reference's defaults (sar_simulator.swift:22-67): 4096 x 4096,
 * B = 100 MHz, Tp = 10 us, fs = 1.2 B, PRF = 1 kHz, v = 100 m/s, fc = 10 GHz,
 * R0 = 20 km; complex-FP16 data; RCMC on (the reference's RDA pipelines
 * always run it, rda_fused_pipeline.swift:158-163). */
MIFFT_API mifft_status mifft_rda_default_params(mifft_rda_params *p);
MIFFT_API mifft_status mifft_rda_create(const mifft_rda_params *p, mifft_rda **rda);
MIFFT_API mifft_status mifft_rda_destroy(mifft_rda *rda);
/* Power-of-two gains applied after range and azimuth compression (they keep
 * complex FP16 in range: 2^-ceil(log2 chirp samples), 2^-ceil(log2 n_azimuth))
 * and the chirp reference length in samples. */
MIFFT_API mifft_status mifft_rda_scales(const mifft_rda *rda, double *range_scale,
                                        double *azimuth_scale, int64_t *chirp_samples);
/* raw: n_azimuth x n_range (pulse-major), image: n_range x n_azimuth
 * (range-bin-major, the layout after the corner turn); raw != image.
 * Asynchronous on `stream`; uses a plan-owned work buffer of one frame. */
MIFFT_API mifft_status mifft_rda_process_device(mifft_rda *rda, const void *raw, void *image,
                                                void *stream);
/* The three steps separately (same layouts; the corner turn is out of place).
 * With rcmc, azimuth compression is azimuth FFT -> RCMC -> filter -> IFFT,
 * through the plan's work frame (in and out may be the same buffer); calls
 * on one plan must not overlap. */
MIFFT_API mifft_status mifft_rda_range_compress_device(const mifft_rda *rda, const void *raw,
                                                       void *out, void *stream);
MIFFT_API mifft_status mifft_rda_corner_turn_device(const mifft_rda *rda, const void *in, void *out,
                                                    void *stream);
MIFFT_API mifft_status mifft_rda_azimuth_compress_device(const mifft_rda *rda, const void *in,
                                                         void *out, void *stream);

/* ------------------------------------------------------- cross-spectral */
typedef struct mifft_xspec mifft_xspec;

/* Analyser for PSD.crossSpectral(channels:fftSize:hopSize:sampleRate:window:)
 * (Sources/MetalFFT/DSP/CrossSpectral.swift:55-173): `channels` (1..16, else
 * MIFFT_ERR_PRECONDITION; the reference traps) signals, frames of fft_size
 * (a supported FFT size). */
MIFFT_API mifft_status mifft_xspec_create(int64_t fft_size, int64_t channels, mifft_xspec **x);
MIFFT_API mifft_status mifft_xspec_destroy(mifft_xspec *x);
/* Frames the analysis uses: pos = 0, hop, 2 hop, ... while pos + fft_size <=
 * samples (CrossSpectral.swift:95-96). */
MIFFT_API int64_t mifft_xspec_frames(const mifft_xspec *x, int64_t samples, int64_t hop);
/* Device data: signals[channels][samples] (FP32, rows back to back), window
 * [fft_size] multiplied into every frame (null: rectangular; the reference's
 * default is .hann, Window.swift:32-36), hop >= 1.  Outputs, each averaged
 * over the frames (CrossSpectralResult, CrossSpectral.swift:11-39;
 * fft_cross_spectral.metal:29-64):
 *   power[channels][fft_size]        |X_c(k)|^2
 *   cross[P][fft_size] (complex)     X_i(k) conj(X_j(k)), pairs (0,1), (0,2),
 *                                    ..., (C-2, C-1), P = C (C-1) / 2
 *   coherence[P][fft_size]           per-frame |S_ij|^2 / max(P_i P_j, 1e-30),
 *                                    averaged (the reference's definition: 1
 *                                    wherever both channels have power)
 *   welch_msc[P][fft_size] (may be null)  |<S_ij>|^2 / max(<P_i><P_j>, 1e-30),
 *                                    the coherence of the averaged spectra
 *                                    (no counterpart in the reference).
 * No frame (samples < fft_size): all outputs zero, as the reference's
 * accumulators.  Asynchronous on `stream`; scratch is analyser-owned, so
 * calls on one analyser must not overlap. */
MIFFT_API mifft_status mifft_xspec_device(mifft_xspec *x, const float *signals, int64_t samples,
                                          int64_t hop, const float *window, float *power,
                                          float *cross, float *coherence, float *welch_msc,
                                          void *stream);

/* Batched framing for STFT.analyze (Sources/MetalFFT/DSP/STFT.swift:70-99,
 * 116-128): every frame of every channel (frames as mifft_xspec_frames)
 * windowed and transformed in one launch instead of one synchronous FFT per
 * frame.  spectra[channels][frames][fft_size] complex64 (STFTFrame.complex);
 * power (may be null) |X|^2 (STFTFrame.magnitudes, Spectrum.swift:25-31);
 * power_db (may be null) db_alpha * log10(|X|^2) clipped below at -120 dB
 * (STFTFrame.magnitudesDB, Spectrum.toDecibels, Spectrum.swift:47-55).
 * toDecibels calls vDSP_vdbcon with flag 1, whose factor is 20 (amplitude
 * convention), although its comment says 10 log10: db_alpha = 20 gives what
 * the reference computes, 10 what its comment describes.  FFTAnalyzer.analyze
 * (FFTAnalyzer.swift:92-116) is the one-frame case (samples = fft_size).  No
 * frame: nothing is written.  Asynchronous on `stream`. */
MIFFT_API mifft_status mifft_stft_device(mifft_xspec *x, const float *signals, int64_t samples,
                                         int64_t hop, const float *window, float *spectra,
                                         float *power, float *power_db, float db_alpha,
                                         void *stream);
/* PSD.welch (Sources/MetalFFT/DSP/PSD.swift:18-59) per channel:
 * psd[channels][fft_size] = mean over frames of |X|^2 / (sum of window^2 *
 * sample_rate) (sample_rate > 0, else MIFFT_ERR_PRECONDITION); zeros when
 * there is no frame.  PSD.coherence (PSD.swift:72-125) is welch_msc of a
 * two-channel mifft_xspec_device.  Asynchronous on `stream`. */
MIFFT_API mifft_status mifft_welch_device(mifft_xspec *x, const float *signals, int64_t samples,
                                          int64_t hop, const float *window, double sample_rate,
                                          float *psd, void *stream);

/* Blocking host-buffer forms of the three analyses above (the reference's
 * calls are synchronous): inputs and outputs are host arrays of the same
 * layouts; the analyser stages them through device memory on its own
 * stream.  A window of NULL is rectangular. */
MIFFT_API mifft_status mifft_xspec_host(mifft_xspec *x, const float *signals, int64_t samples,
                                        int64_t hop, const float *window, float *power,
                                        float *cross, float *coherence, float *welch_msc);
MIFFT_API mifft_status mifft_stft_host(mifft_xspec *x, const float *signals, int64_t samples,
                                       int64_t hop, const float *window, float *spectra,
                                       float *power, float *power_db, float db_alpha);
MIFFT_API mifft_status mifft_welch_host(mifft_xspec *x, const float *signals, int64_t samples,
                                        int64_t hop, const float *window, double sample_rate,
                                        float *psd);

/* ------------------------------------------------------------ diagnostic */
/* No counterpart in the reference.  Copies `bytes` from src to dst (device
 * memory, not overlapping) with the access style of the FFT kernels
 * (16-byte non-temporal loads and stores, one-shot workgroups); the rate it
 * reaches is the achievable HBM streaming ceiling that bench.py reads the
 * transforms' roofline fractions against.  Asynchronous on `stream`. */
MIFFT_API mifft_status mifft_copy_device(const void *src, void *dst, int64_t bytes, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* MIFFT_H */
