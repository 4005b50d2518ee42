/*
 * oracle/cpu_fft.c -- CPU FP32 restatement of MetalFFT's FFT, four-step,
 * overlap-add convolution and correlation algorithms.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker for the HIP library and
 * the CPU baseline timed beside it.  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg load it; the product path never does.
 *
 * Every routine cites the reference file:line whose algorithm it restates.
 * The reference computes twiddles with fast-math sincos per butterfly and
 * chains w2 = w1^2, w3 = w1^3 (Sources/MetalFFT/Resources/fft_multisize.metal:48-55);
 * this restatement reads every twiddle W_N^k from one table generated in FP64
 * and rounded once to FP32 (the same table the GPU library builds), so the
 * FP32 parity check compares arithmetic, not twiddle error.
 *
 * Layout: complex64 interleaved (re, im) float pairs, as SIMD2<Float> in the
 * reference (Sources/MetalFFT/FFT/MetalFFT.swift:7-8).
 */
#define _GNU_SOURCE
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef struct { float re, im; } cf;

static inline cf cadd(cf a, cf b) { cf r = {a.re + b.re, a.im + b.im}; return r; }
static inline cf csub(cf a, cf b) { cf r = {a.re - b.re, a.im - b.im}; return r; }
/* cmul as written in fft_multisize.metal:25-27 */
static inline cf cmul(cf a, cf b) {
    cf r = {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
    return r;
}
/* multiply by -i: (x, y) -> (y, -x), fft_multisize.metal:35 */
static inline cf mul_mi(cf a) { cf r = {a.im, -a.re}; return r; }

/* W_N^k = exp(-2*pi*i*k/N) in FP64, rounded once to FP32.  Sign convention
 * from TWO_PI_OVER_N = -2*pi/N (fft_multisize.metal:70). */
void oracle_twiddle_table(int n, float *tw) {
    for (int k = 0; k < n; ++k) {
        double a = -2.0 * M_PI * (double)k / (double)n;
        tw[2 * k] = (float)cos(a);
        tw[2 * k + 1] = (float)sin(a);
    }
}

/* radix-4 butterfly, fft_multisize.metal:29-40 */
static inline void radix4(cf *x0, cf *x1, cf *x2, cf *x3) {
    cf t0 = cadd(*x0, *x2), t1 = cadd(*x1, *x3);
    cf t2 = csub(*x0, *x2), t3 = csub(*x1, *x3);
    cf t3r = mul_mi(t3);
    *x0 = cadd(t0, t1);
    *x1 = cadd(t2, t3r);
    *x2 = csub(t0, t1);
    *x3 = csub(t2, t3r);
}

/* One radix-4 Stockham pass at stride s (fft_multisize.metal:92-111):
 * butterfly t reads in[t + j*N/4], twiddles x_j by W_N^{j*pos*N/(4s)}
 * (pos = t mod s), writes out[grp*4s + pos + j*s] (grp = t / s). */
static void pass_r4(int n, int s, const cf *in, cf *out, const cf *tw) {
    int q = n / 4;
    int scale = n / (4 * s);
    for (int t = 0; t < q; ++t) {
        int pos = t % s, grp = t / s;
        cf x0 = in[t], x1 = in[t + q], x2 = in[t + 2 * q], x3 = in[t + 3 * q];
        if (s > 1) {
            int k = pos * scale;
            x1 = cmul(x1, tw[k]);
            x2 = cmul(x2, tw[(2 * k) % n]);
            x3 = cmul(x3, tw[(3 * k) % n]);
        }
        radix4(&x0, &x1, &x2, &x3);
        int wr = grp * 4 * s + pos;
        out[wr] = x0; out[wr + s] = x1; out[wr + 2 * s] = x2; out[wr + 3 * s] = x3;
    }
}

/* Final radix-2 pass at stride s = N/2 (fft_multisize.metal:530-547 for
 * N=2048, :898-922 for N=128). */
static void pass_r2(int n, int s, const cf *in, cf *out, const cf *tw) {
    int h = n / 2;
    int scale = n / (2 * s);
    for (int t = 0; t < h; ++t) {
        int pos = t % s, grp = t / s;
        cf x0 = in[t], x1 = in[t + h];
        x1 = cmul(x1, tw[pos * scale]);
        int wr = grp * 2 * s + pos;
        out[wr] = cadd(x0, x1);
        out[wr + s] = csub(x0, x1);
    }
}

/* Radix-4 Stockham FFT for N = 4^a (64, 256, 1024, 4096) or 2*4^a (128, 512,
 * 2048): a radix-4 passes then one radix-2 pass when log2 N is odd
 * (fft_multisize.metal:62-153, :160-281, :288-399, :406-548, :556-687,
 * :759-923).  Ping-pongs between out and scratch; result in out. */
static void stockham_r4(int n, const cf *in, cf *out, cf *scratch, const cf *tw) {
    int lg = 0;
    while ((1 << lg) < n) ++lg;
    int r4 = lg / 2, r2 = lg & 1;
    int npass = r4 + r2;
    /* choose the first destination so that the last pass lands in out */
    cf *bufs[2] = {out, scratch};
    int dst = (npass % 2 == 1) ? 0 : 1;
    const cf *src = in;
    int s = 1;
    for (int p = 0; p < r4; ++p) {
        pass_r4(n, s, src, bufs[dst], tw);
        src = bufs[dst];
        dst ^= 1;
        s *= 4;
    }
    if (r2) pass_r2(n, s, src, bufs[dst], tw);
}

/* Split-radix DIT radix-8 butterfly, same dataflow as
 * Sources/MetalFFT/Resources/fft_4096_batched.metal:36-59. */
static void radix8(cf x[8]) {
    const float h = 0.70710678118654752f;
    cf t0 = cadd(x[0], x[4]), t1 = cadd(x[1], x[5]);
    cf t2 = cadd(x[2], x[6]), t3 = cadd(x[3], x[7]);
    cf t4 = csub(x[0], x[4]), t5 = csub(x[1], x[5]);
    cf t6 = csub(x[2], x[6]), t7 = csub(x[3], x[7]);
    cf t5w = {h * (t5.re + t5.im), h * (t5.im - t5.re)};
    cf t6w = mul_mi(t6);
    cf t7w = {h * (-t7.re + t7.im), h * (-t7.im - t7.re)};
    cf u0 = cadd(t0, t2), u1 = cadd(t1, t3);
    cf u2 = csub(t0, t2), u3 = csub(t1, t3);
    cf u3r = mul_mi(u3);
    cf v0 = cadd(t4, t6w), v1 = cadd(t5w, t7w);
    cf v2 = csub(t4, t6w), v3 = csub(t5w, t7w);
    cf v3r = mul_mi(v3);
    x[0] = cadd(u0, u1); x[4] = csub(u0, u1);
    x[2] = cadd(u2, u3r); x[6] = csub(u2, u3r);
    x[1] = cadd(v0, v1); x[5] = csub(v0, v1);
    x[3] = cadd(v2, v3r); x[7] = csub(v2, v3r);
}

/* Radix-8 Stockham N=4096: 4 passes at strides 1, 8, 64, 512 with twiddle
 * scales 512/64/8/1 (fft_4096_batched.metal:79-203). */
static void stockham_r8_4096(const cf *in, cf *out, cf *scratch, const cf *tw) {
    const int n = 4096, q = 512;
    cf *bufs[2] = {scratch, out};
    const cf *src = in;
    int s = 1;
    for (int p = 0; p < 4; ++p) {
        cf *dst = bufs[p & 1];
        int scale = n / (8 * s);
        for (int t = 0; t < q; ++t) {
            int pos = t % s, grp = t / s;
            cf x[8];
            for (int j = 0; j < 8; ++j) x[j] = src[t + j * q];
            if (s > 1) {
                int k = pos * scale;
                for (int j = 1; j < 8; ++j) x[j] = cmul(x[j], tw[(j * k) % n]);
            }
            radix8(x);
            int wr = grp * 8 * s + pos;
            for (int j = 0; j < 8; ++j) dst[wr + j * s] = x[j];
        }
        src = dst;
        s *= 8;
    }
}

/* Four-step N = n1*n2 (Sources/MetalFFT/FFT/MetalFFT.swift:206-277 with the
 * kernels fft_transpose / fft_twiddle_transpose of
 * Sources/MetalFFT/Resources/fft_multisize.metal:700-749):
 *   0: transpose n2 x n1 -> n1 x n2
 *   1: n1 row FFTs of size n2
 *   2: twiddle W_N^{row*col} + transpose n1 x n2 -> n2 x n1
 *   3: n2 row FFTs of size n1
 *   4: transpose n2 x n1 -> n1 x n2                               */
static void fourstep(int n, int n1, int n2, const cf *in, cf *out, cf *ta, cf *tb,
                     cf *scratch, const cf *tw, const cf *tw1, const cf *tw2) {
    for (int g = 0; g < n; ++g) {            /* ROWS = n2, COLS = n1 */
        int row = g / n1, col = g % n1;
        ta[col * n2 + row] = in[g];
    }
    for (int r = 0; r < n1; ++r) stockham_r4(n2, ta + r * n2, tb + r * n2, scratch, tw2);
    for (int g = 0; g < n; ++g) {            /* N1 = n1, N2 = n2 */
        int row = g / n2, col = g % n2;
        ta[col * n1 + row] = cmul(tb[g], tw[(int)(((int64_t)row * col) % n)]);
    }
    for (int r = 0; r < n2; ++r) stockham_r4(n1, ta + r * n1, tb + r * n1, scratch, tw1);
    for (int g = 0; g < n; ++g) {            /* ROWS = n2, COLS = n1 */
        int row = g / n1, col = g % n1;
        out[col * n2 + row] = tb[g];
    }
}

static int ilog2(int n) { int l = 0; while ((1 << l) < n) ++l; return l; }
/* The reference's sizes (64 ... 16384) and, beyond them, the four-step
 * lengths 32768 ... 2^20 of the GPU library. */
static int supported(int n) { return n >= 64 && n <= (1 << 20) && (n & (n - 1)) == 0; }

/* Size table of Sources/MetalFFT/Internal/MetalContext.swift:48-67 for 8192
 * and 16384.  The reference stops at 16384; longer lengths keep its
 * procedure with n1 = 2^floor(L/2) <= n2 = 2^ceil(L/2) (n = 2^L), the same
 * orientation as its 8192 = 64 x 128 entry. */
static void fourstep_split(int n, int *n1, int *n2) {
    if (n == 8192) { *n1 = 64; *n2 = 128; }
    else if (n == 16384) { *n1 = 128; *n2 = 128; }
    else {
        int l = ilog2(n);
        *n1 = 1 << (l / 2);
        *n2 = n / *n1;
    }
}

/* One forward transform following the reference's dispatch
 * (MetalFFT.swift:137-204): single-pass radix-4 for N <= 4096, radix-8 for
 * the N=4096 batched path, four-step for 8192/16384.
 * kernel: 0 = radix-4 single pass, 1 = radix-8 (N=4096 only). */
static void forward_one(int n, int kernel, const cf *in, cf *out, cf *work,
                        const cf *tw, const cf *tw1, const cf *tw2) {
    if (n <= 4096) {
        if (kernel == 1 && n == 4096) stockham_r8_4096(in, out, work, tw);
        else stockham_r4(n, in, out, work, tw);
    } else {
        int n1, n2;
        fourstep_split(n, &n1, &n2);
        fourstep(n, n1, n2, in, out, work, work + n, work + 2 * n, tw, tw1, tw2);
    }
}

/* Batched forward (inverse=0) or inverse (inverse=1) over `batch` contiguous
 * transforms.  Inverse is the conjugate trick of MetalFFT.swift:56-65:
 * conj on input, forward, conj and multiply by Float(1)/Float(N) on output.
 * kernel selects the radix-8 4096 path (MetalFFT.swift:163-165) when 1.
 * threads > 1 splits the batch across OpenMP threads (CPU baseline only).
 * Returns 0, or -1 for an unsupported size (FFTError.unsupportedFFTSize). */
int oracle_fft(int n, int64_t batch, int inverse, int kernel, const float *inp,
               float *outp, int threads) {
    if (!supported(n)) return -1;
    cf *tw = (cf *)malloc(sizeof(cf) * (size_t)n);
    oracle_twiddle_table(n, (float *)tw);
    cf *tw1 = NULL, *tw2 = NULL;
    if (n > 4096) {
        int n1, n2;
        fourstep_split(n, &n1, &n2);
        tw1 = (cf *)malloc(sizeof(cf) * (size_t)n1);
        tw2 = (cf *)malloc(sizeof(cf) * (size_t)n2);
        oracle_twiddle_table(n1, (float *)tw1);
        oracle_twiddle_table(n2, (float *)tw2);
    }
    const cf *in = (const cf *)inp;
    cf *out = (cf *)outp;
    const float inv_n = 1.0f / (float)n;
    (void)threads;
#ifdef _OPENMP
    if (threads < 1) threads = 1;
#pragma omp parallel num_threads(threads)
#endif
    {
        cf *work = (cf *)malloc(sizeof(cf) * (size_t)n * 4);
        cf *tmp = (cf *)malloc(sizeof(cf) * (size_t)n);
#ifdef _OPENMP
#pragma omp for schedule(static)
#endif
        for (int64_t b = 0; b < batch; ++b) {
            const cf *x = in + b * n;
            cf *y = out + b * n;
            if (inverse) {
                for (int i = 0; i < n; ++i) { tmp[i].re = x[i].re; tmp[i].im = -x[i].im; }
                forward_one(n, kernel, tmp, y, work, tw, tw1, tw2);
                for (int i = 0; i < n; ++i) { y[i].re = y[i].re * inv_n; y[i].im = -y[i].im * inv_n; }
            } else {
                forward_one(n, kernel, x, y, work, tw, tw1, tw2);
            }
        }
        free(work);
        free(tmp);
    }
    free(tw); free(tw1); free(tw2);
    return 0;
}

/* Overlap-add FIR convolution, Sources/MetalFFT/DSP/Convolve.swift:19-76:
 * hop = fftSize - K + 1; each block packs signal[pos:pos+hop] as complex with
 * zero imaginary part, zero-pads to fftSize, forward FFT, multiplies by the
 * kernel spectrum (:58-62), inverse via the conjugate trick, and adds the
 * real part into output[pos + i] for i < fftSize (:67-70).  Output length is
 * S + K - 1.  Returns -1 unless 0 < K < fftSize and fftSize is supported
 * (the reference traps in a precondition, :20-21). */
int oracle_convolve(const float *signal, int64_t S, const float *taps, int K, int fft_size,
                    float *output) {
    if (!supported(fft_size) || K <= 0 || K >= fft_size) return -1;
    int n = fft_size, hop = n - K + 1;
    cf *kc = (cf *)calloc((size_t)n, sizeof(cf));
    cf *ks = (cf *)malloc(sizeof(cf) * (size_t)n);
    cf *blk = (cf *)malloc(sizeof(cf) * (size_t)n);
    cf *spec = (cf *)malloc(sizeof(cf) * (size_t)n);
    cf *res = (cf *)malloc(sizeof(cf) * (size_t)n);
    for (int i = 0; i < K; ++i) kc[i].re = taps[i];
    oracle_fft(n, 1, 0, 0, (const float *)kc, (float *)ks, 1);
    int64_t out_len = S + K - 1;
    memset(output, 0, sizeof(float) * (size_t)out_len);
    for (int64_t pos = 0; pos < S; pos += hop) {
        int64_t chunk = S - pos < hop ? S - pos : hop;
        for (int i = 0; i < n; ++i) {
            blk[i].re = i < chunk ? signal[pos + i] : 0.0f;
            blk[i].im = 0.0f;
        }
        oracle_fft(n, 1, 0, 0, (const float *)blk, (float *)spec, 1);
        for (int i = 0; i < n; ++i) {
            cf s = spec[i], k = ks[i];
            spec[i].re = s.re * k.re - s.im * k.im;
            spec[i].im = s.re * k.im + s.im * k.re;
        }
        oracle_fft(n, 1, 1, 0, (const float *)spec, (float *)res, 1);
        for (int i = 0; i < n; ++i) {
            int64_t o = pos + i;
            if (o < out_len) output[o] += res[i].re;
        }
    }
    free(kc); free(ks); free(blk); free(spec); free(res);
    return 0;
}

/* Circular cross-correlation, Sources/MetalFFT/DSP/Correlate.swift:33-61:
 * a, b truncated or zero-padded to fftSize (:53-61), r = Re IFFT(A * conj B)
 * with the product written as (:40-45).  r[m] = sum_n a[(n+m) mod N] b[n]. */
int oracle_correlate(const float *a, int64_t na, const float *b, int64_t nb, int fft_size,
                     float *out) {
    if (!supported(fft_size)) return -1;
    int n = fft_size;
    cf *pa = (cf *)calloc((size_t)n, sizeof(cf));
    cf *pb = (cf *)calloc((size_t)n, sizeof(cf));
    cf *fa = (cf *)malloc(sizeof(cf) * (size_t)n);
    cf *fb = (cf *)malloc(sizeof(cf) * (size_t)n);
    for (int64_t i = 0; i < na && i < n; ++i) pa[i].re = a[i];
    for (int64_t i = 0; i < nb && i < n; ++i) pb[i].re = b[i];
    oracle_fft(n, 1, 0, 0, (const float *)pa, (float *)fa, 1);
    oracle_fft(n, 1, 0, 0, (const float *)pb, (float *)fb, 1);
    for (int i = 0; i < n; ++i) {
        cf x = fa[i], y = fb[i];
        pa[i].re = x.re * y.re + x.im * y.im;
        pa[i].im = x.im * y.re - x.re * y.im;
    }
    oracle_fft(n, 1, 1, 0, (const float *)pa, (float *)fa, 1);
    for (int i = 0; i < n; ++i) out[i] = fa[i].re;
    free(pa); free(pb); free(fa); free(fb);
    return 0;
}

/* FP64 brute-force DFT, O(N^2), for pinning the FP64 library FFT on small
 * sizes: X[k] = sum_n x[n] exp(-2 pi i n k / N) (sign = fft_multisize.metal:70). */
void oracle_dft_f64(int n, const double *in, double *out, int inverse) {
    double sgn = inverse ? 1.0 : -1.0;
    for (int k = 0; k < n; ++k) {
        double sr = 0.0, si = 0.0;
        for (int j = 0; j < n; ++j) {
            int64_t m = ((int64_t)j * k) % n;
            double a = sgn * 2.0 * M_PI * (double)m / (double)n;
            double c = cos(a), s = sin(a);
            double xr = in[2 * j], xi = in[2 * j + 1];
            sr += xr * c - xi * s;
            si += xr * s + xi * c;
        }
        out[2 * k] = sr;
        out[2 * k + 1] = si;
    }
}

int oracle_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* Raw SAR echo of point targets, following src/radar/sar_simulator.swift:93-166
 * (flat-earth broadside geometry): for pulse az at slow time
 * eta = -(Na/2)/PRF + az/PRF and range sample rg at fast time
 * tau = 2 Rc/c - (Nr/2)/fs + rg/fs, each target (R0, x_t, a) adds
 * a * exp(i (pi Kr dt^2 - 4 pi R(eta)/lambda)) when |dt| <= Tp/2, with
 * R(eta) = sqrt(R0^2 + (v eta - x_t)^2), dt = tau - 2 R(eta)/c.  The target
 * amplitude a is complex (the reference uses real amplitudes).
 * prm = {B, Tp, fs, PRF, v, fc, Rc}; targets = {R0, x_t, re(a), im(a)} each;
 * out: complex128, Na x Nr, accumulated (caller zeroes it). */
void oracle_sar_simulate(int64_t nr, int64_t na, const double *prm, int ntargets,
                         const double *targets, double *out, int threads) {
    const double c = 299792458.0;
    const double B = prm[0], Tp = prm[1], fs = prm[2], prf = prm[3], v = prm[4], fc = prm[5],
                 rc = prm[6];
    const double kr = B / Tp, lambda = c / fc;
    const double t0 = rc * 2.0 / c - (double)(nr / 2) / fs;
    const double eta0 = -(double)(na / 2) / prf;
    (void)threads;
#ifdef _OPENMP
    if (threads < 1) threads = 1;
#pragma omp parallel for schedule(static) num_threads(threads)
#endif
    for (int64_t az = 0; az < na; ++az) {
        const double eta = eta0 + (double)az / prf;
        for (int k = 0; k < ntargets; ++k) {
            const double r0 = targets[4 * k], xt = targets[4 * k + 1];
            const double ar = targets[4 * k + 2], ai = targets[4 * k + 3];
            const double dx = v * eta - xt;
            const double reta = sqrt(r0 * r0 + dx * dx);
            const double tau0 = 2.0 * reta / c;
            /* samples with |tau - tau0| <= Tp/2 */
            int64_t lo = (int64_t)ceil((tau0 - Tp / 2.0 - t0) * fs) - 1;
            int64_t hi = (int64_t)floor((tau0 + Tp / 2.0 - t0) * fs) + 1;
            if (lo < 0) lo = 0;
            if (hi > nr - 1) hi = nr - 1;
            const double rphase = -4.0 * M_PI * reta / lambda;
            for (int64_t rg = lo; rg <= hi; ++rg) {
                const double dt = t0 + (double)rg / fs - tau0;
                if (fabs(dt) > Tp / 2.0) continue;
                const double ph = M_PI * kr * dt * dt + rphase;
                const double cr = cos(ph), ci = sin(ph);
                double *o = out + 2 * (az * nr + rg);
                o[0] += ar * cr - ai * ci;
                o[1] += ar * ci + ai * cr;
            }
        }
    }
}
