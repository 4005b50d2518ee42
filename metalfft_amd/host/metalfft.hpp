// C++ mirror of MetalFFT's Swift API over the libmifft C ABI (include/mifft.h).
//
//   mifft::MetalFFT         MetalFFT            Sources/MetalFFT/FFT/MetalFFT.swift:10-133
//   mifft::MultiChannelFFT  MultiChannelFFT     Sources/MetalFFT/FFT/MultiChannelFFT.swift:19-54
//   mifft::Convolver        Convolver           Sources/MetalFFT/DSP/Convolve.swift:8-87
//   mifft::FusedConvolver   FusedConvolver      Sources/MetalFFT/DSP/FusedConvolver.swift:17-134
//   mifft::Correlator       Correlator          Sources/MetalFFT/DSP/Correlate.swift:7-62
//   mifft::WindowType       WindowType          Sources/MetalFFT/DSP/Window.swift:6-24
//   mifft::STFT, STFTFrame  STFT                Sources/MetalFFT/DSP/STFT.swift:5-99
//   mifft::FFTAnalyzer      FFTAnalyzer         Sources/MetalFFT/DSP/FFTAnalyzer.swift:66-116
//   mifft::PSD              PSD                 Sources/MetalFFT/DSP/PSD.swift:6-134,
//                                               CrossSpectral.swift:11-173
//   mifft::FFTError         FFTError            Sources/MetalFFT/Internal/MetalContext.swift:6-30
//
// Methods throw FFTError where the Swift ones throw; complex values are
// std::complex<float> (same interleaved layout as SIMD2<Float>).
#pragma once

#include <cmath>
#include <complex>
#include <cstdint>
#include <optional>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/mifft.h"

namespace mifft {

using cfloat = std::complex<float>;
using Signal = std::vector<cfloat>;

class FFTError : public std::runtime_error {
   public:
    FFTError(mifft_status s, const std::string& msg, int64_t expected, int64_t got, int64_t index)
        : std::runtime_error(msg), status(s), expected(expected), got(got), batchIndex(index) {}
    mifft_status status;
    int64_t expected, got, batchIndex;  // -1 where the case has no payload

    static void check(mifft_status s) {
        if (s == MIFFT_OK) return;
        int64_t e = -1, g = -1, i = -1;
        mifft_last_error(&e, &g, &i);
        throw FFTError(s, mifft_last_error_message(), e, g, i);
    }
};

inline const float* fp(const cfloat* p) { return reinterpret_cast<const float*>(p); }
inline float* fp(cfloat* p) { return reinterpret_cast<float*>(p); }

class MetalFFT {
   public:
    explicit MetalFFT(int64_t size) : size(size) { FFTError::check(mifft_plan_create(size, &plan_)); }
    ~MetalFFT() { mifft_plan_destroy(plan_); }
    MetalFFT(const MetalFFT&) = delete;
    MetalFFT& operator=(const MetalFFT&) = delete;

    const int64_t size;

    Signal forward(const Signal& input) {
        Signal out((size_t)size);
        FFTError::check(mifft_forward_host(plan_, fp(input.data()), (int64_t)input.size(), fp(out.data())));
        return out;
    }
    // forward(input:output:): resizes `output` when its count differs
    void forward(const cfloat* input, int64_t count, Signal& output) {
        if (count != size)
            FFTError::check(mifft_forward_host(plan_, fp(input), count, nullptr));
        if ((int64_t)output.size() != size) output.assign((size_t)size, cfloat());
        FFTError::check(mifft_forward_host(plan_, fp(input), count, fp(output.data())));
    }
    Signal inverse(const Signal& input) {
        Signal out((size_t)size);
        FFTError::check(mifft_inverse_host(plan_, fp(input.data()), (int64_t)input.size(), fp(out.data())));
        return out;
    }
    std::vector<Signal> forwardBatch(const std::vector<Signal>& input) { return batch(input, false); }
    std::vector<Signal> inverseBatch(const std::vector<Signal>& input) { return batch(input, true); }
    mifft_plan* handle() { return plan_; }

   private:
    std::vector<Signal> batch(const std::vector<Signal>& input, bool inv) {
        if (input.empty()) return {};
        std::vector<Signal> out(input.size(), Signal((size_t)size));
        std::vector<const float*> ins;
        std::vector<float*> outs;
        std::vector<int64_t> counts;
        for (size_t i = 0; i < input.size(); ++i) {
            ins.push_back(fp(input[i].data()));
            outs.push_back(fp(out[i].data()));
            counts.push_back((int64_t)input[i].size());
        }
        auto fn = inv ? mifft_inverse_batch_host : mifft_forward_batch_host;
        FFTError::check(fn(plan_, ins.data(), counts.data(), (int64_t)input.size(), outs.data()));
        return out;
    }
    mifft_plan* plan_ = nullptr;
};

class MultiChannelFFT {
   public:
    MultiChannelFFT(int64_t channels, int64_t size) : channels(channels), size(size), fft_(size) {
        if (channels <= 0)
            throw FFTError(MIFFT_ERR_PRECONDITION, "channels must be > 0", -1, channels, -1);
    }
    const int64_t channels, size;
    std::vector<Signal> forward(const std::vector<Signal>& inputs) {
        check(inputs);
        return fft_.forwardBatch(inputs);
    }
    std::vector<Signal> inverse(const std::vector<Signal>& inputs) {
        check(inputs);
        return fft_.inverseBatch(inputs);
    }

   private:
    void check(const std::vector<Signal>& inputs) const {
        if ((int64_t)inputs.size() != channels)
            throw FFTError(MIFFT_ERR_PRECONDITION,
                           "Expected " + std::to_string(channels) + " channels, got " +
                               std::to_string(inputs.size()),
                           channels, (int64_t)inputs.size(), -1);
    }
    MetalFFT fft_;
};

class Convolver {
   public:
    Convolver(const std::vector<float>& kernel, int64_t fftSize) : Convolver(kernel, fftSize, 0, 0) {}
    ~Convolver() { mifft_convolver_destroy(cv_); }
    Convolver(const Convolver&) = delete;
    Convolver& operator=(const Convolver&) = delete;

    int64_t blockSize() const { return mifft_convolver_block_size(cv_); }
    int64_t hopSize() const { return mifft_convolver_hop_size(cv_); }
    int64_t kernelLen() const { return mifft_convolver_kernel_len(cv_); }

    std::vector<float> apply(const std::vector<float>& signal) {
        std::vector<float> out(signal.size() + (size_t)kernelLen() - 1);
        FFTError::check(mifft_convolver_apply_host(cv_, signal.data(), (int64_t)signal.size(), out.data()));
        return out;
    }
    static std::optional<int64_t> recommendedFFTSize(int64_t kernelCount) {
        const int64_t n = mifft_convolver_recommended_fft_size(kernelCount);
        if (n < 0) return std::nullopt;
        return n;
    }

   protected:
    Convolver(const std::vector<float>& kernel, int64_t fftSize, int fused, int precision) {
        FFTError::check(mifft_convolver_create(kernel.data(), (int64_t)kernel.size(), fftSize, fused,
                                               precision, &cv_));
    }
    mifft_convolver* cv_ = nullptr;
};

class FusedConvolver : public Convolver {
   public:
    enum class Precision { float16Pure = 1, float16Storage = 2, float16Mixed = 3, float32 = 0 };
    static constexpr int64_t supportedSize = 4096;
    explicit FusedConvolver(const std::vector<float>& kernel, int64_t fftSize = 4096,
                            Precision precision = Precision::float32)
        : Convolver(kernel, fftSize, 1, (int)precision), precision(precision) {}
    const Precision precision;
};

class Correlator {
   public:
    explicit Correlator(int64_t fftSize) : fftSize(fftSize) {
        FFTError::check(mifft_correlator_create(fftSize, &c_));
    }
    ~Correlator() { mifft_correlator_destroy(c_); }
    Correlator(const Correlator&) = delete;
    Correlator& operator=(const Correlator&) = delete;
    const int64_t fftSize;

    std::vector<float> cross(const std::vector<float>& a, const std::vector<float>& b) {
        std::vector<float> out((size_t)fftSize);
        FFTError::check(mifft_correlator_cross_host(c_, a.data(), (int64_t)a.size(), b.data(),
                                                    (int64_t)b.size(), out.data()));
        return out;
    }
    std::vector<float> autoCorrelation(const std::vector<float>& x) {  // Swift `auto(_:)`
        std::vector<float> out((size_t)fftSize);
        FFTError::check(mifft_correlator_auto_host(c_, x.data(), (int64_t)x.size(), out.data()));
        return out;
    }

   private:
    mifft_correlator* c_ = nullptr;
};

// WindowType.coefficients (Window.swift:14-23): the vDSP windows from their
// published definitions (periodic, denominator N), the reference's flat-top
// sum and Kaiser.
struct WindowType {
    enum Kind { hann, hamming, blackman, flatTop, kaiser, rectangular } kind = hann;
    double beta = 6.0;  // kaiser only
    static WindowType Kaiser(double beta) { return WindowType{kaiser, beta}; }
    std::vector<float> coefficients(int64_t n) const {
        std::vector<float> w((size_t)n);
        const double pi = 3.14159265358979323846;
        auto i0 = [](double x) {  // power series of I0, as Window.swift:99-110
            double r = 1.0, t = 1.0, h = x / 2.0;
            for (int k = 1; k <= 30; ++k) {
                t *= h * h / (double)(k * k);
                r += t;
                if (t < 1e-12 * r) break;
            }
            return r;
        };
        for (int64_t k = 0; k < n; ++k) {
            const double x = (double)k / (double)n;
            double v = 1.0;
            switch (kind) {
                case hann: v = 0.5 * (1.0 - std::cos(2 * pi * x)); break;
                case hamming: v = 0.54 - 0.46 * std::cos(2 * pi * x); break;
                case blackman: v = 0.42 - 0.5 * std::cos(2 * pi * x) + 0.08 * std::cos(4 * pi * x); break;
                case flatTop:
                    v = 0.21557895 - 0.41663158 * std::cos(2 * pi * x) + 0.277263158 * std::cos(4 * pi * x) -
                        0.083578947 * std::cos(6 * pi * x) + 0.006947368 * std::cos(8 * pi * x);
                    break;
                case kaiser: {
                    const double half = (double)(n - 1) / 2.0, y = ((double)k - half) / half;
                    v = i0(beta * std::sqrt(std::fmax(0.0, 1.0 - y * y))) / i0(beta);
                    break;
                }
                case rectangular: v = 1.0; break;
            }
            w[(size_t)k] = (float)v;
        }
        return w;
    }
};

// Analyser of `channels` signals at one FFT size (mifft_xspec): the frame
// list and scratch shared by STFT, Welch and cross-spectral analysis.
class Analyser {
   public:
    Analyser(int64_t fftSize, int64_t channels) : fftSize(fftSize), channels(channels) {
        FFTError::check(mifft_xspec_create(fftSize, channels, &x_));
    }
    ~Analyser() { mifft_xspec_destroy(x_); }
    Analyser(const Analyser&) = delete;
    Analyser& operator=(const Analyser&) = delete;
    const int64_t fftSize, channels;
    int64_t frames(int64_t samples, int64_t hop) const { return mifft_xspec_frames(x_, samples, hop); }
    mifft_xspec* handle() { return x_; }

   private:
    mifft_xspec* x_ = nullptr;
};

// STFTFrame (STFT.swift:5-15); magnitudesDB as Spectrum.toDecibels computes
// it (vDSP_vdbcon flag 1: 20 log10, floor -120).
struct STFTFrame {
    Signal complex;
    std::vector<float> magnitudes;
    std::vector<float> magnitudesDB() const {
        std::vector<float> d(magnitudes.size());
        for (size_t i = 0; i < d.size(); ++i) d[i] = std::fmax(20.f * std::log10(magnitudes[i]), -120.f);
        return d;
    }
    std::vector<float> phase() const {
        std::vector<float> p(complex.size());
        for (size_t i = 0; i < p.size(); ++i) p[i] = std::atan2(complex[i].imag(), complex[i].real());
        return p;
    }
};

// STFT (STFT.swift:20-99): every frame of the signal in one launch.
class STFT {
   public:
    STFT(int64_t fftSize, int64_t hopSize, WindowType window = WindowType{}, double sampleRate = 44100.0)
        : fftSize(fftSize), hopSize(hopSize), sampleRate(sampleRate), an_(fftSize, 1),
          window_(window.coefficients(fftSize)) {
        if (hopSize < 1) throw FFTError(MIFFT_ERR_PRECONDITION, "hopSize must be >= 1", -1, hopSize, -1);
    }
    const int64_t fftSize, hopSize;
    const double sampleRate;
    int64_t frameCount(int64_t sampleCount) const { return an_.frames(sampleCount, hopSize); }
    double binFrequency(int64_t bin) const { return (double)bin * sampleRate / (double)fftSize; }
    double frameTime(int64_t frameIndex) const { return (double)(frameIndex * hopSize) / sampleRate; }

    std::vector<STFTFrame> analyze(const std::vector<float>& signal) {
        const int64_t f = frameCount((int64_t)signal.size());
        std::vector<STFTFrame> out((size_t)f);
        if (f == 0) return out;
        Signal spec((size_t)(f * fftSize));
        std::vector<float> pw((size_t)(f * fftSize));
        FFTError::check(mifft_stft_host(an_.handle(), signal.data(), (int64_t)signal.size(), hopSize,
                                        window_.data(), fp(spec.data()), pw.data(), nullptr, 20.f));
        for (int64_t i = 0; i < f; ++i) {
            out[(size_t)i].complex.assign(spec.begin() + i * fftSize, spec.begin() + (i + 1) * fftSize);
            out[(size_t)i].magnitudes.assign(pw.begin() + i * fftSize, pw.begin() + (i + 1) * fftSize);
        }
        return out;
    }
    // [time][frequency] dB magnitudes (STFT.swift:82-85)
    std::vector<std::vector<float>> spectrogram(const std::vector<float>& signal) {
        std::vector<std::vector<float>> out;
        for (const STFTFrame& fr : analyze(signal)) out.push_back(fr.magnitudesDB());
        return out;
    }

   private:
    Analyser an_;
    std::vector<float> window_;
};

// FFTAnalyzer.analyze (FFTAnalyzer.swift:92-116): one windowed frame.
class FFTAnalyzer {
   public:
    FFTAnalyzer(int64_t size, double sampleRate, WindowType window = WindowType{})
        : size(size), sampleRate(sampleRate), stft_(size, 1, window, sampleRate) {}
    const int64_t size;
    const double sampleRate;
    STFTFrame analyze(const std::vector<float>& samples, int64_t offset = 0) {
        if (offset == 0 && (int64_t)samples.size() != size)
            throw FFTError(MIFFT_ERR_INVALID_INPUT_SIZE, "samples.count must equal size", size,
                           (int64_t)samples.size(), -1);
        if (offset < 0 || offset + size > (int64_t)samples.size())
            throw FFTError(MIFFT_ERR_PRECONDITION, "offset + size must be <= samples.count", -1, offset, -1);
        std::vector<float> frame(samples.begin() + offset, samples.begin() + offset + size);
        return stft_.analyze(frame).at(0);
    }

   private:
    STFT stft_;
};

// CrossSpectralResult (CrossSpectral.swift:11-39) plus welchMSC, the
// coherence of the averaged spectra.
struct CrossSpectralResult {
    int64_t channels = 0, fftSize = 0;
    std::vector<std::vector<float>> power, coherence, welchMSC;
    std::vector<Signal> crossSpectra;
    int64_t pairCount() const { return channels * (channels - 1) / 2; }
    int64_t pairIndex(int64_t i, int64_t j) const {
        if (!(0 <= i && i < j && j < channels))
            throw FFTError(MIFFT_ERR_PRECONDITION, "pairIndex requires i < j < channels", -1, -1, -1);
        int64_t p = 0;
        for (int64_t a = 0; a < i; ++a) p += channels - 1 - a;
        return p + (j - i - 1);
    }
};

// PSD (PSD.swift:6-134, CrossSpectral.swift:55-173), each over all frames in
// batched launches.
struct PSD {
    static std::vector<float> welch(const std::vector<float>& signal, int64_t fftSize, int64_t hopSize,
                                    double sampleRate, WindowType window = WindowType{}) {
        if (hopSize < 1) throw FFTError(MIFFT_ERR_PRECONDITION, "hopSize must be >= 1", -1, hopSize, -1);
        Analyser an(fftSize, 1);
        const std::vector<float> w = window.coefficients(fftSize);
        std::vector<float> psd((size_t)fftSize);
        FFTError::check(mifft_welch_host(an.handle(), signal.data(), (int64_t)signal.size(), hopSize, w.data(),
                                         sampleRate, psd.data()));
        return psd;
    }
    static CrossSpectralResult crossSpectral(const std::vector<std::vector<float>>& channels, int64_t fftSize,
                                             int64_t hopSize, double /*sampleRate*/ = 0.0,
                                             WindowType window = WindowType{}) {
        const int64_t c = (int64_t)channels.size();
        if (c < 1 || c > 16)
            throw FFTError(MIFFT_ERR_PRECONDITION, "crossSpectral requires 1-16 channels", -1, c, -1);
        const int64_t s = (int64_t)channels[0].size();
        std::vector<float> flat;
        for (const auto& ch : channels) {
            if ((int64_t)ch.size() != s)
                throw FFTError(MIFFT_ERR_PRECONDITION, "all channels must have the same length", s,
                               (int64_t)ch.size(), -1);
            flat.insert(flat.end(), ch.begin(), ch.end());
        }
        Analyser an(fftSize, c);
        const int64_t p = c * (c - 1) / 2, n = fftSize;
        const std::vector<float> w = window.coefficients(n);
        std::vector<float> pw((size_t)(c * n)), cr((size_t)(2 * p * n + 1)), co((size_t)(p * n + 1)),
            ms((size_t)(p * n + 1));
        FFTError::check(mifft_xspec_host(an.handle(), flat.data(), s, hopSize, w.data(), pw.data(), cr.data(),
                                         co.data(), ms.data()));
        CrossSpectralResult r;
        r.channels = c;
        r.fftSize = n;
        for (int64_t i = 0; i < c; ++i) r.power.emplace_back(pw.begin() + i * n, pw.begin() + (i + 1) * n);
        for (int64_t q = 0; q < p; ++q) {
            const cfloat* z = reinterpret_cast<const cfloat*>(cr.data()) + q * n;
            r.crossSpectra.emplace_back(z, z + n);
            r.coherence.emplace_back(co.begin() + q * n, co.begin() + (q + 1) * n);
            r.welchMSC.emplace_back(ms.begin() + q * n, ms.begin() + (q + 1) * n);
        }
        return r;
    }
    // |S_ab|^2 / (S_aa S_bb) per bin, frames over min(len a, len b)
    static std::vector<float> coherence(const std::vector<float>& a, const std::vector<float>& b, int64_t fftSize,
                                        int64_t hopSize, double sampleRate, WindowType window = WindowType{}) {
        const size_t m = a.size() < b.size() ? a.size() : b.size();
        return crossSpectral({std::vector<float>(a.begin(), a.begin() + (int64_t)m),
                              std::vector<float>(b.begin(), b.begin() + (int64_t)m)},
                             fftSize, hopSize, sampleRate, window)
            .welchMSC.at(0);
    }
};

}  // namespace mifft
