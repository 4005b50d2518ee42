// C++ mirror of MetalFFT's Swift API over the libmifft C ABI (include/mifft.h).
//
//   mifft::MetalFFT         MetalFFT            Sources/MetalFFT/FFT/MetalFFT.swift:10-133
//   mifft::MultiChannelFFT  MultiChannelFFT     Sources/MetalFFT/FFT/MultiChannelFFT.swift:19-54
//   mifft::Convolver        Convolver           Sources/MetalFFT/DSP/Convolve.swift:8-87
//   mifft::FusedConvolver   FusedConvolver      Sources/MetalFFT/DSP/FusedConvolver.swift:17-134
//   mifft::Correlator       Correlator          Sources/MetalFFT/DSP/Correlate.swift:7-62
//   mifft::FFTError         FFTError            Sources/MetalFFT/Internal/MetalContext.swift:6-30
//
// Methods throw FFTError where the Swift ones throw; complex values are
// std::complex<float> (same interleaved layout as SIMD2<Float>).
#pragma once

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

}  // namespace mifft
