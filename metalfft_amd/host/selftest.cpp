// Self-test of the C++ mirror (metalfft.hpp) on a GPU, written like the
// reference's XCTest cases (Tests/MetalFFTTests/MetalFFTTests.swift,
// FusedConvolverTests.swift): prints one "PASS"/"FAIL" line per check and
// exits non-zero if any check fails.
#include <cmath>
#include <cstdio>
#include <random>

#include "metalfft.hpp"

using mifft::cfloat;
using mifft::Signal;

static int failures = 0;

static void report(bool ok, const char* name, double value) {
    std::printf("%s %s %.3e\n", ok ? "PASS" : "FAIL", name, value);
    if (!ok) ++failures;
}

static double rel_l2(const Signal& a, const std::vector<std::complex<double>>& b) {
    double num = 0, den = 0;
    for (size_t i = 0; i < a.size(); ++i) {
        num += std::norm(std::complex<double>(a[i]) - b[i]);
        den += std::norm(b[i]);
    }
    return std::sqrt(num / den);
}

// direct FP64 DFT, X[k] = sum_n x[n] exp(-2 pi i n k / N)
static std::vector<std::complex<double>> dft(const Signal& x) {
    const size_t n = x.size();
    std::vector<std::complex<double>> y(n);
    for (size_t k = 0; k < n; ++k) {
        std::complex<double> s = 0;
        for (size_t j = 0; j < n; ++j) {
            const double a = -2.0 * M_PI * (double)((j * k) % n) / (double)n;
            s += std::complex<double>(x[j]) * std::complex<double>(std::cos(a), std::sin(a));
        }
        y[k] = s;
    }
    return y;
}

static Signal random_signal(size_t n, unsigned seed) {
    std::mt19937 g(seed);
    std::uniform_real_distribution<float> u(-1.f, 1.f);
    Signal x(n);
    for (auto& v : x) v = cfloat(u(g), u(g));
    return x;
}

int main() {
    try {
        for (int64_t n = 64; n <= 16384; n *= 2) {  // testAllSizesImpulse
            mifft::MetalFFT fft(n);
            Signal x((size_t)n);
            x[0] = 1;
            Signal y = fft.forward(x);
            double worst = 0;
            for (auto v : y) worst = std::fmax(worst, std::abs(v - cfloat(1, 0)));
            char name[64];
            std::snprintf(name, sizeof name, "impulse_n%lld", (long long)n);
            report(worst < 1e-6, name, worst);
        }
        for (int64_t n : {64, 512, 2048}) {  // random vs direct DFT, round trip
            mifft::MetalFFT fft(n);
            Signal x = random_signal((size_t)n, (unsigned)n);
            Signal y = fft.forward(x);
            char name[64];
            std::snprintf(name, sizeof name, "random_vs_dft_n%lld", (long long)n);
            report(rel_l2(y, dft(x)) < 1e-6, name, rel_l2(y, dft(x)));
            Signal z = fft.inverse(y);
            std::vector<std::complex<double>> xd(x.begin(), x.end());
            std::snprintf(name, sizeof name, "round_trip_n%lld", (long long)n);
            report(rel_l2(z, xd) < 1e-6, name, rel_l2(z, xd));
        }
        {  // testBatchMatchesSingle
            mifft::MetalFFT fft(512);
            Signal a = random_signal(512, 1), b = random_signal(512, 2);
            auto batch = fft.forwardBatch({a, b});
            report(batch[0] == fft.forward(a) && batch[1] == fft.forward(b), "batch_matches_single", 0);
            Signal out;  // forward(input:output:) resizes
            fft.forward(a.data(), (int64_t)a.size(), out);
            report(out == fft.forward(a), "forward_input_output", 0);
        }
        {  // error cases
            bool ok = false;
            try {
                mifft::MetalFFT bad(100);
            } catch (const mifft::FFTError& e) {
                ok = e.status == MIFFT_ERR_UNSUPPORTED_SIZE && e.got == 100;
            }
            report(ok, "unsupported_size_throws", 0);
            ok = false;
            mifft::MetalFFT fft(256);
            try {
                fft.forward(Signal(255));
            } catch (const mifft::FFTError& e) {
                ok = e.status == MIFFT_ERR_INVALID_INPUT_SIZE && e.expected == 256 && e.got == 255;
            }
            report(ok, "invalid_input_size_throws", 0);
            ok = false;
            try {
                fft.forwardBatch({Signal(256), Signal(128)});
            } catch (const mifft::FFTError& e) {
                ok = e.status == MIFFT_ERR_BATCH_INPUT_SIZE && e.batchIndex == 1;
            }
            report(ok, "batch_input_size_throws", 0);
        }
        {  // MultiChannelFFT inverse(forward(x)) == x
            mifft::MultiChannelFFT mc(3, 1024);
            std::vector<Signal> xs = {random_signal(1024, 7), random_signal(1024, 8),
                                      random_signal(1024, 9)};
            auto back = mc.inverse(mc.forward(xs));
            double worst = 0;
            for (int c = 0; c < 3; ++c)
                worst = std::fmax(worst,
                                  rel_l2(back[c], std::vector<std::complex<double>>(xs[c].begin(), xs[c].end())));
            report(worst < 1e-6, "multichannel_round_trip", worst);
        }
        {  // testDeltaFilterIsIdentity
            std::vector<float> delta(64, 0.f);
            delta[0] = 1.f;
            mifft::FusedConvolver conv(delta, 4096);
            std::mt19937 g(3);
            std::uniform_real_distribution<float> u(-1.f, 1.f);
            std::vector<float> s((size_t)conv.hopSize());
            for (auto& v : s) v = u(g);
            auto y = conv.apply(s);
            double num = 0, den = 0;
            for (size_t i = 0; i < s.size(); ++i) {
                num += (y[i] - s[i]) * (double)(y[i] - s[i]);
                den += s[i] * (double)s[i];
            }
            report(std::sqrt(num / den) < 1e-6 && y.size() == s.size() + 63, "delta_filter_identity",
                   std::sqrt(num / den));
            // FP16 storage mode (FusedConvolverTests.swift:43-45 bound 1e-3)
            mifft::FusedConvolver half(delta, 4096, mifft::FusedConvolver::Precision::float16Storage);
            auto yh = half.apply(s);
            num = den = 0;
            for (size_t i = 0; i < s.size(); ++i) {
                num += (yh[i] - s[i]) * (double)(yh[i] - s[i]);
                den += s[i] * (double)s[i];
            }
            report(std::sqrt(num / den) < 1e-3 && std::sqrt(num / den) > 1e-7,
                   "delta_filter_fp16_storage", std::sqrt(num / den));
            report(mifft::Convolver::recommendedFFTSize(1023).value_or(0) == 1024 &&
                       !mifft::Convolver::recommendedFFTSize(16384).has_value(),
                   "recommended_fft_size", 0);
        }
        {  // correlation lag convention r[m] = sum_n a[n+m] b[n]
            mifft::Correlator c(1024);
            std::vector<float> b(1024, 0.f), a(1024, 0.f);
            for (int i = 100; i < 110; ++i) b[(size_t)i] = 1.f;
            for (int i = 0; i < 1024; ++i) a[(size_t)((i + 37) % 1024)] = b[(size_t)i];
            auto r = c.cross(a, b);
            size_t arg = 0;
            for (size_t i = 1; i < r.size(); ++i)
                if (r[i] > r[arg]) arg = i;
            report(arg == 37, "correlator_lag", (double)arg);
        }
        {  // STFT frames vs the direct DFT of the windowed slice; Welch of a constant
            const int64_t n = 256, hop = 96;
            mifft::STFT stft(n, hop, mifft::WindowType{}, 8000.0);
            std::mt19937 g(4);
            std::uniform_real_distribution<float> u(-1.f, 1.f);
            std::vector<float> sig(1000);
            for (auto& v : sig) v = u(g);
            auto frames = stft.analyze(sig);
            const auto w = mifft::WindowType{}.coefficients(n);
            double worst = 0;
            for (size_t f = 0; f < frames.size(); f += 3) {
                Signal seg((size_t)n);
                for (int64_t j = 0; j < n; ++j) seg[(size_t)j] = cfloat(sig[f * hop + (size_t)j] * w[(size_t)j], 0.f);
                worst = std::fmax(worst, rel_l2(Signal(frames[f].complex), dft(seg)));
            }
            report(frames.size() == 8 && worst < 1e-6, "stft_frames_vs_dft", worst);
            std::vector<float> cst((size_t)(7 * n), 0.6f);
            auto psd = mifft::PSD::welch(cst, n, hop, 8000.0);
            const double want0 = 2.0 * 0.36 * n / (3.0 * 8000.0);
            report(std::fabs(psd[0] - want0) < 1e-5 * want0, "welch_constant_closed_form",
                   std::fabs(psd[0] - want0) / want0);
            auto coh = mifft::PSD::coherence(sig, std::vector<float>(sig.size(), 0.f), n, hop, 8000.0);
            double mx = 0;
            for (float v : coh) mx = std::fmax(mx, std::fabs(v));
            report(mx == 0.0, "coherence_zero_signal", mx);
        }
    } catch (const std::exception& e) {
        std::printf("FAIL exception %s\n", e.what());
        return 2;
    }
    std::printf("%s: %d failure(s)\n", failures ? "FAILED" : "OK", failures);
    return failures ? 1 : 0;
}
