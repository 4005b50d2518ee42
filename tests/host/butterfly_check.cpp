// Host check of the in-register butterflies of metalfft_amd/csrc/fft_device.hpp
// (compiled for the CPU: they are __host__ __device__): DFT-2/4/8/16 and the
// twiddle-folded dft_tw<R> against a DFT of the same FP32 inputs in FP64,
// y[k] = sum_j p[j] exp(-2 pi i j k / R), p[j] = w[j] x[j] (w[0] = 1).
// Prints the largest error relative to the output's L2 norm per radix and
// exits 1 if any exceeds 1e-6.
#include <cmath>
#include <complex>
#include <cstdio>
#include <random>

#include "fft_device.hpp"

using cd = std::complex<double>;

template <int R, bool TW>
static double check(std::mt19937& rng, int trials) {
    std::uniform_real_distribution<float> u(-1.f, 1.f);
    double worst = 0.0;
    for (int t = 0; t < trials; ++t) {
        float2 x[R], w[R];
        cd p[R];
        for (int j = 0; j < R; ++j) {
            x[j] = make_float2(u(rng), u(rng));
            const float ang = 6.283185307f * u(rng);
            w[j] = j == 0 ? make_float2(1.f, 0.f) : make_float2(std::cos(ang), std::sin(ang));
            const cd xv(x[j].x, x[j].y), wv(w[j].x, w[j].y);
            p[j] = TW ? wv * xv : xv;
        }
        if constexpr (TW) mifft::dft_tw<R>(x, w); else mifft::dft<R>(x);
        double num = 0.0, den = 0.0;
        for (int k = 0; k < R; ++k) {
            cd y = 0.0;
            for (int j = 0; j < R; ++j) y += p[j] * std::polar(1.0, -2.0 * M_PI * j * k / R);
            num += std::norm(cd(x[k].x, x[k].y) - y);
            den += std::norm(y);
        }
        worst = std::fmax(worst, std::sqrt(num / den));
    }
    return worst;
}

int main() {
    std::mt19937 rng(7);
    const double e[] = {check<2, false>(rng, 2000),  check<4, false>(rng, 2000), check<8, false>(rng, 2000),
                        check<16, false>(rng, 2000), check<4, true>(rng, 2000),  check<8, true>(rng, 2000),
                        check<16, true>(rng, 2000)};
    const char* names[] = {"dft2", "dft4", "dft8", "dft16", "dft_tw4", "dft_tw8", "dft_tw16"};
    int bad = 0;
    for (int i = 0; i < 7; ++i) {
        std::printf("%-9s max rel-L2 %.3e\n", names[i], e[i]);
        if (!(e[i] < 1e-6)) bad = 1;
    }
    return bad;
}
