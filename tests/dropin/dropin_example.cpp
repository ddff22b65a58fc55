// The reference README's usage (README.md:48-127), compiled unchanged in
// spirit against the drop-in headers, plus checks of the results: prints
// "DROPIN OK" and returns 0 when every check passes.
#include <array>
#include <cmath>
#include <complex>
#include <cstdio>
#include <vector>

#include "zldsp_fft_cfft.hpp"
#include "zldsp_fft_rfft.hpp"

static double rel(const std::vector<std::complex<double>>& a, const std::vector<std::complex<double>>& b) {
    double num = 0, den = 0;
    for (size_t i = 0; i < a.size(); ++i) {
        num += std::norm(a[i] - b[i]);
        den += std::norm(b[i]);
    }
    return std::sqrt(num / den);
}

static std::vector<std::complex<double>> dft(const std::vector<std::complex<double>>& x, int sign) {
    const size_t n = x.size();
    std::vector<std::complex<double>> y(n);
    for (size_t k = 0; k < n; ++k) {
        std::complex<double> acc = 0;
        for (size_t j = 0; j < n; ++j) {
            const double a = sign * 2.0 * M_PI * double((j * k) % n) / double(n);
            acc += x[j] * std::complex<double>(std::cos(a), std::sin(a));
        }
        y[k] = acc;
    }
    return y;
}

int main() {
    int bad = 0;
    // a CFFT of size 1024 (2^10) using float
    constexpr size_t order = 10;
    zldsp::fft::CFFT<float> cfft(order);
    const size_t n = cfft.get_size();
    std::vector<std::complex<float>> in_buffer(1 << order);
    std::vector<std::complex<float>> out_buffer(1 << order);
    std::vector<std::complex<double>> x(n);
    for (size_t i = 0; i < n; ++i) {
        in_buffer[i] = {std::sin(0.1f * float(i)), std::cos(0.37f * float(i * i % 97))};
        x[i] = in_buffer[i];
    }
    // forward transform (AoS to AoS)
    cfft.forward(in_buffer.data(), out_buffer.data());
    std::vector<std::complex<double>> y(out_buffer.begin(), out_buffer.end());
    const auto X = dft(x, -1);
    const double e1 = rel(y, X);
    // SoA output of the same transform
    std::vector<float> re(n), im(n);
    cfft.forward(in_buffer.data(), std::array<float*, 2>{re.data(), im.data()});
    for (size_t i = 0; i < n; ++i) bad += (re[i] != out_buffer[i].real() || im[i] != out_buffer[i].imag());
    // backward transform (AoS to AoS): round trip is N x
    std::vector<std::complex<float>> back(n);
    cfft.backward(out_buffer.data(), back.data());
    std::vector<std::complex<double>> b(back.begin(), back.end()), nx(n);
    for (size_t i = 0; i < n; ++i) nx[i] = double(n) * x[i];
    const double e2 = rel(b, nx);

    // an RFFT of size 2048 using float
    zldsp::fft::RFFT<float> rfft(11);
    const size_t nr = rfft.get_size();
    std::vector<float> real(nr), real_back(nr), mag(nr / 2 + 1);
    std::vector<std::complex<float>> bins(nr / 2 + 1);
    std::vector<std::complex<double>> xr(nr);
    for (size_t i = 0; i < nr; ++i) {
        real[i] = std::sin(0.05f * float(i)) + 0.25f * std::cos(1.3f * float(i));
        xr[i] = real[i];
    }
    rfft.forward(real.data(), bins.data());
    rfft.forward_sqr_mag(real.data(), mag.data());
    rfft.backward(bins.data(), real_back.data());
    const auto XR = dft(xr, -1);
    std::vector<std::complex<double>> yb(bins.begin(), bins.end()), xb(XR.begin(), XR.begin() + nr / 2 + 1);
    const double e3 = rel(yb, xb);
    double e4n = 0, e4d = 0, e5n = 0, e5d = 0;
    for (size_t i = 0; i < nr; ++i) {
        e4n += std::pow(real_back[i] - 0.5 * nr * real[i], 2);
        e4d += std::pow(0.5 * nr * real[i], 2);
    }
    for (size_t k = 0; k <= nr / 2; ++k) {
        e5n += std::pow(mag[k] - std::norm(XR[k]), 2);
        e5d += std::pow(std::norm(XR[k]), 2);
    }
    bad += (bins[0].imag() != 0.0f) + (bins[nr / 2].imag() != 0.0f);
    const double e4 = std::sqrt(e4n / e4d), e5 = std::sqrt(e5n / e5d);
    std::printf("cfft fwd %.3e  round trip %.3e  rfft fwd %.3e  rfft round trip %.3e  sqr_mag %.3e  exact-check failures %d\n",
                e1, e2, e3, e4, e5, bad);
    // the same API over double (CFFT<double> / RFFT<double>)
    zldsp::fft::CFFT<double> cfft64(order);
    std::vector<std::complex<double>> in64(x), out64(n), back64(n);
    cfft64.forward(in64.data(), out64.data());
    const double e6 = rel(out64, X);
    cfft64.backward(out64.data(), back64.data());
    const double e7 = rel(back64, nx);
    zldsp::fft::RFFT<double> rfft64(11);
    std::vector<double> real64(real.begin(), real.end()), back_real64(nr), mag64(nr / 2 + 1);
    std::vector<std::complex<double>> bins64(nr / 2 + 1);
    rfft64.forward(real64.data(), bins64.data());
    rfft64.forward_sqr_mag(real64.data(), mag64.data());
    rfft64.backward(bins64.data(), back_real64.data());
    const double e8 = rel(bins64, xb);
    double e9n = 0, e9d = 0;
    for (size_t i = 0; i < nr; ++i) {
        e9n += std::pow(back_real64[i] - 0.5 * nr * real64[i], 2);
        e9d += std::pow(0.5 * nr * real64[i], 2);
    }
    const double e9 = std::sqrt(e9n / e9d);
    bad += (bins64[0].imag() != 0.0) + (bins64[nr / 2].imag() != 0.0);
    for (size_t k = 0; k <= nr / 2; ++k) bad += std::abs(mag64[k] - std::norm(bins64[k])) > 1e-9 * (1 + std::norm(bins64[k]));
    std::printf("double: cfft fwd %.3e  round trip %.3e  rfft fwd %.3e  rfft round trip %.3e\n", e6, e7, e8, e9);
    // invalid use is reported, not undefined
    bool threw = false;
    try {
        zldsp::fft::RFFT<float> zero(0);
    } catch (const std::runtime_error&) {
        threw = true;
    }
    const bool ok = e1 < 1e-6 && e2 < 1e-6 && e3 < 1e-6 && e4 < 1e-6 && e5 < 2e-6 && e6 < 1e-12 && e7 < 1e-12 &&
                    e8 < 1e-12 && e9 < 1e-12 && bad == 0 && threw;
    std::printf(ok ? "DROPIN OK\n" : "DROPIN FAILED\n");
    return ok ? 0 : 1;
}
