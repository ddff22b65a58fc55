// Drop-in zldsp::fft::RFFT<F> for MI355X.
//
// Same public members as the reference class (src/zldsp_fft_rfft.hpp:11-139):
// an explicit order constructor (N = 2^order reals), get_size() = N,
// get_order(), forward (real -> N/2+1 bins, AoS or SoA), backward
// (N/2+1 bins -> N reals, scaled by N/2 like the reference, the imaginary
// parts of bins 0 and N/2 ignored) and forward_sqr_mag (N/2+1 values of F).
// Calls are synchronous and accept host or device pointers.
#pragma once

#include "zldsp_fft_cfft.hpp"

namespace zldsp::fft {

template <typename F>
class RFFT {
    using C = std::complex<F>;

public:
    explicit RFFT(const size_t rfft_order) : plan_(ZLFFT_KIND_RFFT, detail::precision_of<F>(), rfft_order) {}

    [[nodiscard]] size_t get_size() const { return zlfft_plan_get_size(plan_.get()); }

    [[nodiscard]] size_t get_order() const { return zlfft_plan_get_order(plan_.get()); }

    /** real to AoS forward */
    void forward(F* in_buffer, C* out_buffer) {
        detail::check(zlfft_rfft_forward(plan_.get(), in_buffer, ZLFFT_AOS, out_buffer, nullptr));
    }

    /** real to SoA forward */
    void forward(F* in_buffer, std::array<F*, 2> out_buffer) {
        detail::check(zlfft_rfft_forward(plan_.get(), in_buffer, ZLFFT_SOA, out_buffer[0], out_buffer[1]));
    }

    /** AoS to real backward */
    void backward(C* in_buffer, F* out_buffer) {
        detail::check(zlfft_rfft_backward(plan_.get(), ZLFFT_AOS, in_buffer, nullptr, out_buffer));
    }

    /** SoA to real backward */
    void backward(std::array<F*, 2> in_buffer, F* out_buffer) {
        detail::check(zlfft_rfft_backward(plan_.get(), ZLFFT_SOA, in_buffer[0], in_buffer[1], out_buffer));
    }

    /** real to squared magnitude forward */
    void forward_sqr_mag(F* in_buffer, F* out_buffer) {
        detail::check(zlfft_rfft_forward_sqr_mag(plan_.get(), in_buffer, out_buffer));
    }

    /** the underlying C-ABI plan, for batched device calls (zlfft_rfft_*_exec) */
    [[nodiscard]] zlfft_plan native_handle() const { return plan_.get(); }

private:
    detail::Plan plan_;
};

}  // namespace zldsp::fft
