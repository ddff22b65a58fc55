// Drop-in zldsp::fft::CFFT<F> for MI355X.
//
// Same public members as the reference class (src/zldsp_fft_cfft.hpp:11-118):
// an explicit order constructor, get_size(), get_order() and the eight
// forward/backward overloads over AoS (std::complex<F>*) and SoA
// (std::array<F*, 2>) buffers.  Calls are synchronous and accept host or
// device pointers; the work runs on the GPU through libzlfft (zlfft.h).
// Differences: invalid orders and GPU failures throw std::runtime_error (the
// reference has undefined behaviour there).  F is float or double, as in the
// reference; other types fail to compile.
#pragma once

#include <array>
#include <complex>
#include <cstddef>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <utility>

#include "zlfft.h"

namespace zldsp::fft {
namespace detail {

inline void check(int rc) {
    if (rc != ZLFFT_OK) throw std::runtime_error(std::string("zlfft: ") + zlfft_last_error());
}

// Owning, move-only plan handle (the reference's instances are move-only:
// their buffers are AlignedFreeUniquePtr members).
class Plan {
public:
    Plan(int kind, int precision, size_t order) { check(zlfft_plan_create(&h_, kind, precision, order)); }
    ~Plan() {
        if (h_) zlfft_plan_destroy(h_);
    }
    Plan(const Plan&) = delete;
    Plan& operator=(const Plan&) = delete;
    Plan(Plan&& o) noexcept : h_(std::exchange(o.h_, nullptr)) {}
    Plan& operator=(Plan&& o) noexcept {
        if (this != &o) {
            if (h_) zlfft_plan_destroy(h_);
            h_ = std::exchange(o.h_, nullptr);
        }
        return *this;
    }
    zlfft_plan get() const { return h_; }

private:
    zlfft_plan h_ = nullptr;
};

// plan precision of the element type F
template <typename F>
constexpr int precision_of() {
    static_assert(std::is_same_v<F, float> || std::is_same_v<F, double>, "zldsp::fft supports float and double");
    return std::is_same_v<F, double> ? ZLFFT_PREC_F64 : ZLFFT_PREC_F32;
}

}  // namespace detail

template <typename F>
class CFFT {
    using C = std::complex<F>;

public:
    explicit CFFT(const size_t cfft_order) : plan_(ZLFFT_KIND_CFFT, detail::precision_of<F>(), cfft_order) {}

    [[nodiscard]] size_t get_size() const { return zlfft_plan_get_size(plan_.get()); }

    [[nodiscard]] size_t get_order() const { return zlfft_plan_get_order(plan_.get()); }

    /** AoS to AoS forward */
    void forward(C* in_buffer, C* out_buffer) { run(ZLFFT_FORWARD, aos(in_buffer), aos(out_buffer)); }

    /** AoS to AoS backward */
    void backward(C* in_buffer, C* out_buffer) { run(ZLFFT_BACKWARD, aos(in_buffer), aos(out_buffer)); }

    /** AoS to SoA forward */
    void forward(C* in_buffer, std::array<F*, 2> out_buffer) { run(ZLFFT_FORWARD, aos(in_buffer), soa(out_buffer)); }

    /** AoS to SoA backward */
    void backward(C* in_buffer, std::array<F*, 2> out_buffer) {
        run(ZLFFT_BACKWARD, aos(in_buffer), soa(out_buffer));
    }

    /** SoA to AoS forward */
    void forward(std::array<F*, 2> in_buffer, C* out_buffer) { run(ZLFFT_FORWARD, soa(in_buffer), aos(out_buffer)); }

    /** SoA to AoS backward */
    void backward(std::array<F*, 2> in_buffer, C* out_buffer) {
        run(ZLFFT_BACKWARD, soa(in_buffer), aos(out_buffer));
    }

    /** SoA to SoA forward */
    void forward(std::array<F*, 2> in_buffer, std::array<F*, 2> out_buffer) {
        run(ZLFFT_FORWARD, soa(in_buffer), soa(out_buffer));
    }

    /** SoA to SoA backward */
    void backward(std::array<F*, 2> in_buffer, std::array<F*, 2> out_buffer) {
        run(ZLFFT_BACKWARD, soa(in_buffer), soa(out_buffer));
    }

    /** the underlying C-ABI plan, for batched device calls (zlfft_cfft_exec) */
    [[nodiscard]] zlfft_plan native_handle() const { return plan_.get(); }

private:
    struct Side {
        int layout;
        void* p0;
        void* p1;
    };
    static Side aos(C* p) { return {ZLFFT_AOS, p, nullptr}; }
    static Side soa(std::array<F*, 2> p) { return {ZLFFT_SOA, p[0], p[1]}; }

    void run(int dir, Side in, Side out) {
        detail::check(zlfft_cfft_execute(plan_.get(), dir, in.layout, in.p0, in.p1, out.layout, out.p0, out.p1));
    }

    detail::Plan plan_;
};

}  // namespace zldsp::fft
