// Kernel registry shared by the per-order kernel units and the C-ABI layer.
#pragma once

#include "zlfft_device.hpp"

namespace zl {

constexpr int kMaxLdsOrder = 14;  // largest CFFT order kept whole in LDS

// Points-per-thread (log2) and target workgroup size of the LDS kernel for a
// CFFT order.  Orders <= 4 keep a whole transform in one thread.
constexpr int cfg_logr(int logn) { return logn <= 4 ? logn : 4; }
constexpr int cfg_tpb(int) { return 256; }

struct KernelEntry {
    void (*fn)(KArgs);
    int wg;        // threads per workgroup
    int tpw;       // transforms per workgroup
    int logn;      // CFFT order
    int np;        // Stockham passes
    int lr[8];     // log2 radix per pass
    int lw[8];     // log2 width per pass
    int tw_total;  // complex twiddles of all passes
};

// table[kind][lin][lout] for one order; unused combinations stay null
using OrderTable = KernelEntry[4][2][2];

template <class G, int KIND, int LIN, int LOUT>
KernelEntry make_entry() {
    KernelEntry e{};
    e.fn = fft_lds_kernel<G, KIND, LIN, LOUT>;
    e.wg = G::WG;
    e.tpw = G::TPW;
    e.logn = G::logn;
    e.np = G::NP;
    for (int p = 0; p < G::NP && p < 8; ++p) {
        e.lr[p] = G::lr(p);
        e.lw[p] = G::lw(p);
    }
    e.tw_total = G::tw_total();
    return e;
}

}  // namespace zl

#define ZL_DECLARE_ORDER(n) void zl_fill_order_##n(zl::OrderTable& tab);
ZL_DECLARE_ORDER(0)
ZL_DECLARE_ORDER(1)
ZL_DECLARE_ORDER(2)
ZL_DECLARE_ORDER(3)
ZL_DECLARE_ORDER(4)
ZL_DECLARE_ORDER(5)
ZL_DECLARE_ORDER(6)
ZL_DECLARE_ORDER(7)
ZL_DECLARE_ORDER(8)
ZL_DECLARE_ORDER(9)
ZL_DECLARE_ORDER(10)
ZL_DECLARE_ORDER(11)
ZL_DECLARE_ORDER(12)
ZL_DECLARE_ORDER(13)
ZL_DECLARE_ORDER(14)
#undef ZL_DECLARE_ORDER
