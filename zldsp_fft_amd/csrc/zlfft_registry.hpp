// Kernel registry shared by the per-order kernel units and the C-ABI layer.
#pragma once

#include "zlfft_device.hpp"

namespace zl {

constexpr int kMaxLdsOrder = 14;    // largest CFFT order kept whole in LDS (FP32)
constexpr int kMaxLdsOrder64 = 13;  // the same for FP64 (2^14 double2 exceed the LDS)

// Points-per-thread (log2) and target workgroup size of the LDS kernel for a
// CFFT order.  Orders <= 3 keep a whole transform in one thread; orders 4 and
// 5 run 8 points per thread: a 256-thread workgroup then stages half as many
// points through LDS, and twice as many workgroups fit a CU (measured, 2^26
// points per call, against 16 per thread: CFFT order 4 4523 -> 5609 GB/s,
// order 5 3958 -> 5622; RFFT 2^6 bins 3076 -> 4555).
constexpr int cfg_logr(int logn) { return logn <= 3 ? logn : logn <= 5 ? 3 : 4; }
constexpr int cfg_tpb(int) { return 256; }

struct KernelEntry {
    void (*fn)(KArgs);
    void (*stream_fn[2])(KArgs);  // persistent streaming variant [forward, backward], or null
    int wg;        // threads per workgroup
    int tpw;       // transforms per workgroup
    int logn;      // CFFT order
    int np;        // Stockham passes
    int lr[8];     // log2 radix per pass
    int lw[8];     // log2 width per pass
    int tw_total;  // complex twiddles of all passes (direct table)
    // the streaming variant's geometry (the LDS kernel's, except FP64 C2C,
    // whose streaming variant runs 8 points per thread) and its LDS image
    int s_wg, s_tpw, s_np;
    int s_lr[8], s_lw[8];
    int img_total; // complex twiddles of the LDS image
    int img_lo[8]; // per pass: entries of the direct / low factor part
    int img_hi[8]; // per pass: entries of the high factor part (0 if direct)
};

// table[kind][lin][lout] for one order (kinds C2C..C2R, then the framed
// R2C_FRAMED / R2C_SQR_FRAMED); unused combinations stay null
using OrderTable = KernelEntry[6][2][2];

#ifndef ZL_RTW_GLOBAL_MIN
#define ZL_RTW_GLOBAL_MIN 14
#endif

// The persistent streaming variant, for geometries whose direct I/O is
// coalesced and whose LDS (transforms, twiddle image, RFFT twiddles) fits:
// FP32, and FP64 (C2C and the forward real kinds) where two workgroups fit
// one CU.
template <class G, int KIND, int LIN, int LOUT, int FRAMED = 0, class CX = float2>
void add_stream_variant(KernelEntry& e) {
    constexpr int cs = (int)sizeof(CX);
    constexpr int elem = ((KIND == C2C && (LIN == SOA || LOUT == SOA)) || (KIND == C2R && LIN == SOA)) ? cs / 2 : cs;
    constexpr bool direct = G::T * elem >= 128 || G::TPW == 1;
    using TI = TwImage<G, cfg_stream_direct_max<CX>(G::logn, KIND)>;
    // FP32 forward real kinds of inner order >= ZL_RTW_GLOBAL_MIN read the
    // RFFT twiddles from global memory (RTWG)
    constexpr bool rtwg = cs == 8 && !FRAMED && KIND != C2C && G::logn >= ZL_RTW_GLOBAL_MIN;
    constexpr long long stream_lds =
        (long long)cs * (G::TPW * G::LDSN + TI::total() + (KIND == C2C || rtwg ? 1 : G::N / 2) + (FRAMED ? G::N : 0));
    constexpr bool fits = cs == 8 ? stream_lds <= 160 * 1024 : !FRAMED && 2 * stream_lds <= 160 * 1024;
    // C2R holds the frame's bin pairs and bins 0 and M besides the points:
    // at 1024 threads (128 VGPRs) it spilled, so those sizes keep the
    // one-launch kernel
    constexpr bool regs_ok = KIND != C2R || G::WG <= 512;
    if constexpr (direct && G::NP >= 2 && fits && regs_ok) {
        e.stream_fn[0] = fft_stream_kernel<G, KIND, LIN, LOUT, 0, FRAMED, CX, rtwg ? 1 : 0>;
        if constexpr (KIND == C2C) e.stream_fn[1] = fft_stream_kernel<G, KIND, LIN, LOUT, 1, 0, CX>;
        if constexpr (KIND == C2R) e.stream_fn[1] = e.stream_fn[0];  // the direction is the kind's
        e.s_wg = G::WG;
        e.s_tpw = G::TPW;
        e.s_np = G::NP;
        for (int p = 0; p < G::NP && p < 8; ++p) {
            e.s_lr[p] = G::lr(p);
            e.s_lw[p] = G::lw(p);
        }
        e.img_total = TI::total();
        for (int p = 1; p < G::NP && p < 8; ++p) {
            e.img_lo[p] = TI::lo_n(p);
            e.img_hi[p] = TI::hi_n(p);
        }
    }
}

template <class G, int KIND, int LIN, int LOUT, class GS = G>
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
    add_stream_variant<GS, KIND, LIN, LOUT>(e);
    return e;
}

// Entry over complex type CX: FP64 (the LDS kernel, and for C2C and the
// forward real kinds the streaming variant of geometry GS), and the framed
// STFT kernels of both precisions (FP32 with the streaming variant)
template <class G, int KIND, int LIN, int LOUT, class CX = double2, int FRAMED = 0, class GS = G>
KernelEntry make_lds_entry() {
    static_assert((long long)sizeof(CX) * G::TPW * G::LDSN <= 160 * 1024, "geometry exceeds the LDS");
    KernelEntry e{};
    e.fn = fft_lds_kernel<G, KIND, LIN, LOUT, CX, FRAMED>;
    e.wg = G::WG;
    e.tpw = G::TPW;
    e.logn = G::logn;
    e.np = G::NP;
    for (int p = 0; p < G::NP && p < 8; ++p) {
        e.lr[p] = G::lr(p);
        e.lw[p] = G::lw(p);
    }
    e.tw_total = G::tw_total();
    if constexpr (FRAMED && sizeof(CX) == 8) add_stream_variant<G, KIND, LIN, LOUT, FRAMED>(e);
    if constexpr (!FRAMED && sizeof(CX) == 16) add_stream_variant<GS, KIND, LIN, LOUT, 0, CX>(e);
    return e;
}

// ---- multi-pass (HBM) transforms for orders above kMaxLdsOrder

constexpr int kMinMacroLogL = 6;
constexpr int kMaxMacroLogL = 10;
constexpr int kMaxOrder = 28;    // largest CFFT order (FP32)
constexpr int kMaxOrder64 = 27;  // largest CFFT order (FP64)
constexpr int kRaggedMaxOrderTune = 14;  // orders of the ragged tuning table

// columns (butterflies) per workgroup of a radix-2^logl HBM pass: FP32, FP64
// (C columns of 2^logl complex values, padded, fill the LDS)
// (the widest that fits: fewer columns make each SoA access shorter than a
// cache line, measured 8% slower at radix 256 and 35% at radix 512 / 1024;
// radix 512 at 32 columns against 16: orders 17 and 19 on 4 Mi-point SoA
// batches 3.5% and 4.4% faster)
constexpr int cfg_macro_c(int logl) { return logl <= 9 ? 32 : 16; }
constexpr int cfg_macro_c64(int logl) { return logl <= 8 ? 32 : logl == 9 ? 16 : 8; }

// radix (log2) of each HBM pass for a CFFT order above kMaxLdsOrder, first pass first
struct MacroPlan {
    int np;
    int logl[4];
};
constexpr MacroPlan macro_plan(int order) {
    switch (order) {
        case 15: return {2, {8, 7}};
        case 16: return {2, {8, 8}};
        case 17: return {2, {9, 8}};
        case 18: return {2, {10, 8}};  // measured 5% above {9, 9} on 4 Mi-point SoA batches
        case 19: return {2, {10, 9}};
        case 20: return {2, {10, 10}};
        case 21: return {3, {7, 7, 7}};
        case 22: return {3, {8, 7, 7}};
        case 23: return {3, {8, 8, 7}};
        case 24: return {3, {8, 8, 8}};
        case 25: return {3, {9, 8, 8}};
        case 26: return {3, {9, 9, 8}};
        case 27: return {3, {9, 9, 9}};
        case 28: return {3, {10, 9, 9}};
        default: return {0, {0, 0, 0, 0}};
    }
}

// FP64: order 14 also runs as HBM passes
constexpr MacroPlan macro_plan64(int order) {
    if (order == 14) return {2, {7, 7}};
    return order <= kMaxOrder64 ? macro_plan(order) : MacroPlan{0, {0, 0, 0, 0}};
}

struct MacroEntry {
    void (*fn)(MArgs);
    int wg;        // threads per workgroup
    int cols;      // butterflies (columns) per workgroup
    int logl;      // log2 radix of the pass
    int np;        // inner Stockham passes of the radix-L sub-FFT
    int lr[8];
    int lw[8];
    int tw_total;  // inner twiddles
};

// [role][layout]: role 0 = first pass (layout = input), 1 = last pass
// (layout = output), 2 = middle pass (layout unused, AoS).  Layout index
// kInternal (first / last pass only) is AoS on an internal buffer, cached
// for the kernel that reads it next; AOS / SOA are the caller's buffers.
// The real kinds' fused passes: role 3 = last pass of an RFFT forward to bins
// (layout = bins), 4 = last pass to squared magnitudes (AOS only), 5 = first
// pass of an RFFT backward from bins (layout = bins), 6 = first pass of an
// STFT, reading frames of the caller's signal (layout AOS only).
constexpr int kInternal = 2;
constexpr int kRoleR2C = 3, kRoleSqr = 4, kRoleC2R = 5, kRoleFramed = 6;
using MacroTable = MacroEntry[7][3];

template <class GL, int C, int FIRST, int LAST, int LIN, int LOUT, class CX = float2, int NT = 1, int RK = C2C>
MacroEntry make_macro_entry() {
    static_assert((long long)sizeof(CX) * (C * GL::LDSN + (GL::logn <= 7 ? GL::tw_total() : 0)) <= 160 * 1024,
                  "HBM pass exceeds the LDS");
    MacroEntry e{};
    e.fn = fft_macro_pass<GL, C, FIRST, LAST, LIN, LOUT, CX, NT, RK>;
    e.wg = C * GL::T;
    e.cols = C;
    e.logl = GL::logn;
    e.np = GL::NP;
    for (int p = 0; p < GL::NP && p < 8; ++p) {
        e.lr[p] = GL::lr(p);
        e.lw[p] = GL::lw(p);
    }
    e.tw_total = GL::tw_total();
    return e;
}

// fault-injection modes of ZLFFT_INJECT (zl_read_inject, the error-path tests)
constexpr int kInjectNone = 0, kInjectAlloc = 1, kInjectWorkAlloc = 2, kInjectLaunch = 3;

}  // namespace zl

// ZLFFT_INJECT as a kInject* mode (zlfft_api.hip)
int zl_read_inject();

// MFMA split-precision 4096-point kernel for a precision code (one transform
// per workgroup, or the persistent streaming form), or null
void* zl_mfma_kernel(int prec, int backward, int streaming);

// ragged-batch kernels (zlfft_ragged_k.hip): orders up to 13 (large = 0) or
// order 14 (large = 1)
void* zl_ragged_kernel(int large);

#define ZL_DECLARE_MACRO(n)                            \
    void zl_fill_macro_##n(zl::MacroTable& tab);       \
    void zl_fill_macro64_##n(zl::MacroTable& tab);
ZL_DECLARE_MACRO(6)
ZL_DECLARE_MACRO(7)
ZL_DECLARE_MACRO(8)
ZL_DECLARE_MACRO(9)
ZL_DECLARE_MACRO(10)
#undef ZL_DECLARE_MACRO

#define ZL_DECLARE_ORDER(n)                            \
    void zl_fill_order_##n(zl::OrderTable& tab);       \
    void zl_fill_order64_##n(zl::OrderTable& tab);
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
