// One CFFT order's LDS kernels (compiled once per order with -DZL_LOGN=n so
// the build runs in parallel): C2C for the four AoS/SoA combinations and the
// real kinds, whose inner CFFT has this order; FP32 and, up to
// kMaxLdsOrder64, FP64.
#include "zlfft_registry.hpp"

#ifndef ZL_LOGN
#error "compile with -DZL_LOGN=<order>"
#endif

// Points per thread (log2) of the FP32 streaming variants: the LDS
// kernel's, except orders 6 and 7, where 16 points per thread leave 4 or 8
// threads per transform, too few for coalesced direct I/O: there 4 and 8
// points (16 threads, 128-byte rows) let the AoS batches stream (measured,
// 2^26 points: CFFT order 6 3950 -> 5540 GB/s, order 7 3890 -> 5860).
#ifndef ZL_STREAM_LOGR
#define ZL_STREAM_LOGR(n) ((n) == 6 ? 2 : (n) == 7 ? 3 : cfg_logr(n))
#endif
// The FP32 streaming variants at 2^13 points run the radix-2 remainder pass
// last, between the halves of each wave (GeoLP2), saving one of the three
// LDS exchanges.
#ifndef ZL_STREAM_LP2
#define ZL_STREAM_LP2 1
#endif
// The FP32 C2C streaming variants at 2^14 points run 32 points per thread:
// 16 x 32 x 32, two LDS exchanges instead of three, 512 threads at 256 VGPRs
// (measured, 2^26-point AoS batches: 4345 -> 4770 GB/s forward).  The real
// kinds keep 16 points per thread: they share one twiddle image, and the
// squared-magnitude variant spills at 32 (6% slower there).
#ifndef ZL_STREAM_C2C_R32
#define ZL_STREAM_C2C_R32 1
#endif

#define ZL_CAT2(a, b) a##b
#define ZL_CAT(a, b) ZL_CAT2(a, b)

void ZL_CAT(zl_fill_order_, ZL_LOGN)(zl::OrderTable& tab) {
    using namespace zl;
    using G = Geo<ZL_LOGN, cfg_logr(ZL_LOGN), cfg_tpb(ZL_LOGN)>;
    using GS0 = Geo<ZL_LOGN, ZL_STREAM_LOGR(ZL_LOGN), cfg_tpb(ZL_LOGN)>;
#if ZL_LOGN == 13 && ZL_STREAM_LP2
    using GS = GeoLP2<ZL_LOGN, ZL_STREAM_LOGR(ZL_LOGN), cfg_tpb(ZL_LOGN)>;
#else
    using GS = GS0;
#endif
#if ZL_LOGN == 14 && ZL_STREAM_C2C_R32
    using GC = Geo<ZL_LOGN, 5, cfg_tpb(ZL_LOGN)>;
#else
    using GC = GS;
#endif
    tab[C2C][AOS][AOS] = make_entry<G, C2C, AOS, AOS, GC>();
    tab[C2C][AOS][SOA] = make_entry<G, C2C, AOS, SOA, GC>();
    tab[C2C][SOA][AOS] = make_entry<G, C2C, SOA, AOS, GC>();
    tab[C2C][SOA][SOA] = make_entry<G, C2C, SOA, SOA, GC>();
    tab[R2C][AOS][AOS] = make_entry<G, R2C, AOS, AOS, GS>();
    tab[R2C][AOS][SOA] = make_entry<G, R2C, AOS, SOA, GS>();
    tab[R2C_SQR][AOS][AOS] = make_entry<G, R2C_SQR, AOS, AOS, GS>();
    tab[C2R][AOS][AOS] = make_entry<G, C2R, AOS, AOS, GS>();
    tab[C2R][SOA][AOS] = make_entry<G, C2R, SOA, AOS, GS>();
    tab[R2C_FRAMED][AOS][AOS] = make_lds_entry<G, R2C, AOS, AOS, float2, 1>();
    tab[R2C_FRAMED][AOS][SOA] = make_lds_entry<G, R2C, AOS, SOA, float2, 1>();
    tab[R2C_SQR_FRAMED][AOS][AOS] = make_lds_entry<G, R2C_SQR, AOS, AOS, float2, 1>();
}

void ZL_CAT(zl_fill_order64_, ZL_LOGN)(zl::OrderTable& tab) {
    using namespace zl;
#if ZL_LOGN <= 13
    using G = Geo<ZL_LOGN, cfg_logr(ZL_LOGN), cfg_tpb(ZL_LOGN)>;
    // the streaming variants at 8 points per thread: at 16 the C2C forward
    // kernel took 256 VGPRs and scratch at N = 4096 (1 wave per SIMD); at 8
    // it takes 114-128, and two workgroups (2 x (69.6 KB tile + image)) fit
    // one CU's LDS (measured, 8192 x 4096: 0.198-0.200 ms against 0.208-0.210
    // for the one-launch kernel, which stays at 16 points: 3% faster there)
    using GS = Geo<ZL_LOGN, cfg_logr(ZL_LOGN) < 3 ? cfg_logr(ZL_LOGN) : 3, cfg_tpb(ZL_LOGN)>;
    tab[C2C][AOS][AOS] = make_lds_entry<G, C2C, AOS, AOS, double2, 0, GS>();
    tab[C2C][AOS][SOA] = make_lds_entry<G, C2C, AOS, SOA, double2, 0, GS>();
    tab[C2C][SOA][AOS] = make_lds_entry<G, C2C, SOA, AOS, double2, 0, GS>();
    tab[C2C][SOA][SOA] = make_lds_entry<G, C2C, SOA, SOA, double2, 0, GS>();
    tab[R2C][AOS][AOS] = make_lds_entry<G, R2C, AOS, AOS, double2, 0, GS>();
    tab[R2C][AOS][SOA] = make_lds_entry<G, R2C, AOS, SOA, double2, 0, GS>();
    tab[R2C_SQR][AOS][AOS] = make_lds_entry<G, R2C_SQR, AOS, AOS, double2, 0, GS>();
    tab[C2R][AOS][AOS] = make_lds_entry<G, C2R, AOS, AOS, double2, 0, GS>();
    tab[C2R][SOA][AOS] = make_lds_entry<G, C2R, SOA, AOS, double2, 0, GS>();
    tab[R2C_FRAMED][AOS][AOS] = make_lds_entry<G, R2C, AOS, AOS, double2, 1>();
    tab[R2C_FRAMED][AOS][SOA] = make_lds_entry<G, R2C, AOS, SOA, double2, 1>();
    tab[R2C_SQR_FRAMED][AOS][AOS] = make_lds_entry<G, R2C_SQR, AOS, AOS, double2, 1>();
#else
    (void)tab;  // FP64 runs this order as HBM passes
#endif
}
