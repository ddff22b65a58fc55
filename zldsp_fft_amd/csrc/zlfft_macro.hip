// HBM-pass kernels of radix 2^ZL_LOGL for the multi-pass large transforms
// (compiled once per radix with -DZL_LOGL=l).
#include "zlfft_registry.hpp"

#ifndef ZL_LOGL
#error "compile with -DZL_LOGL=<log2 radix>"
#endif

#define ZL_CAT2(a, b) a##b
#define ZL_CAT(a, b) ZL_CAT2(a, b)

// FP32: the radix-1024 and radix-512 passes run with 32 points per thread
// (32 x 32 and 16 x 32, one LDS exchange instead of two): measured 3-8% faster
// on the order-18 and order-20 batches, 1-4% at orders 17, 19 and 25
void ZL_CAT(zl_fill_macro_, ZL_LOGL)(zl::MacroTable& tab) {
    using namespace zl;
    using GL = Geo<ZL_LOGL, ZL_LOGL >= 9 ? 5 : 4, 1>;
    constexpr int C = cfg_macro_c(ZL_LOGL);
    tab[0][AOS] = make_macro_entry<GL, C, 1, 0, AOS, AOS>();
    tab[0][SOA] = make_macro_entry<GL, C, 1, 0, SOA, AOS>();
    tab[0][kInternal] = make_macro_entry<GL, C, 1, 0, AOS, AOS, float2, 0>();
    tab[1][AOS] = make_macro_entry<GL, C, 0, 1, AOS, AOS>();
    tab[1][SOA] = make_macro_entry<GL, C, 0, 1, AOS, SOA>();
    tab[1][kInternal] = make_macro_entry<GL, C, 0, 1, AOS, AOS, float2, 0>();
    tab[2][AOS] = make_macro_entry<GL, C, 0, 0, AOS, AOS>();
    tab[kRoleR2C][AOS] = make_macro_entry<GL, C, 0, 1, AOS, AOS, float2, 1, R2C>();
    tab[kRoleR2C][SOA] = make_macro_entry<GL, C, 0, 1, AOS, SOA, float2, 1, R2C>();
    tab[kRoleSqr][AOS] = make_macro_entry<GL, C, 0, 1, AOS, AOS, float2, 1, R2C_SQR>();
    tab[kRoleC2R][AOS] = make_macro_entry<GL, C, 1, 0, AOS, AOS, float2, 1, C2R>();
    tab[kRoleC2R][SOA] = make_macro_entry<GL, C, 1, 0, SOA, AOS, float2, 1, C2R>();
    tab[kRoleFramed][AOS] = make_macro_entry<GL, C, 1, 0, FRAMED_IN, AOS, float2, 0>();
}

void ZL_CAT(zl_fill_macro64_, ZL_LOGL)(zl::MacroTable& tab) {
    using namespace zl;
    using GL = Geo<ZL_LOGL, 4, 1>;
    constexpr int C = cfg_macro_c64(ZL_LOGL);
    tab[0][AOS] = make_macro_entry<GL, C, 1, 0, AOS, AOS, double2>();
    tab[0][SOA] = make_macro_entry<GL, C, 1, 0, SOA, AOS, double2>();
    tab[0][kInternal] = make_macro_entry<GL, C, 1, 0, AOS, AOS, double2, 0>();
    tab[1][AOS] = make_macro_entry<GL, C, 0, 1, AOS, AOS, double2>();
    tab[1][SOA] = make_macro_entry<GL, C, 0, 1, AOS, SOA, double2>();
    tab[1][kInternal] = make_macro_entry<GL, C, 0, 1, AOS, AOS, double2, 0>();
    tab[2][AOS] = make_macro_entry<GL, C, 0, 0, AOS, AOS, double2>();
    tab[kRoleR2C][AOS] = make_macro_entry<GL, C, 0, 1, AOS, AOS, double2, 1, R2C>();
    tab[kRoleR2C][SOA] = make_macro_entry<GL, C, 0, 1, AOS, SOA, double2, 1, R2C>();
    tab[kRoleSqr][AOS] = make_macro_entry<GL, C, 0, 1, AOS, AOS, double2, 1, R2C_SQR>();
    tab[kRoleC2R][AOS] = make_macro_entry<GL, C, 1, 0, AOS, AOS, double2, 1, C2R>();
    tab[kRoleC2R][SOA] = make_macro_entry<GL, C, 1, 0, SOA, AOS, double2, 1, C2R>();
    tab[kRoleFramed][AOS] = make_macro_entry<GL, C, 1, 0, FRAMED_IN, AOS, double2, 0>();
}
