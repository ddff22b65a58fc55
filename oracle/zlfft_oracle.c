/*
 * CPU oracle for the zldsp::fft hot path (CFFT<F> / RFFT<F>).
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker: tests/, the smoke check and
 * bench.py's CPU-baseline leg load it; the product library never does.
 *
 * It restates, in plain C, the reference's algorithm for float and double:
 *   - twiddles in FP64 rounded to F, cospi(x) = cos(x * pi) as on Linux
 *     (reference src/common/zldsp_fft_common_math.hpp:19-25);
 *   - the pure Stockham DIT plan (radix-8 or radix-4 first pass, radix-4
 *     stages) for orders below the switch order, and the radix-4 DIF macro
 *     stages + micro Stockham + digit-reversed transpose above it
 *     (reference src/common/zldsp_fft_common_init.hpp:181-339,
 *      src/common/zldsp_fft_common_execute.hpp:25-273);
 *   - RFFT pack / post-twiddle / squared magnitude / pre-twiddle
 *     (reference src/zldsp_fft_rfft.hpp:24-138,
 *      src/common/zldsp_fft_common_execute.hpp:317-552).
 * Build with -ffp-contract=off so that only the explicit fma() calls fuse,
 * as Highway's MulAdd/NegMulAdd do on AVX2.
 */
#include <math.h>
#include <pthread.h>
#include <stddef.h>
#include <stdlib.h>
#include <unistd.h>

#ifndef ZLO_MUTATE
#define ZLO_MUTATE 0
#endif

/* reference src/common/zldsp_fft_common_math.hpp:19-25 (non-Apple branch) */
static inline double zlo_cospi(double x) { return cos(x * 3.14159265358979323846); }
static inline double zlo_sinpi(double x) { return sin(x * 3.14159265358979323846); }

#define T float
#define SFX f32
#define FMA fmaf
#include "zlfft_oracle_impl.h"
#undef T
#undef SFX
#undef FMA

#define T double
#define SFX f64
#define FMA fma
#include "zlfft_oracle_impl.h"
#undef T
#undef SFX
#undef FMA

/* get_switch_order (reference src/common/zldsp_fft_common_system.hpp:23-157,
 * Linux branch): cache sizes from sysconf with the reference's defaults. */
static size_t zlo_order_of(size_t n) {
    size_t o = 0;
    if (n == 0) return 0;
    while (n >>= 1) ++o;
    return o;
}

void zlo_switch_order_from(int fsize, size_t l1, size_t l2, size_t l3, int *l1_order, int *sw_order) {
    const size_t max_l1 = zlo_order_of(l1 / (8 * (size_t)fsize));
    size_t sw = zlo_order_of(l2 / (12 * (size_t)fsize));
    if (max_l1 + 4 > sw) sw = max_l1 + 4;
    if (l3 > 64ull * 1024 * 1024) {
        const size_t o3 = zlo_order_of(l3 / (12 * (size_t)fsize));
        if (o3 > sw) sw = o3;
    }
    *l1_order = (int)max_l1;
    *sw_order = (int)sw;
}

void zlo_switch_order(int fsize, int *l1_order, int *sw_order) {
    long v;
    size_t l1 = 32768, l2 = 262144, l3 = 0;
    if ((v = sysconf(_SC_LEVEL1_DCACHE_SIZE)) > 0) l1 = (size_t)v;
    if ((v = sysconf(_SC_LEVEL2_CACHE_SIZE)) > 0) l2 = (size_t)v;
    if ((v = sysconf(_SC_LEVEL3_CACHE_SIZE)) > 0) l3 = (size_t)v;
    zlo_switch_order_from(fsize, l1, l2, l3, l1_order, sw_order);
}

/* the FP64 cospi the oracle uses, exposed so tests can pin it */
double zlo_cospi_f64(double x) { return zlo_cospi(x); }
double zlo_sinpi_f64(double x) { return zlo_sinpi(x); }
int zlo_mutation(void) { return ZLO_MUTATE; }
