/*
 * Type-generic body of the CPU oracle.  TEST INFRASTRUCTURE ONLY: the
 * product (libzlfft) never links, loads or calls this code; only tests/,
 * the smoke check and bench.py's CPU-baseline leg use it, as the checker.
 *
 * Included twice by zlfft_oracle.c with
 *   T    = float | double     (the reference's template parameter F)
 *   SFX  = f32   | f64        (symbol suffix)
 *   FMA  = fmaf  | fma        (Highway MulAdd; NegMulAdd(a,b,c) = FMA(-a,b,c))
 *
 * Everything below is a scalar restatement of the reference's per-lane
 * arithmetic.  A SIMD lane of the reference performs exactly the IEEE
 * operations written here, in this order, with MulAdd/NegMulAdd fused and
 * Mul/Add/Sub unfused (AVX2/NEON build, reference README.md:31-34).  The
 * reference's AoSoA internal layout only moves values around, so the
 * natural-order scalar arrays used here produce the same numbers.
 */

#define CAT2(a, b) a##_##b
#define CAT(a, b) CAT2(a, b)
#define NM(name) CAT(name, SFX)

typedef struct {
    int order;          /* CFFT order n, size N = 2^n */
    size_t n;
    int micro_order;    /* Stockham order (== order when num_macro == 0) */
    int nstages;        /* Stockham stages for micro_order >= 6 */
    int first_radix;    /* 4 or 8 */
    size_t width[40];   /* width w of stage s (stage 0 has w = 1) */
    T *tw_r[40];        /* stage s >= 1: tw[(m-1) * w + k] = W_{4w}^{m k} */
    T *tw_i[40];
    int num_macro;      /* radix-4 DIF macro stages */
    T *mtw_r[20];       /* macro stage s: tw[(m-1) * width + i] = W_{sub_n}^{m i} */
    T *mtw_i[20];
    size_t seg;         /* number of micro segments S = 4^num_macro */
    size_t *rev4;       /* base-4 digit reversal over num_macro digits */
} NM(zlo_plan);

/* ---- layout access: reference src/common/zldsp_fft_common_structure.hpp:46-135 ----
 * Backward transforms swap re<->im on every user-side load and store. */
typedef struct {
    T *aos;   /* interleaved re,im (AoS) or NULL */
    T *re;    /* SoA real */
    T *im;    /* SoA imag */
} NM(zlo_io);

static inline void NM(ld)(const NM(zlo_io) * p, size_t idx, int fwd, T *r, T *i) {
    T a, b;
    if (p->aos) {
        a = p->aos[2 * idx];
        b = p->aos[2 * idx + 1];
    } else {
        a = p->re[idx];
        b = p->im[idx];
    }
    if (fwd) { *r = a; *i = b; } else { *r = b; *i = a; }
}

static inline void NM(st)(const NM(zlo_io) * p, size_t idx, int fwd, T r, T i) {
    T a = fwd ? r : i, b = fwd ? i : r;
    if (p->aos) {
        p->aos[2 * idx] = a;
        p->aos[2 * idx + 1] = b;
    } else {
        p->re[idx] = a;
        p->im[idx] = b;
    }
}

/* ---- twiddles: reference src/common/zldsp_fft_common_init.hpp ---- */

/* W_{4w}^{m k} as the reference computes it: phase = k * (-2 / (4w)),
 * a = phase * m in double, then cospi/sinpi rounded to T
 * (init.hpp:84-90, :138-145, :152-161). */
static void NM(fill_stage_tw)(size_t w, T *tr, T *ti) {
#if ZLO_MUTATE == 1 /* planted mistake: wrong twiddle sign */
    const double phase_step = 2.0 / (double)(w << 2);
#else
    const double phase_step = -2.0 / (double)(w << 2);
#endif
    for (size_t k = 0; k < w; ++k) {
        const double phase = (double)k * phase_step;
        for (int m = 1; m < 4; ++m) {
            const double a = phase * (double)m;
            tr[(m - 1) * w + k] = (T)zlo_cospi(a);
            ti[(m - 1) * w + k] = (T)zlo_sinpi(a);
        }
    }
}

/* DIF macro twiddles, init.hpp:226-274: stage s has sub_n = N >> 2s,
 * width = sub_n / 4, phase_step = -2 / sub_n. */
static void NM(fill_macro_tw)(size_t sub_n, T *tr, T *ti) {
    const size_t width = sub_n >> 2;
    const double phase_step = -2.0 / (double)sub_n;
    for (size_t i = 0; i < width; ++i) {
        const double phase = (double)i * phase_step;
        for (int m = 1; m < 4; ++m) {
            const double a = phase * (double)m;
            tr[(m - 1) * width + i] = (T)zlo_cospi(a);
            ti[(m - 1) * width + i] = (T)zlo_sinpi(a);
        }
    }
}

void NM(zlo_plan_free)(NM(zlo_plan) * p) {
    if (!p) return;
    for (int s = 0; s < 40; ++s) { free(p->tw_r[s]); free(p->tw_i[s]); }
    for (int s = 0; s < 20; ++s) { free(p->mtw_r[s]); free(p->mtw_i[s]); }
    free(p->rev4);
    free(p);
}

/* Stage list of init_stockham_cfft_state (init.hpp:181-215):
 *   odd  order >= 7 (and 5): [R8First, R4 x ..., R4Last]
 *   even order >= 6 (and 4): [R4First, R4Width4, R4 x ..., R4Last]
 * Orders 4 and 5 use the same two-stage structure in the codelets
 * (kernel.hpp:146-276, :289-589).  Widths grow by 4 after the first stage. */
static int NM(init_stockham)(NM(zlo_plan) * p, int order) {
    p->micro_order = order;
    if (order < 4) { p->nstages = 0; return 0; }
    p->first_radix = (order % 2 == 1) ? 8 : 4;
    int st = 1;
    size_t w = (size_t)p->first_radix;
    p->width[0] = 1;
    int remaining = order - (p->first_radix == 8 ? 3 : 2);
    while (remaining > 0) {
        p->width[st] = w;
        p->tw_r[st] = (T *)malloc(sizeof(T) * 3 * w);
        p->tw_i[st] = (T *)malloc(sizeof(T) * 3 * w);
        if (!p->tw_r[st] || !p->tw_i[st]) return -1;
        NM(fill_stage_tw)(w, p->tw_r[st], p->tw_i[st]);
        w <<= 2;
        remaining -= 2;
        ++st;
    }
    p->nstages = st;
    return 0;
}

/* init_cfft_state (init.hpp:304-339): pure Stockham when order < switch_order,
 * otherwise num_macro radix-4 DIF stages around a micro Stockham whose order is
 * max_l1_order or max_l1_order - 1 so that (order - micro) is even. */
NM(zlo_plan) * NM(zlo_plan_create)(int order, int l1_order, int switch_order) {
    if (order < 0 || order > 30) return NULL;
    NM(zlo_plan) *p = (NM(zlo_plan) *)calloc(1, sizeof(NM(zlo_plan)));
    if (!p) return NULL;
    p->order = order;
    p->n = (size_t)1 << order;
    if (order < switch_order) {
        p->num_macro = 0;
        if (NM(init_stockham)(p, order)) goto fail;
    } else {
        int micro = ((order - l1_order) % 2 != 0) ? l1_order - 1 : l1_order;
        if (micro < 6 || micro > order) goto fail; /* the reference's macro path needs a >= 2-stage micro chain */
        p->num_macro = (order - micro) / 2;
        if (NM(init_stockham)(p, micro)) goto fail;
        for (int s = 0; s < p->num_macro; ++s) {
            const size_t sub_n = p->n >> (2 * s);
            p->mtw_r[s] = (T *)malloc(sizeof(T) * 3 * (sub_n >> 2));
            p->mtw_i[s] = (T *)malloc(sizeof(T) * 3 * (sub_n >> 2));
            if (!p->mtw_r[s] || !p->mtw_i[s]) goto fail;
            NM(fill_macro_tw)(sub_n, p->mtw_r[s], p->mtw_i[s]);
        }
        /* init_bit_reversal_table, init.hpp:282-296 */
        p->seg = (size_t)1 << (order - micro);
        p->rev4 = (size_t *)malloc(sizeof(size_t) * p->seg);
        if (!p->rev4) goto fail;
        for (size_t c = 0; c < p->seg; ++c) {
            size_t rev = 0, tmp = c;
            for (int d = 0; d < p->num_macro; ++d) {
                rev = (rev << 2) | (tmp & 3);
                tmp >>= 2;
            }
            p->rev4[c] = rev;
        }
    }
    return p;
fail:
    NM(zlo_plan_free)(p);
    return NULL;
}

/* ---- butterflies: reference src/common/zldsp_fft_common_radix.hpp ---- */

/* complex multiply as written in radix4_aosoa (radix.hpp:115-129):
 * re = NegMulAdd(i, wi, r*wr), im = MulAdd(i, wr, r*wi) */
static inline void NM(cmul)(T r, T i, T wr, T wi, T *orr, T *oi) {
    *orr = FMA(-i, wi, r * wr);
    *oi = FMA(i, wr, r * wi);
}

/* radix4_first_pass_fused_aosoa (radix.hpp:169-239) and its AoSoA twin
 * radix4_first_pass_aosoa (radix.hpp:852-923): width-1 radix-4, no twiddles.
 * out[4j + q] = sum_m x[j + m N/4] (-i)^{mq}. */
static void NM(r4_first)(const NM(zlo_io) * in, int fwd, T *or_, T *oi, size_t n) {
    const size_t q = n >> 2;
    for (size_t j = 0; j < q; ++j) {
        T x0r, x0i, x1r, x1i, x2r, x2i, x3r, x3i;
        NM(ld)(in, j, fwd, &x0r, &x0i);
        NM(ld)(in, j + 2 * q, fwd, &x2r, &x2i);
        const T t0r = x0r + x2r, t0i = x0i + x2i;
        const T t1r = x0r - x2r, t1i = x0i - x2i;
        NM(ld)(in, j + q, fwd, &x1r, &x1i);
        NM(ld)(in, j + 3 * q, fwd, &x3r, &x3i);
        const T t2r = x1r + x3r, t2i = x1i + x3i;
        const T t3r = x1r - x3r, t3i = x1i - x3i;
        or_[4 * j + 0] = t0r + t2r; oi[4 * j + 0] = t0i + t2i;
        or_[4 * j + 2] = t0r - t2r; oi[4 * j + 2] = t0i - t2i;
        or_[4 * j + 1] = t1r + t3i; oi[4 * j + 1] = t1i - t3r;
        or_[4 * j + 3] = t1r - t3i; oi[4 * j + 3] = t1i + t3r;
    }
}

/* radix8_first_pass_fused_aosoa (radix.hpp:476-657), radix8_first_pass_aosoa
 * (radix.hpp:932-1116) and callback_order_3 (kernel.hpp:97-133):
 * out[8j + q] = sum_m x[j + m N/8] W_8^{mq}, split as even/odd radix-4 with
 * W_8 applied through 1/sqrt(2). */
static void NM(r8_first)(const NM(zlo_io) * in, int fwd, T *or_, T *oi, size_t n) {
    const size_t e = n >> 3;
    const T c = (T)(1.0 / 1.41421356237309504880168872420969808);
    for (size_t j = 0; j < e; ++j) {
        T ar, ai, br, bi;
        NM(ld)(in, j, fwd, &ar, &ai);
        NM(ld)(in, j + 4 * e, fwd, &br, &bi);
        const T t0r = ar + br, t0i = ai + bi, t1r = ar - br, t1i = ai - bi;
        NM(ld)(in, j + 2 * e, fwd, &ar, &ai);
        NM(ld)(in, j + 6 * e, fwd, &br, &bi);
        const T t2r = ar + br, t2i = ai + bi, t3r = ar - br, t3i = ai - bi;
        const T y00r = t0r + t2r, y00i = t0i + t2i;
        const T y02r = t0r - t2r, y02i = t0i - t2i;
        const T y01r = t1r + t3i, y01i = t1i - t3r;
        const T y03r = t1r - t3i, y03i = t1i + t3r;

        NM(ld)(in, j + e, fwd, &ar, &ai);
        NM(ld)(in, j + 5 * e, fwd, &br, &bi);
        const T u0r = ar + br, u0i = ai + bi, u1r = ar - br, u1i = ai - bi;
        NM(ld)(in, j + 3 * e, fwd, &ar, &ai);
        NM(ld)(in, j + 7 * e, fwd, &br, &bi);
        const T u2r = ar + br, u2i = ai + bi, u3r = ar - br, u3i = ai - bi;
        const T y10r = u0r + u2r, y10i = u0i + u2i;
        const T y12r = u0r - u2r, y12i = u0i - u2i;
        const T y11r = u1r + u3i, y11i = u1i - u3r;
        const T y13r = u1r - u3i, y13i = u1i + u3r;

        const T v1r = (y11r + y11i) * c, v1i = (y11i - y11r) * c;
        const T v2r = y12i, v2i = -y12r;
        const T v3r = (y13i - y13r) * c, v3i = (-(y13r + y13i)) * c;

        T *o_r = or_ + 8 * j, *o_i = oi + 8 * j;
        o_r[0] = y00r + y10r; o_i[0] = y00i + y10i;
        o_r[4] = y00r - y10r; o_i[4] = y00i - y10i;
        o_r[1] = y01r + v1r;  o_i[1] = y01i + v1i;
        o_r[5] = y01r - v1r;  o_i[5] = y01i - v1i;
        o_r[2] = y02r + v2r;  o_i[2] = y02i + v2i;
        o_r[6] = y02r - v2r;  o_i[6] = y02i - v2i;
        o_r[3] = y03r + v3r;  o_i[3] = y03i + v3i;
        o_r[7] = y03r - v3r;  o_i[7] = y03i - v3i;
    }
}

/* Core Stockham DIT radix-4 stage, radix4_aosoa (radix.hpp:77-158) and
 * radix4_width4_aosoa (radix.hpp:249-377), which has the same arithmetic:
 * i in [0, N/4), k = i mod w, t_m = x[i + m N/4] W_{4w}^{m k},
 * out[4(i - k) + k + q w] = sum_m t_m (-i)^{mq}.
 * With out_io != NULL this is radix4_last_pass_fused_aosoa (radix.hpp:390-465),
 * storing to the user layout with the backward swap. */
static void NM(r4_stage)(const T *ir, const T *ii, T *or_, T *oi, const NM(zlo_io) * out_io, int fwd,
                         size_t n, size_t w, const T *twr, const T *twi) {
    const size_t q = n >> 2;
    for (size_t i = 0; i < q; ++i) {
#if ZLO_MUTATE == 2 /* planted mistake: wrong twiddle index */
        const size_t k = (i + 1) & (w - 1);
#else
        const size_t k = i & (w - 1);
#endif
        T t1r, t1i, t3r, t3i, t2r, t2i;
        NM(cmul)(ir[i + q], ii[i + q], twr[k], twi[k], &t1r, &t1i);
        NM(cmul)(ir[i + 3 * q], ii[i + 3 * q], twr[2 * w + k], twi[2 * w + k], &t3r, &t3i);
        const T s2r = t1r + t3r, s2i = t1i + t3i;
        const T s3r = t1r - t3r, s3i = t1i - t3i;
        NM(cmul)(ir[i + 2 * q], ii[i + 2 * q], twr[w + k], twi[w + k], &t2r, &t2i);
        const T r0 = ir[i], i0 = ii[i];
        const T s0r = r0 + t2r, s0i = i0 + t2i;
        const T s1r = r0 - t2r, s1i = i0 - t2i;
        const size_t o = ((i & ~(w - 1)) << 2) + k;
        if (out_io) {
            NM(st)(out_io, o, fwd, s0r + s2r, s0i + s2i);
            NM(st)(out_io, o + 2 * w, fwd, s0r - s2r, s0i - s2i);
            NM(st)(out_io, o + w, fwd, s1r + s3i, s1i - s3r);
            NM(st)(out_io, o + 3 * w, fwd, s1r - s3i, s1i + s3r);
        } else {
            or_[o] = s0r + s2r;         oi[o] = s0i + s2i;
#if ZLO_MUTATE == 3 /* planted mistake: swapped outputs 1 and 3 */
            or_[o + 3 * w] = s1r + s3i; oi[o + 3 * w] = s1i - s3r;
            or_[o + w] = s1r - s3i;     oi[o + w] = s1i + s3r;
            continue;
#endif
            or_[o + w] = s1r + s3i;     oi[o + w] = s1i - s3r;
            or_[o + 2 * w] = s0r - s2r; oi[o + 2 * w] = s0i - s2i;
            or_[o + 3 * w] = s1r - s3i; oi[o + 3 * w] = s1i + s3r;
        }
    }
}

/* order 0..3 codelets, kernel.hpp:18-133 */
static void NM(small)(int order, const NM(zlo_io) * in, const NM(zlo_io) * out, int fwd) {
    if (order == 0) {
        T r, i;
        NM(ld)(in, 0, fwd, &r, &i);
        NM(st)(out, 0, fwd, r, i);
    } else if (order == 1) {
        T r0, i0, r1, i1;
        NM(ld)(in, 0, fwd, &r0, &i0);
        NM(ld)(in, 1, fwd, &r1, &i1);
        NM(st)(out, 0, fwd, r0 + r1, i0 + i1);
        NM(st)(out, 1, fwd, r0 - r1, i0 - i1);
    } else {
        T br[8], bi[8];
        if (order == 2) NM(r4_first)(in, fwd, br, bi, 4); else NM(r8_first)(in, fwd, br, bi, 8);
        const size_t n = (size_t)1 << order;
        for (size_t k = 0; k < n; ++k) NM(st)(out, k, fwd, br[k], bi[k]);
    }
}

/* execute_stockham_cfft (execute.hpp:25-144) for a pure-Stockham plan.
 * ws must hold 4 N values of T. */
static void NM(stockham)(const NM(zlo_plan) * p, int order, const NM(zlo_io) * in, const NM(zlo_io) * out,
                         int fwd, T *ws) {
    const size_t n = (size_t)1 << order;
    if (order < 4) { NM(small)(order, in, out, fwd); return; }
    T *ar = ws, *ai = ws + n, *br = ws + 2 * n, *bi = ws + 3 * n;
    if (p->first_radix == 4) NM(r4_first)(in, fwd, ar, ai, n); else NM(r8_first)(in, fwd, ar, ai, n);
    for (int s = 1; s < p->nstages; ++s) {
        const int last = (s == p->nstages - 1);
        NM(r4_stage)(ar, ai, br, bi, last ? out : NULL, fwd, n, p->width[s], p->tw_r[s], p->tw_i[s]);
        T *t = ar; ar = br; br = t;
        t = ai; ai = bi; bi = t;
    }
}

/* radix4_first_pass_dif_fused_aosoa (radix.hpp:669-751) and
 * radix4_dif_aosoa_inplace (radix.hpp:761-843): DIF radix-4 on blocks of
 * 4*width, y_q = sum_m x[i + m width] (-i)^{mq}, out[i + q width] = y_q W^{q i}. */
static void NM(dif_block)(T *r, T *im, size_t width, const T *twr, const T *twi) {
    for (size_t i = 0; i < width; ++i) {
        const T r0 = r[i], i0 = im[i], r2 = r[i + 2 * width], i2 = im[i + 2 * width];
        const T t0r = r0 + r2, t0i = i0 + i2, t1r = r0 - r2, t1i = i0 - i2;
        const T r1 = r[i + width], i1 = im[i + width], r3 = r[i + 3 * width], i3 = im[i + 3 * width];
        const T t2r = r1 + r3, t2i = i1 + i3, t3r = r1 - r3, t3i = i1 - i3;
        r[i] = t0r + t2r;
        im[i] = t0i + t2i;
        {
            const T y2r = t0r - t2r, y2i = t0i - t2i;
            NM(cmul)(y2r, y2i, twr[width + i], twi[width + i], &r[i + 2 * width], &im[i + 2 * width]);
        }
        {
            const T y1r = t1r + t3i, y1i = t1i - t3r;
#if ZLO_MUTATE == 5 /* planted mistake: DIF twiddle applied to the wrong output */
            NM(cmul)(y1r, y1i, twr[2 * width + i], twi[2 * width + i], &r[i + width], &im[i + width]);
#else
            NM(cmul)(y1r, y1i, twr[i], twi[i], &r[i + width], &im[i + width]);
#endif
        }
        {
            const T y3r = t1r - t3i, y3i = t1i + t3r;
            NM(cmul)(y3r, y3i, twr[2 * width + i], twi[2 * width + i], &r[i + 3 * width], &im[i + 3 * width]);
        }
    }
}

/* execute_cfft (execute.hpp:156-273).  ws must hold 8 N values of T. */
static void NM(execute)(const NM(zlo_plan) * p, const NM(zlo_io) * in, const NM(zlo_io) * out, int fwd, T *ws) {
    if (p->num_macro == 0) { NM(stockham)(p, p->order, in, out, fwd, ws); return; }
    const size_t n = p->n, m = (size_t)1 << p->micro_order;
    T *mr = ws, *mi = ws + n;          /* macro space (natural order per element) */
    T *xr = ws + 2 * n, *xi = ws + 3 * n; /* matrix of micro results, row c holds segment c */
    T *micro_ws = ws + 4 * n; /* 4 * 2^micro values */
    /* first DIF pass reads the user layout with the backward swap */
    for (size_t i = 0; i < n; ++i) NM(ld)(in, i, fwd, &mr[i], &mi[i]);
    for (int s = 0; s < p->num_macro; ++s) {
        const size_t width = n >> (2 * s + 2);
        for (size_t blk = 0; blk < n; blk += 4 * width)
            NM(dif_block)(mr + blk, mi + blk, width, p->mtw_r[s], p->mtw_i[s]);
    }
    /* micro Stockham on segment rev4(c), forward (direction already in data),
     * then X[k S + c] = row_c[k] stored with the swap (execute.hpp:200-271) */
    for (size_t c = 0; c < p->seg; ++c) {
        const size_t l = p->rev4[c];
        NM(zlo_io) src = {NULL, mr + l * m, mi + l * m};
        NM(zlo_io) dst = {NULL, xr + c * m, xi + c * m};
        NM(stockham)(p, p->micro_order, &src, &dst, 1, micro_ws);
    }
    for (size_t c = 0; c < p->seg; ++c)
        for (size_t k = 0; k < m; ++k) NM(st)(out, k * p->seg + c, fwd, xr[c * m + k], xi[c * m + k]);
}

static size_t NM(ws_len)(const NM(zlo_plan) * p) { return 8 * p->n + 64; }

/* ---- public CFFT entry (reference CFFT<F>::forward/backward, zldsp_fft_cfft.hpp:38-117) ----
 * layout 0 = AoS (in0 interleaved), 1 = SoA (in0 = re, in1 = im). */
int NM(zlo_cfft)(int order, int fwd, int l1_order, int switch_order, int in_layout, T *in0, T *in1,
                 int out_layout, T *out0, T *out1) {
    NM(zlo_plan) *p = NM(zlo_plan_create)(order, l1_order, switch_order);
    if (!p) return -1;
    T *ws = (T *)malloc(sizeof(T) * NM(ws_len)(p));
    if (!ws) { NM(zlo_plan_free)(p); return -2; }
    NM(zlo_io) in = {in_layout == 0 ? in0 : NULL, in0, in1};
    NM(zlo_io) out = {out_layout == 0 ? out0 : NULL, out0, out1};
    NM(execute)(p, &in, &out, fwd, ws);
    free(ws);
    NM(zlo_plan_free)(p);
    return 0;
}

/* ---- RFFT: reference zldsp_fft_rfft.hpp and execute.hpp:317-552 ---- */

/* generate_rfft_pre_post_twiddles (init.hpp:347-388): for k = 1..M/2,
 * a = k * (2 / N), wc = (T)(0.5 cospi(a)), ws = (T)(0.5 sinpi(a) + 0.5). */
static void NM(rfft_tw)(size_t k, size_t nreal, T *wc, T *wsn) {
    const double a = (double)k * (2.0 / (double)nreal);
#if ZLO_MUTATE == 4 /* planted mistake: dropped +1/2 term of ws */
    *wc = (T)(0.5 * zlo_cospi(a));
    *wsn = (T)(0.5 * zlo_sinpi(a));
#else
    *wc = (T)(0.5 * zlo_cospi(a));
    *wsn = (T)(0.5 * zlo_sinpi(a) + 0.5);
#endif
}

/* forward post-processing, execute.hpp:317-385 (mode 0: complex out to `out`)
 * and execute.hpp:480-552 (mode 1: squared magnitude to sq).  The bin M/2 is
 * written twice by the reference, last by the "M - k" formula. */
static void NM(rfft_post)(size_t mh, const T *zr, const T *zi, const NM(zlo_io) * out, T *sq, int mode) {
    const size_t nreal = 2 * mh;
    const T r0 = zr[0], i0 = zi[0];
    if (mode == 0) {
        NM(st)(out, 0, 1, r0 + i0, (T)0.0);
        NM(st)(out, mh, 1, r0 - i0, (T)0.0);
    } else {
        const T dc = r0 + i0, ny = r0 - i0;
        sq[0] = dc * dc;
        sq[mh] = ny * ny;
    }
    for (size_t k = 1; k <= mh / 2; ++k) {
        const T r1 = zr[k], i1 = zi[k], r2 = zr[mh - k], i2 = zi[mh - k];
        T wc, wsn;
        NM(rfft_tw)(k, nreal, &wc, &wsn);
        const T si = i1 + i2, dr = r1 - r2;
        const T xr_k = FMA(-dr, wsn, FMA(si, wc, r1));
        const T xi_k = FMA(-dr, wc, FMA(-si, wsn, i1));
        const T xr_m = FMA(dr, wsn, FMA(-si, wc, r2));
        const T xi_m = FMA(-dr, wc, FMA(-si, wsn, i2));
        if (mode == 0) {
            NM(st)(out, k, 1, xr_k, xi_k);
            NM(st)(out, mh - k, 1, xr_m, xi_m);
        } else {
            sq[k] = FMA(xr_k, xr_k, xi_k * xi_k);
            sq[mh - k] = FMA(xr_m, xr_m, xi_m * xi_m);
        }
    }
}

/* backward pre-processing, execute.hpp:398-468.  Imag parts of bins 0 and
 * N/2 are never read. */
static void NM(rfft_pre)(size_t mh, const NM(zlo_io) * in, T *zr, T *zi) {
    const size_t nreal = 2 * mh;
    T r0, i0, rm, im;
    NM(ld)(in, 0, 1, &r0, &i0);
    NM(ld)(in, mh, 1, &rm, &im);
    zr[0] = (r0 + rm) * (T)0.5;
    zi[0] = (r0 - rm) * (T)0.5;
    for (size_t k = 1; k <= mh / 2; ++k) {
        T r1, i1, r2, i2;
        NM(ld)(in, k, 1, &r1, &i1);
        NM(ld)(in, mh - k, 1, &r2, &i2);
        T wc, wsn;
        NM(rfft_tw)(k, nreal, &wc, &wsn);
        const T si = i1 + i2, dr = r1 - r2;
        const T zr_k = FMA(-dr, wsn, FMA(-si, wc, r1));
        const T zi_k = FMA(dr, wc, FMA(-si, wsn, i1));
        const T zr_m = FMA(dr, wsn, FMA(si, wc, r2));
        const T zi_m = FMA(dr, wc, FMA(-si, wsn, i2));
        zr[k] = zr_k;
        zi[k] = zi_k;
        zr[mh - k] = zr_m;
        zi[mh - k] = zi_m;
    }
}

/* mode: 0 forward (real -> complex, out layout), 1 forward_sqr_mag,
 * 2 backward (complex in layout -> real).  rfft order >= 1. */
int NM(zlo_rfft)(int rorder, int mode, int l1_order, int switch_order, int layout, T *a0, T *a1, T *real,
                 T *sq) {
    if (rorder < 1) return -1;
    const int corder = rorder - 1;
    const size_t mh = (size_t)1 << corder;
    NM(zlo_plan) *p = NM(zlo_plan_create)(corder, l1_order, switch_order);
    if (!p) return -1;
    T *ws = (T *)malloc(sizeof(T) * NM(ws_len)(p));
    T *zr = (T *)malloc(sizeof(T) * mh), *zi = (T *)malloc(sizeof(T) * mh);
    if (!ws || !zr || !zi) { free(ws); free(zr); free(zi); NM(zlo_plan_free)(p); return -2; }
    NM(zlo_io) cplx = {layout == 0 ? a0 : NULL, a0, a1};
    NM(zlo_io) realv = {real, NULL, NULL};
    NM(zlo_io) ztmp = {NULL, zr, zi};
    if (mode == 0 || mode == 1) {
        /* RFFT::execute_forward / forward_sqr_mag (rfft.hpp:104-124) */
        NM(execute)(p, &realv, &ztmp, 1, ws);
        NM(rfft_post)(mh, zr, zi, &cplx, sq, mode);
    } else {
        /* RFFT::execute_backward (rfft.hpp:132-138) */
        NM(rfft_pre)(mh, &cplx, zr, zi);
        NM(execute)(p, &ztmp, &realv, 0, ws);
    }
    free(ws); free(zr); free(zi);
    NM(zlo_plan_free)(p);
    return 0;
}

/* ---- batched, multi-threaded CFFT for the CPU baseline ----
 * One plan (read-only twiddles) and one workspace per thread; each thread
 * takes a contiguous slice of the batch, as separate reference instances would. */
typedef struct {
    const NM(zlo_plan) * p;
    int fwd, in_layout, out_layout;
    T *in0, *in1, *out0, *out1;
    size_t b0, b1;
} NM(zlo_job);

static void *NM(zlo_worker)(void *arg) {
    NM(zlo_job) *j = (NM(zlo_job) *)arg;
    const size_t n = j->p->n;
    T *ws = (T *)malloc(sizeof(T) * NM(ws_len)(j->p));
    if (!ws) return (void *)1;
    for (size_t b = j->b0; b < j->b1; ++b) {
        NM(zlo_io) in, out;
        if (j->in_layout == 0) { in.aos = j->in0 + 2 * b * n; in.re = in.im = NULL; }
        else { in.aos = NULL; in.re = j->in0 + b * n; in.im = j->in1 + b * n; }
        if (j->out_layout == 0) { out.aos = j->out0 + 2 * b * n; out.re = out.im = NULL; }
        else { out.aos = NULL; out.re = j->out0 + b * n; out.im = j->out1 + b * n; }
        NM(execute)(j->p, &in, &out, j->fwd, ws);
    }
    free(ws);
    return NULL;
}

int NM(zlo_cfft_batch)(int order, size_t batch, int fwd, int l1_order, int switch_order, int in_layout, T *in0,
                       T *in1, int out_layout, T *out0, T *out1, int nthreads) {
    NM(zlo_plan) *p = NM(zlo_plan_create)(order, l1_order, switch_order);
    if (!p) return -1;
    if (nthreads < 1) nthreads = 1;
    if (nthreads > 256) nthreads = 256;
    pthread_t th[256];
    NM(zlo_job) jobs[256];
    int rc = 0;
    for (int t = 0; t < nthreads; ++t) {
        jobs[t].p = p;
        jobs[t].fwd = fwd;
        jobs[t].in_layout = in_layout;
        jobs[t].out_layout = out_layout;
        jobs[t].in0 = in0; jobs[t].in1 = in1; jobs[t].out0 = out0; jobs[t].out1 = out1;
        jobs[t].b0 = batch * (size_t)t / (size_t)nthreads;
        jobs[t].b1 = batch * (size_t)(t + 1) / (size_t)nthreads;
        if (pthread_create(&th[t], NULL, NM(zlo_worker), &jobs[t])) { rc = -3; nthreads = t; break; }
    }
    for (int t = 0; t < nthreads; ++t) {
        void *r = NULL;
        pthread_join(th[t], &r);
        if (r) rc = -2;
    }
    NM(zlo_plan_free)(p);
    return rc;
}

#undef NM
#undef CAT
#undef CAT2
