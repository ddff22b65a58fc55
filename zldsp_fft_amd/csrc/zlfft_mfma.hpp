// Matrix-core (MFMA) variant of the 4096-point CFFT: each radix-16 Stockham
// pass is a 16x16 complex DFT-matrix product on FP16 or BF16 MFMA with FP32
// accumulate, with the operands optionally split into 2 (hi + lo) or 3
// (hi + mid + lo) parts to win back precision.  Twiddles between the passes
// stay in FP32 (VALU), from the same LDS image as fft_stream_kernel.
//
// v_mfma_f32_16x16x32_{f16,bf16} lane layout (checked on gfx950):
//   A lane l holds A[l % 16][8 (l / 16) + e], e < 8
//   B lane l holds B[8 (l / 16) + e][l % 16]
//   D lane l holds D[4 (l / 16) + r][l % 16], r < 4
// A group of 16 butterflies b (columns) is one product.  The K index
// interleaves real and imaginary parts, k = 2 m + c, so lane l's B operand is
// the 4 complex inputs m = 4 (l / 16) .. + 3 of butterfly l % 16, and
//   Yr = A_r B with A_r[q][2m] = Re F[q][m], A_r[q][2m+1] = -Im F[q][m]
//   Yi = A_i B with A_i[q][2m] = Im F[q][m], A_i[q][2m+1] =  Re F[q][m]
// for F[q][m] = exp(-2 pi i q m / 16).  Lane l then holds the 4 complex
// outputs q = 4 (l / 16) + r of butterfly l % 16.
#pragma once

#include "zlfft_device.hpp"

namespace zl {

enum MfmaPrec : int { MF_F16 = 0, MF_BF16 = 1 };

typedef _Float16 zl_h8 __attribute__((ext_vector_type(8)));
typedef __bf16 zl_b8 __attribute__((ext_vector_type(8)));
typedef float zl_f4 __attribute__((ext_vector_type(4)));

template <int P>
struct MfOp;
template <>
struct MfOp<MF_F16> {
    using V = zl_h8;
    using S = _Float16;
    static __device__ __forceinline__ zl_f4 mma(V a, V b, zl_f4 c) {
        return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
    }
};
template <>
struct MfOp<MF_BF16> {
    using V = zl_b8;
    using S = __bf16;
    static __device__ __forceinline__ zl_f4 mma(V a, V b, zl_f4 c) {
        return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
    }
};

typedef _Float16 zl_h2 __attribute__((ext_vector_type(2)));
typedef __bf16 zl_b2 __attribute__((ext_vector_type(2)));

// upper 16 bits of a and b packed (a low, b high): the BF16 values of two
// FP32 numbers whose low 16 bits are zero
__device__ __forceinline__ unsigned bf16_pack_upper(float a, float b) {
    return __builtin_amdgcn_perm(__builtin_bit_cast(unsigned, b), __builtin_bit_cast(unsigned, a), 0x07060302u);
}
__device__ __forceinline__ float bf16_trunc(float x) {
    return __builtin_bit_cast(float, __builtin_bit_cast(unsigned, x) & 0xffff0000u);
}

// Splits 8 FP32 values into NS parts of the MFMA input type, v = sum of the
// parts to about NS x (mantissa bits) bits.
//  FP16: parts are round-to-nearest, each the FP16 rounding of the previous
//        remainder (exact in FP32).
//  BF16: with three parts, the first two are truncations of the remainder to
//        its top 16 bits (exactly BF16 values, remainders exact in FP32) and
//        the last is rounded to nearest; with two, both round to nearest.
// neg1 must be -1.0f, but not a compile-time constant: with a constant the
// compiler rewrites fma(h, -1, x) as x - h and loses the single
// mixed-precision FMA (v_fma_mix) that computes the FP16 remainder.
template <int P, int NS>
__device__ __forceinline__ void split8(const float* v, typename MfOp<P>::V (&parts)[NS], float neg1) {
    if constexpr (P == MF_F16) {
        static_assert(NS <= 2, "FP16 splits into at most two parts");
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const _Float16 h = (_Float16)v[e];
            parts[0][e] = h;
            if constexpr (NS == 2) parts[1][e] = (_Float16)__builtin_fmaf((float)h, neg1, v[e]);
        }
    } else {
#pragma unroll
        for (int e = 0; e < 8; e += 2) {
            float ra = v[e], rb = v[e + 1];
#pragma unroll
            for (int s = 0; s < NS; ++s) {
                zl_b2 pk;
                if (s + 1 < NS && NS >= 3) {
                    const float ta = bf16_trunc(ra), tb = bf16_trunc(rb);
                    pk = __builtin_bit_cast(zl_b2, bf16_pack_upper(ta, tb));
                    ra -= ta;
                    rb -= tb;
                } else {
                    pk = zl_b2{(__bf16)ra, (__bf16)rb};
                    if (s + 1 < NS) {  // two parts: round-to-nearest hi, remainder exact in FP32
                        ra -= (float)pk.x;
                        rb -= (float)pk.y;
                    }
                }
                parts[s][e] = pk.x;
                parts[s][e + 1] = pk.y;
            }
        }
    }
}

// cos(2 pi j / 16), j = 0..15, in double
__device__ __forceinline__ double cos16(int j) {
    constexpr double c1 = 0.92387953251128675613, c2 = 0.70710678118654752440, c3 = 0.38268343236508977173;
    const double t[16] = {1.0, c1, c2, c3, 0.0, -c3, -c2, -c1, -1.0, -c1, -c2, -c3, 0.0, c3, c2, c1};
    return t[j & 15];
}

// A operands (DFT-16 matrix, split in NS parts) of this lane
template <int P, int NS>
__device__ __forceinline__ void dft16_operands(int lane, typename MfOp<P>::V (&ar)[NS], typename MfOp<P>::V (&ai)[NS],
                                               float neg1) {
    const int q = lane & 15, h = lane >> 4;
    float fr[8], fi[8];
#pragma unroll
    for (int e = 0; e < 8; ++e) {
        const int m = 4 * h + (e >> 1);
        const int qm = (q * m) & 15;
        const double re = cos16(qm), im = -cos16(qm - 4);  // Im exp(-2 pi i qm/16) = -sin = -cos(x - pi/2)
        // exact split of the FP64 value: parts of (re, im) in the operand type
        fr[e] = (e & 1) ? (float)(-im) : (float)re;
        fi[e] = (e & 1) ? (float)re : (float)im;
    }
    split8<P, NS>(fr, ar, neg1);
    split8<P, NS>(fi, ai, neg1);
}

// Y = F X for one group of 16 butterflies; x: this lane's 4 complex inputs
// (m = 4h + i), returns this lane's 4 complex outputs (q = 4h + r).
template <int P, int NS>
__device__ __forceinline__ void dft16_mfma(const float2 (&x)[4], float2 (&y)[4], const typename MfOp<P>::V (&ar)[NS],
                                           const typename MfOp<P>::V (&ai)[NS], float neg1) {
    using OP = MfOp<P>;
    float v[8];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        v[2 * i] = x[i].x;
        v[2 * i + 1] = x[i].y;
    }
    typename OP::V bp[NS];
    split8<P, NS>(v, bp, neg1);
    zl_f4 cr = {0.f, 0.f, 0.f, 0.f}, ci = {0.f, 0.f, 0.f, 0.f};
    // terms of total split order < NS, smallest first
#pragma unroll
    for (int ord = NS - 1; ord >= 0; --ord) {
#pragma unroll
        for (int ia = 0; ia <= ord; ++ia) {
            const int ib = ord - ia;
            cr = OP::mma(ar[ia], bp[ib], cr);
            ci = OP::mma(ai[ia], bp[ib], ci);
        }
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) y[r] = make_float2(cr[r], ci[r]);
}

// W_{16 w}^{m k} from the LDS twiddle image of geometry G, pass P
// (for FP16 the image holds W / 16: its direct and low-factor entries are
// divided by 16; m = 0 gives 1 / 16 likewise)
template <class G, int P, int PREC>
__device__ __forceinline__ float2 img_twiddle(const float2* ltw, int m, int k) {
    using TI = TwImage<G>;
    constexpr int w = 1 << G::lw(P);
    const int mm = m > 0 ? m - 1 : 0;
    float2 t;
    if constexpr (w <= 16) {
        t = ltw[TI::off(P) + mm * w + k];
    } else {
        t = cmul(ltw[TI::off(P) + mm * 16 + (k & 15)], ltw[TI::off(P) + TI::lo_n(P) + mm * (w / 16) + (k >> 4)]);
    }
    return m > 0 ? t : make_float2(PREC == MF_F16 ? 0.0625f : 1.f, 0.f);
}

// 4096-point CFFT, AoS in and out, one transform per workgroup of 256 threads
// (4 waves, each owning 4 groups of 16 butterflies per pass); twiddles from
// the LDS image of the plan (see TwImage).
// FP16 range: the input is scaled by 2^(15 - e), e the exponent of its largest
// |component|, so pass inputs stay below 2^15 and every split part is a normal
// FP16 number; each pass can grow values by at most 16, and the FP16 twiddle
// image holds W / 16, so the next pass's inputs are again below 2^15.  All
// scalings are powers of two (exact) and are undone at the store.
template <int P, int NS, int BWD>
__global__ void __launch_bounds__(256) fft4096_mfma_kernel(KArgs a) {
    using G = Geo<12, 4, 256>;
    using OP = MfOp<P>;
    constexpr int N = 4096;
    using TI = TwImage<G>;
    __shared__ float2 lds[G::LDSN];
    __shared__ float2 ltw[TI::total()];
    __shared__ float wmax[4];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int b = lane & 15, h = lane >> 4;
    typename OP::V ar[NS], ai[NS];
    const float neg1 = a.batch < 0 ? 1.0f : -1.0f;  // -1 (batch >= 0), unknown to the compiler
    dft16_operands<P, NS>(lane, ar, ai, neg1);
    for (int e = tid; e < TI::total(); e += 256) ltw[e] = a.tw[e];

    const long long f = blockIdx.x;
    const float2* in = reinterpret_cast<const float2*>(a.in0) + f * N;
    float2* out = reinterpret_cast<float2*>(a.out0) + f * N;
    float2 x[4][4], y[4][4];
#pragma unroll
    for (int gg = 0; gg < 4; ++gg) {
        const int i = 16 * (wave * 4 + gg) + b;
#pragma unroll
        for (int mi = 0; mi < 4; ++mi) {
            const float2 t = ld_nt(in + i + 256 * (4 * h + mi));
            x[gg][mi] = BWD ? swapri(t) : t;
        }
    }
    int sh = 0;
    if constexpr (P == MF_F16) {
        float mx = 0.f;
#pragma unroll
        for (int gg = 0; gg < 4; ++gg)
#pragma unroll
            for (int mi = 0; mi < 4; ++mi) mx = fmaxf(mx, fmaxf(fabsf(x[gg][mi].x), fabsf(x[gg][mi].y)));
#pragma unroll
        for (int o = 32; o >= 1; o >>= 1) mx = fmaxf(mx, __shfl_xor(mx, o));
        if (lane == 0) wmax[wave] = mx;
    }
    __syncthreads();  // twiddle image, per-wave maxima
    if constexpr (P == MF_F16) {
        const float mx = fmaxf(fmaxf(wmax[0], wmax[1]), fmaxf(wmax[2], wmax[3]));
        int ex = 0;
        (void)frexpf(mx, &ex);  // mx = f 2^ex, f in [0.5, 1)
        sh = mx > 0.f ? 15 - ex : 0;
#pragma unroll
        for (int gg = 0; gg < 4; ++gg)
#pragma unroll
            for (int mi = 0; mi < 4; ++mi) x[gg][mi] = make_float2(ldexpf(x[gg][mi].x, sh), ldexpf(x[gg][mi].y, sh));
    }
    // net power of two to undo at the store: the input scale and two 1/16 twiddle factors
    const int unscale = P == MF_F16 ? 8 - sh : 0;
#pragma unroll
    for (int p = 0; p < 3; ++p) {
        if (p > 0) {
#pragma unroll
            for (int gg = 0; gg < 4; ++gg) {
                const int i = 16 * (wave * 4 + gg) + b;
#pragma unroll
                for (int mi = 0; mi < 4; ++mi) {
                    const int m = 4 * h + mi;
                    const float2 t = lds[G::pad(i + 256 * m)];
                    const float2 w = p == 1 ? img_twiddle<G, 1, P>(ltw, m, i & 15) : img_twiddle<G, 2, P>(ltw, m, i);
                    x[gg][mi] = cmul(t, w);
                }
            }
        }
#pragma unroll
        for (int gg = 0; gg < 4; ++gg) dft16_mfma<P, NS>(x[gg], y[gg], ar, ai, neg1);
        if (p < 2) {
            __syncthreads();  // all reads of the previous LDS contents are done
            const int lw = 4 * p;
#pragma unroll
            for (int gg = 0; gg < 4; ++gg) {
                const int i = 16 * (wave * 4 + gg) + b;
                const int base = ((i >> lw) << (lw + 4)) + (i & ((1 << lw) - 1));
#pragma unroll
                for (int r = 0; r < 4; ++r) lds[G::pad(base + ((4 * h + r) << lw))] = y[gg][r];
            }
            __syncthreads();
        } else {
#pragma unroll
            for (int gg = 0; gg < 4; ++gg) {
                const int i = 16 * (wave * 4 + gg) + b;
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    float2 o = y[gg][r];
                    if constexpr (P == MF_F16) o = make_float2(ldexpf(o.x, unscale), ldexpf(o.y, unscale));
                    st_nt(out + i + 256 * (4 * h + r), BWD ? swapri(o) : o);
                }
            }
        }
    }
}

}  // namespace zl
