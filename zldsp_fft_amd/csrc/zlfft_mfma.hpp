// Matrix-core (MFMA) variant of the 4096-point CFFT: each radix-16 Stockham
// pass is a 16x16 complex DFT-matrix product on FP16 or BF16 MFMA with FP32
// accumulate, with the operands optionally split into 2 (hi + lo) or 3
// (hi + mid + lo) parts to win back precision.  Twiddles between the passes
// stay in FP32 (VALU), from an LDS copy of the plan's pass twiddles.
//
// v_mfma_f32_16x16x32_{f16,bf16} lane layout (checked on gfx950):
//   A lane l holds A[l % 16][8 (l / 16) + e], e < 8
//   B lane l holds B[8 (l / 16) + e][l % 16]
//   D lane l holds D[4 (l / 16) + r][l % 16], r < 4
// A group of 16 butterflies b (columns) is one product.  The K index
// interleaves real and imaginary parts, k = 2 m + c, so lane l's B operand is
// the 4 complex inputs m = 4 (l / 16) .. + 3 of butterfly l % 16.  The rows
// interleave real and imaginary parts of the outputs too: with
// F[q][m] = exp(-2 pi i q m / 16), A_lo holds outputs q = 0..7 and A_hi
// q = 8..15 as
//   A[2 q'][2m] = Re F[q][m],  A[2 q'][2m+1] = -Im F[q][m]     (Re y_q)
//   A[2 q'+1][2m] = Im F[q][m], A[2 q'+1][2m+1] = Re F[q][m]   (Im y_q)
// (q' = q mod 8), so lane l holds the complex outputs q = 2 (l / 16) and
// 2 (l / 16) + 1 of each product, with the real and imaginary parts in
// neighbouring accumulator registers, ready to store as pairs.
#pragma once

#include "zlfft_device.hpp"

namespace zl {

#ifndef ZL_MFMA_STREAM_WAVES
#define ZL_MFMA_STREAM_WAVES 2
#endif

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

typedef float zl_fv2 __attribute__((ext_vector_type(2)));

// FP16 split of data: hi = v truncated to 11 significant bits (the
// round-toward-zero packed conversion v_cvt_pkrtz_f16_f32, exact truncation in
// FP16's normal range) and lo = v - hi, exact in FP32 (one v_fma_mix_f32 with
// hi as its FP16 operand), rounded to FP16 by a packed conversion: two VALU
// operations per value instead of three for the form that masks the low 13
// mantissa bits and converts both parts (the same parts in FP16's normal
// range; measured on the 16384 x 4096 batch: 13% fewer VALU instructions,
// 0.5% faster).  Unsplit (NS = 1): v rounded to nearest.
// v - (float)h for the FP16 value h in the low (HI = 0) or high (HI = 1) half
// of hp: one v_fma_mix_f32 (FP16 operand converted inside the FMA, exact here)
template <int HI>
__device__ __forceinline__ float sub_f16(unsigned hp, float v) {
    float r;
    if constexpr (HI)
        asm("v_fma_mix_f32 %0, -%1, 1.0, %2 op_sel:[1,0,0] op_sel_hi:[1,0,0]" : "=v"(r) : "v"(hp), "v"(v));
    else
        asm("v_fma_mix_f32 %0, -%1, 1.0, %2 op_sel:[0,0,0] op_sel_hi:[1,0,0]" : "=v"(r) : "v"(hp), "v"(v));
    return r;
}

template <int NS>
__device__ __forceinline__ void split8_trunc(const float* v, zl_h8 (&parts)[NS]) {
#pragma unroll
    for (int e = 0; e < 8; e += 2) {
        if constexpr (NS == 2) {
            const auto hp = __builtin_amdgcn_cvt_pkrtz(v[e], v[e + 1]);
            const unsigned hb = __builtin_bit_cast(unsigned, hp);
            parts[0][e] = hp[0];
            parts[0][e + 1] = hp[1];
            const zl_h2 lp = __builtin_convertvector((zl_fv2){sub_f16<0>(hb, v[e]), sub_f16<1>(hb, v[e + 1])}, zl_h2);
            parts[1][e] = lp.x;
            parts[1][e + 1] = lp.y;
        } else {
            const zl_h2 hp = __builtin_convertvector((zl_fv2){v[e], v[e + 1]}, zl_h2);
            parts[0][e] = hp.x;
            parts[0][e + 1] = hp.y;
        }
    }
}

// Splits 8 FP32 values into NS parts of the MFMA input type, v = sum of the
// parts to about NS x (mantissa bits) bits.
//  FP16: data (RNE = false) as split8_trunc; constants (RNE = true) with
//        round-to-nearest parts, each the FP16 rounding of the previous
//        remainder (exact in FP32).
//  BF16: with three parts, the first two are truncations of the remainder to
//        its top 16 bits (exactly BF16 values, remainders exact in FP32) and
//        the last is rounded to nearest; with two, both round to nearest.
template <int P, int NS, bool RNE = false>
__device__ __forceinline__ void split8(const float* v, typename MfOp<P>::V (&parts)[NS]) {
    if constexpr (P == MF_F16) {
        static_assert(NS <= 2, "FP16 splits into at most two parts");
        if constexpr (RNE) {
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                const _Float16 h = (_Float16)v[e];
                parts[0][e] = h;
                if constexpr (NS == 2) parts[1][e] = (_Float16)(v[e] - (float)h);
            }
        } else {
            split8_trunc<NS>(v, parts);
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

// FP16 split of the scaled values s v (s a power of two, so s v is exact).
template <int NS>
__device__ __forceinline__ void split8_scaled(const float* v, float s, zl_h8 (&parts)[NS]) {
    float w[8];
#pragma unroll
    for (int e = 0; e < 8; ++e) w[e] = v[e] * s;
    split8_trunc<NS>(w, parts);
}

// cos(2 pi j / 16), j = 0..15, in double
__device__ __forceinline__ double cos16(int j) {
    constexpr double c1 = 0.92387953251128675613, c2 = 0.70710678118654752440, c3 = 0.38268343236508977173;
    const double t[16] = {1.0, c1, c2, c3, 0.0, -c3, -c2, -c1, -1.0, -c1, -c2, -c3, 0.0, c3, c2, c1};
    return t[j & 15];
}

// A operands (DFT-16 matrix rows of outputs 0..7 and 8..15, split in NS
// parts) of this lane
template <int P, int NS>
__device__ __forceinline__ void dft16_operands(int lane, typename MfOp<P>::V (&alo)[NS],
                                               typename MfOp<P>::V (&ahi)[NS]) {
    const int row = lane & 15, h = lane >> 4;
    const int c = row & 1;  // 0: real part of the output, 1: imaginary part
    float flo[8], fhi[8];
#pragma unroll
    for (int e = 0; e < 8; ++e) {
        const int m = 4 * h + (e >> 1);
#pragma unroll
        for (int half = 0; half < 2; ++half) {
            const int q = 8 * half + (row >> 1);
            const int qm = (q * m) & 15;
            const double re = cos16(qm), im = -cos16(qm - 4);  // Im exp(-2 pi i qm/16) = -sin = -cos(x - pi/2)
            // exact split of the FP64 value: parts in the operand type
            const float v = c == 0 ? ((e & 1) ? (float)(-im) : (float)re) : ((e & 1) ? (float)re : (float)im);
            (half ? fhi : flo)[e] = v;
        }
    }
    split8<P, NS, true>(flo, alo);
    split8<P, NS, true>(fhi, ahi);
}

// Y = F X for one group of 16 butterflies; x: this lane's 4 complex inputs
// (m = 4h + i), returns this lane's 4 complex outputs y[r], q = mf_out_q(h, r).
// SCALED (FP16 only): the inputs are multiplied by the power of two s first,
// inside the split.
template <int P, int NS, bool SCALED = false>
__device__ __forceinline__ void dft16_mfma(const float2 (&x)[4], float2 (&y)[4], const typename MfOp<P>::V (&alo)[NS],
                                           const typename MfOp<P>::V (&ahi)[NS], float s = 1.0f) {
    using OP = MfOp<P>;
    float v[8];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        v[2 * i] = x[i].x;
        v[2 * i + 1] = x[i].y;
    }
    typename OP::V bp[NS];
    if constexpr (SCALED) {
        static_assert(P == MF_F16, "scaled split is FP16 only");
        split8_scaled<NS>(v, s, bp);
    } else {
        split8<P, NS>(v, bp);
    }
    zl_f4 dlo = {0.f, 0.f, 0.f, 0.f}, dhi = {0.f, 0.f, 0.f, 0.f};
    // terms of total split order < NS, smallest first
#pragma unroll
    for (int ord = NS - 1; ord >= 0; --ord) {
#pragma unroll
        for (int ia = 0; ia <= ord; ++ia) {
            const int ib = ord - ia;
            dlo = OP::mma(alo[ia], bp[ib], dlo);
            dhi = OP::mma(ahi[ia], bp[ib], dhi);
        }
    }
    y[0] = make_float2(dlo[0], dlo[1]);
    y[1] = make_float2(dlo[2], dlo[3]);
    y[2] = make_float2(dhi[0], dhi[1]);
    y[3] = make_float2(dhi[2], dhi[3]);
}

// output index q of y[r] in lane group h (dft16_mfma): 2 h + (r & 1) + 8 (r >> 1)
constexpr int mf_out_dq(int r) { return (r & 1) + 8 * (r >> 1); }

// largest |component| of this wave's inputs into wmax[wave]: bit patterns
// with the sign cleared compare as unsigned integers in the order of the
// magnitudes (the swap of the backward direction does not change it)
__device__ __forceinline__ void mfma_wave_absmax(const float2 (&x)[4][4], float* wmax) {
    unsigned mb = 0u;
#pragma unroll
    for (int gg = 0; gg < 4; ++gg)
#pragma unroll
        for (int mi = 0; mi < 4; ++mi)
            mb = max(mb, max(__float_as_uint(x[gg][mi].x) & 0x7fffffffu, __float_as_uint(x[gg][mi].y) & 0x7fffffffu));
    float mx = __uint_as_float(mb);
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) mx = fmaxf(mx, __shfl_xor(mx, o));
    if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = mx;
}

// scale exponent 15 - e, e the exponent of the transform's largest |component|
// (0 for an all-zero or NaN input)
__device__ __forceinline__ int mfma_scale_exponent(const float* wmax) {
    const float m4 = fmaxf(fmaxf(wmax[0], wmax[1]), fmaxf(wmax[2], wmax[3]));
    int ex = 0;
    (void)frexpf(m4, &ex);  // m4 = f 2^ex, f in [0.5, 1)
    return m4 > 0.f ? 15 - ex : 0;
}

// One 4096-point transform by a workgroup of 256 threads (4 waves, each
// owning 4 groups of 16 butterflies per pass): x holds this lane's inputs
// x[gg][mi] = in[i + 256 (4 h + mi)], i = 16 (4 wave + gg) + b, already
// swapped for the backward direction; the outputs go to out (non-temporal).
// Twiddles come from ltw, an LDS copy of the plan's pass twiddles.
// FP16 range: the input is scaled by 2^(15 - e), e the exponent of its largest
// |component|, so pass inputs stay below 2^15 and every split part is a normal
// FP16 number; each pass can grow values by at most 16, and the FP16 twiddle
// table holds W / 16, so the next pass's inputs are again below 2^15.  All
// scalings are powers of two (exact) and are undone at the store.
// The caller synchronises the workgroup before the call (LDS twiddles ready,
// previous transform's LDS readers done).  MAXDONE: the caller has already
// stored this transform's per-wave maxima in wmax (mfma_wave_absmax) before
// that synchronisation.
template <int P, int NS, int BWD, bool MAXDONE>
__device__ __forceinline__ void mfma4096_transform(float2 (&x)[4][4], float2* lds, const float2* ltw, float* wmax,
                                                   const typename MfOp<P>::V (&alo)[NS],
                                                   const typename MfOp<P>::V (&ahi)[NS], float2* out) {
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int b = lane & 15, h = lane >> 4;
    // LDS and twiddle addresses as a per-lane base plus compile-time offsets
    // (b < 16, so the one-per-16 padding e + (e >> 4) separates):
    //  pass input e = i + 256 m, i = 16 (4 wave + gg) + b, m = 4 h + mi:
    //    pad(e) = b + 17 (4 wave + gg) + 272 m
    //  outputs q = 2 h + mf_out_dq(r):
    //  pass 0 output 16 i + q:                        pad = 17 i + q
    //  pass 1 output 256 (4 wave + gg) + b + 16 q:
    //    pad = 272 (4 wave + gg) + b + 17 q
    //  twiddles: pass 1 [16 m + b], pass 2 [kMfmaTwPass2 + 256 m + i]
    const int rd = b + 68 * wave + 1088 * h;     // + 17 gg + 272 mi
    const int wr0 = 1088 * wave + 17 * b + 2 * h;  // + 272 gg + mf_out_dq(r)
    const int wr1 = 1088 * wave + b + 34 * h;      // + 272 gg + 17 mf_out_dq(r)
    const int tw1 = 64 * h + b;                    // + 16 mi
    const int tw2 = kMfmaTwPass2 + 1024 * h + 64 * wave + b;  // + 256 mi + 16 gg
    float2 y[4][4];
    int sh = 0;
    float s = 1.0f;  // 2^sh, applied inside the first pass's split
    if constexpr (P == MF_F16) {
        if (!MAXDONE) {
            mfma_wave_absmax(x, wmax);
            __syncthreads();  // per-wave maxima
        }
        sh = mfma_scale_exponent(wmax);
        if (sh > 120) {  // tiny inputs: the part of the scale 2^sh a float cannot hold goes in first
#pragma unroll
            for (int gg = 0; gg < 4; ++gg)
#pragma unroll
                for (int mi = 0; mi < 4; ++mi)
                    x[gg][mi] = make_float2(ldexpf(x[gg][mi].x, sh - 120), ldexpf(x[gg][mi].y, sh - 120));
            s = ldexpf(1.0f, 120);
        } else {
            s = ldexpf(1.0f, sh);
        }
    }
    // net power of two to undo at the store: the input scale and two 1/16
    // twiddle factors (one v_ldexp_f32 per value, exact for every result)
    const int unscale = P == MF_F16 ? 8 - sh : 0;
#pragma unroll
    for (int p = 0; p < 3; ++p) {
        if (p > 0) {
#pragma unroll
            for (int gg = 0; gg < 4; ++gg) {
#pragma unroll
                for (int mi = 0; mi < 4; ++mi) {
                    const float2 t = lds[rd + 17 * gg + 272 * mi];
                    const float2 w = p == 1 ? ltw[tw1 + 16 * mi] : ltw[tw2 + 256 * mi + 16 * gg];
                    x[gg][mi] = cmul(t, w);
                }
            }
        }
        if (P == MF_F16 && p == 0) {
#pragma unroll
            for (int gg = 0; gg < 4; ++gg) dft16_mfma<P, NS, P == MF_F16>(x[gg], y[gg], alo, ahi, s);
        } else {
#pragma unroll
            for (int gg = 0; gg < 4; ++gg) dft16_mfma<P, NS>(x[gg], y[gg], alo, ahi);
        }
        if (p < 2) {
            __syncthreads();  // all reads of the previous LDS contents are done
#pragma unroll
            for (int gg = 0; gg < 4; ++gg) {
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    if (p == 0)
                        lds[wr0 + 272 * gg + mf_out_dq(r)] = y[gg][r];
                    else
                        lds[wr1 + 272 * gg + 17 * mf_out_dq(r)] = y[gg][r];
                }
            }
            __syncthreads();
        } else {
            // output i + 256 q, i = 64 wave + 16 gg + b, q = 2 h + mf_out_dq(r);
            // byte offsets as 32-bit unsigned values
            char* obase = reinterpret_cast<char*>(out);
            const unsigned so = 8u * (64u * wave + b + 512u * h);
#pragma unroll
            for (int gg = 0; gg < 4; ++gg) {
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    float2 o = y[gg][r];
                    if constexpr (P == MF_F16) o = make_float2(ldexpf(o.x, unscale), ldexpf(o.y, unscale));
                    st_nt(reinterpret_cast<float2*>(obase + (so + 8u * (16u * gg + 256u * mf_out_dq(r)))),
                          BWD ? swapri(o) : o);
                }
            }
        }
    }
}

// this lane's inputs of transform `in` (4096 complex, AoS; the same pointer
// in every lane)
__device__ __forceinline__ void mfma4096_load(float2 (&x)[4][4], const float2* in) {
    const unsigned lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const unsigned b = lane & 15, h = lane >> 4;
    // byte offsets as 32-bit unsigned values: scalar base + vector offset
    const char* base = reinterpret_cast<const char*>(in);
    const unsigned lo = 8u * (64u * wave + b + 1024u * h);
#pragma unroll
    for (int gg = 0; gg < 4; ++gg)
#pragma unroll
        for (int mi = 0; mi < 4; ++mi)
            x[gg][mi] = ld_nt(reinterpret_cast<const float2*>(base + (lo + 8u * (16u * gg + 256u * mi))));
}

// 4096-point CFFT, AoS in and out, one transform per workgroup.
template <int P, int NS, int BWD>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(ZL_MFMA_STREAM_WAVES))) fft4096_mfma_kernel(KArgs a) {
    using G = Geo<12, 4, 256>;
    __shared__ float2 lds[G::LDSN];
    __shared__ float2 ltw[kMfmaTwTotal];
    __shared__ float wmax[4];
    const int tid = threadIdx.x;
    typename MfOp<P>::V alo[NS], ahi[NS];
    dft16_operands<P, NS>(tid & 63, alo, ahi);
    for (int e = tid; e < kMfmaTwTotal; e += 256) ltw[e] = static_cast<const float2*>(a.tw)[e];
    const long long f = blockIdx.x;
    float2 x[4][4];
    mfma4096_load(x, reinterpret_cast<const float2*>(a.in0) + f * 4096);
    if constexpr (BWD) {
#pragma unroll
        for (int gg = 0; gg < 4; ++gg)
#pragma unroll
            for (int mi = 0; mi < 4; ++mi) x[gg][mi] = swapri(x[gg][mi]);
    }
    __syncthreads();  // LDS twiddles
    mfma4096_transform<P, NS, BWD, false>(x, lds, ltw, wmax, alo, ahi, reinterpret_cast<float2*>(a.out0) + f * 4096);
}

// Persistent form: one workgroup per resident slot loops over transforms
// f = blockIdx.x, + gridDim.x, ...; the next transform's input is loaded into
// registers (non-temporal) while the current one computes, and the DFT
// operands and the LDS twiddles are set up once per workgroup.  A lane with
// no next transform prefetches from the sink instead, so the loop has no
// branch around its loads.
template <int P, int NS, int BWD>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(ZL_MFMA_STREAM_WAVES)))
fft4096_mfma_stream_kernel(KArgs a) {
    using G = Geo<12, 4, 256>;
    __shared__ float2 lds[G::LDSN];
    __shared__ float2 ltw[kMfmaTwTotal];
    __shared__ float wmax[4];
    const int tid = threadIdx.x;
    // twiddle table: all loads first, then the first prefetch, then the LDS
    // writes (one memory round trip at the start, not one per table row)
    constexpr int NT = (kMfmaTwTotal + 255) / 256;
    float2 twr[NT];
#pragma unroll
    for (int i = 0; i < NT; ++i) {
        const int e = tid + 256 * i;
        if (e < kMfmaTwTotal) twr[i] = static_cast<const float2*>(a.tw)[e];
    }
    typename MfOp<P>::V alo[NS], ahi[NS];
    dft16_operands<P, NS>(tid & 63, alo, ahi);
    const float2* in = reinterpret_cast<const float2*>(a.in0);
    const float2* sink = reinterpret_cast<const float2*>(a.sink);
    // source of transform t: the input, or the sink past the batch end.  A
    // 32-bit unsigned compare (batch < 2^31, checked by the host) stays in
    // scalar registers, so the pointer select does too.
    const unsigned nbatch = (unsigned)a.batch;
    auto src = [&](long long t) { return (unsigned)t < nbatch ? in + t * 4096 : sink; };
    long long f = blockIdx.x;
    // two prefetch buffers used in turn (the loop body is unrolled twice), so
    // the prefetched registers are never copied
    float2 pa[4][4], pb[4][4];
    mfma4096_load(pa, src(f));
#pragma unroll
    for (int i = 0; i < NT; ++i) {
        const int e = tid + 256 * i;
        if (e < kMfmaTwTotal) ltw[e] = twr[i];
    }
    // FP16: the per-wave maxima of a transform are taken from the prefetch
    // registers at the end of the previous iteration, so the synchronisation
    // at the top of the next one publishes them and the transform needs none
    // of its own for the scale
    if constexpr (P == MF_F16) mfma_wave_absmax(pa, wmax);
    auto one = [&](long long f, float2 (&cur)[4][4], float2 (&nxt)[4][4]) {
        if constexpr (BWD) {
#pragma unroll
            for (int gg = 0; gg < 4; ++gg)
#pragma unroll
                for (int mi = 0; mi < 4; ++mi) cur[gg][mi] = swapri(cur[gg][mi]);
        }
        // the prefetch is issued at raised wave priority, so its loads leave
        // ahead of the other waves' matrix and vector work: measured 2-4%
        // faster for the FP16 kernels and the BF16 2-way split, 1-3% slower
        // for BF16 unsplit and 3-way split (and 5% slower in the plain
        // radix-16 kernel, which does not use it)
        constexpr bool prio = P == MF_F16 || NS == 2;
        if constexpr (prio) __builtin_amdgcn_s_setprio(3);
        mfma4096_load(nxt, src(f + gridDim.x));
        if constexpr (prio) __builtin_amdgcn_s_setprio(0);
        __syncthreads();  // LDS twiddles and maxima ready / previous transform's LDS readers done
        mfma4096_transform<P, NS, BWD, true>(cur, lds, ltw, wmax, alo, ahi,
                                             reinterpret_cast<float2*>(a.out0) + f * 4096);
        // every wave read wmax before the transform's first LDS exchange barrier
        if constexpr (P == MF_F16) mfma_wave_absmax(nxt, wmax);
    };
    if (f < a.batch) {
        one(f, pa, pb);
        for (f += gridDim.x; f < a.batch; f += gridDim.x) {
            one(f, pb, pa);
            f += gridDim.x;
            if (f >= a.batch) break;
            one(f, pa, pb);
        }
    }
}

}  // namespace zl
