// Device side of libzlfft: register codelets, LDS-exchanged Stockham passes
// and the fused RFFT epilogue/prologue, for gfx950 (wave64).
//
// One kernel instantiation handles one CFFT order n (N = 2^n) for one kind
// (C2C, real->complex, real->|X|^2, complex->real) and one pair of user
// layouts.  Each transform is owned by T = N/R threads holding R points in
// registers; TPW transforms share a workgroup.  Pass p of the Stockham
// decomposition has radix r_p (the first pass takes the remainder radix) and
// width w_p = r_0 ... r_{p-1}.  Thread t computes butterflies i = t + T b and
// always holds register j = element t + T j between passes, so the first
// pass reads and the last pass writes user memory in coalesced order; only
// the N-point exchanges between passes go through LDS.
//
// Arithmetic conventions follow the reference (zldsp::fft):
//   forward X[k] = sum_n x[n] e^{-2 pi i nk/N}, unnormalised
//     (reference src/common/zldsp_fft_common_radix.hpp:149-156);
//   backward = swap(re, im) -> forward -> swap(re, im)
//     (reference src/common/zldsp_fft_common_structure.hpp:46-135);
//   complex multiply re = fma(-ai, wi, ar*wr), im = fma(ai, wr, ar*wi)
//     (reference src/common/zldsp_fft_common_radix.hpp:115-129).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace zl {

enum Kind : int { C2C = 0, R2C = 1, R2C_SQR = 2, C2R = 3 };
// kernel-table slots of the framed (STFT front-end) real-to-complex kernels
enum FramedKind : int { R2C_FRAMED = 4, R2C_SQR_FRAMED = 5 };
enum Layout : int { AOS = 0, SOA = 1 };
// input "layout" of the first HBM pass of an STFT: frames of a real signal
constexpr int FRAMED_IN = 2;

constexpr int cdiv(int a, int b) { return (a + b - 1) / b; }

// Geometry of one LDS-resident transform.  LOGN: CFFT order; LOGR_: log2 of
// points per thread; TPB: target workgroup size.
template <int LOGN, int LOGR_, int TPB>
struct Geo {
    static constexpr int logn = LOGN;
    static constexpr int LOGR = LOGN < LOGR_ ? LOGN : LOGR_;
    static constexpr int N = 1 << LOGN;
    static constexpr int R = 1 << LOGR;
    static constexpr int T = N / R;
    static constexpr int TPW = T >= TPB ? 1 : TPB / T;
    static constexpr int WG = T * TPW;
    static constexpr int NP = LOGN == 0 ? 0 : cdiv(LOGN, LOGR);
    static constexpr bool LP2 = false;  // see GeoLP2
    // log2 radix of pass p: the first pass takes the remainder
    static constexpr int lr(int p) { return p == 0 ? LOGN - (NP - 1) * LOGR : LOGR; }
    // log2 width of pass p
    static constexpr int lw(int p) { return p == 0 ? 0 : lr(0) + (p - 1) * LOGR; }
    // offset (in complex entries) of pass p's twiddle block; block p holds
    // W_{r w}^{m k} at [(m - 1) w + k], m = 1..r-1, k = 0..w-1
    static constexpr int tw_off(int p) {
        int o = 0;
        for (int q = 1; q < p; ++q) o += ((1 << lr(q)) - 1) << lw(q);
        return o;
    }
    static constexpr int tw_total() { return tw_off(NP); }
    // LDS row per transform: one pad slot per 16 complex values
    static constexpr int LDSN = N + (N >> 4) + 1;
    static constexpr int pad(int e) { return e + (e >> 4); }
};

// The same transform with its radix-2 remainder pass run last, between the
// two halves of each wave (run_pass_img, v_permlane32_swap), instead of first
// with an LDS exchange after it: one LDS exchange fewer.  Persistent streaming
// kernels at 2^13 points (16 x 16 x 16 x 2).
template <int LOGN, int LOGR_, int TPB>
struct GeoLP2 : Geo<LOGN, LOGR_, TPB> {
    using B = Geo<LOGN, LOGR_, TPB>;
    static constexpr bool LP2 = true;
    static_assert(B::NP >= 3 && LOGN - (B::NP - 1) * B::LOGR == 1 && B::T >= 64 && B::TPW == 1,
                  "lane-pair radix 2: the remainder is 2 and a transform spans whole waves");
    static constexpr int lr(int p) { return p == B::NP - 1 ? 1 : B::LOGR; }
    static constexpr int lw(int p) { return B::LOGR * p; }
    static constexpr int tw_off(int p) {
        int o = 0;
        for (int q = 1; q < p; ++q) o += ((1 << lr(q)) - 1) << lw(q);
        return o;
    }
    static constexpr int tw_total() { return tw_off(B::NP); }
};

// ---------------------------------------------------------------- codelets

// Complex element types: float2 (the FP32 path) and double2 (the FP64 path,
// CFFT<double> / RFFT<double>).  Every codelet, pass and the LDS, multi-pass
// and RFFT kernels are written once over CX; the FP32 instantiations compile
// to the same operations as FP32-only code.
template <class CX>
struct CxT;
template <>
struct CxT<float2> {
    using S = float;
    static __device__ __forceinline__ float2 mk(float x, float y) { return make_float2(x, y); }
    static __device__ __forceinline__ float fma(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
    static constexpr float kH = 0.70710678118654752440f;    // 1/sqrt(2)
    static constexpr float kC16 = 0.92387953251128675613f;  // cos(pi/8)
    static constexpr float kS16 = 0.38268343236508977173f;  // sin(pi/8)
};
template <>
struct CxT<double2> {
    using S = double;
    static __device__ __forceinline__ double2 mk(double x, double y) { return make_double2(x, y); }
    static __device__ __forceinline__ double fma(double a, double b, double c) { return __builtin_fma(a, b, c); }
    static constexpr double kH = 0.70710678118654752440;
    static constexpr double kC16 = 0.92387953251128675613;
    static constexpr double kS16 = 0.38268343236508977173;
};

template <class CX>
__device__ __forceinline__ CX cadd(CX a, CX b) { return CxT<CX>::mk(a.x + b.x, a.y + b.y); }
template <class CX>
__device__ __forceinline__ CX csub(CX a, CX b) { return CxT<CX>::mk(a.x - b.x, a.y - b.y); }
// a * (-i)
template <class CX>
__device__ __forceinline__ CX mul_mi(CX a) { return CxT<CX>::mk(a.y, -a.x); }
template <class CX>
__device__ __forceinline__ CX cmul(CX a, CX w) {
    return CxT<CX>::mk(CxT<CX>::fma(-a.y, w.y, a.x * w.x), CxT<CX>::fma(a.y, w.x, a.x * w.y));
}
template <class CX>
__device__ __forceinline__ CX swapri(CX a) { return CxT<CX>::mk(a.y, a.x); }

// x * W_16^e for the exponents a radix-16 codelet needs
template <int E, class CX>
__device__ __forceinline__ CX w16(CX a) {
    using T = CxT<CX>;
    constexpr int e = E & 15;
    if constexpr (e == 0) return a;
    else if constexpr (e == 4) return mul_mi(a);
    else if constexpr (e == 8) return T::mk(-a.x, -a.y);
    else if constexpr (e == 12) return T::mk(-a.y, a.x);
    else if constexpr (e == 2) return T::mk((a.x + a.y) * T::kH, (a.y - a.x) * T::kH);
    else if constexpr (e == 6) return T::mk((a.y - a.x) * T::kH, (-(a.x + a.y)) * T::kH);
    else if constexpr (e == 1) return cmul(a, T::mk(T::kC16, -T::kS16));
    else if constexpr (e == 3) return cmul(a, T::mk(T::kS16, -T::kC16));
    else if constexpr (e == 9) return cmul(a, T::mk(-T::kC16, T::kS16));
    else {
        static_assert(e == 0, "unsupported W16 exponent");
        return a;
    }
}

// In-place forward DFT-r on x[0..r), natural-order output.
//  r = 4: radix-4 butterfly in the reference's operation order
//         (reference src/common/zldsp_fft_common_radix.hpp:184-235)
//  r = 8: two radix-4 plus W_8 via 1/sqrt(2)
//         (reference src/common/zldsp_fft_common_radix.hpp:499-560)
//  r = 16: 4 x 4: DFT-4 over x[j + 4m], twiddle W16^{j q}, DFT-4 over j.
//  r = 32: 2 x 16 (the radix-1024 HBM passes: 32 x 32).
template <int r, class CX>
__device__ __forceinline__ void dft(CX* x) {
    if constexpr (r == 1) {
    } else if constexpr (r == 2) {
        const CX a = x[0], b = x[1];
        x[0] = cadd(a, b);
        x[1] = csub(a, b);
    } else if constexpr (r == 4) {
        const CX t0 = cadd(x[0], x[2]), t1 = csub(x[0], x[2]);
        const CX t2 = cadd(x[1], x[3]), t3 = csub(x[1], x[3]);
        x[0] = cadd(t0, t2);
        x[2] = csub(t0, t2);
        x[1] = CxT<CX>::mk(t1.x + t3.y, t1.y - t3.x);
        x[3] = CxT<CX>::mk(t1.x - t3.y, t1.y + t3.x);
    } else if constexpr (r == 8) {
        CX e[4] = {x[0], x[2], x[4], x[6]};
        CX o[4] = {x[1], x[3], x[5], x[7]};
        dft<4>(e);
        dft<4>(o);
        o[1] = w16<2>(o[1]);
        o[2] = w16<4>(o[2]);
        o[3] = w16<6>(o[3]);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            x[q] = cadd(e[q], o[q]);
            x[q + 4] = csub(e[q], o[q]);
        }
    } else if constexpr (r == 16) {
        CX b[4][4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            b[j][0] = x[j];
            b[j][1] = x[j + 4];
            b[j][2] = x[j + 8];
            b[j][3] = x[j + 12];
            dft<4>(b[j]);
        }
        b[1][1] = w16<1>(b[1][1]);
        b[1][2] = w16<2>(b[1][2]);
        b[1][3] = w16<3>(b[1][3]);
        b[2][1] = w16<2>(b[2][1]);
        b[2][2] = w16<4>(b[2][2]);
        b[2][3] = w16<6>(b[2][3]);
        b[3][1] = w16<3>(b[3][1]);
        b[3][2] = w16<6>(b[3][2]);
        b[3][3] = w16<9>(b[3][3]);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            CX c[4] = {b[0][q], b[1][q], b[2][q], b[3][q]};
            dft<4>(c);
            x[q] = c[0];
            x[q + 4] = c[1];
            x[q + 8] = c[2];
            x[q + 12] = c[3];
        }
    } else if constexpr (r == 32) {
        // 2 x 16: DFT-16 over the even and the odd inputs, W_32^k on the odd
        // half, then radix-2 butterflies
        using T = CxT<CX>;
        using S = typename T::S;
        // cos(pi j / 16), j = 0..8; W_32^k = (cos(pi k / 16), -sin(pi k / 16))
        constexpr double c[9] = {1.0, 0.98078528040323044913, 0.92387953251128675613, 0.83146961230254523708,
                                 0.70710678118654752440, 0.55557023301960222474, 0.38268343236508977173,
                                 0.19509032201612826785, 0.0};
        CX e[16], o[16];
#pragma unroll
        for (int j = 0; j < 16; ++j) {
            e[j] = x[2 * j];
            o[j] = x[2 * j + 1];
        }
        dft<16>(e);
        dft<16>(o);
#pragma unroll
        for (int k = 1; k < 16; ++k) {
            if (k == 8) {
                o[k] = mul_mi(o[k]);
            } else if (k == 4) {
                o[k] = w16<2>(o[k]);
            } else if (k == 12) {
                o[k] = w16<6>(o[k]);
            } else {
                const double cr = k < 8 ? c[k] : -c[16 - k], ci = k < 8 ? c[8 - k] : c[k - 8];  // cos, sin of pi k/16
                o[k] = cmul(o[k], T::mk((S)cr, (S)-ci));
            }
        }
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            x[q] = cadd(e[q], o[q]);
            x[q + 16] = csub(e[q], o[q]);
        }
    } else {
        static_assert(r == 1, "unsupported radix");
    }
}

// ---------------------------------------------------------------- passes

// Pass P of geometry G on registers v; slds is this transform's LDS row.
// Non-last passes leave the next pass's inputs in v (exchanged through LDS).
template <class G, int P, class CX>
__device__ __forceinline__ void run_pass(CX (&v)[G::R], CX* slds, const CX* __restrict__ tw, int t) {
    constexpr int lr = G::lr(P), r = 1 << lr, nb = G::R / r, lw = G::lw(P), w = 1 << lw;
    constexpr bool last = (P == G::NP - 1);
#pragma unroll
    for (int b = 0; b < nb; ++b) {
        CX x[r];
#pragma unroll
        for (int m = 0; m < r; ++m) x[m] = v[b + m * nb];
        if constexpr (P > 0) {
            const int i = t + G::T * b;
            const int k = i & (w - 1);
            const CX* twp = tw + G::tw_off(P) + k;
#pragma unroll
            for (int m = 1; m < r; ++m) x[m] = cmul(x[m], twp[(m - 1) * w]);
        }
        dft<r>(x);
#pragma unroll
        for (int q = 0; q < r; ++q) v[b + q * nb] = x[q];
    }
    if constexpr (!last) {
        __syncthreads();  // every reader of the previous LDS contents is done
#pragma unroll
        for (int b = 0; b < nb; ++b) {
            const int i = t + G::T * b;
            const int base = ((i >> lw) << (lw + lr)) + (i & (w - 1));
#pragma unroll
            for (int q = 0; q < r; ++q) slds[G::pad(base + (q << lw))] = v[b + q * nb];
        }
        __syncthreads();
#pragma unroll
        for (int j = 0; j < G::R; ++j) v[j] = slds[G::pad(t + G::T * j)];
        run_pass<G, P + 1>(v, slds, tw, t);
    }
}

template <class G, class CX>
__device__ __forceinline__ void run_passes(CX (&v)[G::R], CX* slds, const CX* __restrict__ tw, int t) {
    if constexpr (G::NP > 0) run_pass<G, 0>(v, slds, tw, t);
}

// ---------------------------------------------------------------- kernel

// Workgroup order across the chip's 8 XCDs.  The dispatcher hands workgroup b
// to XCD b mod 8, so consecutive workgroups land in different L2 caches.
// xcd_block renumbers them so that XCD x runs the consecutive logical
// indices [x n / 8, (x + 1) n / 8) in dispatch order: neighbours that share
// cache lines and DRAM pages run in the same L2.  A bijection on [0, n).
// Used by the HBM passes (with their non-temporal user I/O it gains about
// 10%); the persistent streaming kernels keep the plain order, which
// measured faster for them.
#ifndef ZL_XCD_REMAP
#define ZL_XCD_REMAP 1
#endif
__device__ __forceinline__ unsigned xcd_block(unsigned b, unsigned n) {
#if ZL_XCD_REMAP
    constexpr unsigned X = 8;
    const unsigned q = n / X, r = n % X;
    const unsigned x = b % X, k = b / X;  // XCD x runs q + (x < r) workgroups
    return x < r ? x * (q + 1) + k : r * (q + 1) + (x - r) * q + k;
#else
    (void)n;
    return b;
#endif
}

struct KArgs {
    const void* in0;   // AoS: float2*, SoA: real part, real input: float*
    const void* in1;   // SoA: imag part
    void* out0;
    void* out1;
    const void* tw;    // Stockham pass twiddles of this geometry (complex type of the kernel)
    const void* rtw;   // RFFT (wc, ws) for k = 1..M/2, at [k - 1]
    long long batch;
    int backward;      // C2C only
    float2* sink;      // kSinkBytes of scratch that inactive lanes of the
                       // streaming kernel read and write instead of user memory
    // framed real input (STFT front-end, FRAMED kernels): frame f of the batch
    // is signal[f * hop + n], n < N (in0 = signal), times window[n] (in1 =
    // window, or null for none)
    long long hop;
    int db;            // R2C_SQR: write 10 log10(max(|X|^2, db_floor)) instead of |X|^2
    float db_floor;
    int cached_out;    // streaming RFFT epilogue: cached instead of non-temporal stores
};

// Squared magnitude, or its level in dB when db is set (STFT front-end).
// FP32: 10 log10(q) = (10 log10 2) log2(q) with the hardware base-2 logarithm
// (v_log_f32) for normal q, the library log10f for a subnormal floor; the
// tests hold the levels to 1e-4 dB of the FP64 value of the same power.
template <class S>
__device__ __forceinline__ S power_out(S p, int db, float db_floor) {
    if (!db) return p;
    const S q = p > (S)db_floor ? p : (S)db_floor;
    if constexpr (sizeof(S) == 4)
        return q >= 1.17549435e-38f ? 3.01029995663981195f * __builtin_amdgcn_logf(q) : 10.0f * log10f(q);
    else
        return 10.0 * log10(q);
}

// Complex element m (real samples 2m, 2m + 1) of frame `frame` of a framed
// real input, windowed.
template <class CX>
__device__ __forceinline__ CX framed_pair(const KArgs& a, long long frame, long long m) {
    using S = typename CxT<CX>::S;
    const S* x = static_cast<const S*>(a.in0) + frame * a.hop + 2 * m;
    S x0 = x[0], x1 = x[1];
    if (a.in1) {
        const S* w = static_cast<const S*>(a.in1) + 2 * m;
        x0 *= w[0];
        x1 *= w[1];
    }
    return CxT<CX>::mk(x0, x1);
}

// Twiddle table of the MFMA 4096-point kernels, rows m = 0..15 so that every
// lane addresses it from a base plus compile-time offsets: pass 1 holds
// W_256^{m k} at [16 m + k] (k < 16), pass 2 W_4096^{m k} at
// [kMfmaTwPass2 + 256 m + k] (k < 256); divided by 16 for the FP16 kernels.
// (Factoring pass 2 as W_256^{m a} W_4096^{m b}, k = 16 a + b, cuts the table
// to 4 KiB and allows 3 workgroups per CU, but its extra complex product per
// value measured slower: 1.5% with the earlier FP16 split, 7% with the current
// one, and 7% at 4 workgroups per CU, capped at 128 VGPRs.)
constexpr int kMfmaTwPass2 = 256;
constexpr int kMfmaTwTotal = kMfmaTwPass2 + 16 * 256;

// size of KArgs::sink: two arrays of the largest LDS transform (2^14 complex)
constexpr long long kSinkBytes = 2ll * 8 * (1 << 14);

template <class CX = float2>
__device__ __forceinline__ CX ld_c(const void* p0, const void* p1, int lay, long long e) {
    using S = typename CxT<CX>::S;
    if (lay == AOS) return reinterpret_cast<const CX*>(p0)[e];
    return CxT<CX>::mk(reinterpret_cast<const S*>(p0)[e], reinterpret_cast<const S*>(p1)[e]);
}
template <class CX>
__device__ __forceinline__ void st_c(void* p0, void* p1, int lay, long long e, CX v) {
    using S = typename CxT<CX>::S;
    if (lay == AOS) {
        reinterpret_cast<CX*>(p0)[e] = v;
    } else {
        reinterpret_cast<S*>(p0)[e] = v.x;
        reinterpret_cast<S*>(p1)[e] = v.y;
    }
}

// Non-temporal global access: every input element is read once and every
// output element written once per launch, so neither is kept in cache.
typedef float zl_f2v __attribute__((ext_vector_type(2)));
typedef double zl_d2v __attribute__((ext_vector_type(2)));

__device__ __forceinline__ float ld_nt(const float* p) { return __builtin_nontemporal_load(p); }
__device__ __forceinline__ double ld_nt(const double* p) { return __builtin_nontemporal_load(p); }
__device__ __forceinline__ float2 ld_nt(const float2* p) {
    const zl_f2v t = __builtin_nontemporal_load(reinterpret_cast<const zl_f2v*>(p));
    return make_float2(t.x, t.y);
}
__device__ __forceinline__ double2 ld_nt(const double2* p) {
    const zl_d2v t = __builtin_nontemporal_load(reinterpret_cast<const zl_d2v*>(p));
    return make_double2(t.x, t.y);
}
__device__ __forceinline__ void st_nt(float* p, float v) { __builtin_nontemporal_store(v, p); }
__device__ __forceinline__ void st_nt(double* p, double v) { __builtin_nontemporal_store(v, p); }
__device__ __forceinline__ void st_nt(float2* p, float2 v) {
    const zl_f2v t = {v.x, v.y};
    __builtin_nontemporal_store(t, reinterpret_cast<zl_f2v*>(p));
}
__device__ __forceinline__ void st_nt(double2* p, double2 v) {
    const zl_d2v t = {v.x, v.y};
    __builtin_nontemporal_store(t, reinterpret_cast<zl_d2v*>(p));
}
template <class CX = float2>
__device__ __forceinline__ CX ld_c_nt(const void* p0, const void* p1, int lay, long long e) {
    using S = typename CxT<CX>::S;
    if (lay == AOS) return ld_nt(reinterpret_cast<const CX*>(p0) + e);
    return CxT<CX>::mk(ld_nt(reinterpret_cast<const S*>(p0) + e), ld_nt(reinterpret_cast<const S*>(p1) + e));
}
template <class CX>
__device__ __forceinline__ void st_c_nt(void* p0, void* p1, int lay, long long e, CX v) {
    using S = typename CxT<CX>::S;
    if (lay == AOS) {
        st_nt(reinterpret_cast<CX*>(p0) + e, v);
    } else {
        st_nt(reinterpret_cast<S*>(p0) + e, v.x);
        st_nt(reinterpret_cast<S*>(p1) + e, v.y);
    }
}

// Fused LDS-resident transform.
//   C2C    : user layout LIN -> user layout LOUT, direction at run time
//   R2C    : N real (= M complex, AoS view) -> M+1 bins in layout LOUT
//            (reference RFFT::forward, src/zldsp_fft_rfft.hpp:68-79, 118-124)
//   R2C_SQR: N real -> M+1 squared magnitudes
//            (reference RFFT::forward_sqr_mag, src/zldsp_fft_rfft.hpp:104-109)
//   C2R    : M+1 bins in layout LIN -> N real
//            (reference RFFT::backward, src/zldsp_fft_rfft.hpp:86-97, 132-138)
// For the real kinds G describes the inner CFFT of size M = N/2.
// FRAMED (R2C and R2C_SQR only): the real input is framed and windowed at
// load time (STFT front-end, see KArgs::hop).
template <class G, int KIND, int LIN, int LOUT, class CX = float2, int FRAMED = 0>
__global__ void __launch_bounds__(G::WG) fft_lds_kernel(KArgs a) {
    static_assert(!FRAMED || KIND == R2C || KIND == R2C_SQR, "framed input is real");
    using T = CxT<CX>;
    using S = typename T::S;
    __shared__ CX lds[G::TPW * G::LDSN];
    const int tid = threadIdx.x;
    const int slot = tid / G::T;
    const int t = tid - slot * G::T;
    const long long f = (long long)blockIdx.x * G::TPW + slot;
    const bool active = f < a.batch;
    CX* slds = lds + slot * G::LDSN;
    constexpr int M = G::N;  // CFFT size
    constexpr int ELEM_IN = (KIND == C2C && LIN == SOA) ? (int)sizeof(S) : (int)sizeof(CX);
    constexpr int ELEM_OUT = (KIND == C2C && LOUT == SOA) ? (int)sizeof(S) : (int)sizeof(CX);
    constexpr bool direct_in = G::T * ELEM_IN >= 128 || G::TPW == 1;
    constexpr bool direct_out = G::T * ELEM_OUT >= 128 || G::TPW == 1;
    const bool bwd = (KIND == C2C) ? (a.backward != 0) : (KIND == C2R);
    // global access: C2C reads each element once and writes it once per
    // launch, non-temporal (measured 3% faster on the 4096-point FP32 and FP64
    // batches, 14% on a 512-transform batch, 4% at orders 6-7); the real kinds
    // keep cached access (their squared-magnitude batches measured slower
    // non-temporal)
    constexpr bool nt = KIND == C2C;
    auto gld = [](const void* p0, const void* p1, int lay, long long e) {
        if constexpr (nt) return ld_c_nt<CX>(p0, p1, lay, e);
        else return ld_c<CX>(p0, p1, lay, e);
    };
    auto gst = [](void* p0, void* p1, int lay, long long e, CX x) {
        if constexpr (nt) st_c_nt(p0, p1, lay, e, x);
        else st_c(p0, p1, lay, e, x);
    };

    CX v[G::R];

    // ------------------------------------------------------------ input
    if constexpr (KIND == C2R) {
        // pre-twiddle prologue (reference src/common/zldsp_fft_common_execute.hpp:398-468):
        // Z written to LDS already swapped for the backward CFFT.
        const long long fb = f * (M + 1);
        if (active) {
            if (t == 0) {
                const S r0 = gld(a.in0, a.in1, LIN, fb).x;
                const S rm = gld(a.in0, a.in1, LIN, fb + M).x;
                slds[0] = swapri(T::mk((r0 + rm) * S(0.5), (r0 - rm) * S(0.5)));
            }
            for (int k = 1 + t; k <= M / 2; k += G::T) {
                const CX x1 = gld(a.in0, a.in1, LIN, fb + k);
                const CX x2 = gld(a.in0, a.in1, LIN, fb + (M - k));
                const CX w = static_cast<const CX*>(a.rtw)[k - 1];  // (wc, ws)
                const S si = x1.y + x2.y, dr = x1.x - x2.x;
                const S zr_m = T::fma(dr, w.y, T::fma(si, w.x, x2.x));
                const S zi_m = T::fma(dr, w.x, T::fma(-si, w.y, x2.y));
                if (k != M - k) {
                    const S zr_k = T::fma(-dr, w.y, T::fma(-si, w.x, x1.x));
                    const S zi_k = T::fma(dr, w.x, T::fma(-si, w.y, x1.y));
                    slds[G::pad(k)] = T::mk(zi_k, zr_k);
                }
                slds[G::pad(M - k)] = T::mk(zi_m, zr_m);
            }
        }
        __syncthreads();
#pragma unroll
        for (int j = 0; j < G::R; ++j) v[j] = slds[G::pad(t + G::T * j)];
    } else {
        constexpr int lin = (KIND == C2C) ? LIN : AOS;  // real input viewed as M complex
        if constexpr (direct_in) {
            const long long fb = f * M;
            if (active) {
                if constexpr (FRAMED) {
#pragma unroll
                    for (int j = 0; j < G::R; ++j) v[j] = framed_pair<CX>(a, f, t + G::T * j);
                } else {
#pragma unroll
                    for (int j = 0; j < G::R; ++j) v[j] = gld(a.in0, a.in1, lin, fb + t + G::T * j);
                }
            } else {
#pragma unroll
                for (int j = 0; j < G::R; ++j) v[j] = T::mk(S(0), S(0));
            }
        } else {
            // coalesced cooperative copy of the workgroup's contiguous block
            const long long wb = (long long)blockIdx.x * G::TPW * M;
            const long long wn = (a.batch - (long long)blockIdx.x * G::TPW) * M;
#pragma unroll
            for (int j = 0; j < G::R; ++j) {
                const int e = tid + G::WG * j;
                const int s = e / M, pos = e - s * M;
                CX val = T::mk(S(0), S(0));
                if (e < wn) {
                    if constexpr (FRAMED)
                        val = framed_pair<CX>(a, (long long)blockIdx.x * G::TPW + s, pos);
                    else
                        val = gld(a.in0, a.in1, lin, wb + e);
                }
                lds[s * G::LDSN + G::pad(pos)] = val;
            }
            __syncthreads();
#pragma unroll
            for (int j = 0; j < G::R; ++j) v[j] = slds[G::pad(t + G::T * j)];
        }
        if (bwd) {
#pragma unroll
            for (int j = 0; j < G::R; ++j) v[j] = swapri(v[j]);
        }
    }

    // ------------------------------------------------------------ passes
    run_passes<G>(v, slds, static_cast<const CX*>(a.tw), t);

    // ------------------------------------------------------------ output
    if constexpr (KIND == R2C || KIND == R2C_SQR) {
        // post-twiddle epilogue (reference src/common/zldsp_fft_common_execute.hpp:317-385, 480-552)
        __syncthreads();
#pragma unroll
        for (int j = 0; j < G::R; ++j) slds[G::pad(t + G::T * j)] = v[j];
        __syncthreads();
        // few threads per frame (T complex stores under 128 bytes): the
        // outputs are written over Z in LDS (the thread of pair (k, M - k)
        // owns both slots; X[M] goes to the row's spare slot pad(M)), then the
        // workgroup stores its frames' contiguous output block cooperatively
        constexpr bool stage_out = G::T * (int)sizeof(CX) < 128 && G::TPW > 1;
        if constexpr (stage_out) {
            if (active) {
                if (t == 0) {
                    const CX z0 = slds[0];
                    const S dc = z0.x + z0.y, ny = z0.x - z0.y;
                    if constexpr (KIND == R2C) {
                        slds[0] = T::mk(dc, S(0));
                        slds[G::pad(M)] = T::mk(ny, S(0));
                    } else {
                        slds[0] = T::mk(power_out(dc * dc, a.db, a.db_floor), S(0));
                        slds[G::pad(M)] = T::mk(power_out(ny * ny, a.db, a.db_floor), S(0));
                    }
                }
                for (int k = 1 + t; k <= M / 2; k += G::T) {
                    const CX z1 = slds[G::pad(k)];
                    const CX z2 = slds[G::pad(M - k)];
                    const CX w = static_cast<const CX*>(a.rtw)[k - 1];
                    const S si = z1.y + z2.y, dr = z1.x - z2.x;
                    const S xr_m = T::fma(dr, w.y, T::fma(-si, w.x, z2.x));
                    const S xi_m = T::fma(-dr, w.x, T::fma(-si, w.y, z2.y));
                    const S xr_k = T::fma(-dr, w.y, T::fma(si, w.x, z1.x));
                    const S xi_k = T::fma(-dr, w.x, T::fma(-si, w.y, z1.y));
                    if constexpr (KIND == R2C) {
                        if (k != M - k) slds[G::pad(k)] = T::mk(xr_k, xi_k);
                        slds[G::pad(M - k)] = T::mk(xr_m, xi_m);
                    } else {
                        if (k != M - k)
                            slds[G::pad(k)] = T::mk(power_out(T::fma(xr_k, xr_k, xi_k * xi_k), a.db, a.db_floor), S(0));
                        slds[G::pad(M - k)] = T::mk(power_out(T::fma(xr_m, xr_m, xi_m * xi_m), a.db, a.db_floor), S(0));
                    }
                }
            }
            __syncthreads();
            const long long f0 = (long long)blockIdx.x * G::TPW;
            const long long nf = a.batch - f0 < G::TPW ? a.batch - f0 : G::TPW;
            const int wn = (int)nf * (M + 1);
            for (int e = tid; e < wn; e += G::WG) {
                const int sl = e / (M + 1), k = e - sl * (M + 1);
                const CX x = lds[sl * G::LDSN + G::pad(k)];
                if constexpr (KIND == R2C)
                    gst(a.out0, a.out1, LOUT, f0 * (M + 1) + e, x);
                else
                    reinterpret_cast<S*>(a.out0)[f0 * (M + 1) + e] = x.x;
            }
        } else if (active) {
            const long long fb = f * (M + 1);
            if (t == 0) {
                const CX z0 = slds[0];
                const S dc = z0.x + z0.y, ny = z0.x - z0.y;
                if constexpr (KIND == R2C) {
                    gst(a.out0, a.out1, LOUT, fb, T::mk(dc, S(0)));
                    gst(a.out0, a.out1, LOUT, fb + M, T::mk(ny, S(0)));
                } else {
                    S* o = reinterpret_cast<S*>(a.out0);
                    o[fb] = power_out(dc * dc, a.db, a.db_floor);
                    o[fb + M] = power_out(ny * ny, a.db, a.db_floor);
                }
            }
            for (int k = 1 + t; k <= M / 2; k += G::T) {
                const CX z1 = slds[G::pad(k)];
                const CX z2 = slds[G::pad(M - k)];
                const CX w = static_cast<const CX*>(a.rtw)[k - 1];
                const S si = z1.y + z2.y, dr = z1.x - z2.x;
                const S xr_m = T::fma(dr, w.y, T::fma(-si, w.x, z2.x));
                const S xi_m = T::fma(-dr, w.x, T::fma(-si, w.y, z2.y));
                const S xr_k = T::fma(-dr, w.y, T::fma(si, w.x, z1.x));
                const S xi_k = T::fma(-dr, w.x, T::fma(-si, w.y, z1.y));
                if constexpr (KIND == R2C) {
                    if (k != M - k) gst(a.out0, a.out1, LOUT, fb + k, T::mk(xr_k, xi_k));
                    gst(a.out0, a.out1, LOUT, fb + (M - k), T::mk(xr_m, xi_m));
                } else {
                    S* o = reinterpret_cast<S*>(a.out0);
                    if (k != M - k) o[fb + k] = power_out(T::fma(xr_k, xr_k, xi_k * xi_k), a.db, a.db_floor);
                    o[fb + (M - k)] = power_out(T::fma(xr_m, xr_m, xi_m * xi_m), a.db, a.db_floor);
                }
            }
        }
    } else {
        constexpr int lout = (KIND == C2C) ? LOUT : AOS;  // real output viewed as M complex
        if (bwd) {
#pragma unroll
            for (int j = 0; j < G::R; ++j) v[j] = swapri(v[j]);
        }
        if constexpr (direct_out) {
            if (active) {
                const long long fb = f * M;
#pragma unroll
                for (int j = 0; j < G::R; ++j) gst(a.out0, a.out1, lout, fb + t + G::T * j, v[j]);
            }
        } else {
            __syncthreads();
#pragma unroll
            for (int j = 0; j < G::R; ++j) slds[G::pad(t + G::T * j)] = v[j];
            __syncthreads();
            const long long wb = (long long)blockIdx.x * G::TPW * M;
            const long long wn = (a.batch - (long long)blockIdx.x * G::TPW) * M;
#pragma unroll
            for (int j = 0; j < G::R; ++j) {
                const int e = tid + G::WG * j;
                const int s = e / M, pos = e - s * M;
                if (e < wn) gst(a.out0, a.out1, lout, wb + e, lds[s * G::LDSN + G::pad(pos)]);
            }
        }
    }
}

// ---------------------------------------------------------------- streaming kernel

// Largest pass width whose twiddles the streaming kernel's LDS image holds
// directly (above it: two factors).  The direct table costs LDS, but saves a
// complex product per value: measured 1.2% faster for the 4096-point C2C
// batch (2 workgroups per CU instead of 3) and 1-3% at orders 9-10; slower
// for the RFFT and STFT kernels, whose LDS also holds the RFFT twiddles and
// the window, and for orders 13-14, where the table outgrows the LDS budget.
constexpr int cfg_tw_direct_max(int logn, int kind) { return kind == C2C && logn >= 9 && logn <= 12 ? 4096 : 16; }
// The same for the streaming kernel over complex type CX: FP64 keeps every
// pass above width 16 factored, so that two workgroups' tiles and images fit
// the LDS of one CU.
template <class CX>
constexpr int cfg_stream_direct_max(int logn, int kind) { return sizeof(CX) == 8 ? cfg_tw_direct_max(logn, kind) : 16; }

// LDS image of the Stockham twiddles of geometry G.  Pass p (radix r, width w)
// needs W_{r w}^{m k}, m = 1..r-1, k < w.  For w <= DMAX the image holds them
// directly at [(m - 1) w + k]; for w > DMAX it holds two factors,
// lo[(m - 1) 16 + b] = W_{r w}^{m b} and hi[(m - 1) (w / 16) + a] = W_{r w}^{16 m a},
// and W^{m k} = lo * hi with k = 16 a + b, so the image stays a few KiB.
template <class G, int DMAX = 16>
struct TwImage {
    static constexpr int lo_n(int p) {
        return ((1 << G::lr(p)) - 1) * ((1 << G::lw(p)) <= DMAX ? (1 << G::lw(p)) : 16);
    }
    static constexpr int hi_n(int p) {
        return (1 << G::lw(p)) <= DMAX ? 0 : ((1 << G::lr(p)) - 1) * ((1 << G::lw(p)) / 16);
    }
    static constexpr int off(int p) {
        int o = 0;
        for (int q = 1; q < p; ++q) o += lo_n(q) + hi_n(q);
        return o;
    }
    static constexpr int total() { return G::NP > 1 ? off(G::NP) : 0; }
};

// Lane-pair radix 2 (G::LP2): the butterfly a thread runs in the pass before
// the last, chosen so that butterflies i and i + T/2, whose outputs the
// radix-2 pass combines, sit in lanes l and l + 32 of one wave.
template <class G>
__device__ __forceinline__ int lp2_index(int t) {
    return (t & 31) + ((t >> 6) << 5) + (((t >> 5) & 1) ? G::T / 2 : 0);
}

// Output position of v[q] after the passes of a lane-pair geometry: that of
// the pass before the last (the radix-2 pass keeps each lane's positions).
template <class G>
__device__ __forceinline__ int lp2_pos(int t, int q) {
    constexpr int lw = G::lw(G::NP - 2), w = 1 << lw;
    const int i = lp2_index<G>(t);
    return ((i >> lw) << (lw + G::LOGR)) + (i & (w - 1)) + q * w;
}

// Pass P with twiddles from the LDS image `ltw` (see TwImage).
template <class G, int P, int DMAX, class CX = float2>
__device__ __forceinline__ void run_pass_img(CX (&v)[G::R], CX* slds, const CX* ltw, int t) {
    constexpr int lr = G::lr(P), r = 1 << lr, nb = G::R / r, lw = G::lw(P), w = 1 << lw;
    constexpr bool last = (P == G::NP - 1);
    // the pass before a lane-pair radix 2: butterflies remapped (lp2_index)
    constexpr bool lp2 = G::LP2 && P == G::NP - 2;
    using TI = TwImage<G, DMAX>;
    static_assert(!lp2 || nb == 1, "lane-pair radix 2 after a full-radix pass");
#pragma unroll
    for (int b = 0; b < nb; ++b) {
        CX x[r];
#pragma unroll
        for (int m = 0; m < r; ++m) x[m] = v[b + m * nb];
        if constexpr (P > 0) {
            const int i = (lp2 ? lp2_index<G>(t) : t) + G::T * b;
            const int k = i & (w - 1);
            if constexpr (w <= DMAX) {
                const CX* twp = ltw + TI::off(P) + k;
#pragma unroll
                for (int m = 1; m < r; ++m) x[m] = cmul(x[m], twp[(m - 1) * w]);
            } else {
                const CX* lo = ltw + TI::off(P) + (k & 15);
                const CX* hi = ltw + TI::off(P) + TI::lo_n(P) + (k >> 4);
#pragma unroll
                for (int m = 1; m < r; ++m) x[m] = cmul(x[m], cmul(lo[(m - 1) * 16], hi[(m - 1) * (w / 16)]));
            }
        }
        dft<r>(x);
#pragma unroll
        for (int q = 0; q < r; ++q) v[b + q * nb] = x[q];
    }
    if constexpr (lp2) {
        // the last pass, radix 2 at width N/2: lane l (l < 32) holds y[k],
        // lane l + 32 holds y[k + N/2], for k = lp2_pos(t, q); the lower lane
        // keeps y[k] + W y[k + N/2], the upper one y[k] - W y[k + N/2], with
        // W = W_N^k, both at their own positions
        static_assert(sizeof(CX) == 8, "lane-pair radix 2: FP32");
        constexpr int PL = G::NP - 1, wl = 1 << G::lw(PL);
        const bool upper = (t >> 5) & 1;
        const CX* lo = ltw + TI::off(PL);
        const CX* hi = ltw + TI::off(PL) + TI::lo_n(PL);
#pragma unroll
        for (int q = 0; q < G::R; ++q) {
            const auto sx = __builtin_amdgcn_permlane32_swap(__float_as_uint(v[q].x), __float_as_uint(v[q].x),
                                                             false, false);
            const auto sy = __builtin_amdgcn_permlane32_swap(__float_as_uint(v[q].y), __float_as_uint(v[q].y),
                                                             false, false);
            const CX a = CxT<CX>::mk(__uint_as_float(sx[0]), __uint_as_float(sy[0]));
            const CX bb = CxT<CX>::mk(__uint_as_float(sx[1]), __uint_as_float(sy[1]));
            const int k = lp2_pos<G>(t, q) & (wl - 1);
            static_assert(wl > DMAX, "lane-pair radix 2: factored twiddles");
            const CX wb = cmul(bb, cmul(lo[k & 15], hi[k >> 4]));
            v[q] = upper ? csub(a, wb) : cadd(a, wb);
        }
    } else if constexpr (!last) {
        __syncthreads();
#pragma unroll
        for (int b = 0; b < nb; ++b) {
            const int i = t + G::T * b;
            const int base = ((i >> lw) << (lw + lr)) + (i & (w - 1));
#pragma unroll
            for (int q = 0; q < r; ++q) slds[G::pad(base + (q << lw))] = v[b + q * nb];
        }
        __syncthreads();
        // the next pass's inputs x[i + (N / r') m] of its butterfly i
        const int tn = (G::LP2 && P + 1 == G::NP - 2) ? lp2_index<G>(t) : t;
#pragma unroll
        for (int j = 0; j < G::R; ++j) v[j] = slds[G::pad(tn + G::T * j)];
        run_pass_img<G, P + 1, DMAX, CX>(v, slds, ltw, t);
    }
}

// Persistent streaming C2C / R2C kernel (direction BWD fixed at compile time)
// for geometries whose direct I/O is
// coalesced (T * element size >= 128 bytes or one transform per workgroup).
// Each workgroup loops over groups of TPW transforms; while it computes one
// group the next group's input is already in flight (register prefetch with
// non-temporal loads), and twiddles come from an LDS image loaded once, so no
// global load waits behind the prefetch.  Same arithmetic as fft_lds_kernel.
// FRAMED (R2C / R2C_SQR): the STFT front-end, as in fft_lds_kernel: frame f is
// signal[f hop + n] (in0) times window[n] (in1, or none); the window sits in
// LDS, the frames are read through the cache (frames overlap when hop < N;
// measured equal or faster than non-temporal loads), and R2C_SQR applies the
// dB option.
#ifndef ZL_STREAM_MIN_WAVES
#define ZL_STREAM_MIN_WAVES 1
#endif
// RTWG (R2C, R2C_SQR): the RFFT twiddles are read from global memory (the
// L2 holds the table) instead of an LDS copy, for sizes whose LDS copy would
// not fit beside the transform tile or would halve the resident workgroups.
template <class G, int KIND, int LIN, int LOUT, int BWD, int FRAMED = 0, class CX = float2, int RTWG = 0>
__global__ void __launch_bounds__(G::WG) __attribute__((amdgpu_waves_per_eu(ZL_STREAM_MIN_WAVES, 8)))
fft_stream_kernel(KArgs a) {
    static_assert(KIND == C2C || KIND == R2C || KIND == R2C_SQR || KIND == C2R,
                  "streaming kernel handles C2C, R2C and C2R");
    static_assert(!FRAMED || KIND != C2C, "framed input is real");
    static_assert(sizeof(CX) == 8 || !FRAMED, "FP64 streaming kernel: not framed");
    using XT = CxT<CX>;
    using S = typename XT::S;
    constexpr int DMAX = cfg_stream_direct_max<CX>(G::logn, KIND);
    using TI = TwImage<G, DMAX>;
    __shared__ CX lds[G::TPW * G::LDSN];
    __shared__ CX ltw[TI::total() > 0 ? TI::total() : 1];
    static_assert(!RTWG || KIND != C2C, "RFFT twiddles: real kinds only");
    __shared__ CX lrtw[KIND == C2C || RTWG ? 1 : (G::N / 2 > 0 ? G::N / 2 : 1)];
    const int tid = threadIdx.x;
    const int slot = tid / G::T;
    const int t = tid - slot * G::T;
    CX* slds = lds + slot * G::LDSN;
    constexpr int M = G::N;
    constexpr int lin = (KIND == C2C) ? LIN : AOS;
    constexpr bool bwd = (KIND == C2C) && BWD != 0;  // direction fixed per instantiation
    const long long ngroups = (a.batch + G::TPW - 1) / G::TPW;
    static_assert(KIND != C2R || (!FRAMED && G::R >= 2), "C2R: plain frames of whole bin pairs");

    // LDS tables (twiddle image, RFFT twiddles, STFT window): all their loads
    // are issued first, then the first group's prefetch, then the LDS writes,
    // so the kernel's start waits for one memory round trip in all, not one
    // per row of each table (measured: the 4096-frame STFT 7% faster)
    constexpr int NI = (TI::total() + G::WG - 1) / G::WG;
    constexpr int NR = KIND == C2C || RTWG ? 0 : (M / 2 + G::WG - 1) / G::WG;
    constexpr int NW = FRAMED ? (M + G::WG - 1) / G::WG : 0;
    CX im[NI > 0 ? NI : 1], rr[NR > 0 ? NR : 1];
    float2 ww[NW > 0 ? NW : 1];
#pragma unroll
    for (int i = 0; i < NI; ++i) {
        const int e = tid + G::WG * i;
        if (e < TI::total()) im[i] = static_cast<const CX*>(a.tw)[e];
    }
#pragma unroll
    for (int i = 0; i < NR; ++i) {
        const int e = tid + G::WG * i;
        if (e < M / 2) rr[i] = static_cast<const CX*>(a.rtw)[e];
    }
    if constexpr (FRAMED) {
        // the window as M pairs (1 without a window: x * 1 = x)
        const float* w = static_cast<const float*>(a.in1);
#pragma unroll
        for (int i = 0; i < NW; ++i) {
            const int e = tid + G::WG * i;
            if (e < M) ww[i] = w ? make_float2(w[2 * e], w[2 * e + 1]) : make_float2(1.f, 1.f);
        }
    }

    // Lanes without a transform (past the batch end, or past the last group)
    // load from and store to the sink instead: every path of the loop then
    // issues the same memory operations, with no branch around them, so the
    // waits for the prefetch never wait for the stores issued after it.
    const void* sk0 = a.sink;
    const void* sk1 = reinterpret_cast<const float*>(a.sink) + (1 << 14);
    // prefetch of frame / transform f into p (the sink where !ok); a framed
    // lane reads floats [0, 2 M) of the sink, M <= 2^14 (the sink holds 2^16)
    // C2R: a frame is M + 1 bins in layout LIN; lane t prefetches the bin
    // pairs (k, M - k), k = 1 + t + T i for i < R / 2 (each k in 1..M/2 once),
    // into p[i] and p[R / 2 + i], and bins 0 and M into pe (every lane of the
    // frame alike)
    CX pe[KIND == C2R ? 2 : 1];
    auto prefetch = [&](CX (&p)[G::R], long long f, bool ok) {
        if constexpr (KIND == C2R) {
            const void* b0 = ok ? a.in0 : sk0;
            const void* b1 = ok ? a.in1 : sk1;
            const long long eo = ok ? f * (M + 1) : 0;
#pragma unroll
            for (int i = 0; i < G::R / 2; ++i) {
                const int k = 1 + t + G::T * i;
                p[i] = ld_c_nt<CX>(b0, b1, LIN, eo + k);
                p[G::R / 2 + i] = ld_c_nt<CX>(b0, b1, LIN, eo + (M - k));
            }
            pe[0] = ld_c_nt<CX>(b0, b1, LIN, eo);
            pe[1] = ld_c_nt<CX>(b0, b1, LIN, eo + M);
        } else if constexpr (FRAMED) {
            const float* x = static_cast<const float*>(ok ? a.in0 : sk0) + (ok ? f * a.hop : 0);
#pragma unroll
            for (int j = 0; j < G::R; ++j) {
                const int m = t + G::T * j;
                p[j] = make_float2(x[2 * m], x[2 * m + 1]);
            }
        } else {
            const void* b0 = ok ? a.in0 : sk0;
            const void* b1 = ok ? a.in1 : sk1;
            const long long eo = ok ? f * M : 0;
#pragma unroll
            for (int j = 0; j < G::R; ++j) p[j] = ld_c_nt<CX>(b0, b1, lin, eo + t + G::T * j);
        }
    };
    // the window in LDS (kept in registers instead it cost 32 VGPRs and a
    // third of the occupancy)
    __shared__ float2 lwin[FRAMED ? M : 1];
    CX p[G::R];
    long long g = blockIdx.x;
    {
        const long long f = g * G::TPW + slot;
        prefetch(p, f, g < ngroups && f < a.batch);
    }
#pragma unroll
    for (int i = 0; i < NI; ++i) {
        const int e = tid + G::WG * i;
        if (e < TI::total()) ltw[e] = im[i];
    }
#pragma unroll
    for (int i = 0; i < NR; ++i) {
        const int e = tid + G::WG * i;
        if (e < M / 2) lrtw[e] = rr[i];
    }
#pragma unroll
    for (int i = 0; i < NW; ++i) {
        const int e = tid + G::WG * i;
        if (e < M) lwin[e] = ww[i];
    }
    // One group: take the prefetched input, prefetch the next group, transform,
    // store.  The first call is peeled off the loop so the loop is entered
    // only from a copy of its own body: both ways into the loop header then
    // have the prefetch loads followed by the stores in flight, and the
    // header's wait covers the loads only.
    auto group = [&](long long g) {
        const long long f = g * G::TPW + slot;
        const bool active = f < a.batch;
        CX v[G::R];
        CX u[KIND == C2R ? G::R : 1], ue[KIND == C2R ? 2 : 1];
        if constexpr (KIND == C2R) {
#pragma unroll
            for (int j = 0; j < G::R; ++j) u[j] = p[j];
            ue[0] = pe[0];
            ue[1] = pe[1];
        } else {
#pragma unroll
            for (int j = 0; j < G::R; ++j) v[j] = bwd ? swapri(p[j]) : p[j];
        }
        {
            const long long fn = (g + gridDim.x) * G::TPW + slot;
            prefetch(p, fn, g + gridDim.x < ngroups && fn < a.batch);
        }
        __syncthreads();  // LDS image ready (first group) / previous group's LDS readers done
        if constexpr (KIND == C2R) {
            // pre-twiddle prologue of fft_lds_kernel (reference
            // src/common/zldsp_fft_common_execute.hpp:398-468) on the
            // prefetched bin pairs: Z written to LDS already swapped for the
            // backward CFFT; thread 0 of the frame writes Z[0] from bins 0, M
            const CX* rt = RTWG ? static_cast<const CX*>(a.rtw) : lrtw;
            if (t == 0)
                slds[0] = swapri(XT::mk((ue[0].x + ue[1].x) * S(0.5), (ue[0].x - ue[1].x) * S(0.5)));
#pragma unroll
            for (int i = 0; i < G::R / 2; ++i) {
                const int k = 1 + t + G::T * i;
                const CX x1 = u[i], x2 = u[G::R / 2 + i];
                const CX w = rt[k - 1];  // (wc, ws)
                const S si = x1.y + x2.y, dr = x1.x - x2.x;
                const S zr_m = XT::fma(dr, w.y, XT::fma(si, w.x, x2.x));
                const S zi_m = XT::fma(dr, w.x, XT::fma(-si, w.y, x2.y));
                if (k != M - k) {
                    const S zr_k = XT::fma(-dr, w.y, XT::fma(-si, w.x, x1.x));
                    const S zi_k = XT::fma(dr, w.x, XT::fma(-si, w.y, x1.y));
                    slds[G::pad(k)] = XT::mk(zi_k, zr_k);
                }
                slds[G::pad(M - k)] = XT::mk(zi_m, zr_m);
            }
            __syncthreads();
#pragma unroll
            for (int j = 0; j < G::R; ++j) v[j] = slds[G::pad(t + G::T * j)];
        }
        if constexpr (FRAMED) {
#pragma unroll
            for (int j = 0; j < G::R; ++j) {
                const float2 wj = lwin[t + G::T * j];
                v[j] = make_float2(v[j].x * wj.x, v[j].y * wj.y);
            }
        }
        if constexpr (G::NP > 0) run_pass_img<G, 0, DMAX>(v, slds, ltw, t);
        // output position of v[j]: t + T j, or the lane-pair geometry's lp2_pos;
        // C2R writes its N reals as M pairs, swapped back from the backward CFFT
        if constexpr (KIND == C2C || KIND == C2R) {
            constexpr int lout = KIND == C2C ? LOUT : AOS;
            constexpr bool sw = bwd || KIND == C2R;
            void* o0 = active ? a.out0 : a.sink;
            void* o1 = active ? a.out1 : const_cast<void*>(sk1);
            const long long eo = active ? f * M : 0;
            if constexpr (G::LP2) {
#pragma unroll
                for (int j = 0; j < G::R; ++j)
                    st_c_nt(o0, o1, lout, eo + lp2_pos<G>(t, j), sw ? swapri(v[j]) : v[j]);
            } else {
#pragma unroll
                for (int j = 0; j < G::R; ++j) st_c_nt(o0, o1, lout, eo + t + G::T * j, sw ? swapri(v[j]) : v[j]);
            }
        } else {
            // post-twiddle epilogue, the arithmetic of fft_lds_kernel's.  A
            // frame's M + 1 outputs start off a cache-line boundary; with
            // cached_out the stores go through the cache, which merges the
            // lines that neighbouring frames share (the host picks it by batch
            // size).  Every lane issues the same stores in straight-line code:
            // bins k and M - k for k = 1 + t + T i, i < R / 2 (each k in
            // 1..M/2 once), and DC and Nyquist, which every lane of the frame
            // computes from Z[0] and stores alike.  Bin M/2 gets the M - k
            // formula's value (the reference's last store) from both of its
            // stores.  Lanes past the batch store to the sink.  A loop of
            // run-time length or a store under a divergent branch here made
            // the next group's wait for its prefetched input drain all of
            // these stores first.
            __syncthreads();
            if constexpr (G::LP2) {
#pragma unroll
                for (int j = 0; j < G::R; ++j) slds[G::pad(lp2_pos<G>(t, j))] = v[j];
            } else {
#pragma unroll
                for (int j = 0; j < G::R; ++j) slds[G::pad(t + G::T * j)] = v[j];
            }
            __syncthreads();
            const long long fb = active ? f * (M + 1) : 0;
            void* const o0 = active ? a.out0 : a.sink;
            void* const o1 = active ? a.out1 : const_cast<void*>(sk1);
            // pw: squared magnitude -> output value (|X|^2, or its dB level in
            // the framed kernels); cached_tag: cached or non-temporal stores
            auto epilogue = [&](auto pw, auto cached_tag) {
                constexpr bool cached = decltype(cached_tag)::value;
                auto st2 = [&](CX* q, CX x) {
                    if constexpr (cached) *q = x; else st_nt(q, x);
                };
                auto st1 = [&](S* q, S x) {
                    if constexpr (cached) *q = x; else st_nt(q, x);
                };
                const CX z0 = slds[0];
                const S dc = z0.x + z0.y, ny = z0.x - z0.y;
                // bins k and M - k: put(k, X[k] or, for k = M / 2, X[M - k], X[M - k])
                auto bins = [&](auto put) {
#pragma unroll
                    for (int i = 0; i < G::R / 2; ++i) {
                        // t through an empty asm per bin pair: the compiler
                        // cannot hoist a pair's reads above the previous pair,
                        // so the pairs' values are not all live at once
                        int tt = t;
                        asm volatile("" : "+v"(tt));
                        const int k = 1 + tt + G::T * i;
                        const CX z1 = slds[G::pad(k)];
                        const CX z2 = slds[G::pad(M - k)];
                        CX w;
                        if constexpr (RTWG) w = static_cast<const CX*>(a.rtw)[k - 1]; else w = lrtw[k - 1];
                        const S si = z1.y + z2.y, dr = z1.x - z2.x;
                        const CX xm = XT::mk(XT::fma(dr, w.y, XT::fma(-si, w.x, z2.x)),
                                             XT::fma(-dr, w.x, XT::fma(-si, w.y, z2.y)));
                        const CX xk = XT::mk(XT::fma(-dr, w.y, XT::fma(si, w.x, z1.x)),
                                             XT::fma(-dr, w.x, XT::fma(-si, w.y, z1.y)));
                        put(k, k != M - k ? xk : xm, xm);
                    }
                };
                if constexpr (KIND == R2C && LOUT == AOS) {
                    CX* const o = static_cast<CX*>(o0) + fb;
                    st2(o, XT::mk(dc, S(0)));
                    st2(o + M, XT::mk(ny, S(0)));
                    bins([&](int k, CX xk, CX xm) {
                        st2(o + k, xk);
                        st2(o + (M - k), xm);
                    });
                } else if constexpr (KIND == R2C) {
                    S* const q0 = static_cast<S*>(o0) + fb;
                    S* const q1 = static_cast<S*>(o1) + fb;
                    st1(q0, dc);
                    st1(q1, S(0));
                    st1(q0 + M, ny);
                    st1(q1 + M, S(0));
                    bins([&](int k, CX xk, CX xm) {
                        st1(q0 + k, xk.x);
                        st1(q1 + k, xk.y);
                        st1(q0 + (M - k), xm.x);
                        st1(q1 + (M - k), xm.y);
                    });
                } else {
                    S* const q = static_cast<S*>(o0) + fb;
                    st1(q, pw(dc * dc));
                    st1(q + M, pw(ny * ny));
                    bins([&](int k, CX xk, CX xm) {
                        st1(q + k, pw(XT::fma(xk.x, xk.x, xk.y * xk.y)));
                        st1(q + (M - k), pw(XT::fma(xm.x, xm.x, xm.y * xm.y)));
                    });
                }
            };
            auto with_pw = [&](auto pw) {
                if (a.cached_out) epilogue(pw, std::true_type{}); else epilogue(pw, std::false_type{});
            };
            if (FRAMED && KIND == R2C_SQR && a.db)
                with_pw([&](S q) { return power_out(q, 1, a.db_floor); });
            else
                with_pw([](S q) { return q; });
        }
    };
    if (g < ngroups) {
        group(g);
        for (g += gridDim.x; g < ngroups; g += gridDim.x) group(g);
    }
}

// ---------------------------------------------------------------- large N

#ifndef ZL_MACRO_NT_MIN_ROW
#define ZL_MACRO_NT_MIN_ROW 128
#endif
#ifndef ZL_MACRO_NT
#define ZL_MACRO_NT 1
#endif
// RFFT twiddles of the fused real passes.  The first pass of a backward
// transform takes them with the data (measured on 2^22-2^23-point batches,
// orders 16-22: 1-6% faster than after the mirror exchange).  The last pass of
// a forward transform takes them right after its sub-FFT, ahead of the
// exchange's barriers: 0.3-6% faster than after the exchange, while taking
// them with the data or before the sub-FFT was 4-9% slower (registers held
// across the sub-FFT cost occupancy).

// Arguments of one HBM pass of a large transform.
struct MArgs {
    const void* in0;
    const void* in1;
    void* out0;
    void* out1;
    const void* tw;     // inner Stockham twiddles of the radix-L sub-FFT (geometry GL)
    const void* twa;    // macro twiddle table A[a] = W_{L w}^{a 2^h}
    const void* twb;    // macro twiddle table B[b] = W_{L w}^{b}, b < 2^h
                        // (all three of the kernel's complex type)
    long long batch;
    int logn;           // log2 N of the whole transform
    int logw;           // log2 width w of this pass (product of earlier radices)
    int logh;           // split of the two-level macro twiddle
    int backward;       // swap re/im at the user-side load (first pass) and store (last pass)
    const void* rtw;    // real kinds: RFFT twiddles (wc, ws) for k = 1..N/2, at [k - 1]
    int db;             // R2C_SQR: write 10 log10(max(|X|^2, db_floor)) instead of |X|^2
    float db_floor;
    // first pass of an STFT (LIN = FRAMED_IN): transform b is the 2N reals
    // signal[b hop + r] (in0 = signal), times window[r] (null: none)
    long long hop;
    const void* window;
    // FP32 real kinds: RFFT twiddles from FP64 factors instead of the table,
    // double2 W_{2N}^e = (cos, -sin): A[e >> rtwh] at rtwf, B[e & (2^rtwh - 1)]
    // at rtwf + rtwb, and the steps W^{N/16} at rtwf + rtws, W^{N/32} at
    // rtwf + rtws + 2
    const void* rtwf;
    int rtwh;
    int rtwb;
    int rtws;
};

// One pass of a multi-pass Stockham DIT transform of N = 2^logn points:
// radix L = GL::N, width w, N/L butterflies i = (j, k), k = i mod w.
// Butterfly i reads x[i + m N/L] (m < L), multiplies by W_{L w}^{m k}, takes a
// length-L DFT in LDS and writes y_q to out[j L w + k + q w].  A workgroup owns
// C butterflies of one transform ("columns"): lanes run over columns first so
// every global access covers consecutive elements.
// FIRST: reads the user layout LIN (w = 1, no macro twiddle) and writes its
// output columns (L consecutive elements each) through LDS; LAST: writes the
// user layout LOUT.  Intermediate buffers are AoS float2.
// NT: the first pass's input / the last pass's output is the caller's buffer,
// touched once (non-temporal access); 0 for an internal buffer that the next
// kernel reads back (the STFT's frames).
//
// RK, the real transforms (an N-point CFFT inside a 2N-point RFFT, reference
// src/common/zldsp_fft_common_execute.hpp:317-552):
//   R2C / R2C_SQR (last pass): the RFFT post-twiddle (or |X|^2) is applied to
//     the pass's outputs, which go out as the N + 1 bins of each frame (layout
//     LOUT; squared magnitudes: one real each);
//   C2R (first pass): the pass reads the caller's N + 1 bins (layout LIN) and
//     applies the pre-twiddle before its sub-FFT.
// Both need element k with element N - k.  With cols = N / L columns, element
// k = i + m cols (column i, row m) mirrors to N - k = (cols - i) + (L - 1 - m)
// cols, and column 0's row m to row (L - m) mod L of itself.  A workgroup of
// these kinds therefore owns C/2 columns a0 .. a0 + C/2 - 1 (lanes c < C/2)
// and their mirror columns cols - a0 - c' (lanes C/2 + c'); in the first
// workgroup lane C/2 takes column cols/2, which mirrors onto itself as column
// 0 does.  Each lane exchanges its column with the mirror lane's through LDS.
// Per element, the formula is the one the reference's pair loop uses for that
// element (k < N - k: its "k" output, otherwise its "N - k" output), with the
// twiddle of min(k, N - k), so every bin is rounded as in the one-kernel path.
template <class GL, int C, int FIRST, int LAST, int LIN, int LOUT, class CX = float2, int NT = 1, int RK = C2C>
__global__ void __launch_bounds__(C * GL::T) fft_macro_pass(MArgs a) {
    static_assert(RK == C2C || (RK == C2R && FIRST && !LAST) || ((RK == R2C || RK == R2C_SQR) && LAST && !FIRST),
                  "real kinds fuse into the first (C2R) or last (R2C) pass");
    using T = CxT<CX>;
    using S = typename T::S;
    constexpr bool PAIR = RK != C2C;
    const CX* twa = static_cast<const CX*>(a.twa);
    const CX* twb = static_cast<const CX*>(a.twb);
    __shared__ CX lds[C * GL::LDSN];
    // Up to radix 128 the inner pass's twiddles are copied to LDS first.
    // Read from global memory after the first LDS exchange, their loads
    // queue behind the transform's data loads (measured, 4 Mi-point batches:
    // orders 15, 21 and 22, whose passes are radix 128 and 256, 3-8% faster
    // with the radix-128 tables in LDS; radix 512 and 1024 (orders 19 and 20)
    // 8-10% slower, radix 256 unchanged).
    constexpr bool lds_tw = GL::logn <= 7 && GL::tw_total() > 0;
    __shared__ CX ltw[lds_tw ? GL::tw_total() : 1];
    constexpr int L = GL::N, TT = GL::T, R = GL::R, WG = C * TT, H = C / 2;
    const int tid = threadIdx.x;
    constexpr int NTR = lds_tw ? (GL::tw_total() + WG - 1) / WG : 0;
    CX twr[NTR > 0 ? NTR : 1];
#pragma unroll
    for (int r = 0; r < NTR; ++r) {
        const int e = tid + WG * r;
        if (e < GL::tw_total()) twr[r] = static_cast<const CX*>(a.tw)[e];
    }
    const int c = tid % C;
    const int t = tid / C;
    const long long n = 1ll << a.logn;
    const long long cols = n >> GL::logn;  // butterflies per transform
    const long long cblocks = cols / C;
    const long long bid = xcd_block(blockIdx.x, gridDim.x);
    const long long b = bid / cblocks;
    const long long bi = bid - b * cblocks;
    // column of LDS slot / lane s
    auto column = [&](int s) -> long long {
        if constexpr (PAIR) {
            if (s < H) return bi * H + s;
            return bi == 0 && s == H ? cols / 2 : cols - bi * H - (s - H);
        } else {
            return bi * C + s;
        }
    };
    const long long i = column(c);
    // frame bases: the caller's bins hold N + 1 elements per frame
    const long long base = b * n;
    const long long ibase = (FIRST && RK == C2R) ? b * (n + 1) : base;
    constexpr int lin = FIRST ? LIN : AOS;
    constexpr int lout = LAST ? LOUT : AOS;
    CX* slds = lds + c * GL::LDSN;
    // the mirror lane's slot: the other half of the workgroup, or itself for
    // the self-mirrored columns 0 and cols/2
    CX* mlds = lds + (PAIR ? (bi == 0 && (c == 0 || c == H) ? c : (c < H ? c + H : c - H)) : c) * GL::LDSN;

    // The user's input (first pass) and output (last pass) are touched once:
    // non-temporal, so they do not evict the intermediate buffer, which the
    // next pass reads back from the cache (measured: the 4 Mi-point SoA
    // transforms of orders 16-22 gain 5-25%).
    constexpr bool nt_in = FIRST && NT && ZL_MACRO_NT;
    // A row of the last pass's output is C consecutive elements; below a
    // whole 128-byte line (SoA at 16 columns: 64 bytes per array) the
    // non-temporal stores reach memory as partial lines: measured 1.27x the
    // output bytes in WRITE_SIZE for the radix-1024 SoA pass.  Such rows go
    // through the cache, which merges the halves of a line that neighbouring
    // column blocks (renumbered onto one XCD by xcd_block) write.  The bins
    // of the real kinds start off a line boundary in every frame but the
    // first and also go through the cache (measured: non-temporal bins 4-11%
    // slower).
    constexpr int row_bytes = C * ((LAST && LOUT == SOA) ? (int)sizeof(CX) / 2 : (int)sizeof(CX));
    constexpr bool nt_out = LAST && NT && ZL_MACRO_NT && row_bytes >= ZL_MACRO_NT_MIN_ROW && !PAIR;
    CX v[R];
    if constexpr (lin == FRAMED_IN) {
        // element m of frame b: reals 2m and 2m + 1 of the frame, windowed as
        // the one-launch STFT kernel does; frames overlap, so cached loads
        const S* sig = static_cast<const S*>(a.in0) + b * a.hop;
        const S* win = static_cast<const S*>(a.window);
#pragma unroll
        for (int j = 0; j < R; ++j) {
            const long long m2 = 2 * (i + (long long)(t + TT * j) * cols);
            S x0 = sig[m2], x1 = sig[m2 + 1];
            if (win) {
                x0 *= win[m2];
                x1 *= win[m2 + 1];
            }
            v[j] = T::mk(x0, x1);
        }
    } else if constexpr (nt_in) {
#pragma unroll
        for (int j = 0; j < R; ++j) v[j] = ld_c_nt<CX>(a.in0, a.in1, lin, ibase + i + (long long)(t + TT * j) * cols);
    } else {
#pragma unroll
        for (int j = 0; j < R; ++j) v[j] = ld_c<CX>(a.in0, a.in1, lin, ibase + i + (long long)(t + TT * j) * cols);
    }
    // real kinds: both passes work on the pairs (k, N - k) whose first
    // element a lane holds, rows q < L/2 (j < R/2, k < N/2), with the RFFT
    // twiddle rtw[k - 1] (32-bit indices: N < 2^31)
    const CX* rtw = static_cast<const CX*>(a.rtw);
    const int ni = (int)n, ii = (int)i, ci = (int)cols;
    constexpr bool rtw_early = PAIR && FIRST;
    CX rw[rtw_early ? R / 2 : 1];
    // the twiddles of this lane's pairs, w[j] for k_j = k_0 + j N/R.  FP32:
    // W^{k_0} from the FP64 factors, then one FP64 complex product by W^{N/R}
    // per j, rounded to float as (0.5 cos, 0.5 sin + 0.5): the table's value
    // but for ties closer than the recurrence's error (R/2 products, about
    // 1e-15).  No lane reads the N/2-entry table, which for a single 2^22-point
    // transform is half the size of its input.  FP64: the table.
    auto twiddles = [&](CX (&w)[R / 2]) {
        if constexpr (sizeof(CX) == 8) {
            const double2* fa = static_cast<const double2*>(a.rtwf);
            const int k0 = ii + t * ci;
            const double2 wa = fa[k0 >> a.rtwh], wb = fa[a.rtwb + (k0 & ((1 << a.rtwh) - 1))];
            const double2 st = fa[a.rtws + (R == 16 ? 0 : 2)];  // W^{N/(2R)}: R = 16 or 32
            double c = wa.x * wb.x - wa.y * wb.y;
            double sn = wa.x * wb.y + wa.y * wb.x;  // -sin
#pragma unroll
            for (int j = 0; j < R / 2; ++j) {
                w[j] = T::mk((float)(0.5 * c), (float)(0.5 * -sn + 0.5));
                const double c2 = c * st.x - sn * st.y;
                sn = c * st.y + sn * st.x;
                c = c2;
            }
        } else {
#pragma unroll
            for (int j = 0; j < R / 2; ++j) {
                const int k = ii + (t + TT * j) * ci;
                w[j] = rtw[k > 0 ? k - 1 : 0];
            }
        }
    };
    // bin N/2: W^{N/2} of the real size 2N is (0, -1), so (wc, ws) = (0, 1)
    auto tw_half = [&]() -> CX {
        if constexpr (sizeof(CX) == 8) return T::mk(S(0), S(1));
        else return rtw[ni / 2 - 1];
    };
    if constexpr (rtw_early) twiddles(rw);
    if constexpr (FIRST && RK == C2R) {
        // RFFT pre-twiddle (reference src/common/zldsp_fft_common_execute.hpp:398-468):
        // Z[0] = ((X0 + XN) / 2, (X0 - XN) / 2) from the real parts of bins 0
        // and N; otherwise, for k = 1..N/2, with x1 = X[k], x2 = X[N - k]:
        //   Z[k]     = (x1.x - si wc - dr ws, x1.y - si ws + dr wc)
        //   Z[N - k] = (x2.x + si wc + dr ws, x2.y - si ws + dr wc)
        // si = x1.y + x2.y, dr = x1.x - x2.x, (wc, ws) = rtw[k - 1].
        // A lane computes both values of each pair whose first element it
        // holds (rows q < L/2): X[N - k] is read from the mirror lane's rows
        // q >= L/2 in LDS, and Z[N - k] is written back in its place.
        const S rn = ld_c<CX>(a.in0, a.in1, lin, ibase + n).x;  // bin N (one cache line per frame)
#pragma unroll
        for (int j = R / 2; j < R; ++j) slds[GL::pad(t + TT * j)] = v[j];
        __syncthreads();  // rows q >= L/2 of every column in LDS
#pragma unroll
        for (int j = 0; j < R / 2; ++j) {
            const int q = t + TT * j;
            const int k = ii + q * ci;
            const CX x1 = v[j];
            if (k == 0) {
                v[j] = T::mk((x1.x + rn) * S(0.5), (x1.x - rn) * S(0.5));
            } else {
                CX* pm = mlds + GL::pad(ii == 0 ? L - q : L - 1 - q);
                const CX x2 = *pm;
                const CX w = rw[rtw_early ? j : 0];
                const S si = x1.y + x2.y, dr = x1.x - x2.x;
                v[j] = T::mk(T::fma(-dr, w.y, T::fma(-si, w.x, x1.x)), T::fma(dr, w.x, T::fma(-si, w.y, x1.y)));
                *pm = T::mk(T::fma(dr, w.y, T::fma(si, w.x, x2.x)), T::fma(dr, w.x, T::fma(-si, w.y, x2.y)));
            }
        }
        // column 0's row L/2 (lane t = 0, j = R/2) is bin N/2, its own
        // mirror: the pair formula's second value, as in the reference
        CX zh = v[R / 2];
        if (ii == 0 && t == 0) {
            const CX w = tw_half();
            const S si = zh.y + zh.y, dr = zh.x - zh.x;
            zh = T::mk(T::fma(dr, w.y, T::fma(si, w.x, zh.x)), T::fma(dr, w.x, T::fma(-si, w.y, zh.y)));
        }
        __syncthreads();  // every pair's Z[N - k] in LDS
#pragma unroll
        for (int j = R / 2; j < R; ++j) v[j] = slds[GL::pad(t + TT * j)];
        if (ii == 0 && t == 0) v[R / 2] = zh;
        __syncthreads();  // LDS reads done before the sub-FFT reuses the LDS
    }
    if (FIRST && a.backward) {
#pragma unroll
        for (int j = 0; j < R; ++j) v[j] = swapri(v[j]);
    }
    if constexpr (!FIRST) {
        // two-level macro twiddle W_{L w}^{m k}: e = m k mod (L w), W^e = A[e >> h] B[e & (2^h - 1)]
        const long long k = i & ((1ll << a.logw) - 1);
        const long long emask = (1ll << (a.logw + GL::logn)) - 1;
        const long long hmask = (1ll << a.logh) - 1;
#pragma unroll
        for (int j = 0; j < R; ++j) {
            const long long m = t + TT * j;
            const long long e = (m * k) & emask;
            const CX w = cmul(twa[e >> a.logh], twb[e & hmask]);
            v[j] = cmul(v[j], w);
        }
    }
    if constexpr (lds_tw) {
#pragma unroll
        for (int r = 0; r < NTR; ++r) {
            const int e = tid + WG * r;
            if (e < GL::tw_total()) ltw[e] = twr[r];
        }
        __syncthreads();  // LDS twiddles ready
        run_passes<GL>(v, slds, ltw, t);  // v[j] = y_q with q = t + T j
    } else {
        run_passes<GL>(v, slds, static_cast<const CX*>(a.tw), t);
    }
    if (LAST && a.backward) {
#pragma unroll
        for (int j = 0; j < R; ++j) v[j] = swapri(v[j]);
    }
    if constexpr (FIRST) {
        // w = 1: column i's outputs are out[i L + q], L consecutive elements
        __syncthreads();
#pragma unroll
        for (int j = 0; j < R; ++j) slds[GL::pad(t + TT * j)] = v[j];
        __syncthreads();
#pragma unroll
        for (int j = 0; j < R; ++j) {
            const int e = tid + WG * j;
            const int cc = e / L, q = e - cc * L;
            st_c(a.out0, a.out1, lout, base + column(cc) * L + q, lds[cc * GL::LDSN + GL::pad(q)]);
        }
    } else if constexpr (PAIR) {
        // last pass of an RFFT: cols = w, element k = i + q w.  Post-twiddle
        // (reference src/common/zldsp_fft_common_execute.hpp:317-385, 480-552):
        // X[0] = (Z0.x + Z0.y, 0), X[N] = (Z0.x - Z0.y, 0); otherwise, for
        // k = 1..N/2, with z1 = Z[k], z2 = Z[N - k]:
        //   X[k]     = (z1.x + si wc - dr ws, z1.y - si ws - dr wc)
        //   X[N - k] = (z2.x - si wc + dr ws, z2.y - si ws - dr wc)
        // si = z1.y + z2.y, dr = z1.x - z2.x, (wc, ws) = rtw[k - 1].
        // A lane computes both bins of each pair whose first element it holds
        // (rows q < L/2, k < N/2): Z[k] is its own, Z[N - k] comes from the
        // mirror lane's rows q >= L/2, the only rows that go through LDS.
        CX w2[R / 2];  // twiddles of the pairs, in flight during the exchange
        twiddles(w2);
        __syncthreads();  // the sub-FFT's LDS reads done
#pragma unroll
        for (int j = R / 2; j < R; ++j) slds[GL::pad(t + TT * j)] = v[j];
        __syncthreads();
        // frame b's bins, addressed by 32-bit offsets (N < 2^31)
        const long long fb = b * (n + 1);
        void* o0 = RK == R2C && LOUT == AOS ? static_cast<void*>(static_cast<CX*>(a.out0) + fb)
                                            : static_cast<void*>(static_cast<S*>(a.out0) + fb);
        void* o1 = RK == R2C && LOUT == SOA ? static_cast<void*>(static_cast<S*>(a.out1) + fb) : nullptr;
        // bin k from its value (re, im); squared magnitudes take re^2 for the
        // real-valued DC and Nyquist bins
        auto put = [&](int k, S re, S im, bool real) {
            if constexpr (RK == R2C) {
                st_c(o0, o1, lout, k, T::mk(re, im));
            } else {
                static_cast<S*>(o0)[k] = power_out(real ? re * re : T::fma(re, re, im * im), a.db, a.db_floor);
            }
        };
        auto pair = [&](int k, CX z1, CX z2, CX w) {
            const S si = z1.y + z2.y, dr = z1.x - z2.x;
            put(k, T::fma(-dr, w.y, T::fma(si, w.x, z1.x)), T::fma(-dr, w.x, T::fma(-si, w.y, z1.y)), false);
            put(ni - k, T::fma(dr, w.y, T::fma(-si, w.x, z2.x)), T::fma(-dr, w.x, T::fma(-si, w.y, z2.y)), false);
        };
#pragma unroll
        for (int j = 0; j < R / 2; ++j) {
            const int q = t + TT * j;
            const int k = ii + q * ci;
            const CX z1 = v[j];
            if (k == 0) {
                put(0, z1.x + z1.y, S(0), true);
                put(ni, z1.x - z1.y, S(0), true);
            } else {
                const CX z2 = mlds[GL::pad(ii == 0 ? L - q : L - 1 - q)];
                pair(k, z1, z2, w2[j]);
            }
        }
        // column 0's row L/2 (lane t = 0, j = R/2) is bin N/2, its own mirror;
        // its value is the pair formula's second bin, as in the reference
        if (ii == 0 && t == 0) {
            const CX z = v[R / 2];
            const CX w = tw_half();
            const S si = z.y + z.y, dr = z.x - z.x;
            put(ni / 2, T::fma(dr, w.y, T::fma(-si, w.x, z.x)), T::fma(-dr, w.x, T::fma(-si, w.y, z.y)), false);
        }
    } else {
        const long long w = 1ll << a.logw;
        const long long ob = base + ((i >> a.logw) << (a.logw + GL::logn)) + (i & (w - 1));
        if constexpr (nt_out) {
#pragma unroll
            for (int j = 0; j < R; ++j) st_c_nt(a.out0, a.out1, lout, ob + (long long)(t + TT * j) * w, v[j]);
        } else {
#pragma unroll
            for (int j = 0; j < R; ++j) st_c(a.out0, a.out1, lout, ob + (long long)(t + TT * j) * w, v[j]);
        }
    }
}

}  // namespace zl
