// Ragged batches: many transforms of different orders, kinds, directions and
// layouts in one persistent launch.
//
// The host sorts the transforms into work items of one (CFFT order, variant):
// up to TPW = WG / T transforms (T = 2^n / R threads each, R = 8, 16 or 32
// points per thread, chosen per order) that one workgroup of WG threads
// handles together.  It deals the items to the workgroups longest first, each
// to the workgroup with the least work so far (by bytes moved), so each
// workgroup owns a fixed run of items and no queue is needed.  A workgroup streams through its run the way the
// batched streaming kernel streams through a batch: the input of the next item
// is loaded into registers while the current one is transformed and stored,
// and the descriptors two items ahead.  The Stockham and RFFT twiddles come
// from one small LDS table (W_{2^15}^e as two factors), so no global load sits
// between the prefetch and its use.
#pragma once

#include "zlfft_device.hpp"

namespace zl {

constexpr int kRaggedMinOrder = 3;  // CFFT orders handled by the ragged kernels
constexpr int kRaggedMaxOrder = 14;
// Two kernels.  A batch without order-14 transforms runs in 256-thread
// workgroups (4 waves of up to 256 VGPRs, LDS tiles of at most 73 KB: two
// workgroups per CU); a batch with them, whose 128 KiB transforms fill a CU's
// LDS, in 512-thread workgroups, one per CU.  Measured on 10,000 transforms
// of orders 6-13: the small kernel at 1.17x the rate of the large one (order
// 12 alone 1.28x, order 13 1.16x).  Every order but 14 must therefore fit
// 256 threads per transform: ragged_wg(n).
constexpr int kRaggedLargeOrder = 14;
constexpr int kRaggedSmallWG = 256, kRaggedLargeWG = 512;
constexpr int ragged_wg(int n) { return n >= kRaggedLargeOrder ? kRaggedLargeWG : kRaggedSmallWG; }

// One transform of a ragged batch (device copy of the user descriptor).
struct RDesc {
    const void* in0;
    const void* in1;
    void* out0;
    void* out1;
};

// Variant of a work item (kind, complex-side layout, direction):
//   0..3: C2C, layout (key >> 1) & 1, backward key & 1
//   4, 5: R2C, layout key - 4;   6: R2C_SQR;   7, 8: C2R, layout key - 7
constexpr int kRaggedKeys = 9;
constexpr int ragged_key(int kind, int layout, int backward) {
    return kind == C2C ? layout * 2 + backward : kind == R2C ? 4 + layout : kind == R2C_SQR ? 6 : 7 + layout;
}

// A work item: descriptors [first, first + count) share (corder, key).
struct RItem {
    int first;
    int count;
    int corder;
    int key;
};

// Twiddle table: W_{2^15}^e = exp(-2 pi i e / 2^15) for every e < 2^15 (2^15
// is the real size of the largest RFFT, twice the largest CFFT) as two
// factors, W^e = lo[e mod 256] * hi[e / 256].  lo is stored with one pad slot
// per 16 entries: lanes reading entries 16 apart then hit different banks.
constexpr int kMtBits = 15;
constexpr int kMtLoBits = 8;
constexpr int kMtLo = 1 << kMtLoBits;
constexpr int kMtLoPadded = kMtLo + (kMtLo >> 4);
constexpr int kMtHi = 1 << (kMtBits - kMtLoBits);
constexpr int kMtTotal = kMtLoPadded + kMtHi;

// bytes of the sink: inputs and outputs of lanes without a transform
// (in0 / out0 at offset 0, in1 / out1 at kRaggedSinkHalf)
constexpr long long kRaggedSinkHalf = 8ll * ((1 << kRaggedMaxOrder) + 64);
constexpr long long kRaggedSinkBytes = 2 * kRaggedSinkHalf;

struct RArgs {
    const RDesc* desc;    // descriptors; entry `sink` points at the sink
    const RItem* items;   // work items, each workgroup's run contiguous
    const int* runs;      // workgroup b: items [runs[4b], runs[4b+1]) at 32 points
                          // per thread, then up to runs[4b+2] at 16, then up to
                          // runs[4b+3] at 8
    const float2* mt;     // twiddle table (kMtTotal entries)
    int sink;             // index of the sink descriptor
};

// Points per thread (log2) the ragged kernels support for CFFT order n: 8
// for orders 3-7, 16 for orders 4-12, 32 for orders 8-14, with at most a
// workgroup's threads per transform (each pair is a pass chain of its own in
// a kernel, so the small-order / large-point pairs no order is tuned to are
// left out)
constexpr int kRaggedMinLogr = 3;
constexpr int kRaggedMaxLogr = 5;
constexpr bool ragged_logr_ok(int n, int logr) {
    return ((logr == 3 && n >= 3 && n <= 7) || (logr == 4 && n >= 4 && n <= 13) || (logr == 5 && n >= 8 && n <= 14)) &&
           n - logr >= 0 && (1 << (n - logr)) <= ragged_wg(n);
}

// LDS of a ragged workgroup of WG threads for orders [MINO, MAXO]:
// TPW = WG / T transforms of 2^n points (T = 2^n / R threads each) with one
// pad slot per 16, the largest over the orders and point counts handled.
template <int WG, int MINO, int MAXO>
constexpr int ragged_lds_elems() {
    int m = 0;
    for (int n = MINO; n <= MAXO; ++n)
        for (int logr = kRaggedMinLogr; logr <= kRaggedMaxLogr; ++logr) {
            if (!ragged_logr_ok(n, logr)) continue;
            const int N = 1 << n, T = N >> logr, tpw = WG / T;
            const int ldsn = N + (N >> 4) + 1;
            if (tpw * ldsn > m) m = tpw * ldsn;
        }
    return m;
}
static_assert(8ll * (ragged_lds_elems<kRaggedSmallWG, kRaggedMinOrder, kRaggedLargeOrder - 1>() + kMtTotal) <=
                  80 * 1024,
              "two small ragged workgroups per CU");
static_assert(8ll * (ragged_lds_elems<kRaggedLargeWG, kRaggedMinOrder, kRaggedMaxOrder>() + kMtTotal) <=
                  160 * 1024,
              "ragged LDS exceeds the CU's");

__device__ __forceinline__ int pad16(int e) { return e + (e >> 4); }

// t, opaque to the compiler: LDS addresses derived from it are recomputed at
// each exchange instead of being computed once and held in registers for the
// whole transform (R of them, with T known only at run time)
__device__ __forceinline__ int rg_opaque(int t) {
    asm volatile("" : "+v"(t));
    return t;
}

// Global-memory access through the descriptors' pointers.  They are loaded
// from memory, so the compiler cannot tell they address global memory and
// would emit flat loads, which count against the LDS wait counter as well:
// every LDS wait of the transform would then wait for the prefetch.  The
// casts make them global loads and stores.
#define ZL_GLOBAL __attribute__((address_space(1)))
__device__ __forceinline__ float2 rg_ld2(const void* p, int e) {  // complex e, non-temporal
    const zl_f2v t = __builtin_nontemporal_load((const ZL_GLOBAL zl_f2v*)p + e);
    return make_float2(t.x, t.y);
}
__device__ __forceinline__ float rg_ld1(const void* p, int e) {
    return __builtin_nontemporal_load((const ZL_GLOBAL float*)p + e);
}
__device__ __forceinline__ void rg_st2(void* p, int e, float2 v) {  // non-temporal
    const zl_f2v t = {v.x, v.y};
    __builtin_nontemporal_store(t, (ZL_GLOBAL zl_f2v*)p + e);
}
__device__ __forceinline__ void rg_st1(void* p, int e, float v) { __builtin_nontemporal_store(v, (ZL_GLOBAL float*)p + e); }
__device__ __forceinline__ ZL_GLOBAL float* rg_gf(void* p) { return (ZL_GLOBAL float*)p; }
__device__ __forceinline__ ZL_GLOBAL zl_f2v* rg_gf2(void* p) { return (ZL_GLOBAL zl_f2v*)p; }
__device__ __forceinline__ void rg_put(ZL_GLOBAL zl_f2v* p, float2 v) { *p = zl_f2v{v.x, v.y}; }

// W_{2^15}^e from the LDS table
__device__ __forceinline__ float2 rg_tw(const float2* mt, int e) {
    const int l = e & (kMtLo - 1);
    return cmul(mt[pad16(l)], mt[kMtLoPadded + (e >> kMtLoBits)]);
}

// Pass geometry of CFFT order NL at 2^LOGR points per thread, all compile
// time: T = 2^NL / R threads per transform; a first pass of radix 2^lr0, then
// radix 2^LP passes.  At 32 points per thread LP is 5 where radix-32 passes
// save a pass over radix 16 (orders 9, 10, 13, 14), else 4: the thread then
// runs two radix-16 butterflies per pass (measured per order, 2^24 points:
// order 14 0.093 vs 0.100 ms, orders 9-10 0.083 vs 0.091, order 12 0.101 vs
// 0.095 with radix 32 and 16).  Each (order, points) pair the kernel runs
// has its own pass chain, so LDS offsets and twiddle exponents are constants
// or shifts of t, as in the batched kernels.
template <int NL, int LOGR>
struct RGeo {
    static constexpr int R = 1 << LOGR;
    static constexpr int T = 1 << (NL - LOGR);
    static constexpr int LP = LOGR <= 4 ? LOGR : (NL + 4) / 5 < (NL + 3) / 4 ? 5 : 4;
    static constexpr int NP = (NL + LP - 1) / LP;
    static constexpr int LR0 = NL - LP * (NP - 1);
    static constexpr int lr(int p) { return p == 0 ? LR0 : LP; }
    static constexpr int lw(int p) { return p == 0 ? 0 : LR0 + LP * (p - 1); }
};

// Pass P of geometry G on this thread's R registers (nb = R / r butterflies
// i = t + T b), twiddle W_{r w}^{m k} = W_{2^15}^{(m k) << (15 - lr - lw)},
// then the LDS exchange to the next pass's inputs and the next pass.
template <class G, int P>
__device__ __forceinline__ void rgc_pass(float2 (&v)[G::R], float2* slds, const float2* mt, int t) {
    constexpr int LR = G::lr(P), r = 1 << LR, nb = G::R / r, lw = G::lw(P), w = 1 << lw;
#pragma unroll
    for (int b = 0; b < nb; ++b) {
        float2 x[r];
#pragma unroll
        for (int m = 0; m < r; ++m) x[m] = v[b + m * nb];
        if constexpr (P > 0) {
            constexpr int sh = kMtBits - LR - lw;
            const int k = (t + G::T * b) & (w - 1);
#pragma unroll
            for (int m = 1; m < r; ++m) x[m] = cmul(x[m], rg_tw(mt, (m * k) << sh));
        }
        dft<r>(x);
#pragma unroll
        for (int q = 0; q < r; ++q) v[b + q * nb] = x[q];
        // one butterfly at a time: interleaved, the nb butterflies'
        // temporaries would be live together on top of the points and the
        // prefetch
        if constexpr (nb > 1) __builtin_amdgcn_sched_barrier(0);
    }
    if constexpr (P + 1 < G::NP) {
        __syncthreads();
#pragma unroll
        for (int b = 0; b < nb; ++b) {
            const int i = t + G::T * b;
            const int base = ((i >> lw) << (lw + LR)) + (i & (w - 1));
#pragma unroll
            for (int q = 0; q < r; ++q) slds[pad16(base + (q << lw))] = v[b + q * nb];
        }
        __syncthreads();
#pragma unroll
        for (int j = 0; j < G::R; ++j) v[j] = slds[pad16(t + G::T * j)];
        rgc_pass<G, P + 1>(v, slds, mt, t);
    }
}

// the pass chain of order n (one compare per order in [NL, MAXO] the kernel
// runs at 2^LOGR points per thread)
template <int LOGR, int NL, int MAXO>
__device__ __forceinline__ void rgc_passes(int n, float2 (&v)[1 << LOGR], float2* slds, const float2* mt, int t) {
    if constexpr (NL <= MAXO) {
        if constexpr (ragged_logr_ok(NL, LOGR)) {
            if (n == NL) {
                rgc_pass<RGeo<NL, LOGR>, 0>(v, slds, mt, t);
                return;
            }
        }
        rgc_passes<LOGR, NL + 1, MAXO>(n, v, slds, mt, t);
    }
}

// Kind / layout / direction of a variant key.
struct RVariant {
    int kind, lay, bwd;
};
__device__ __forceinline__ RVariant rg_variant(int key) {
    RVariant r;
    r.kind = key < 4 ? C2C : key < 6 ? R2C : key == 6 ? R2C_SQR : C2R;
    r.lay = key < 4 ? (key >> 1) : key < 6 ? key - 4 : key == 6 ? AOS : key - 7;
    r.bwd = key < 4 ? (key & 1) : r.kind == C2R;
    return r;
}

// This lane's descriptor for item `it`: its slot's, or the sink's when the
// slot is past the item's count.
template <int LOGR>
__device__ __forceinline__ RDesc rg_desc(const RArgs& a, const RItem& it, int tid) {
    const int slot = tid >> (it.corder - LOGR);
    return a.desc[slot < it.count ? it.first + slot : a.sink];
}

// Loads this lane's input of item `it` into p (element t + T j of its
// transform, as complex: an RFFT's real input is read as 2^n pairs) and, for
// C2R, bin 2^n into px.
template <int LOGR>
__device__ __forceinline__ void rg_prefetch(float2 (&p)[1 << LOGR], float2& px, const RItem& it, const RDesc& d,
                                            int tid) {
    constexpr int R = 1 << LOGR;
    const int logt = it.corder - LOGR;
    const int T = 1 << logt;
    const int t = tid & (T - 1);
    const RVariant vr = rg_variant(it.key);
    const int lin = vr.kind == C2C || vr.kind == C2R ? vr.lay : AOS;
    if (lin == AOS) {
#pragma unroll
        for (int j = 0; j < R; ++j) p[j] = rg_ld2(d.in0, t + T * j);
        if (vr.kind == C2R) px = rg_ld2(d.in0, 1 << it.corder);
    } else {
#pragma unroll
        for (int j = 0; j < R; ++j) p[j] = make_float2(rg_ld1(d.in0, t + T * j), rg_ld1(d.in1, t + T * j));
        if (vr.kind == C2R) px = make_float2(rg_ld1(d.in0, 1 << it.corder), rg_ld1(d.in1, 1 << it.corder));
    }
}

// One work item: `count` transforms of CFFT order n with variant `key`, one
// per slot of T = 2^n / R threads, from the input in v (and px).  One code
// path for every order: the pass list (first pass radix 2^(n - LOGR(NP-1)),
// then radix R) is walked at run time.  Same arithmetic as fft_lds_kernel
// apart from the twiddles' rounding.  Lanes past the item's count run on the
// sink's data and store to the sink, so every lane issues the same stores.
// after_passes() runs once the passes are done, before the epilogue's stores:
// loads it issues are ahead of the stores in the memory counter, so waiting
// for them later does not wait for this item's stores.
template <int LOGR, int MINO, int MAXO, class AfterPasses>
__device__ __forceinline__ void rg_transform(float2 (&v)[1 << LOGR], float2 px, const RItem& it, const RDesc& d,
                                             float2* lds, const float2* mt, int tid, AfterPasses after_passes) {
    constexpr int R = 1 << LOGR;
    const int n = it.corder;
    const int M = 1 << n;
    const int logt = n - LOGR;
    const int T = 1 << logt;
    const int slot = tid >> logt;
    const int t = tid & (T - 1);
    const int ldsn = M + (M >> 4) + 1;
    float2* slds = lds + slot * ldsn;
    const RVariant vr = rg_variant(it.key);
    const int kind = vr.kind, lay = vr.lay;

    if (kind == C2R) {
        // pre-twiddle (reference src/common/zldsp_fft_common_execute.hpp:398-468)
        // in place in LDS: thread t owns the bin pairs (k, M - k)
        __syncthreads();
        {
            const int tw = rg_opaque(t);
#pragma unroll
            for (int j = 0; j < R; ++j) slds[pad16(tw + T * j)] = v[j];
        }
        __syncthreads();
        if (t == 0) {
            const float r0 = slds[0].x;
            const float rm = px.x;
            slds[0] = swapri(make_float2((r0 + rm) * 0.5f, (r0 - rm) * 0.5f));
        }
        const int rsh = kMtBits - 1 - n;  // W_{2M}^k = W_{2^15}^{k << rsh}
#pragma unroll
        for (int i = 0; i < R / 2; ++i) {
            const int k = 1 + t + T * i;  // 1..M/2, each once
            const float2 x1 = slds[pad16(k)];
            const float2 x2 = slds[pad16(M - k)];
            const float2 c = rg_tw(mt, k << rsh);
            const float2 w = make_float2(0.5f * c.x, 0.5f - 0.5f * c.y);
            const float si = x1.y + x2.y, dr = x1.x - x2.x;
            const float zr_m = __builtin_fmaf(dr, w.y, __builtin_fmaf(si, w.x, x2.x));
            const float zi_m = __builtin_fmaf(dr, w.x, __builtin_fmaf(-si, w.y, x2.y));
            if (k != M - k) {
                const float zr_k = __builtin_fmaf(-dr, w.y, __builtin_fmaf(-si, w.x, x1.x));
                const float zi_k = __builtin_fmaf(dr, w.x, __builtin_fmaf(-si, w.y, x1.y));
                slds[pad16(k)] = make_float2(zi_k, zr_k);
            }
            slds[pad16(M - k)] = make_float2(zi_m, zr_m);
        }
        __syncthreads();
        const int tr = rg_opaque(t);
#pragma unroll
        for (int j = 0; j < R; ++j) v[j] = slds[pad16(tr + T * j)];
    } else if (vr.bwd) {
#pragma unroll
        for (int j = 0; j < R; ++j) v[j] = swapri(v[j]);
    }

    rgc_passes<LOGR, MINO, MAXO>(n, v, slds, mt, t);
    after_passes();

    if (kind == R2C || kind == R2C_SQR) {
        // post-twiddle (reference src/common/zldsp_fft_common_execute.hpp:317-385,
        // 480-552); every lane issues the same stores (lanes past the item's
        // count through the sink descriptor)
        __syncthreads();
        {
            const int tw = rg_opaque(t);
#pragma unroll
            for (int j = 0; j < R; ++j) slds[pad16(tw + T * j)] = v[j];
        }
        __syncthreads();
        const float2 z0 = slds[0];
        const float dc = z0.x + z0.y, ny = z0.x - z0.y;
        const int rsh = kMtBits - 1 - n;
        // bins k and M - k for k = 1 + t + T i: put(k, X[k] or, for
        // k = M / 2, X[M - k], X[M - k]); bin M/2 thus gets the M - k
        // formula's value (the reference's last store) from both stores
        auto bins = [&](auto put) {
#pragma unroll
            for (int i = 0; i < R / 2; ++i) {
                const int k = 1 + t + T * i;
                const float2 z1 = slds[pad16(k)];
                const float2 z2 = slds[pad16(M - k)];
                const float2 c = rg_tw(mt, k << rsh);
                const float2 w = make_float2(0.5f * c.x, 0.5f - 0.5f * c.y);
                const float si = z1.y + z2.y, dr = z1.x - z2.x;
                const float2 xm = make_float2(__builtin_fmaf(dr, w.y, __builtin_fmaf(-si, w.x, z2.x)),
                                              __builtin_fmaf(-dr, w.x, __builtin_fmaf(-si, w.y, z2.y)));
                const float2 xk = make_float2(__builtin_fmaf(-dr, w.y, __builtin_fmaf(si, w.x, z1.x)),
                                              __builtin_fmaf(-dr, w.x, __builtin_fmaf(-si, w.y, z1.y)));
                put(k, k != M - k ? xk : xm, xm);
            }
        };
        // every lane of a transform computes DC and Nyquist from Z[0] and
        // stores them alike, so no store sits under a branch
        if (kind == R2C && lay == AOS) {
            ZL_GLOBAL zl_f2v* const o = rg_gf2(d.out0);
            rg_put(o, make_float2(dc, 0.f));
            rg_put(o + M, make_float2(ny, 0.f));
            bins([&](int k, float2 xk, float2 xm) {
                rg_put(o + k, xk);
                rg_put(o + (M - k), xm);
            });
        } else if (kind == R2C) {
            ZL_GLOBAL float* const o0 = rg_gf(d.out0);
            ZL_GLOBAL float* const o1 = rg_gf(d.out1);
            o0[0] = dc;
            o1[0] = 0.f;
            o0[M] = ny;
            o1[M] = 0.f;
            bins([&](int k, float2 xk, float2 xm) {
                o0[k] = xk.x;
                o1[k] = xk.y;
                o0[M - k] = xm.x;
                o1[M - k] = xm.y;
            });
        } else {
            ZL_GLOBAL float* const o = rg_gf(d.out0);
            o[0] = dc * dc;
            o[M] = ny * ny;
            bins([&](int k, float2 xk, float2 xm) {
                o[k] = __builtin_fmaf(xk.x, xk.x, xk.y * xk.y);
                o[M - k] = __builtin_fmaf(xm.x, xm.x, xm.y * xm.y);
            });
        }
    } else {
        // C2C, and C2R's real output as 2^n pairs
        const int lout = kind == C2C ? lay : AOS;
        if (vr.bwd) {
#pragma unroll
            for (int j = 0; j < R; ++j) v[j] = swapri(v[j]);
        }
        const int ts = rg_opaque(t);
        if (lout == AOS) {
#pragma unroll
            for (int j = 0; j < R; ++j) rg_st2(d.out0, ts + T * j, v[j]);
        } else {
#pragma unroll
            for (int j = 0; j < R; ++j) {
                rg_st1(d.out0, ts + T * j, v[j].x);
                rg_st1(d.out1, ts + T * j, v[j].y);
            }
        }
    }
}

}  // namespace zl
