// Ragged batches: many transforms of different orders, kinds, directions and
// layouts in one persistent launch.
//
// The host sorts the transforms into work items of one (CFFT order, variant):
// up to TPW = RWG / T transforms (T = 2^n / 16 threads each, 16 points per
// thread) that one workgroup of RWG threads handles together.  Items are
// queued longest first; each persistent workgroup takes the next item from a
// device counter until the queue is empty, so every workgroup ends when the
// queue does.  Pointers come from each transform's descriptor.
#pragma once

#include "zlfft_device.hpp"

namespace zl {

constexpr int RWG = 1024;          // threads per ragged workgroup
constexpr int kRaggedMinOrder = 3;  // CFFT orders handled by the ragged kernel
constexpr int kRaggedMaxOrder = 14;

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

struct RArgs {
    const RDesc* desc;
    const RItem* items;
    int nitems;
    unsigned* counter;             // next item; zeroed before each launch
    const float2* tw;              // Stockham twiddles of all orders, back to back
    const float2* rtw;             // RFFT twiddles of all orders, back to back
    long long tw_off[kRaggedMaxOrder + 1];
    long long rtw_off[kRaggedMaxOrder + 1];
    int logr[kRaggedMaxOrder + 1]; // points per thread (log2: 3 or 4) of each order
};

// Points per thread (log2) the ragged kernel supports for CFFT order n: 16
// for every order, 8 where 2^n / 8 threads fit one workgroup.
constexpr bool ragged_logr_ok(int n, int logr) {
    return (logr == 4 && n >= 4) || (logr == 3 && n >= 3 && (1 << (n - 3)) <= RWG);
}

// LDS of a ragged workgroup: TPW = RWG / T transforms of 2^n points
// (T = 2^n / R threads each) with one pad slot per 16, the largest over the
// orders and point counts handled.
constexpr int ragged_lds_elems() {
    int m = 0;
    for (int n = kRaggedMinOrder; n <= kRaggedMaxOrder; ++n)
        for (int logr = 3; logr <= 4; ++logr) {
            if (!ragged_logr_ok(n, logr)) continue;
            const int N = 1 << n, T = N >> logr, tpw = T >= RWG ? 1 : RWG / T;
            const int ldsn = N + (N >> 4) + 1;
            if (tpw * ldsn > m) m = tpw * ldsn;
        }
    return m;
}

__device__ __forceinline__ int pad16(int e) { return e + (e >> 4); }

// Stockham pass with radix r = 2^LR (compile time) and run-time width w = 2^lw
// on this thread's R registers (nb = R / r butterflies i = t + T b); tw
// points at the pass's twiddles W_{rw}^{mk} at [(m - 1) w + k].  Leaves the
// outputs in v (v[b + q nb] = y_q of butterfly b).
template <int R, int LR>
__device__ __forceinline__ void rg_butterflies(float2 (&v)[R], int t, int T, int lw, const float2* __restrict__ tw) {
    constexpr int r = 1 << LR, nb = R / r;
    const int w = 1 << lw;
#pragma unroll
    for (int b = 0; b < nb; ++b) {
        float2 x[r];
#pragma unroll
        for (int m = 0; m < r; ++m) x[m] = v[b + m * nb];
        if (lw > 0) {
            const int k = (t + T * b) & (w - 1);
#pragma unroll
            for (int m = 1; m < r; ++m) x[m] = cmul(x[m], tw[(m - 1) * w + k]);
        }
        dft<r>(x);
#pragma unroll
        for (int q = 0; q < r; ++q) v[b + q * nb] = x[q];
    }
}

// LDS exchange after a pass of radix 2^lr and width 2^lw: outputs to their
// Stockham positions, then reload v[j] = element t + T j.
template <int R, int LR>
__device__ __forceinline__ void rg_exchange(float2 (&v)[R], float2* slds, int t, int T, int lw) {
    constexpr int r = 1 << LR, nb = R / r;
    const int w = 1 << lw;
    __syncthreads();
#pragma unroll
    for (int b = 0; b < nb; ++b) {
        const int i = t + T * b;
        const int base = ((i >> lw) << (lw + LR)) + (i & (w - 1));
#pragma unroll
        for (int q = 0; q < r; ++q) slds[pad16(base + (q << lw))] = v[b + q * nb];
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < R; ++j) v[j] = slds[pad16(t + T * j)];
}

// One work item of a ragged batch: `count` transforms of CFFT order n with
// variant `key`, one per slot of T = 2^n / R threads.  One code path for
// every order: the pass list (first pass radix 2^(n - LOGR(NP-1)), then radix
// R) is walked at run time, so the persistent kernel holds one transform body
// per point count R instead of one per order.  Same arithmetic as
// fft_lds_kernel.
template <int LOGR>
__device__ __forceinline__ void ragged_item(const RDesc* __restrict__ desc, const RItem& it,
                                            const float2* __restrict__ tw, const float2* __restrict__ rtw,
                                            float2* lds) {
    constexpr int R = 1 << LOGR;
    const int n = it.corder;
    const int M = 1 << n;
    const int logt = n - LOGR;
    const int T = 1 << logt;
    const int tid = threadIdx.x;
    const int slot = tid >> logt;
    const int t = tid & (T - 1);
    const bool active = slot < it.count;
    const int ldsn = M + (M >> 4) + 1;
    float2* slds = lds + slot * ldsn;
    const int key = it.key;
    const int kind = key < 4 ? C2C : key < 6 ? R2C : key == 6 ? R2C_SQR : C2R;
    const int lay = key < 4 ? (key >> 1) : key < 6 ? key - 4 : key == 6 ? AOS : key - 7;
    const bool bwd = key < 4 ? (key & 1) != 0 : kind == C2R;
    RDesc d{};
    if (active) d = desc[it.first + slot];

    float2 v[R];
    if (kind == C2R) {
        if (active) {
            if (t == 0) {
                const float r0 = ld_c(d.in0, d.in1, lay, 0).x;
                const float rm = ld_c(d.in0, d.in1, lay, M).x;
                slds[0] = swapri(make_float2((r0 + rm) * 0.5f, (r0 - rm) * 0.5f));
            }
            for (int k = 1 + t; k <= M / 2; k += T) {
                const float2 x1 = ld_c(d.in0, d.in1, lay, k);
                const float2 x2 = ld_c(d.in0, d.in1, lay, M - k);
                const float2 w = rtw[k - 1];
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
        }
        __syncthreads();
#pragma unroll
        for (int j = 0; j < R; ++j) v[j] = slds[pad16(t + T * j)];
    } else {
        const int lin = kind == C2C ? lay : AOS;
        if (active) {
#pragma unroll
            for (int j = 0; j < R; ++j) {
                const float2 x = ld_c(d.in0, d.in1, lin, t + T * j);
                v[j] = bwd ? swapri(x) : x;
            }
        } else {
#pragma unroll
            for (int j = 0; j < R; ++j) v[j] = make_float2(0.f, 0.f);
        }
    }

    // passes: first radix 2^lr0, then radix R
    const int np = (n + LOGR - 1) / LOGR;
    const int lr0 = n - LOGR * (np - 1);
    switch (lr0) {
        case 1: rg_butterflies<R, 1>(v, t, T, 0, tw); break;
        case 2: rg_butterflies<R, 2>(v, t, T, 0, tw); break;
        case 3: rg_butterflies<R, 3>(v, t, T, 0, tw); break;
        default: rg_butterflies<R, LOGR>(v, t, T, 0, tw); break;
    }
    int lw = lr0;
    const float2* twp = tw;
    if (np > 1) {
        switch (lr0) {
            case 1: rg_exchange<R, 1>(v, slds, t, T, 0); break;
            case 2: rg_exchange<R, 2>(v, slds, t, T, 0); break;
            case 3: rg_exchange<R, 3>(v, slds, t, T, 0); break;
            default: rg_exchange<R, LOGR>(v, slds, t, T, 0); break;
        }
    }
    for (int p = 1; p < np; ++p) {
        rg_butterflies<R, LOGR>(v, t, T, lw, twp);
        twp += (R - 1) << lw;
        if (p + 1 < np) rg_exchange<R, LOGR>(v, slds, t, T, lw);
        lw += LOGR;
    }

    if (kind == R2C || kind == R2C_SQR) {
        __syncthreads();
#pragma unroll
        for (int j = 0; j < R; ++j) slds[pad16(t + T * j)] = v[j];
        __syncthreads();
        if (active) {
            if (t == 0) {
                const float2 z0 = slds[0];
                const float dc = z0.x + z0.y, ny = z0.x - z0.y;
                if (kind == R2C) {
                    st_c(d.out0, d.out1, lay, 0, make_float2(dc, 0.f));
                    st_c(d.out0, d.out1, lay, M, make_float2(ny, 0.f));
                } else {
                    float* o = reinterpret_cast<float*>(d.out0);
                    o[0] = dc * dc;
                    o[M] = ny * ny;
                }
            }
            for (int k = 1 + t; k <= M / 2; k += T) {
                const float2 z1 = slds[pad16(k)];
                const float2 z2 = slds[pad16(M - k)];
                const float2 w = rtw[k - 1];
                const float si = z1.y + z2.y, dr = z1.x - z2.x;
                const float xr_m = __builtin_fmaf(dr, w.y, __builtin_fmaf(-si, w.x, z2.x));
                const float xi_m = __builtin_fmaf(-dr, w.x, __builtin_fmaf(-si, w.y, z2.y));
                const float xr_k = __builtin_fmaf(-dr, w.y, __builtin_fmaf(si, w.x, z1.x));
                const float xi_k = __builtin_fmaf(-dr, w.x, __builtin_fmaf(-si, w.y, z1.y));
                if (kind == R2C) {
                    if (k != M - k) st_c(d.out0, d.out1, lay, k, make_float2(xr_k, xi_k));
                    st_c(d.out0, d.out1, lay, M - k, make_float2(xr_m, xi_m));
                } else {
                    float* o = reinterpret_cast<float*>(d.out0);
                    if (k != M - k) o[k] = __builtin_fmaf(xr_k, xr_k, xi_k * xi_k);
                    o[M - k] = __builtin_fmaf(xr_m, xr_m, xi_m * xi_m);
                }
            }
        }
    } else {
        const int lout = kind == C2C ? lay : AOS;
        if (active) {
#pragma unroll
            for (int j = 0; j < R; ++j) st_c(d.out0, d.out1, lout, t + T * j, bwd ? swapri(v[j]) : v[j]);
        }
    }
    __syncthreads();  // the LDS is free for the next item
}

}  // namespace zl
