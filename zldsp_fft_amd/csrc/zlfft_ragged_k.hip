// The ragged-batch kernel (one translation unit: it dispatches over every
// order and kind at run time).
#include "zlfft_ragged.hpp"

namespace zl {

// A workgroup's run of items at 2^LOGR points per thread, items [beg, end).
// With 8 or 16 points per thread, step i holds item i's input in p (loaded
// during step i - 1), loads item i + 2's descriptors and item i + 1's input,
// then transforms and stores item i.  The first step is peeled off the loop so
// the loop header is reached only from a copy of the loop body: both ways in
// then have the same loads and stores in flight, and the header's wait covers
// the input loads only.  With 32 points per thread the passes alone take most
// of the 256 VGPRs, so item i + 1's input is loaded after item i's passes,
// ahead of item i's stores (its descriptors one step ahead): the next step's
// wait for it then leaves those stores in flight.
template <int LOGR>
__device__ __forceinline__ void rg_run(const RArgs& a, int beg, int end, float2* lds, float2* mt,
                                       const float2 (&mt_reg)[2], bool fill_mt) {
    constexpr int R = 1 << LOGR;
    constexpr bool prefetch = LOGR <= 4;
    const int tid = threadIdx.x;
    // items past the run: no transform (every lane on the sink)
    auto item = [&](int i) {
        const RItem r = a.items[i < end ? i : end - 1];
        return i < end ? r : RItem{0, 0, LOGR, 0};
    };
    auto fill = [&]() {
        if (fill_mt) {
#pragma unroll
            for (int i = 0; i < 2; ++i) {
                const int e = tid + RWG * i;
                if (e < kMtTotal) mt[e] = mt_reg[i];
            }
            __syncthreads();
        }
    };
    RItem ic = item(beg);
    RItem in = item(beg + 1);
    RDesc dc = rg_desc<LOGR>(a, ic, tid);
    RDesc dn = rg_desc<LOGR>(a, in, tid);
    if constexpr (prefetch) {
        float2 p[R], px = make_float2(0.f, 0.f);
        rg_prefetch<LOGR>(p, px, ic, dc, tid);
        fill();
        auto step = [&](int i) __attribute__((always_inline)) {
            float2 v[R];
#pragma unroll
            for (int j = 0; j < R; ++j) v[j] = p[j];
            const float2 vx = px;
            const RItem i2 = item(i + 2);
            const RDesc d2 = rg_desc<LOGR>(a, i2, tid);
            rg_prefetch<LOGR>(p, px, in, dn, tid);
            rg_transform<LOGR>(v, vx, ic, dc, lds, mt, tid, []() {});
            ic = in;
            dc = dn;
            in = i2;
            dn = d2;
        };
        step(beg);
        for (int i = beg + 1; i < end; ++i) step(i);
    } else {
        float2 p[R], px = make_float2(0.f, 0.f);
        rg_prefetch<LOGR>(p, px, ic, dc, tid);
        fill();
        for (int i = beg; i < end; ++i) {
            float2 v[R];
#pragma unroll
            for (int j = 0; j < R; ++j) v[j] = p[j];
            const float2 vx = px;
            // the loads stay between the passes and the stores
            rg_transform<LOGR>(v, vx, ic, dc, lds, mt, tid, [&]() {
                __builtin_amdgcn_sched_barrier(0);
                rg_prefetch<LOGR>(p, px, in, dn, tid);
                __builtin_amdgcn_sched_barrier(0);
            });
            ic = in;
            dc = dn;
            in = item(i + 2);
            dn = rg_desc<LOGR>(a, in, tid);
        }
    }
}

__global__ void __launch_bounds__(RWG) fft_ragged_kernel(RArgs a) {
    __shared__ float2 lds[ragged_lds_elems()];
    __shared__ float2 mt[kMtTotal];
    const int b = blockIdx.x;
    const int r0 = a.runs[4 * b], r1 = a.runs[4 * b + 1], r2 = a.runs[4 * b + 2], r3 = a.runs[4 * b + 3];
    const int tid = threadIdx.x;
    // the twiddle table (kMtTotal <= 2 RWG entries) is loaded here and
    // written to LDS after the first run's first prefetch is issued
    float2 mt_reg[2];
#pragma unroll
    for (int i = 0; i < 2; ++i) {
        const int e = tid + RWG * i;
        mt_reg[i] = e < kMtTotal ? a.mt[e] : make_float2(0.f, 0.f);
    }
    bool fill = true;
    if (r0 < r1) {
        rg_run<5>(a, r0, r1, lds, mt, mt_reg, fill);
        fill = false;
    }
    if (r1 < r2) {
        rg_run<4>(a, r1, r2, lds, mt, mt_reg, fill);
        fill = false;
    }
    if (r2 < r3) rg_run<3>(a, r2, r3, lds, mt, mt_reg, fill);
}

}  // namespace zl

void* zl_ragged_kernel() { return reinterpret_cast<void*>(zl::fft_ragged_kernel); }
