// The ragged-batch kernels (one translation unit: each dispatches over its
// orders and every kind at run time).
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
template <int LOGR, int WG, int MINO, int MAXO>
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
                const int e = tid + WG * i;
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
            rg_transform<LOGR, MINO, MAXO>(v, vx, ic, dc, lds, mt, tid, []() {});
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
            rg_transform<LOGR, MINO, MAXO>(v, vx, ic, dc, lds, mt, tid, [&]() {
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

// One kernel of a ragged batch: workgroups of WG threads, CFFT orders
// [MINO, MAXO].
template <int WG, int MINO, int MAXO>
__global__ void __launch_bounds__(WG) fft_ragged_kernel(RArgs a) {
    static_assert(kMtTotal <= 2 * WG, "the twiddle table loads in two rounds");
    __shared__ float2 lds[ragged_lds_elems<WG, MINO, MAXO>()];
    __shared__ float2 mt[kMtTotal];
    const int b = blockIdx.x;
    const int r0 = a.runs[4 * b], r1 = a.runs[4 * b + 1], r2 = a.runs[4 * b + 2], r3 = a.runs[4 * b + 3];
    const int tid = threadIdx.x;
    // the twiddle table (kMtTotal <= 2 WG entries) is loaded here and
    // written to LDS after the first run's first prefetch is issued
    float2 mt_reg[2];
#pragma unroll
    for (int i = 0; i < 2; ++i) {
        const int e = tid + WG * i;
        mt_reg[i] = e < kMtTotal ? a.mt[e] : make_float2(0.f, 0.f);
    }
    bool fill = true;
    // each run only where an order of this kernel takes its point count
    constexpr bool has5 = MAXO >= 8, has4 = MINO <= 12 && MAXO >= 4, has3 = MINO <= 7;
    if constexpr (has5) {
        if (r0 < r1) {
            rg_run<5, WG, MINO, MAXO>(a, r0, r1, lds, mt, mt_reg, fill);
            fill = false;
        }
    }
    if constexpr (has4) {
        if (r1 < r2) {
            rg_run<4, WG, MINO, MAXO>(a, r1, r2, lds, mt, mt_reg, fill);
            fill = false;
        }
    }
    if constexpr (has3) {
        if (r2 < r3) rg_run<3, WG, MINO, MAXO>(a, r2, r3, lds, mt, mt_reg, fill);
    }
}

}  // namespace zl

// The large kernel runs every order (a batch with order-14 transforms runs
// in it whole); the small one orders up to 13.
void* zl_ragged_kernel(int large) {
    using namespace zl;
    return large ? reinterpret_cast<void*>(fft_ragged_kernel<kRaggedLargeWG, kRaggedMinOrder, kRaggedMaxOrder>)
                 : reinterpret_cast<void*>(fft_ragged_kernel<kRaggedSmallWG, kRaggedMinOrder, kRaggedLargeOrder - 1>);
}
