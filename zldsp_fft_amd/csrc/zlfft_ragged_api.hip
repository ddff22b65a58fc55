// C-ABI of ragged batches: descriptor table, longest-first work queue,
// per-order twiddles and the captured graph.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>
#include <vector>

#include "../../include/zlfft.h"
#include "zlfft_ragged.hpp"
#include "zlfft_registry.hpp"

// shared with zlfft_api.hip
int zl_fail(int code, const std::string& msg);
int zl_tuned_ragged_logr(int order);

namespace {

int hipf(hipError_t e, const char* what) {
    return zl_fail(ZLFFT_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

#define RG_HIP(call)                                  \
    do {                                              \
        hipError_t e_ = (call);                       \
        if (e_ != hipSuccess) return hipf(e_, #call); \
    } while (0)

// Stockham twiddles of a 2^n-point transform with a first pass of radix
// 2^(n - logr(NP-1)) and radix-2^logr passes after it: pass p (width w) holds
// W_{r w}^{m k} at [(m - 1) w + k]; FP64 on the GPU, rounded to FP32.
__global__ void rg_gen_twiddles(float2* out, int n, int logr) {
    const int np = (n + logr - 1) / logr, lr0 = n - logr * (np - 1);
    long long off = 0;
    int lw = lr0;
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    for (int p = 1; p < np; ++p) {
        const long long cnt = ((1ll << logr) - 1) << lw;
        if (idx < cnt) {
            const long long w = 1ll << lw, m = idx / w + 1, k = idx % w, rw = 1ll << (lw + logr);
            const long long e = (m * k) & (rw - 1);
            double s, c;
            sincospi(-2.0 * (double)e / (double)rw, &s, &c);
            out[off + idx] = make_float2((float)c, (float)s);
        }
        off += cnt;
        lw += logr;
    }
}

// RFFT (wc, ws) for real size 2^(n+1), k = 1..2^n / 2
// (reference src/common/zldsp_fft_common_init.hpp:374-386)
__global__ void rg_gen_rfft_twiddles(float2* out, int n) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    const int cnt = (1 << n) / 2;
    if (idx >= cnt) return;
    const double a = (double)(idx + 1) * (2.0 / (double)(2ll << n));
    double s, c;
    sincospi(a, &s, &c);
    out[idx] = make_float2((float)(0.5 * c), (float)(0.5 * s + 0.5));
}

long long tw_count(int n, int logr) {
    const int np = (n + logr - 1) / logr, lr0 = n - logr * (np - 1);
    long long c = 0;
    int lw = lr0;
    for (int p = 1; p < np; ++p, lw += logr) c += ((1ll << logr) - 1) << lw;
    return c;
}

}  // namespace

struct zlfft_ragged_s {
    zl::RDesc* d_desc = nullptr;
    zl::RItem* d_items = nullptr;
    unsigned* d_counter = nullptr;
    float2* d_tw = nullptr;
    float2* d_rtw = nullptr;
    int nitems = 0;
    hipStream_t stream = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

extern "C" {

int zlfft_ragged_destroy(zlfft_ragged r) {
    if (!r) return zl_fail(ZLFFT_ERR_NULL, "ragged handle is NULL");
    if (r->stream) (void)hipStreamSynchronize(r->stream);
    if (r->exec) (void)hipGraphExecDestroy(r->exec);
    if (r->graph) (void)hipGraphDestroy(r->graph);
    if (r->d_desc) (void)hipFree(r->d_desc);
    if (r->d_items) (void)hipFree(r->d_items);
    if (r->d_counter) (void)hipFree(r->d_counter);
    if (r->d_tw) (void)hipFree(r->d_tw);
    if (r->d_rtw) (void)hipFree(r->d_rtw);
    if (r->stream) (void)hipStreamDestroy(r->stream);
    delete r;
    return ZLFFT_OK;
}

int zlfft_ragged_create(zlfft_ragged* out, const zlfft_ragged_desc* descs, size_t count) {
    return zlfft_ragged_create_config(out, descs, count, nullptr);
}

int zlfft_ragged_create_config(zlfft_ragged* out, const zlfft_ragged_desc* descs, size_t count, const int* logr) {
    using namespace zl;
    int lr[kRaggedMaxOrder + 1];
    for (int n = 0; n <= kRaggedMaxOrder; ++n) {
        lr[n] = logr ? logr[n] : zl_tuned_ragged_logr(n);
        if (n >= kRaggedMinOrder && !ragged_logr_ok(n, lr[n]))
            return zl_fail(ZLFFT_ERR_INVALID_ARG, "points per thread 2^" + std::to_string(lr[n]) +
                                                      " not available for order " + std::to_string(n));
    }
    if (!out) return zl_fail(ZLFFT_ERR_NULL, "ragged out-pointer is NULL");
    *out = nullptr;
    if (count && !descs) return zl_fail(ZLFFT_ERR_NULL, "descriptor array is NULL");
    if (count > (size_t)0x7fffffff) return zl_fail(ZLFFT_ERR_INVALID_ARG, "too many descriptors");
    // validate, and compute each transform's (corder, key)
    struct Tmp {
        int corder, key;
        RDesc d;
    };
    std::vector<Tmp> t(count);
    for (size_t i = 0; i < count; ++i) {
        const zlfft_ragged_desc& s = descs[i];
        const std::string at = " (descriptor " + std::to_string(i) + ")";
        if (s.layout != ZLFFT_AOS && s.layout != ZLFFT_SOA) return zl_fail(ZLFFT_ERR_INVALID_ARG, "bad layout" + at);
        int kind, corder;
        if (s.kind == ZLFFT_KIND_CFFT) {
            if (s.direction != ZLFFT_FORWARD && s.direction != ZLFFT_BACKWARD)
                return zl_fail(ZLFFT_ERR_INVALID_ARG, "bad CFFT direction" + at);
            kind = C2C;
            corder = s.order;
        } else if (s.kind == ZLFFT_KIND_RFFT) {
            kind = s.direction == ZLFFT_FORWARD ? R2C
                 : s.direction == ZLFFT_FORWARD_SQR_MAG ? R2C_SQR
                 : s.direction == ZLFFT_BACKWARD ? C2R : -1;
            if (kind < 0) return zl_fail(ZLFFT_ERR_INVALID_ARG, "bad RFFT direction" + at);
            corder = s.order - 1;
        } else {
            return zl_fail(ZLFFT_ERR_INVALID_ARG, "unknown kind" + at);
        }
        if (corder < kRaggedMinOrder || corder > kRaggedMaxOrder)
            return zl_fail(ZLFFT_ERR_INVALID_ORDER, "order outside the ragged range" + at);
        const bool soa_in = (kind == C2C || kind == C2R) && s.layout == ZLFFT_SOA;
        const bool soa_out = (kind == C2C || kind == R2C) && s.layout == ZLFFT_SOA;
        if (!s.in0 || !s.out0 || (soa_in && !s.in1) || (soa_out && !s.out1))
            return zl_fail(ZLFFT_ERR_NULL, "NULL data pointer" + at);
        t[i].corder = corder;
        t[i].key = ragged_key(kind, s.layout, s.direction == ZLFFT_BACKWARD ? 1 : 0);
        t[i].d = RDesc{s.in0, s.in1, s.out0, s.out1};
    }
    // longest transforms first, then by variant: items of one (corder, key)
    std::stable_sort(t.begin(), t.end(), [](const Tmp& a, const Tmp& b) {
        return a.corder != b.corder ? a.corder > b.corder : a.key < b.key;
    });
    std::vector<RDesc> hd(count);
    std::vector<RItem> hi;
    for (size_t i = 0; i < count; ++i) hd[i] = t[i].d;
    for (size_t i = 0; i < count;) {
        const int n = t[i].corder, key = t[i].key;
        const int T = 1 << (n - lr[n]), tpw = T >= RWG ? 1 : RWG / T;
        size_t j = i;
        while (j < count && (int)(j - i) < tpw && t[j].corder == n && t[j].key == key) ++j;
        hi.push_back(RItem{(int)i, (int)(j - i), n, key});
        i = j;
    }

    zlfft_ragged r = new (std::nothrow) zlfft_ragged_s();
    if (!r) return zl_fail(ZLFFT_ERR_ALLOC, "allocation failed");
    auto cleanup = [&](int rc) {
        zlfft_ragged_destroy(r);
        return rc;
    };
    r->nitems = (int)hi.size();
    hipError_t e = hipStreamCreateWithFlags(&r->stream, hipStreamNonBlocking);
    if (e != hipSuccess) return cleanup(hipf(e, "hipStreamCreateWithFlags"));
    RArgs a{};
    long long ntw = 0, nrtw = 0;
    for (int n = kRaggedMinOrder; n <= kRaggedMaxOrder; ++n) {
        a.tw_off[n] = ntw;
        a.rtw_off[n] = nrtw;
        a.logr[n] = lr[n];
        ntw += tw_count(n, lr[n]);
        nrtw += std::max(1ll, (1ll << n) / 2);
    }
    if (hipMalloc(&r->d_desc, sizeof(RDesc) * std::max<size_t>(count, 1)) != hipSuccess ||
        hipMalloc(&r->d_items, sizeof(RItem) * std::max<size_t>(hi.size(), 1)) != hipSuccess ||
        hipMalloc(&r->d_counter, sizeof(unsigned)) != hipSuccess ||
        hipMalloc(&r->d_tw, sizeof(float2) * (size_t)ntw) != hipSuccess ||
        hipMalloc(&r->d_rtw, sizeof(float2) * (size_t)nrtw) != hipSuccess)
        return cleanup(zl_fail(ZLFFT_ERR_ALLOC, "ragged hipMalloc failed"));
    if (count) {
        e = hipMemcpyAsync(r->d_desc, hd.data(), sizeof(RDesc) * count, hipMemcpyHostToDevice, r->stream);
        if (e == hipSuccess)
            e = hipMemcpyAsync(r->d_items, hi.data(), sizeof(RItem) * hi.size(), hipMemcpyHostToDevice, r->stream);
        if (e != hipSuccess) return cleanup(hipf(e, "hipMemcpyAsync"));
    }
    for (int n = kRaggedMinOrder; n <= kRaggedMaxOrder; ++n) {
        const long long c = tw_count(n, lr[n]);
        if (c > 0)
            rg_gen_twiddles<<<(unsigned)((c + 255) / 256), 256, 0, r->stream>>>(r->d_tw + a.tw_off[n], n, lr[n]);
        rg_gen_rfft_twiddles<<<((1 << n) / 2 + 255) / 256, 256, 0, r->stream>>>(r->d_rtw + a.rtw_off[n], n);
    }
    e = hipStreamSynchronize(r->stream);
    if (e != hipSuccess) return cleanup(hipf(e, "ragged setup"));

    a.desc = r->d_desc;
    a.items = r->d_items;
    a.nitems = r->nitems;
    a.counter = r->d_counter;
    a.tw = r->d_tw;
    a.rtw = r->d_rtw;
    int dev = 0, cus = 0, per_cu = 0;
    e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (e == hipSuccess) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, zl_ragged_kernel(), RWG, 0);
    if (e != hipSuccess) return cleanup(hipf(e, "occupancy query"));
    const int grid = std::max(1, std::min(per_cu * cus, r->nitems));

    // capture: reset the queue counter (a kernel, not a memset node), then the
    // persistent kernel
    e = hipStreamBeginCapture(r->stream, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) return cleanup(hipf(e, "hipStreamBeginCapture"));
    void* rp[] = {&r->d_counter};
    e = hipLaunchKernel(zl_ragged_reset_kernel(), dim3(1), dim3(64), rp, 0, r->stream);
    void* kp[] = {&a};
    if (e == hipSuccess) e = hipLaunchKernel(zl_ragged_kernel(), dim3(grid), dim3(RWG), kp, 0, r->stream);
    hipGraph_t g = nullptr;
    const hipError_t e2 = hipStreamEndCapture(r->stream, &g);
    r->graph = g;
    if (e != hipSuccess) return cleanup(hipf(e, "ragged capture"));
    if (e2 != hipSuccess) return cleanup(hipf(e2, "hipStreamEndCapture"));
    e = hipGraphInstantiate(&r->exec, r->graph, nullptr, nullptr, 0);
    if (e != hipSuccess) return cleanup(hipf(e, "hipGraphInstantiate"));
    *out = r;
    return ZLFFT_OK;
}

int zlfft_ragged_exec(zlfft_ragged r, void* stream) {
    if (!r) return zl_fail(ZLFFT_ERR_NULL, "ragged handle is NULL");
    if (r->nitems == 0) return ZLFFT_OK;
    RG_HIP(hipGraphLaunch(r->exec, reinterpret_cast<hipStream_t>(stream)));
    return ZLFFT_OK;
}

}  // extern "C"
