// C-ABI of ragged batches: descriptor table, longest-first work queue,
// per-order twiddles and the captured graph.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <queue>
#include <utility>
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

// The twiddle table of the ragged kernel: lo[e] = W_{2^15}^e (e < 256) at
// pad16(e), then hi[a] = W_{2^15}^{256 a} (a < 128); FP64 on the GPU, rounded
// to FP32.  Pad slots hold 0.
__global__ void rg_gen_table(float2* out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= zl::kMtTotal) return;
    long long e = -1;
    if (i < zl::kMtLoPadded) {
        if ((i % 17) != 16) e = i - i / 17;  // pad16(e) = e + e / 16 = i
    } else {
        e = (long long)(i - zl::kMtLoPadded) << zl::kMtLoBits;
    }
    if (e < 0) {
        out[i] = make_float2(0.f, 0.f);
        return;
    }
    double s, c;
    sincospi(-2.0 * (double)e / (double)(1 << zl::kMtBits), &s, &c);
    out[i] = make_float2((float)c, (float)s);
}

// points per thread (log2) of CFFT order n without a tuned value: 16, or 32
// where 16 would need more than a workgroup's threads, 8 for order 3
int default_logr(int n) { return n <= 3 ? 3 : zl::ragged_logr_ok(n, 4) ? 4 : 5; }

// bytes one transform of the item's (corder, key) moves: read once, written once
long long item_bytes(int corder, int key) {
    const long long M = 1ll << corder;
    if (key < 4) return 16 * M;                  // C2C
    if (key < 6) return 8 * M + 8 * (M + 1);     // R2C
    if (key == 6) return 8 * M + 4 * (M + 1);    // R2C_SQR
    return 8 * (M + 1) + 8 * M;                  // C2R
}

}  // namespace

struct zlfft_ragged_s {
    int inject = zl::kInjectNone;  // fault injection (zl_read_inject)
    zl::RDesc* d_desc = nullptr;
    zl::RItem* d_items = nullptr;
    int* d_runs = nullptr;
    float2* d_mt = nullptr;
    char* d_sink = nullptr;
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
    if (r->d_runs) (void)hipFree(r->d_runs);
    if (r->d_mt) (void)hipFree(r->d_mt);
    if (r->d_sink) (void)hipFree(r->d_sink);
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
        if (logr) {
            lr[n] = logr[n];
            if (n >= kRaggedMinOrder && !ragged_logr_ok(n, lr[n]))
                return zl_fail(ZLFFT_ERR_INVALID_ARG, "points per thread 2^" + std::to_string(lr[n]) +
                                                          " not available for order " + std::to_string(n));
        } else {
            // a tuned value the kernel cannot run (a table from another
            // build) falls back to the default
            lr[n] = zl_tuned_ragged_logr(n);
            if (n >= kRaggedMinOrder && !ragged_logr_ok(n, lr[n])) lr[n] = default_logr(n);
        }
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
    // items of one (corder, key): descriptors sorted by order, then variant
    std::stable_sort(t.begin(), t.end(), [](const Tmp& a, const Tmp& b) {
        return a.corder != b.corder ? a.corder > b.corder : a.key < b.key;
    });
    // The kernel: a batch with order-14 transforms runs in 512-thread
    // workgroups (one per CU, every order), one without in 256-thread
    // workgroups (two per CU).  Splitting a batch between the two kernels was
    // measured slower on the 10,000-transform mix of orders 6-14 (141.8
    // against 129.0 us; as parallel graph branches 151.8): order 14 alone
    // leaves the large kernel's workgroups unevenly loaded.
    const int cls = count > 0 && t[0].corder >= kRaggedLargeOrder ? 1 : 0;  // t is sorted by order, largest first
    const int wg = cls ? kRaggedLargeWG : kRaggedSmallWG;
    std::vector<RDesc> hd(count + 1);
    std::vector<RItem> hi;
    for (size_t i = 0; i < count; ++i) hd[i] = t[i].d;
    for (size_t i = 0; i < count;) {
        const int n = t[i].corder, key = t[i].key;
        const int T = 1 << (n - lr[n]), tpw = wg / T;
        size_t j = i;
        while (j < count && (int)(j - i) < tpw && t[j].corder == n && t[j].key == key) ++j;
        hi.push_back(RItem{(int)i, (int)(j - i), n, key});
        i = j;
    }

    int dev = 0, cus = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (e != hipSuccess) return hipf(e, "device query");

    // Deal the items to the workgroups: largest first, each to the workgroup
    // with the least bytes so far (plus a fixed cost per item).  A
    // workgroup's run is ordered by points per thread (one loop each in the
    // kernel), then largest first.
    int per_cu = 0;
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, zl_ragged_kernel(cls), wg, 0);
    if (e != hipSuccess) return hipf(e, "occupancy query");
    const int grid = std::max(1, std::min(std::max(per_cu, 1) * cus, (int)hi.size()));
    constexpr long long kItemCost = 16384;
    std::vector<long long> cost(hi.size());
    std::vector<int> order(hi.size());
    for (size_t i = 0; i < hi.size(); ++i) {
        cost[i] = hi[i].count * item_bytes(hi[i].corder, hi[i].key) + kItemCost;
        order[i] = (int)i;
    }
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return cost[x] > cost[y]; });
    using Load = std::pair<long long, int>;
    std::priority_queue<Load, std::vector<Load>, std::greater<Load>> heap;
    for (int g = 0; g < grid; ++g) heap.push({0, g});
    std::vector<std::vector<int>> run(grid);
    for (int i : order) {
        Load l = heap.top();
        heap.pop();
        run[l.second].push_back(i);
        l.first += cost[i];
        heap.push(l);
    }
    std::vector<RItem> items;
    items.reserve(hi.size());
    std::vector<int> runs(4 * (size_t)grid);
    for (int g = 0; g < grid; ++g) {
        std::vector<int>& rg = run[g];
        std::stable_sort(rg.begin(), rg.end(), [&](int x, int y) {
            const int lx = lr[hi[x].corder], ly = lr[hi[y].corder];
            return lx != ly ? lx > ly : cost[x] > cost[y];
        });
        size_t k = 0;
        runs[4 * g] = (int)items.size();
        for (int L = kRaggedMaxLogr; L >= kRaggedMinLogr; --L) {
            while (k < rg.size() && lr[hi[rg[k]].corder] == L) items.push_back(hi[rg[k++]]);
            runs[4 * g + 1 + (kRaggedMaxLogr - L)] = (int)items.size();
        }
    }

    zlfft_ragged r = new (std::nothrow) zlfft_ragged_s();
    if (!r) return zl_fail(ZLFFT_ERR_ALLOC, "allocation failed");
    r->inject = zl_read_inject();
    auto cleanup = [&](int rc) {
        zlfft_ragged_destroy(r);
        return rc;
    };
    r->nitems = (int)items.size();
    e = hipStreamCreateWithFlags(&r->stream, hipStreamNonBlocking);
    if (e != hipSuccess) return cleanup(hipf(e, "hipStreamCreateWithFlags"));
    if (r->inject == kInjectAlloc || hipMalloc(&r->d_desc, sizeof(RDesc) * hd.size()) != hipSuccess ||
        hipMalloc(&r->d_items, sizeof(RItem) * std::max<size_t>(items.size(), 1)) != hipSuccess ||
        hipMalloc(&r->d_runs, sizeof(int) * runs.size()) != hipSuccess ||
        hipMalloc(&r->d_mt, sizeof(float2) * kMtTotal) != hipSuccess ||
        hipMalloc(&r->d_sink, kRaggedSinkBytes) != hipSuccess)
        return cleanup(zl_fail(ZLFFT_ERR_ALLOC, "ragged hipMalloc failed"));
    // the sink descriptor: lanes without a transform read and write there
    hd[count] = RDesc{r->d_sink, r->d_sink + kRaggedSinkHalf, r->d_sink, r->d_sink + kRaggedSinkHalf};
    e = hipMemsetAsync(r->d_sink, 0, kRaggedSinkBytes, r->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(r->d_desc, hd.data(), sizeof(RDesc) * hd.size(), hipMemcpyHostToDevice, r->stream);
    if (e == hipSuccess && !items.empty())
        e = hipMemcpyAsync(r->d_items, items.data(), sizeof(RItem) * items.size(), hipMemcpyHostToDevice, r->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(r->d_runs, runs.data(), sizeof(int) * runs.size(), hipMemcpyHostToDevice, r->stream);
    if (e != hipSuccess) return cleanup(hipf(e, "ragged setup copies"));
    rg_gen_table<<<(kMtTotal + 255) / 256, 256, 0, r->stream>>>(r->d_mt);
    e = hipStreamSynchronize(r->stream);
    if (e != hipSuccess) return cleanup(hipf(e, "ragged setup"));

    RArgs a{};
    a.desc = r->d_desc;
    a.items = r->d_items;
    a.runs = r->d_runs;
    a.mt = r->d_mt;
    a.sink = (int)count;

    // capture: the persistent kernel (each workgroup runs its own items, so
    // there is no queue to reset between launches)
    e = hipStreamBeginCapture(r->stream, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) return cleanup(hipf(e, "hipStreamBeginCapture"));
    void* kp[] = {&a};
    e = hipLaunchKernel(zl_ragged_kernel(cls), dim3(grid), dim3(wg), kp, 0, r->stream);
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
    RG_HIP(r->inject == zl::kInjectLaunch ? hipErrorLaunchFailure
                                          : hipGraphLaunch(r->exec, reinterpret_cast<hipStream_t>(stream)));
    return ZLFFT_OK;
}

}  // extern "C"
