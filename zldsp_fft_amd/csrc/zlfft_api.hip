// C-ABI of libzlfft: device plans (kernel choice + twiddles generated on the
// GPU in FP64), batched asynchronous execution on device pointers, and the
// synchronous single-transform calls used by the drop-in C++ classes.
#include <hip/hip_runtime.h>

#include <dlfcn.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/zlfft.h"
#include "zlfft_registry.hpp"

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}

}  // namespace

// shared with the other C-ABI translation units
int zl_fail(int code, const std::string& msg) { return fail(code, msg); }

// Fault injection, for the tests of the error paths: ZLFFT_INJECT, read when a
// plan or ragged batch is created.  "alloc": that creation's device
// allocations fail; "work_alloc": the plan is created, its later allocations
// (multi-pass workspace, host-pipeline staging) fail; "launch": its kernel
// launches fail.  Each failing call returns the code a real failure would
// (ZLFFT_ERR_ALLOC, ZLFFT_ERR_HIP) and no kernel is launched in a failing state.
int zl_read_inject() {
    const char* e = getenv("ZLFFT_INJECT");
    if (!e) return zl::kInjectNone;
    if (!strcmp(e, "alloc")) return zl::kInjectAlloc;
    if (!strcmp(e, "work_alloc")) return zl::kInjectWorkAlloc;
    if (!strcmp(e, "launch")) return zl::kInjectLaunch;
    return zl::kInjectNone;
}

// ---- tuning table -------------------------------------------------------
// Text file of "<key> <order> <value>" lines written by
// zldsp_fft_amd/autotune.py: $ZLFFT_TUNE_FILE, else zlfft_tune.txt next to
// this library.  Missing entries keep the built-in defaults.
namespace {
struct Tune {
    int ragged_logr[zl::kRaggedMaxOrderTune + 1];
};
Tune g_tune;
std::once_flag g_tune_once;

void load_tune() {
    for (int n = 0; n <= zl::kRaggedMaxOrderTune; ++n) g_tune.ragged_logr[n] = n <= 3 ? 3 : 4;
    std::string path;
    if (const char* env = getenv("ZLFFT_TUNE_FILE")) {
        path = env;
    } else {
        Dl_info info;
        if (dladdr(reinterpret_cast<void*>(&zlfft_version), &info) && info.dli_fname) {
            path = info.dli_fname;
            const size_t slash = path.rfind('/');
            path = (slash == std::string::npos ? std::string(".") : path.substr(0, slash)) + "/zlfft_tune.txt";
        }
    }
    FILE* f = path.empty() ? nullptr : fopen(path.c_str(), "r");
    if (!f) return;
    char key[64];
    int order = 0, value = 0;
    char line[256];
    while (fgets(line, sizeof(line), f)) {
        if (line[0] == '#') continue;
        if (sscanf(line, "%63s %d %d", key, &order, &value) != 3) continue;
        if (!strcmp(key, "ragged_logr") && order >= 0 && order <= zl::kRaggedMaxOrderTune &&
            value >= 3 && value <= 5)
            g_tune.ragged_logr[order] = value;
    }
    fclose(f);
}
}  // namespace

// ---- scratch of the allocation-free entry point ---------------------------
// The multi-pass intermediates of zlfft_cfft_exec_noalloc live in this array
// (one copy per GPU, allocated with the code object and held until the
// process ends).  Host code never uses its name as an address: the device
// address comes from hipGetSymbolAddress, once per device, with that device
// current, and is kept in g_scratch_ptr.
constexpr size_t kScratchBytes = 64ull << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices];
}  // namespace

int zl_tuned_ragged_logr(int order) {
    std::call_once(g_tune_once, load_tune);
    if (order < 0 || order > zl::kRaggedMaxOrderTune) return 4;
    return g_tune.ragged_logr[order];
}

namespace {

int hip_fail(hipError_t e, const char* what) {
    return fail(ZLFFT_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

#define ZL_HIP(call)                                   \
    do {                                               \
        hipError_t e_ = (call);                        \
        if (e_ != hipSuccess) return hip_fail(e_, #call); \
    } while (0)

zl::OrderTable g_tab[zl::kMaxLdsOrder + 1];
zl::MacroTable g_macro[zl::kMaxMacroLogL + 1];
zl::OrderTable g_tab64[zl::kMaxLdsOrder64 + 1];  // FP64
zl::MacroTable g_macro64[zl::kMaxMacroLogL + 1];
std::once_flag g_tab_once;

void fill_tables() {
    zl_fill_order_0(g_tab[0]);
    zl_fill_order_1(g_tab[1]);
    zl_fill_order_2(g_tab[2]);
    zl_fill_order_3(g_tab[3]);
    zl_fill_order_4(g_tab[4]);
    zl_fill_order_5(g_tab[5]);
    zl_fill_order_6(g_tab[6]);
    zl_fill_order_7(g_tab[7]);
    zl_fill_order_8(g_tab[8]);
    zl_fill_order_9(g_tab[9]);
    zl_fill_order_10(g_tab[10]);
    zl_fill_order_11(g_tab[11]);
    zl_fill_order_12(g_tab[12]);
    zl_fill_order_13(g_tab[13]);
    zl_fill_order_14(g_tab[14]);
    zl_fill_macro_6(g_macro[6]);
    zl_fill_macro_7(g_macro[7]);
    zl_fill_macro_8(g_macro[8]);
    zl_fill_macro_9(g_macro[9]);
    zl_fill_macro_10(g_macro[10]);
    zl_fill_order64_0(g_tab64[0]);
    zl_fill_order64_1(g_tab64[1]);
    zl_fill_order64_2(g_tab64[2]);
    zl_fill_order64_3(g_tab64[3]);
    zl_fill_order64_4(g_tab64[4]);
    zl_fill_order64_5(g_tab64[5]);
    zl_fill_order64_6(g_tab64[6]);
    zl_fill_order64_7(g_tab64[7]);
    zl_fill_order64_8(g_tab64[8]);
    zl_fill_order64_9(g_tab64[9]);
    zl_fill_order64_10(g_tab64[10]);
    zl_fill_order64_11(g_tab64[11]);
    zl_fill_order64_12(g_tab64[12]);
    zl_fill_order64_13(g_tab64[13]);
    zl_fill_macro64_6(g_macro64[6]);
    zl_fill_macro64_7(g_macro64[7]);
    zl_fill_macro64_8(g_macro64[8]);
    zl_fill_macro64_9(g_macro64[9]);
    zl_fill_macro64_10(g_macro64[10]);
}

// Stockham twiddles of one pass: out[(m - 1) w + k] = W_{r w}^{m k}, computed
// in FP64 (exact dyadic angle, sincospi) and rounded to FP32, as the reference
// computes its tables in double and casts (src/common/zldsp_fft_common_init.hpp:84-90).
template <class CX>
__global__ void gen_pass_twiddles(CX* out, int lr, int lw, int count) {
    using S = typename zl::CxT<CX>::S;
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= count) return;
    const int w = 1 << lw;
    const long long m = idx / w + 1, k = idx % w;
    const long long rw = 1ll << (lr + lw);
    const long long e = (m * k) & (rw - 1);
    double s, c;
    sincospi(-2.0 * (double)e / (double)rw, &s, &c);
    out[idx] = zl::CxT<CX>::mk((S)c, (S)s);
}

// LDS twiddle image of one Stockham pass (radix 2^lr, width 2^lw): the
// direct table for w <= 16, else lo[(m-1)16 + b] = W_{rw}^{m b} followed by
// hi[(m-1)(w/16) + a] = W_{rw}^{16 m a}; FP64 on the GPU, rounded to FP32.
template <class CX>
__global__ void gen_image_twiddles(CX* out, int lr, int lw, int lo_n, int hi_n) {
    using S = typename zl::CxT<CX>::S;
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= lo_n + hi_n) return;
    const long long w = 1ll << lw, rw = 1ll << (lr + lw);
    long long e;
    if (idx < lo_n) {
        const long long s = hi_n == 0 ? w : 16;  // direct table, or the low factor
        e = (idx / s + 1) * (idx % s);
    } else {
        const long long j = idx - lo_n, s = w / 16;
        e = 16 * (j / s + 1) * (j % s);
    }
    e &= rw - 1;
    double sn, cs;
    sincospi(-2.0 * (double)e / (double)rw, &sn, &cs);
    out[idx] = zl::CxT<CX>::mk((S)cs, (S)sn);
}

// MFMA kernels' twiddle table (layout at kMfmaTwTotal): W_rw^{m k} for rows
// m = 0..15, FP64 on the GPU, rounded to FP32, times `scale` (1, or 1/16 for
// the FP16 kernels: exact), the same values as the pass table's.
__global__ void gen_mfma_twiddles(float2* out, float scale) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= zl::kMfmaTwTotal) return;
    const bool p1 = idx < zl::kMfmaTwPass2;
    const int j = p1 ? idx : idx - zl::kMfmaTwPass2;
    const int lk = p1 ? 4 : 8;  // log2 of the row length
    const long long m = j >> lk, k = j & ((1 << lk) - 1);
    const long long rw = p1 ? 256 : 4096;
    const long long e = (m * k) & (rw - 1);
    double sn, cs;
    sincospi(-2.0 * (double)e / (double)rw, &sn, &cs);
    out[idx] = make_float2((float)cs * scale, (float)sn * scale);
}

// Macro twiddle table of an HBM pass: out[idx] = W_{2^log_lw}^{idx << shift},
// FP64 on the GPU, rounded to FP32.
template <class CX>
__global__ void gen_macro_twiddles(CX* out, int log_lw, int shift, int count) {
    using S = typename zl::CxT<CX>::S;
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= count) return;
    const long long mod = 1ll << log_lw;
    const long long e = ((long long)idx << shift) & (mod - 1);
    double s, c;
    sincospi(-2.0 * (double)e / (double)mod, &s, &c);
    out[idx] = zl::CxT<CX>::mk((S)c, (S)s);
}

// RFFT pre/post twiddles, reference src/common/zldsp_fft_common_init.hpp:374-386:
// a = k * (2 / N); wc = (F)(0.5 cospi(a)), ws = (F)(0.5 sinpi(a) + 0.5), k = 1..N/4.
template <class CX>
__global__ void gen_rfft_twiddles(CX* out, long long nreal, int count) {
    using S = typename zl::CxT<CX>::S;
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= count) return;
    const double a = (double)(idx + 1) * (2.0 / (double)nreal);
    double s, c;
    sincospi(a, &s, &c);
    out[idx] = zl::CxT<CX>::mk((S)(0.5 * c), (S)(0.5 * s + 0.5));
}

}  // namespace

struct zlfft_plan_s {
    int inject = zl::kInjectNone;  // fault injection (zl_read_inject)
    int kind = ZLFFT_KIND_CFFT;
    int precision = ZLFFT_PREC_F32;
    int order = 0;   // user order
    int corder = 0;  // inner CFFT order
    int device = 0;
    bool f64 = false;           // FP64 plan (CFFT<double> / RFFT<double>)
    size_t ss = 4;              // bytes per real scalar (4 or 8)
    const zl::OrderTable* tab = nullptr;  // LDS kernels (corder <= kMaxLdsOrder[64])
    void* d_tw = nullptr;       // pass twiddles, complex type of the plan
    float2* d_img = nullptr;    // LDS twiddle image of the streaming kernels of the plan's kind
    float2* d_sink = nullptr;   // zl::kSinkBytes of scratch for their idle lanes
    float2* d_twm = nullptr;    // MFMA kernels' twiddle table (gen_mfma_twiddles)
    int cus = 0;                // compute units of the device
    int stream_blocks[6][2][2] = {};  // resident workgroups per CU of each streaming kernel
    int mfma_blocks = 0;              // resident workgroups per CU of the persistent MFMA kernel
    void* d_rtw = nullptr;      // RFFT (wc, ws), complex type of the plan
    double2* d_rtwf = nullptr;  // FP32 multi-pass RFFT: FP64 twiddle factors and steps (MArgs::rtwf)
    int rtw_h = 0, rtw_b = 0, rtw_s = 0;
    int n_rtw = 0;
    // multi-pass transforms (corder > kMaxLdsOrder)
    bool macro = false;
    int mp = 0;                 // HBM passes
    int logl[4] = {}, logw[4] = {}, logh[4] = {};
    void* d_itw[4] = {};        // inner Stockham twiddles per pass
    void* d_twa[4] = {};        // macro twiddle tables per pass (p >= 1)
    void* d_twb[4] = {};
    void* d_work = nullptr;     // intermediate buffers, grown on demand
    size_t work_bytes = 0;
    // the workspace is used on one stream at a time: a launch on another
    // stream than the last one waits for the last one's work (work_ev)
    hipEvent_t work_ev = nullptr;
    hipStream_t work_stream = nullptr;
    bool work_ev_set = false;
    // the synchronous calls on host or device buffers first wait for the work
    // queued on the legacy default stream (caller_ev recorded there)
    hipEvent_t caller_ev = nullptr;
    // calls on host pointers: a pipeline of slots (see run_host)
    struct HostSlot {
        void* pin = nullptr;        // pinned host staging of one chunk
        void* dev = nullptr;        // device staging of one chunk
        size_t bytes = 0;
        hipStream_t stream = nullptr;      // copy-in and kernels
        hipStream_t stream_out = nullptr;  // copy-out, after `computed`
        hipEvent_t computed = nullptr;
        hipEvent_t done = nullptr;
    } slot[4];
    hipStream_t stream = nullptr;
};

namespace {

// device allocation of a plan; fails as out of memory under the fault
// injection `when` (kInjectAlloc at creation, kInjectWorkAlloc later)
template <class T>
hipError_t plan_malloc(zlfft_plan p, T** ptr, size_t bytes, int when) {
    if (p->inject == when) return hipErrorOutOfMemory;
    return hipMalloc(reinterpret_cast<void**>(ptr), bytes);
}

// kernel launch of a plan's call (fails under kInjectLaunch)
hipError_t plan_launch(zlfft_plan p, const void* fn, dim3 grid, dim3 block, void** params, hipStream_t stream) {
    if (p->inject == zl::kInjectLaunch) return hipErrorLaunchFailure;
    return hipLaunchKernel(fn, grid, block, params, 0, stream);
}

// the same check before a launch written with the triple-chevron syntax
hipError_t plan_launch_check(zlfft_plan p) {
    return p->inject == zl::kInjectLaunch ? hipErrorLaunchFailure : hipSuccess;
}

int check_plan(zlfft_plan p) {
    if (!p) return fail(ZLFFT_ERR_NULL, "plan is NULL");
    return ZLFFT_OK;
}

bool capturing(void* stream) {
    hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(reinterpret_cast<hipStream_t>(stream), &st) != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    return st != hipStreamCaptureStatusNone;
}

// waits on the host for the last queued work on the workspace
int wait_work(zlfft_plan p) {
    if (p->work_ev_set) ZL_HIP(hipEventSynchronize(p->work_ev));
    return ZLFFT_OK;
}

int ensure_work(zlfft_plan p, size_t bytes) {
    if (p->work_bytes >= bytes) return ZLFFT_OK;
    if (p->d_work) {
        // the previous workspace may still be in use by queued work
        if (int rc = wait_work(p)) return rc;
        (void)hipFree(p->d_work);
        p->d_work = nullptr;
        p->work_bytes = 0;
    }
    hipError_t e = plan_malloc(p, &p->d_work, bytes, zl::kInjectWorkAlloc);
    if (e != hipSuccess) return fail(ZLFFT_ERR_ALLOC, std::string("workspace hipMalloc: ") + hipGetErrorString(e));
    p->work_bytes = bytes;
    return ZLFFT_OK;
}

// Multi-pass CFFT of order p->corder on `batch` transforms: pass 0 reads the
// user layout lin, the last pass writes lout, intermediates ping-pong in work.
// nt_io: bit 0 when the input, bit 1 when the output is the caller's buffer
// (touched once: non-temporal access) rather than an internal one.
// rk: C2C, or the real kind whose pre-twiddle (C2R: the first pass reads the
// caller's bins) or post-twiddle (R2C / R2C_SQR: the last pass writes bins or
// squared magnitudes, db / db_floor as for the STFT) the passes include.
// lin == FRAMED_IN: the first pass reads frames of the signal in0 (hop,
// window) for the STFT.
int macro_cfft(zlfft_plan p, int backward, int lin, const void* in0, const void* in1, int lout, void* out0,
               void* out1, size_t batch, void* stream, void* work, int nt_io, int rk = zl::C2C, int db = 0,
               float db_floor = 0.f, long long hop = 0, const void* window = nullptr) {
    const long long n = 1ll << p->corder;
    void* tmp[2] = {work, static_cast<char*>(work) + (size_t)batch * (size_t)n * 2 * p->ss};
    const void* s0 = in0;
    const void* s1 = in1;
    for (int q = 0; q < p->mp; ++q) {
        const bool last = q == p->mp - 1;  // writes the caller's output
        int role = q == 0 ? 0 : (last ? 1 : 2);
        int lay = role == 0 ? ((nt_io & 1) ? lin : zl::kInternal)
                            : (role == 1 ? ((nt_io & 2) ? lout : zl::kInternal) : zl::AOS);
        if (role == 0 && lin == zl::FRAMED_IN) {
            role = zl::kRoleFramed;
            lay = zl::AOS;
        } else if (role == 0 && rk == zl::C2R) {
            role = zl::kRoleC2R;
            lay = lin;
        } else if (role == 1 && (rk == zl::R2C || rk == zl::R2C_SQR)) {
            role = rk == zl::R2C ? zl::kRoleR2C : zl::kRoleSqr;
            lay = rk == zl::R2C ? lout : zl::AOS;
        }
        const zl::MacroEntry& e = (p->f64 ? g_macro64 : g_macro)[p->logl[q]][role][lay];
        if (!e.fn) return fail(ZLFFT_ERR_UNSUPPORTED, "no HBM-pass kernel for this layout");
        void* d0 = last ? out0 : tmp[q & 1];
        void* d1 = last ? out1 : nullptr;
        zl::MArgs a{};
        a.in0 = s0;
        a.in1 = s1;
        a.out0 = d0;
        a.out1 = d1;
        a.tw = p->d_itw[q];
        a.twa = p->d_twa[q];
        a.twb = p->d_twb[q];
        a.batch = (long long)batch;
        a.logn = p->corder;
        a.logw = p->logw[q];
        a.logh = p->logh[q];
        a.backward = backward;
        a.rtw = p->d_rtw;
        a.rtwf = p->d_rtwf;
        a.rtwh = p->rtw_h;
        a.rtwb = p->rtw_b;
        a.rtws = p->rtw_s;
        a.db = db;
        a.db_floor = db_floor;
        a.hop = hop;
        a.window = window;
        const size_t grid = batch * (size_t)(n >> p->logl[q]) / (size_t)e.cols;
        if (grid > 0x7fffffffull) return fail(ZLFFT_ERR_INVALID_ARG, "batch too large for one launch");
        void* params[] = {&a};
        ZL_HIP(plan_launch(p, reinterpret_cast<const void*>(e.fn), dim3((unsigned)grid), dim3(e.wg), params,
                           reinterpret_cast<hipStream_t>(stream)));
        s0 = d0;
        s1 = d1;
    }
    return ZLFFT_OK;
}

size_t macro_work_bytes(zlfft_plan p, size_t batch) {
    const size_t n = (size_t)1 << p->corder;
    const size_t bufs = p->mp > 2 ? 2 : 1;
    return bufs * batch * n * 2 * p->ss;
}

// The multi-pass transform of one call, on a workspace the caller has sized
// (ensure_work) and ordered for `stream` (work_acquire).  lin == FRAMED_IN
// (forward real kinds): in0 is the STFT's signal, framed by hop and window.
int launch_macro_body(zlfft_plan p, int kind, int lin, int lout, const void* in0, const void* in1, void* out0,
                      void* out1, size_t batch, int backward, void* stream, int db, float db_floor,
                      long long hop = 0, const void* window = nullptr) {
    void* work = p->d_work;
    if (kind == zl::C2C) return macro_cfft(p, backward, lin, in0, in1, lout, out0, out1, batch, stream, work, 3);
    // RFFT: N reals as N/2 complex; the post-twiddle runs in the last pass,
    // the pre-twiddle in the first
    if (kind == zl::R2C || kind == zl::R2C_SQR)
        return macro_cfft(p, 0, lin == zl::FRAMED_IN ? lin : zl::AOS, in0, nullptr, lout, out0, out1, batch, stream,
                          work, 3, kind, db, db_floor, hop, window);
    return macro_cfft(p, 1, lin, in0, in1, zl::AOS, out0, nullptr, batch, stream, work, 3, zl::C2R);
}

// The workspace serves one stream at a time: a call on another stream than
// the previous multi-pass call's waits, on the device, for that call's work
// (work_ev, recorded by work_release).  Not inside a stream capture, where
// the caller orders the captured work.
int work_acquire(zlfft_plan p, void* stream) {
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    if (p->work_ev_set && p->work_stream != s && !capturing(stream)) ZL_HIP(hipStreamWaitEvent(s, p->work_ev, 0));
    return ZLFFT_OK;
}

int work_release(zlfft_plan p, void* stream) {
    if (capturing(stream)) return ZLFFT_OK;
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    ZL_HIP(hipEventRecord(p->work_ev, s));
    p->work_stream = s;
    p->work_ev_set = true;
    return ZLFFT_OK;
}

int launch_macro(zlfft_plan p, int kind, int lin, int lout, const void* in0, const void* in1, void* out0, void* out1,
                 size_t batch, int backward, void* stream, int db = 0, float db_floor = 0.f, long long hop = 0,
                 const void* window = nullptr) {
    if (batch == 0) return ZLFFT_OK;
    if (int rc = ensure_work(p, macro_work_bytes(p, batch))) return rc;
    if (int rc = work_acquire(p, stream)) return rc;
    const int rc = launch_macro_body(p, kind, lin, lout, in0, in1, out0, out1, batch, backward, stream, db, db_floor,
                                     hop, window);
    if (rc) return rc;
    return work_release(p, stream);
}

int launch_mfma(zlfft_plan p, int lin, int lout, const void* in0, void* out0, size_t batch, int backward,
                void* stream) {
    if (lin != ZLFFT_AOS || lout != ZLFFT_AOS)
        return fail(ZLFFT_ERR_UNSUPPORTED, "the MFMA precisions take AoS input and output");
    // more transforms than resident workgroups: the persistent form
    const long long resident = (long long)p->mfma_blocks * p->cus;
    const bool streaming = resident > 0 && p->d_sink && (long long)batch > resident;
    void* fn = zl_mfma_kernel(p->precision, backward, streaming ? 1 : 0);
    if (!fn) return fail(ZLFFT_ERR_UNSUPPORTED, "no MFMA kernel for this precision");
    if (batch == 0) return ZLFFT_OK;
    if (batch > 0x7fffffffull) return fail(ZLFFT_ERR_INVALID_ARG, "batch too large for one launch");
    zl::KArgs a{};
    a.in0 = in0;
    a.out0 = out0;
    a.tw = p->d_twm;
    a.sink = p->d_sink;
    a.batch = (long long)batch;
    void* params[] = {&a};
    const unsigned grid = (unsigned)(streaming ? resident : (long long)batch);
    ZL_HIP(plan_launch(p, fn, dim3(grid), dim3(256), params, reinterpret_cast<hipStream_t>(stream)));
    return ZLFFT_OK;
}

// Streaming RFFT epilogue store policy from the output size of a call: cached
// stores above ZLFFT_RFFT_CACHED_MB (default 64 MiB), non-temporal below
// (measured: the 4096-frame analyzer batch is faster non-temporal, batches of
// tens of thousands of frames cached).
int rfft_cached_out(size_t out_bytes) {
    static const size_t threshold = [] {
        size_t mb = 64;
        if (const char* e = getenv("ZLFFT_RFFT_CACHED_MB")) mb = (size_t)atol(e);
        return mb << 20;
    }();
    return out_bytes > threshold ? 1 : 0;
}

// ZLFFT_NO_STREAM (tuning): LDS kernels only, no persistent streaming kernels
bool no_stream_kernels() {
    static const bool no_stream = getenv("ZLFFT_NO_STREAM") != nullptr;
    return no_stream;
}

int launch(zlfft_plan p, int kind, int lin, int lout, const void* in0, const void* in1, void* out0, void* out1,
           size_t batch, int backward, void* stream) {
    if (p->macro) return launch_macro(p, kind, lin, lout, in0, in1, out0, out1, batch, backward, stream);
    if (p->precision != ZLFFT_PREC_F32 && !p->f64) return launch_mfma(p, lin, lout, in0, out0, batch, backward, stream);
    const zl::KernelEntry& e = (*p->tab)[kind][lin][lout];
    if (!e.fn) return fail(ZLFFT_ERR_UNSUPPORTED, "no kernel for this kind/layout combination");
    if (batch == 0) return ZLFFT_OK;
    zl::KArgs a{};
    a.in0 = in0;
    a.in1 = in1;
    a.out0 = out0;
    a.out1 = out1;
    a.tw = p->d_tw;
    a.rtw = p->d_rtw;
    a.batch = (long long)batch;
    a.backward = backward;
    const size_t groups = (batch + e.tpw - 1) / e.tpw;
    if (groups > 0x7fffffffull) return fail(ZLFFT_ERR_INVALID_ARG, "batch too large for one launch");
    void* params[] = {&a};
    const int resident = (e.stream_fn[0] && !no_stream_kernels()) ? p->stream_blocks[kind][lin][lout] * p->cus : 0;
    if (resident > 0 && (batch + e.s_tpw - 1) / e.s_tpw > (size_t)resident) {
        // persistent streaming kernel: one workgroup per resident slot
        a.tw = p->d_img;
        a.sink = p->d_sink;
        a.cached_out = rfft_cached_out(batch * (((size_t)1 << p->corder) + 1) * 2 * p->ss);
        ZL_HIP(plan_launch(p, reinterpret_cast<const void*>(e.stream_fn[backward ? 1 : 0]), dim3((unsigned)resident),
                           dim3(e.s_wg), params, reinterpret_cast<hipStream_t>(stream)));
        return ZLFFT_OK;
    }
    ZL_HIP(plan_launch(p, reinterpret_cast<const void*>(e.fn), dim3((unsigned)groups), dim3(e.wg), params,
                           reinterpret_cast<hipStream_t>(stream)));
    return ZLFFT_OK;
}

bool valid_layout(int l) { return l == ZLFFT_AOS || l == ZLFFT_SOA; }

// Where a user buffer lives, for the calls that accept host pointers.
enum MemKind { MEM_PAGEABLE = 0, MEM_PINNED = 1, MEM_DEVICE = 2 };

MemKind mem_kind(const void* ptr) {
    hipPointerAttribute_t attr;
    std::memset(&attr, 0, sizeof(attr));
    hipError_t e = hipPointerGetAttributes(&attr, ptr);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return MEM_PAGEABLE;
    }
    if (attr.type == hipMemoryTypeDevice || attr.type == hipMemoryTypeManaged) return MEM_DEVICE;
    if (attr.type == hipMemoryTypeHost) return MEM_PINNED;
    return MEM_PAGEABLE;
}

// One side (input or output) of a call on host or device buffers: up to two
// arrays, `bytes` per transform each.
struct Buf {
    const void* p[2];
    size_t bytes[2];
};

constexpr size_t kAlign = 256;
constexpr int kMaxSlots = 4;

// Pipeline shape: bytes of inputs + outputs per chunk and slots in turn
// (ZLFFT_HOST_CHUNK_MB and ZLFFT_HOST_SLOTS override the defaults).
struct HostShape {
    size_t chunk_bytes = 32u << 20;
    int slots = 3;
};

// memcpy of a large pageable block on several host threads (the single-thread
// copy is slower than the link)
void par_memcpy(void* dst, const void* src, size_t bytes) {
    constexpr size_t kMinPerThread = 4u << 20;
    const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    const size_t nt = std::min<size_t>({8, hw, std::max<size_t>(1, bytes / kMinPerThread)});
    if (nt <= 1) {
        std::memcpy(dst, src, bytes);
        return;
    }
    const size_t part = (bytes + nt - 1) / nt;
    std::vector<std::thread> th;
    th.reserve(nt - 1);
    for (size_t i = 1; i < nt; ++i) {
        const size_t o = i * part;
        if (o >= bytes) break;
        th.emplace_back([=] { std::memcpy(static_cast<char*>(dst) + o, static_cast<const char*>(src) + o,
                                          std::min(part, bytes - o)); });
    }
    std::memcpy(dst, src, std::min(part, bytes));
    for (auto& t : th) t.join();
}
HostShape host_shape() {
    static const HostShape s = [] {
        HostShape h;
        if (const char* e = getenv("ZLFFT_HOST_CHUNK_MB")) {
            const long v = atol(e);
            if (v >= 1 && v <= 1024) h.chunk_bytes = (size_t)v << 20;
        }
        if (const char* e = getenv("ZLFFT_HOST_SLOTS")) {
            const int v = atoi(e);
            if (v >= 1 && v <= kMaxSlots) h.slots = v;
        }
        return h;
    }();
    return s;
}

size_t round_up(size_t v) { return (v + kAlign - 1) / kAlign * kAlign; }

int ensure_slot(zlfft_plan p, int i, size_t bytes) {
    auto& sl = p->slot[i];
    if (!sl.stream) {
        ZL_HIP(hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking));
        ZL_HIP(hipStreamCreateWithFlags(&sl.stream_out, hipStreamNonBlocking));
        ZL_HIP(hipEventCreateWithFlags(&sl.computed, hipEventDisableTiming));
        ZL_HIP(hipEventCreateWithFlags(&sl.done, hipEventDisableTiming));
    }
    if (sl.bytes >= bytes) return ZLFFT_OK;
    ZL_HIP(hipStreamSynchronize(sl.stream));
    ZL_HIP(hipStreamSynchronize(sl.stream_out));
    if (sl.pin) (void)hipHostFree(sl.pin);
    if (sl.dev) (void)hipFree(sl.dev);
    sl.pin = sl.dev = nullptr;
    sl.bytes = 0;
    if (p->inject == zl::kInjectWorkAlloc || hipHostMalloc(&sl.pin, bytes, hipHostMallocDefault) != hipSuccess ||
        hipMalloc(&sl.dev, bytes) != hipSuccess)
        return fail(ZLFFT_ERR_ALLOC, "host pipeline staging allocation failed");
    sl.bytes = bytes;
    return ZLFFT_OK;
}

// Runs body(in0, in1, out0, out1, count, stream) over `batch` transforms whose
// buffers may be pageable host, pinned host or device memory, and returns when
// the outputs are in place.  All-device calls run directly on the plan stream.
// Otherwise the batch goes through in chunks (32 MiB of inputs + outputs by
// default), taking the slots in turn (three by default; each is pinned
// staging + device staging + a stream for copy-in and kernels + a stream for
// copy-out + events): the copy-in of
// one chunk, the kernels of the previous one and the copy-out of the one
// before overlap (the copy engines and the GPU run concurrently).  Pageable
// arrays are copied through the pinned staging by the host; pinned arrays are
// transferred directly; device arrays are used in place.  Multi-pass plans
// share one workspace, so they use one slot.
template <class Body>
int run_host_chunks(zlfft_plan p, Buf in, Buf out, size_t batch, Body body);

// run_host_chunks, and on failure waits for whatever it already queued, so
// that no copy into the staging outlives the call
template <class Body>
int run_host(zlfft_plan p, Buf in, Buf out, size_t batch, Body body) {
    const int rc = run_host_chunks(p, in, out, batch, body);
    if (rc != ZLFFT_OK) {
        for (auto& sl : p->slot) {
            if (sl.stream) (void)hipStreamSynchronize(sl.stream);
            if (sl.stream_out) (void)hipStreamSynchronize(sl.stream_out);
        }
        (void)hipStreamSynchronize(p->stream);
    }
    return rc;
}

template <class Body>
int run_host_chunks(zlfft_plan p, Buf in, Buf out, size_t batch, Body body) {
    MemKind kin[2] = {MEM_DEVICE, MEM_DEVICE}, kout[2] = {MEM_DEVICE, MEM_DEVICE};
    bool all_dev = true, any_dev = false;
    for (int k = 0; k < 2; ++k) {
        if (in.p[k]) all_dev &= (kin[k] = mem_kind(in.p[k])) == MEM_DEVICE;
        if (out.p[k]) all_dev &= (kout[k] = mem_kind(out.p[k])) == MEM_DEVICE;
        any_dev |= (in.p[k] && kin[k] == MEM_DEVICE) || (out.p[k] && kout[k] == MEM_DEVICE);
    }
    // device buffers of the caller: the library's streams do not synchronise
    // with the legacy default stream, so every stream that touches them first
    // waits for the work queued there (a kernel still producing the input, or
    // still reading the output's old contents)
    if (any_dev) ZL_HIP(hipEventRecord(p->caller_ev, nullptr));
    auto after_caller = [&](hipStream_t s) -> int {
        if (any_dev) ZL_HIP(hipStreamWaitEvent(s, p->caller_ev, 0));
        return ZLFFT_OK;
    };
    if (all_dev) {
        if (int rc = after_caller(p->stream)) return rc;
        int rc = body(in.p[0], in.p[1], const_cast<void*>(out.p[0]), const_cast<void*>(out.p[1]), batch, p->stream);
        if (rc) return rc;
        ZL_HIP(hipStreamSynchronize(p->stream));
        return ZLFFT_OK;
    }
    const size_t per = in.bytes[0] + in.bytes[1] + out.bytes[0] + out.bytes[1];
    const HostShape shape = host_shape();
    const size_t chunk = std::max<size_t>(1, std::min<size_t>(batch, shape.chunk_bytes / std::max<size_t>(per, 1)));
    // slot layout: in0 | in1 | out0 | out1, each kAlign-aligned
    size_t off[4], o = 0;
    const size_t bsz[4] = {in.bytes[0], in.bytes[1], out.bytes[0], out.bytes[1]};
    for (int k = 0; k < 4; ++k) {
        off[k] = o;
        o += round_up(chunk * bsz[k]);
    }
    const int nslots = p->macro ? 1 : shape.slots;
    for (int i = 0; i < nslots; ++i) {
        if (int rc = ensure_slot(p, i, o)) return rc;
        if (int rc = after_caller(p->slot[i].stream)) return rc;
        if (int rc = after_caller(p->slot[i].stream_out)) return rc;
    }
    size_t pend_first[kMaxSlots] = {}, pend_n[kMaxSlots] = {};
    // waits for a slot's chunk and moves its outputs from pinned staging to pageable user memory
    auto finish = [&](int i) -> int {
        if (!pend_n[i]) return ZLFFT_OK;
        auto& sl = p->slot[i];
        ZL_HIP(hipEventSynchronize(sl.done));
        for (int k = 0; k < 2; ++k)
            if (out.p[k] && kout[k] == MEM_PAGEABLE)
                par_memcpy(static_cast<char*>(const_cast<void*>(out.p[k])) + pend_first[i] * out.bytes[k],
                           static_cast<char*>(sl.pin) + off[2 + k], pend_n[i] * out.bytes[k]);
        pend_n[i] = 0;
        return ZLFFT_OK;
    };
    int c = 0;
    for (size_t first = 0; first < batch; first += chunk, ++c) {
        const size_t n = std::min(chunk, batch - first);
        const int i = c % nslots;
        if (int rc = finish(i)) return rc;
        auto& sl = p->slot[i];
        char* pin = static_cast<char*>(sl.pin);
        char* dev = static_cast<char*>(sl.dev);
        const void* din[2] = {nullptr, nullptr};
        void* dout[2] = {nullptr, nullptr};
        for (int k = 0; k < 2; ++k) {
            if (!in.p[k]) continue;
            const char* src = static_cast<const char*>(in.p[k]) + first * in.bytes[k];
            const size_t nb = n * in.bytes[k];
            if (kin[k] == MEM_DEVICE) {
                din[k] = src;
            } else {
                const void* from = src;
                if (kin[k] == MEM_PAGEABLE) {
                    par_memcpy(pin + off[k], src, nb);
                    from = pin + off[k];
                }
                ZL_HIP(hipMemcpyAsync(dev + off[k], from, nb, hipMemcpyHostToDevice, sl.stream));
                din[k] = dev + off[k];
            }
        }
        for (int k = 0; k < 2; ++k) {
            if (!out.p[k]) continue;
            dout[k] = kout[k] == MEM_DEVICE ? static_cast<char*>(const_cast<void*>(out.p[k])) + first * out.bytes[k]
                                            : static_cast<void*>(dev + off[2 + k]);
        }
        if (int rc = body(din[0], din[1], dout[0], dout[1], n, sl.stream)) return rc;
        // copy-out on a stream of its own, so the two copy directions of
        // neighbouring chunks run at the same time
        ZL_HIP(hipEventRecord(sl.computed, sl.stream));
        ZL_HIP(hipStreamWaitEvent(sl.stream_out, sl.computed, 0));
        for (int k = 0; k < 2; ++k) {
            if (!out.p[k] || kout[k] == MEM_DEVICE) continue;
            void* to = kout[k] == MEM_PINNED ? static_cast<char*>(const_cast<void*>(out.p[k])) + first * out.bytes[k]
                                             : static_cast<void*>(pin + off[2 + k]);
            ZL_HIP(hipMemcpyAsync(to, dev + off[2 + k], n * out.bytes[k], hipMemcpyDeviceToHost, sl.stream_out));
        }
        ZL_HIP(hipEventRecord(sl.done, sl.stream_out));
        pend_first[i] = first;
        pend_n[i] = n;
    }
    // the older chunk first
    for (int j = 0; j < nslots; ++j)
        if (int rc = finish((c + j) % nslots)) return rc;
    return ZLFFT_OK;
}

// The device address of `scratch` on the current device (looked up once per
// device).  The calling thread's current device must be the plan's.
int scratch_ptr(zlfft_plan p, void** out) {
    int dev = -1;
    ZL_HIP(hipGetDevice(&dev));
    if (dev != p->device) return fail(ZLFFT_ERR_INVALID_ARG, "the plan's device is not the current device");
    if (dev < 0 || dev >= kMaxDevices) return fail(ZLFFT_ERR_UNSUPPORTED, "device index beyond the scratch table");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* ptr = nullptr;
        ZL_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        g_scratch_ptr[dev] = ptr;
    }
    *out = g_scratch_ptr[dev];
    return ZLFFT_OK;
}

}  // namespace

extern "C" {

const char* zlfft_last_error(void) { return g_err.c_str(); }

int zlfft_version(void) { return 100; }

int zlfft_plan_create(zlfft_plan* plan, int kind, int precision, size_t order) {
    g_err.clear();
    if (!plan) return fail(ZLFFT_ERR_NULL, "plan out-pointer is NULL");
    *plan = nullptr;
    if (kind != ZLFFT_KIND_CFFT && kind != ZLFFT_KIND_RFFT) return fail(ZLFFT_ERR_INVALID_ARG, "unknown plan kind");
    if (precision < ZLFFT_PREC_F32 || precision > ZLFFT_PREC_F64)
        return fail(ZLFFT_ERR_UNSUPPORTED, "unknown precision");
    const bool f64 = precision == ZLFFT_PREC_F64;
    const bool mfma = precision != ZLFFT_PREC_F32 && !f64;
    if (mfma && (kind != ZLFFT_KIND_CFFT || order != 12))
        return fail(ZLFFT_ERR_UNSUPPORTED, "the MFMA precisions cover the 4096-point CFFT (order 12) only");
    if (kind == ZLFFT_KIND_RFFT && order < 1)
        return fail(ZLFFT_ERR_INVALID_ORDER, "RFFT order must be >= 1 (the reference underflows at 0)");
    const size_t corder = kind == ZLFFT_KIND_RFFT ? order - 1 : order;
    if (corder > (size_t)(f64 ? zl::kMaxOrder64 : zl::kMaxOrder))
        return fail(ZLFFT_ERR_INVALID_ORDER, "order " + std::to_string(order) + " exceeds the supported maximum");
    std::call_once(g_tab_once, fill_tables);

    zlfft_plan p = new (std::nothrow) zlfft_plan_s();
    if (!p) return fail(ZLFFT_ERR_ALLOC, "plan allocation failed");
    p->inject = zl_read_inject();
    p->kind = kind;
    p->precision = precision;
    p->order = (int)order;
    p->corder = (int)corder;
    p->f64 = f64;
    p->ss = f64 ? 8 : 4;
    p->macro = corder > (size_t)(f64 ? zl::kMaxLdsOrder64 : zl::kMaxLdsOrder);
    p->tab = p->macro ? nullptr : f64 ? &g_tab64[corder] : &g_tab[corder];
    const size_t csz = 2 * p->ss;  // bytes per complex value
    // the table entry whose twiddles the plan holds: CFFT plans launch only
    // C2C kernels, RFFT plans only the real kinds (which share one geometry)
    const int own_kind = kind == ZLFFT_KIND_RFFT ? zl::R2C : zl::C2C;
    // twiddle generators for the plan's complex type
    auto gen_pass = [&](void* out, size_t off, int lr, int lw, int cnt) {
        const unsigned g = (unsigned)((cnt + 255) / 256);
        if (f64)
            gen_pass_twiddles<<<g, 256, 0, p->stream>>>(static_cast<double2*>(out) + off, lr, lw, cnt);
        else
            gen_pass_twiddles<<<g, 256, 0, p->stream>>>(static_cast<float2*>(out) + off, lr, lw, cnt);
    };
    auto gen_macro = [&](void* out, int log_lw, int shift, int cnt) {
        const unsigned g = (unsigned)((cnt + 255) / 256);
        if (f64)
            gen_macro_twiddles<<<g, 256, 0, p->stream>>>(static_cast<double2*>(out), log_lw, shift, cnt);
        else
            gen_macro_twiddles<<<g, 256, 0, p->stream>>>(static_cast<float2*>(out), log_lw, shift, cnt);
    };
    auto cleanup = [&](int rc) {
        zlfft_plan_destroy(p);
        return rc;
    };
    hipError_t e = hipGetDevice(&p->device);
    if (e != hipSuccess) return cleanup(hip_fail(e, "hipGetDevice"));
    e = hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking);
    if (e != hipSuccess) return cleanup(hip_fail(e, "hipStreamCreateWithFlags"));

    e = hipEventCreateWithFlags(&p->work_ev, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&p->caller_ev, hipEventDisableTiming);
    if (e != hipSuccess) return cleanup(hip_fail(e, "hipEventCreateWithFlags"));
    e = hipDeviceGetAttribute(&p->cus, hipDeviceAttributeMultiprocessorCount, p->device);
    if (e != hipSuccess) return cleanup(hip_fail(e, "hipDeviceGetAttribute"));

    if (p->macro) {
        const zl::MacroPlan mpl = f64 ? zl::macro_plan64((int)corder) : zl::macro_plan((int)corder);
        p->mp = mpl.np;
        int lw = 0;
        for (int q = 0; q < mpl.np; ++q) {
            p->logl[q] = mpl.logl[q];
            p->logw[q] = lw;
            const zl::MacroEntry& me = (f64 ? g_macro64 : g_macro)[mpl.logl[q]][0][0];
            if (me.tw_total > 0) {
                e = plan_malloc(p, &p->d_itw[q], csz * (size_t)me.tw_total, zl::kInjectAlloc);
                if (e != hipSuccess) return cleanup(fail(ZLFFT_ERR_ALLOC, "twiddle hipMalloc failed"));
                size_t off = 0;
                for (int r = 1; r < me.np; ++r) {
                    const int cnt = ((1 << me.lr[r]) - 1) << me.lw[r];
                    gen_pass(p->d_itw[q], off, me.lr[r], me.lw[r], cnt);
                    off += (size_t)cnt;
                }
            }
            if (q > 0) {
                const int log_lw = lw + mpl.logl[q];
                const int h = (log_lw + 1) / 2;
                p->logh[q] = h;
                const int na = 1 << (log_lw - h), nb = 1 << h;
                e = plan_malloc(p, &p->d_twa[q], csz * (size_t)na, zl::kInjectAlloc);
                if (e == hipSuccess) e = plan_malloc(p, &p->d_twb[q], csz * (size_t)nb, zl::kInjectAlloc);
                if (e != hipSuccess) return cleanup(fail(ZLFFT_ERR_ALLOC, "twiddle hipMalloc failed"));
                gen_macro(p->d_twa[q], log_lw, h, na);
                gen_macro(p->d_twb[q], log_lw, 0, nb);
            }
            lw += mpl.logl[q];
        }
        e = hipGetLastError();
        if (e != hipSuccess) return cleanup(hip_fail(e, "twiddle generation"));
    }
    if (!p->macro) {
        for (int k = 0; k < 6; ++k)
            for (int i = 0; i < 2; ++i)
                for (int o = 0; o < 2; ++o) {
                    const zl::KernelEntry& ke = (*p->tab)[k][i][o];
                    if (!ke.stream_fn[0]) continue;
                    int nb = 0;
                    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(
                        &nb, reinterpret_cast<const void*>(ke.stream_fn[0]), ke.s_wg, 0);
                    if (e != hipSuccess) return cleanup(hip_fail(e, "hipOccupancyMaxActiveBlocksPerMultiprocessor"));
                    p->stream_blocks[k][i][o] = nb;
                }
        if (mfma) {
            e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&p->mfma_blocks, zl_mfma_kernel(precision, 0, 1), 256, 0);
            if (e != hipSuccess) return cleanup(hip_fail(e, "hipOccupancyMaxActiveBlocksPerMultiprocessor"));
        }
        const zl::KernelEntry& ke = (*p->tab)[zl::C2C][0][0];
        if (ke.stream_fn[0] || (*p->tab)[zl::R2C][0][0].stream_fn[0]) {
            e = plan_malloc(p, &p->d_sink, (size_t)zl::kSinkBytes, zl::kInjectAlloc);
            if (e != hipSuccess) return cleanup(fail(ZLFFT_ERR_ALLOC, "sink hipMalloc failed"));
        }
        // the LDS twiddle image of the plan's kind: C2C for CFFT plans, the
        // real kinds' (one geometry and image for all of them) for RFFT plans
        const zl::KernelEntry& ki = (*p->tab)[own_kind][0][0];
        if (ki.img_total > 0) {
            e = plan_malloc(p, &p->d_img, csz * (size_t)ki.img_total, zl::kInjectAlloc);
            if (e != hipSuccess) return cleanup(fail(ZLFFT_ERR_ALLOC, "twiddle hipMalloc failed"));
            size_t off = 0;
            for (int q = 1; q < ki.s_np; ++q) {
                const int cnt = ki.img_lo[q] + ki.img_hi[q];
                const unsigned g = (unsigned)((cnt + 255) / 256);
                if (f64)
                    gen_image_twiddles<<<g, 256, 0, p->stream>>>(reinterpret_cast<double2*>(p->d_img) + off,
                                                                  ki.s_lr[q], ki.s_lw[q], ki.img_lo[q], ki.img_hi[q]);
                else
                    gen_image_twiddles<<<g, 256, 0, p->stream>>>(p->d_img + off, ki.s_lr[q], ki.s_lw[q], ki.img_lo[q],
                                                                  ki.img_hi[q]);
                off += (size_t)cnt;
            }
        }
    }
    // pass twiddles of the plan's kind (the LDS kernels')
    const zl::KernelEntry& ent = p->macro ? zl::KernelEntry{} : (*p->tab)[own_kind][0][0];
    if (ent.tw_total > 0) {
        e = plan_malloc(p, &p->d_tw, csz * (size_t)ent.tw_total, zl::kInjectAlloc);
        if (e != hipSuccess) return cleanup(fail(ZLFFT_ERR_ALLOC, "twiddle hipMalloc failed"));
        size_t off = 0;
        for (int q = 1; q < ent.np; ++q) {
            const int cnt = ((1 << ent.lr[q]) - 1) << ent.lw[q];
            gen_pass(p->d_tw, off, ent.lr[q], ent.lw[q], cnt);
            off += (size_t)cnt;
        }
        if (mfma) {
            e = plan_malloc(p, &p->d_twm, sizeof(float2) * (size_t)zl::kMfmaTwTotal, zl::kInjectAlloc);
            if (e != hipSuccess) return cleanup(fail(ZLFFT_ERR_ALLOC, "twiddle hipMalloc failed"));
            const bool f16 = precision == ZLFFT_PREC_F16_MFMA || precision == ZLFFT_PREC_F16_MFMA_SPLIT;
            gen_mfma_twiddles<<<(zl::kMfmaTwTotal + 255) / 256, 256, 0, p->stream>>>(p->d_twm, f16 ? 0.0625f : 1.0f);
        }
        e = hipGetLastError();
        if (e != hipSuccess) return cleanup(hip_fail(e, "twiddle generation"));
    }
    if (kind == ZLFFT_KIND_RFFT) {
        const long long nreal = 1ll << order;
        p->n_rtw = (int)(nreal / 4);
        if (p->n_rtw > 0) {
            e = plan_malloc(p, &p->d_rtw, csz * (size_t)p->n_rtw, zl::kInjectAlloc);
            if (e != hipSuccess) return cleanup(fail(ZLFFT_ERR_ALLOC, "rfft twiddle hipMalloc failed"));
            const unsigned g = (unsigned)((p->n_rtw + 255) / 256);
            if (f64)
                gen_rfft_twiddles<<<g, 256, 0, p->stream>>>(static_cast<double2*>(p->d_rtw), nreal, p->n_rtw);
            else
                gen_rfft_twiddles<<<g, 256, 0, p->stream>>>(static_cast<float2*>(p->d_rtw), nreal, p->n_rtw);
            e = hipGetLastError();
            if (e != hipSuccess) return cleanup(hip_fail(e, "rfft twiddle generation"));
        }
        if (p->macro && !f64) {
            // FP64 W_N^e for the fused passes' twiddles: A[k >> h], B[k & (2^h - 1)]
            // for k = 0..N/4, then [idx 0, 1] of W^{N/32 idx} and of W^{N/64 idx}
            // (the steps N/(2 R) of the real size N for R = 16 and 32 points per thread)
            const int lk = (int)order - 2;  // log2 (N/4)
            const int h = (lk + 1) / 2;
            const int na = (1 << (lk - h)) + 1, nb = 1 << h;
            e = plan_malloc(p, &p->d_rtwf, sizeof(double2) * (size_t)(na + nb + 4), zl::kInjectAlloc);
            if (e != hipSuccess) return cleanup(fail(ZLFFT_ERR_ALLOC, "rfft twiddle hipMalloc failed"));
            p->rtw_h = h;
            p->rtw_b = na;
            p->rtw_s = na + nb + 1;  // the W^{N/32} entry; W^{N/64} two further on
            gen_macro_twiddles<<<(na + 255) / 256, 256, 0, p->stream>>>(p->d_rtwf, (int)order, h, na);
            gen_macro_twiddles<<<(nb + 255) / 256, 256, 0, p->stream>>>(p->d_rtwf + na, (int)order, 0, nb);
            gen_macro_twiddles<<<1, 256, 0, p->stream>>>(p->d_rtwf + na + nb, (int)order, (int)order - 5, 2);
            gen_macro_twiddles<<<1, 256, 0, p->stream>>>(p->d_rtwf + na + nb + 2, (int)order, (int)order - 6, 2);
            e = hipGetLastError();
            if (e != hipSuccess) return cleanup(hip_fail(e, "rfft twiddle generation"));
        }
    }
    e = hipStreamSynchronize(p->stream);
    if (e != hipSuccess) return cleanup(hip_fail(e, "plan synchronize"));
    *plan = p;
    return ZLFFT_OK;
}

int zlfft_plan_destroy(zlfft_plan p) {
    if (!p) return fail(ZLFFT_ERR_NULL, "plan is NULL");
    if (p->stream) (void)hipStreamSynchronize(p->stream);
    if (p->work_ev_set) (void)hipEventSynchronize(p->work_ev);  // multi-pass work queued on a caller's stream
    if (p->d_tw) (void)hipFree(p->d_tw);
    if (p->d_img) (void)hipFree(p->d_img);
    if (p->d_sink) (void)hipFree(p->d_sink);
    if (p->d_twm) (void)hipFree(p->d_twm);
    if (p->d_rtw) (void)hipFree(p->d_rtw);
    if (p->d_rtwf) (void)hipFree(p->d_rtwf);
    for (auto& sl : p->slot) {
        if (sl.stream) (void)hipStreamSynchronize(sl.stream);
        if (sl.stream_out) (void)hipStreamSynchronize(sl.stream_out);
        if (sl.pin) (void)hipHostFree(sl.pin);
        if (sl.dev) (void)hipFree(sl.dev);
        if (sl.computed) (void)hipEventDestroy(sl.computed);
        if (sl.done) (void)hipEventDestroy(sl.done);
        if (sl.stream) (void)hipStreamDestroy(sl.stream);
        if (sl.stream_out) (void)hipStreamDestroy(sl.stream_out);
    }
    if (p->d_work) (void)hipFree(p->d_work);
    if (p->work_ev) (void)hipEventDestroy(p->work_ev);
    if (p->caller_ev) (void)hipEventDestroy(p->caller_ev);
    for (int q = 0; q < 4; ++q) {
        if (p->d_itw[q]) (void)hipFree(p->d_itw[q]);
        if (p->d_twa[q]) (void)hipFree(p->d_twa[q]);
        if (p->d_twb[q]) (void)hipFree(p->d_twb[q]);
    }
    if (p->stream) (void)hipStreamDestroy(p->stream);
    delete p;
    return ZLFFT_OK;
}

size_t zlfft_plan_get_size(zlfft_plan p) { return p ? (size_t)1 << p->order : 0; }

size_t zlfft_plan_get_order(zlfft_plan p) { return p ? (size_t)p->order : 0; }

int zlfft_cfft_exec(zlfft_plan p, int direction, int in_layout, const void* in0, const void* in1, int out_layout,
                    void* out0, void* out1, size_t batch, void* stream) {
    if (int rc = check_plan(p)) return rc;
    if (p->kind != ZLFFT_KIND_CFFT) return fail(ZLFFT_ERR_INVALID_ARG, "plan is not a CFFT plan");
    if (direction != ZLFFT_FORWARD && direction != ZLFFT_BACKWARD)
        return fail(ZLFFT_ERR_INVALID_ARG, "direction must be ZLFFT_FORWARD or ZLFFT_BACKWARD");
    if (!valid_layout(in_layout) || !valid_layout(out_layout)) return fail(ZLFFT_ERR_INVALID_ARG, "bad layout");
    if (batch && (!in0 || !out0 || (in_layout == ZLFFT_SOA && !in1) || (out_layout == ZLFFT_SOA && !out1)))
        return fail(ZLFFT_ERR_NULL, "NULL data pointer");
    return launch(p, zl::C2C, in_layout, out_layout, in0, in1, out0, out1, batch, direction == ZLFFT_BACKWARD,
                  stream);
}

int zlfft_cfft_exec_noalloc(zlfft_plan p, int direction, int in_layout, const void* in0, const void* in1,
                            int out_layout, void* out0, void* out1, size_t batch, void* stream) {
    if (int rc = check_plan(p)) return rc;
    if (p->kind != ZLFFT_KIND_CFFT) return fail(ZLFFT_ERR_INVALID_ARG, "plan is not a CFFT plan");
    if (direction != ZLFFT_FORWARD && direction != ZLFFT_BACKWARD)
        return fail(ZLFFT_ERR_INVALID_ARG, "direction must be ZLFFT_FORWARD or ZLFFT_BACKWARD");
    if (!valid_layout(in_layout) || !valid_layout(out_layout)) return fail(ZLFFT_ERR_INVALID_ARG, "bad layout");
    if (batch && (!in0 || !out0 || (in_layout == ZLFFT_SOA && !in1) || (out_layout == ZLFFT_SOA && !out1)))
        return fail(ZLFFT_ERR_NULL, "NULL data pointer");
    // one-kernel plans use no scratch: the same launch as zlfft_cfft_exec
    if (!p->macro)
        return launch(p, zl::C2C, in_layout, out_layout, in0, in1, out0, out1, batch, direction == ZLFFT_BACKWARD,
                      stream);
    if (batch == 0) return ZLFFT_OK;
    const size_t need = macro_work_bytes(p, batch);
    if (need > kScratchBytes)
        return fail(ZLFFT_ERR_INVALID_ARG, "batch needs " + std::to_string(need) +
                                               " bytes of intermediates; the allocation-free call holds at most " +
                                               std::to_string(kScratchBytes));
    void* work = nullptr;
    if (int rc = scratch_ptr(p, &work)) return rc;
    // every intermediate element is written before it is read: nothing to clear
    return macro_cfft(p, direction == ZLFFT_BACKWARD, in_layout, in0, in1, out_layout, out0, out1, batch, stream,
                      work, 3);
}

int zlfft_rfft_forward_exec(zlfft_plan p, const void* in, int out_layout, void* out0, void* out1, size_t batch,
                            void* stream) {
    if (int rc = check_plan(p)) return rc;
    if (p->kind != ZLFFT_KIND_RFFT) return fail(ZLFFT_ERR_INVALID_ARG, "plan is not an RFFT plan");
    if (!valid_layout(out_layout)) return fail(ZLFFT_ERR_INVALID_ARG, "bad layout");
    if (batch && (!in || !out0 || (out_layout == ZLFFT_SOA && !out1))) return fail(ZLFFT_ERR_NULL, "NULL data pointer");
    return launch(p, zl::R2C, 0, out_layout, in, nullptr, out0, out1, batch, 0, stream);
}

int zlfft_rfft_forward_sqr_mag_exec(zlfft_plan p, const void* in, void* out, size_t batch, void* stream) {
    if (int rc = check_plan(p)) return rc;
    if (p->kind != ZLFFT_KIND_RFFT) return fail(ZLFFT_ERR_INVALID_ARG, "plan is not an RFFT plan");
    if (batch && (!in || !out)) return fail(ZLFFT_ERR_NULL, "NULL data pointer");
    return launch(p, zl::R2C_SQR, 0, 0, in, nullptr, out, nullptr, batch, 0, stream);
}

int zlfft_rfft_backward_exec(zlfft_plan p, int in_layout, const void* in0, const void* in1, void* out, size_t batch,
                             void* stream) {
    if (int rc = check_plan(p)) return rc;
    if (p->kind != ZLFFT_KIND_RFFT) return fail(ZLFFT_ERR_INVALID_ARG, "plan is not an RFFT plan");
    if (!valid_layout(in_layout)) return fail(ZLFFT_ERR_INVALID_ARG, "bad layout");
    if (batch && (!in0 || !out || (in_layout == ZLFFT_SOA && !in1))) return fail(ZLFFT_ERR_NULL, "NULL data pointer");
    return launch(p, zl::C2R, in_layout, 0, in0, in1, out, nullptr, batch, 1, stream);
}

int zlfft_cfft_execute_batch(zlfft_plan p, int direction, int in_layout, const void* in0, const void* in1,
                             int out_layout, void* out0, void* out1, size_t batch) {
    if (int rc = check_plan(p)) return rc;
    if (p->kind != ZLFFT_KIND_CFFT) return fail(ZLFFT_ERR_INVALID_ARG, "plan is not a CFFT plan");
    if (!valid_layout(in_layout) || !valid_layout(out_layout)) return fail(ZLFFT_ERR_INVALID_ARG, "bad layout");
    if (batch == 0) return ZLFFT_OK;
    if (!in0 || !out0 || (in_layout == ZLFFT_SOA && !in1) || (out_layout == ZLFFT_SOA && !out1))
        return fail(ZLFFT_ERR_NULL, "NULL data pointer");
    const size_t n = (size_t)1 << p->order, ss = p->ss;
    const bool si = in_layout == ZLFFT_SOA, so = out_layout == ZLFFT_SOA;
    const Buf bi{{in0, si ? in1 : nullptr}, {si ? ss * n : 2 * ss * n, si ? ss * n : 0}};
    const Buf bo{{out0, so ? out1 : nullptr}, {so ? ss * n : 2 * ss * n, so ? ss * n : 0}};
    return run_host(p, bi, bo, batch, [&](const void* i0, const void* i1, void* o0, void* o1, size_t cnt, hipStream_t s) {
        return zlfft_cfft_exec(p, direction, in_layout, i0, i1, out_layout, o0, o1, cnt, s);
    });
}

int zlfft_cfft_execute(zlfft_plan p, int direction, int in_layout, const void* in0, const void* in1, int out_layout,
                       void* out0, void* out1) {
    return zlfft_cfft_execute_batch(p, direction, in_layout, in0, in1, out_layout, out0, out1, 1);
}

int zlfft_rfft_forward_batch(zlfft_plan p, const void* in, int out_layout, void* out0, void* out1, size_t batch) {
    if (int rc = check_plan(p)) return rc;
    if (p->kind != ZLFFT_KIND_RFFT) return fail(ZLFFT_ERR_INVALID_ARG, "plan is not an RFFT plan");
    if (!valid_layout(out_layout)) return fail(ZLFFT_ERR_INVALID_ARG, "bad layout");
    if (batch == 0) return ZLFFT_OK;
    if (!in || !out0 || (out_layout == ZLFFT_SOA && !out1)) return fail(ZLFFT_ERR_NULL, "NULL data pointer");
    const size_t n = (size_t)1 << p->order, nb = n / 2 + 1, ss = p->ss;
    const bool so = out_layout == ZLFFT_SOA;
    const Buf bi{{in, nullptr}, {ss * n, 0}};
    const Buf bo{{out0, so ? out1 : nullptr}, {so ? ss * nb : 2 * ss * nb, so ? ss * nb : 0}};
    return run_host(p, bi, bo, batch, [&](const void* i0, const void*, void* o0, void* o1, size_t cnt, hipStream_t s) {
        return zlfft_rfft_forward_exec(p, i0, out_layout, o0, o1, cnt, s);
    });
}

int zlfft_rfft_forward(zlfft_plan p, const void* in, int out_layout, void* out0, void* out1) {
    return zlfft_rfft_forward_batch(p, in, out_layout, out0, out1, 1);
}

int zlfft_rfft_forward_sqr_mag_batch(zlfft_plan p, const void* in, void* out, size_t batch) {
    if (int rc = check_plan(p)) return rc;
    if (p->kind != ZLFFT_KIND_RFFT) return fail(ZLFFT_ERR_INVALID_ARG, "plan is not an RFFT plan");
    if (batch == 0) return ZLFFT_OK;
    if (!in || !out) return fail(ZLFFT_ERR_NULL, "NULL data pointer");
    const size_t n = (size_t)1 << p->order, nb = n / 2 + 1, ss = p->ss;
    const Buf bi{{in, nullptr}, {ss * n, 0}};
    const Buf bo{{out, nullptr}, {ss * nb, 0}};
    return run_host(p, bi, bo, batch, [&](const void* i0, const void*, void* o0, void*, size_t cnt, hipStream_t s) {
        return zlfft_rfft_forward_sqr_mag_exec(p, i0, o0, cnt, s);
    });
}

int zlfft_rfft_forward_sqr_mag(zlfft_plan p, const void* in, void* out) {
    return zlfft_rfft_forward_sqr_mag_batch(p, in, out, 1);
}

int zlfft_rfft_backward_batch(zlfft_plan p, int in_layout, const void* in0, const void* in1, void* out, size_t batch) {
    if (int rc = check_plan(p)) return rc;
    if (p->kind != ZLFFT_KIND_RFFT) return fail(ZLFFT_ERR_INVALID_ARG, "plan is not an RFFT plan");
    if (!valid_layout(in_layout)) return fail(ZLFFT_ERR_INVALID_ARG, "bad layout");
    if (batch == 0) return ZLFFT_OK;
    if (!in0 || !out || (in_layout == ZLFFT_SOA && !in1)) return fail(ZLFFT_ERR_NULL, "NULL data pointer");
    const size_t n = (size_t)1 << p->order, nb = n / 2 + 1, ss = p->ss;
    const bool si = in_layout == ZLFFT_SOA;
    const Buf bi{{in0, si ? in1 : nullptr}, {si ? ss * nb : 2 * ss * nb, si ? ss * nb : 0}};
    const Buf bo{{out, nullptr}, {ss * n, 0}};
    return run_host(p, bi, bo, batch, [&](const void* i0, const void* i1, void* o0, void*, size_t cnt, hipStream_t s) {
        return zlfft_rfft_backward_exec(p, in_layout, i0, i1, o0, cnt, s);
    });
}

int zlfft_rfft_backward(zlfft_plan p, int in_layout, const void* in0, const void* in1, void* out) {
    return zlfft_rfft_backward_batch(p, in_layout, in0, in1, out, 1);
}

int zlfft_rfft_stft_exec(zlfft_plan p, const void* signal, size_t hop, const void* window, size_t nframes, int output,
                         int out_layout, void* out0, void* out1, float db_floor, void* stream) {
    if (int rc = check_plan(p)) return rc;
    if (p->kind != ZLFFT_KIND_RFFT) return fail(ZLFFT_ERR_INVALID_ARG, "plan is not an RFFT plan");
    if (output != ZLFFT_STFT_BINS && output != ZLFFT_STFT_POWER && output != ZLFFT_STFT_DB)
        return fail(ZLFFT_ERR_INVALID_ARG, "output must be ZLFFT_STFT_BINS, _POWER or _DB");
    if (output == ZLFFT_STFT_BINS && !valid_layout(out_layout)) return fail(ZLFFT_ERR_INVALID_ARG, "bad layout");
    const int lout = output == ZLFFT_STFT_BINS ? out_layout : ZLFFT_AOS;
    if (output == ZLFFT_STFT_DB && !(db_floor > 0.f)) return fail(ZLFFT_ERR_INVALID_ARG, "db_floor must be > 0");
    if (hop > ((size_t)1 << 40)) return fail(ZLFFT_ERR_INVALID_ARG, "hop too large");
    if (nframes && (!signal || !out0 || (lout == ZLFFT_SOA && !out1))) return fail(ZLFFT_ERR_NULL, "NULL data pointer");
    if (nframes == 0) return ZLFFT_OK;
    const int db = output == ZLFFT_STFT_DB;
    const int kind = output == ZLFFT_STFT_BINS ? zl::R2C : zl::R2C_SQR;
    if (!p->macro) {
        // one fused launch: framing and window at load time
        const zl::KernelEntry& e = (*p->tab)[kind == zl::R2C ? zl::R2C_FRAMED : zl::R2C_SQR_FRAMED][zl::AOS][lout];
        if (!e.fn) return fail(ZLFFT_ERR_UNSUPPORTED, "no framed kernel for this order");
        const size_t groups = (nframes + e.tpw - 1) / e.tpw;
        if (groups > 0x7fffffffull) return fail(ZLFFT_ERR_INVALID_ARG, "too many frames for one launch");
        zl::KArgs a{};
        a.in0 = signal;
        a.in1 = window;
        a.out0 = out0;
        a.out1 = out1;
        a.tw = p->d_tw;
        a.rtw = p->d_rtw;
        a.batch = (long long)nframes;
        a.hop = (long long)hop;
        a.db = db;
        a.db_floor = db_floor;
        void* params[] = {&a};
        const int fk = kind == zl::R2C ? zl::R2C_FRAMED : zl::R2C_SQR_FRAMED;
        const int resident = (e.stream_fn[0] && !no_stream_kernels()) ? p->stream_blocks[fk][zl::AOS][lout] * p->cus : 0;
        if (resident > 0 && (nframes + e.s_tpw - 1) / e.s_tpw > (size_t)resident) {
            // persistent streaming kernel, as in launch()
            a.tw = p->d_img;
            a.sink = p->d_sink;
            a.cached_out = rfft_cached_out(nframes * (((size_t)1 << p->corder) + 1) * 2 * p->ss);
            ZL_HIP(plan_launch(p, reinterpret_cast<const void*>(e.stream_fn[0]), dim3((unsigned)resident),
                               dim3(e.s_wg), params, reinterpret_cast<hipStream_t>(stream)));
            return ZLFFT_OK;
        }
        ZL_HIP(plan_launch(p, reinterpret_cast<const void*>(e.fn), dim3((unsigned)groups), dim3(e.wg), params,
                           reinterpret_cast<hipStream_t>(stream)));
        return ZLFFT_OK;
    }
    // multi-pass sizes: the first HBM pass reads the frames (hop, window)
    // straight from the signal
    return launch_macro(p, kind, zl::FRAMED_IN, lout, signal, nullptr, out0, out1, nframes, 0, stream, db, db_floor,
                        (long long)hop, window);
}

int zlfft_plan_copy_rfft_twiddles(zlfft_plan p, float* host_out, size_t n_floats) {
    if (int rc = check_plan(p)) return rc;
    if (p->kind != ZLFFT_KIND_RFFT) return fail(ZLFFT_ERR_INVALID_ARG, "plan is not an RFFT plan");
    if (!host_out) return fail(ZLFFT_ERR_NULL, "host_out is NULL");
    if (p->f64) return fail(ZLFFT_ERR_UNSUPPORTED, "FP64 plans hold double twiddles; this call copies floats");
    const size_t need = 2 * (size_t)p->n_rtw;
    if (n_floats < need) return fail(ZLFFT_ERR_INVALID_ARG, "buffer too small");
    if (need) ZL_HIP(hipMemcpy(host_out, p->d_rtw, need * sizeof(float), hipMemcpyDeviceToHost));
    return ZLFFT_OK;
}

}  // extern "C"
