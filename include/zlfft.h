/*
 * libzlfft: C-ABI of the MI355X (gfx950) FFT engine behind the zldsp::fft
 * CFFT<F> / RFFT<F> API.
 *
 * The reference (ZL-Audio/zldsp_fft) is a header-only C++ library with no C
 * ABI; its interface is the two class templates in src/zldsp_fft_cfft.hpp and
 * src/zldsp_fft_rfft.hpp.  Each entry point below names the reference member
 * it replaces.  Plain pointers and sizes only; `stream` is a hipStream_t
 * passed as void* (NULL = the default stream).
 *
 * Semantics (identical to the reference):
 *   forward  X[k] = sum_n x[n] e^{-2 pi i n k / N}, unnormalised;
 *   backward = unnormalised inverse, so backward(forward(x)) = N x (CFFT);
 *   RFFT forward: N reals -> N/2+1 bins, bins 0 and N/2 have imag exactly 0;
 *   RFFT backward: N/2+1 bins -> N reals = (N/2) irfft(X), with the
 *     imaginary parts of bins 0 and N/2 ignored;
 *   outputs in natural order; inputs are never written.
 * Layouts: ZLFFT_AOS = interleaved (std::complex<F>*), ZLFFT_SOA = separate
 * real and imaginary arrays (std::array<F*, 2>), reference
 * src/common/zldsp_fft_common_structure.hpp:8-33.
 *
 * Batched device entry points (`*_exec`) take device pointers, run
 * asynchronously on `stream`, and process `batch` transforms stored back to
 * back (CFFT: N complex per transform; RFFT real side: N floats; RFFT complex
 * side: N/2+1 complex; squared magnitude: N/2+1 floats).
 * CFFT orders up to 14 (RFFT up to 15) run as one LDS-resident kernel; larger
 * ones run as 2-3 HBM passes through a plan-owned workspace that the first
 * call of a given batch size allocates (hipMalloc) and later calls reuse.
 * Synchronous entry points (no `_exec` suffix) process one transform, accept
 * host or device pointers, and return when the result is in `out`; they are
 * what the drop-in C++ classes in zldsp_fft_cfft.hpp / zldsp_fft_rfft.hpp call.
 *
 * Errors: every int-returning function returns ZLFFT_OK (0) or a ZLFFT_ERR_*
 * code; zlfft_last_error() describes the last failure of the calling thread.
 * The reference has no error paths (its invalid inputs are undefined
 * behaviour); here they are reported.
 *
 * Threading: a plan may be used by one thread at a time (the reference's
 * instances are not shareable either, src/common/zldsp_fft_common_init.hpp:24-43);
 * distinct plans are independent.
 * Streams: the multi-pass workspace of a plan is handed from stream to stream:
 * a call on another stream than the plan's previous multi-pass call waits
 * (on the device) for that call's work, except inside a stream capture,
 * where the caller orders the captured work.  The synchronous and *_batch
 * calls wait for the work queued on the legacy default stream before they
 * touch device buffers.
 */
#ifndef ZLFFT_H
#define ZLFFT_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* The library is built with hidden symbol visibility; these declarations are
 * its exported interface. */
#if defined(__GNUC__) || defined(__clang__)
#pragma GCC visibility push(default)
#endif

#define ZLFFT_OK 0
#define ZLFFT_ERR_INVALID_ORDER 1
#define ZLFFT_ERR_NULL 2
#define ZLFFT_ERR_ALLOC 3
#define ZLFFT_ERR_HIP 4
#define ZLFFT_ERR_INVALID_ARG 5
#define ZLFFT_ERR_UNSUPPORTED 6

/* plan kinds */
#define ZLFFT_KIND_CFFT 0
#define ZLFFT_KIND_RFFT 1

/* arithmetic of a plan.  ZLFFT_PREC_F32 is the plain path (CFFT<float> /
 * RFFT<float>), for every kind, order and layout; ZLFFT_PREC_F64 is the same
 * over double (CFFT<double> / RFFT<double>, reference src/zldsp_fft_cfft.hpp:11,
 * src/zldsp_fft_rfft.hpp:11: the classes are templates on F), CFFT orders up to
 * ZLFFT_MAX_CFFT_ORDER_F64; its buffers hold double (AoS: std::complex<double>).
 * The MFMA precisions run each radix-16 pass of the
 * 4096-point CFFT (order 12, AoS in and out) as a 16x16 DFT-matrix product on
 * the matrix cores with FP32 accumulate; the operands are FP16 or BF16, whole
 * or split into 2 (hi + lo) or 3 (hi + mid + lo) parts.  Other orders, kinds or
 * layouts with an MFMA precision return ZLFFT_ERR_UNSUPPORTED. */
#define ZLFFT_PREC_F32 0
#define ZLFFT_PREC_F16_MFMA 1
#define ZLFFT_PREC_F16_MFMA_SPLIT 2
#define ZLFFT_PREC_BF16_MFMA 3
#define ZLFFT_PREC_BF16_MFMA_SPLIT 4
#define ZLFFT_PREC_BF16_MFMA_SPLIT3 5
#define ZLFFT_PREC_F64 6

/* user layouts */
#define ZLFFT_AOS 0
#define ZLFFT_SOA 1

/* CFFT directions */
#define ZLFFT_FORWARD 0
#define ZLFFT_BACKWARD 1

/* largest supported order: CFFT 2^ZLFFT_MAX_CFFT_ORDER points (RFFT one more) */
#define ZLFFT_MAX_CFFT_ORDER 28
#define ZLFFT_MAX_CFFT_ORDER_F64 27

typedef struct zlfft_plan_s *zlfft_plan;

/* Replaces the constructors CFFT<F>::CFFT(size_t order)
 * (reference src/zldsp_fft_cfft.hpp:19-23) and RFFT<F>::RFFT(size_t order)
 * (reference src/zldsp_fft_rfft.hpp:24-53).  Builds the device plan: kernel
 * choice and twiddle tables generated on the GPU in FP64 (rounded to FP32 for
 * the FP32 and MFMA precisions).
 * kind: ZLFFT_KIND_CFFT (size 2^order complex) or ZLFFT_KIND_RFFT
 * (size 2^order real, order >= 1). */
int zlfft_plan_create(zlfft_plan *plan, int kind, int precision, size_t order);

/* Replaces the destructor (device memory is released). */
int zlfft_plan_destroy(zlfft_plan plan);

/* Replace get_size() / get_order() (reference src/zldsp_fft_cfft.hpp:25-31,
 * src/zldsp_fft_rfft.hpp:55-61).  0 for a NULL plan. */
size_t zlfft_plan_get_size(zlfft_plan plan);
size_t zlfft_plan_get_order(zlfft_plan plan);

/* Batched device-pointer CFFT: the eight CFFT<F>::forward/backward overloads
 * (reference src/zldsp_fft_cfft.hpp:38-103) over `batch` transforms.
 * in1 / out1 are the imaginary arrays for ZLFFT_SOA and ignored for AoS. */
int zlfft_cfft_exec(zlfft_plan plan, int direction, int in_layout, const void *in0, const void *in1,
                    int out_layout, void *out0, void *out1, size_t batch, void *stream);

/* As zlfft_cfft_exec, allocating nothing.  No counterpart in the reference:
 * for callers that cannot call hipMalloc (code under hipGraph stream capture,
 * serving loops, harnesses that forbid allocation).  Multi-pass plans (CFFT
 * orders above 14) keep their intermediates in a library-owned device array
 * of at most 64 MiB per GPU (allocated with the library's code object, held
 * until the process ends) instead of the plan workspace: a call whose
 * intermediates exceed it (FP32: 2 * batch * 2^order * 8 bytes for orders
 * 21-28, half that for orders 15-20; FP64 twice as much) returns
 * ZLFFT_ERR_INVALID_ARG.  That array is shared by every call on the GPU, so
 * calls that overlap on one GPU (other streams, other plans, other threads)
 * race on it; calls queued on one stream do not.  One-kernel plans (orders up
 * to 14) need no scratch and run exactly as zlfft_cfft_exec.  The plan's
 * device must be the calling thread's current device. */
int zlfft_cfft_exec_noalloc(zlfft_plan plan, int direction, int in_layout, const void *in0, const void *in1,
                            int out_layout, void *out0, void *out1, size_t batch, void *stream);

/* Real-side buffers (`in` of forward, `out` of backward and sqr-mag) hold F
 * values: float for ZLFFT_PREC_F32 plans, double for ZLFFT_PREC_F64 plans. */

/* Batched device-pointer RFFT<F>::forward (reference src/zldsp_fft_rfft.hpp:68-79). */
int zlfft_rfft_forward_exec(zlfft_plan plan, const void *in, int out_layout, void *out0, void *out1,
                            size_t batch, void *stream);

/* Batched device-pointer RFFT<F>::forward_sqr_mag (reference src/zldsp_fft_rfft.hpp:104-109). */
int zlfft_rfft_forward_sqr_mag_exec(zlfft_plan plan, const void *in, void *out, size_t batch, void *stream);

/* Batched device-pointer RFFT<F>::backward (reference src/zldsp_fft_rfft.hpp:86-97). */
int zlfft_rfft_backward_exec(zlfft_plan plan, int in_layout, const void *in0, const void *in1, void *out,
                             size_t batch, void *stream);

/* Synchronous single-transform forms of the four calls above for host or
 * device pointers (the reference's calling convention: the call returns with
 * the result in out). */
int zlfft_cfft_execute(zlfft_plan plan, int direction, int in_layout, const void *in0, const void *in1,
                       int out_layout, void *out0, void *out1);
int zlfft_rfft_forward(zlfft_plan plan, const void *in, int out_layout, void *out0, void *out1);
int zlfft_rfft_forward_sqr_mag(zlfft_plan plan, const void *in, void *out);
int zlfft_rfft_backward(zlfft_plan plan, int in_layout, const void *in0, const void *in1, void *out);

/* Batched forms of the synchronous calls: `batch` transforms back to back in
 * host or device memory, returning when every output is in place.  Host
 * batches run as a pipeline of chunks through pinned staging on two streams,
 * so host-to-device copies, kernels and device-to-host copies overlap; pinned
 * host buffers (hipHostMalloc) are transferred directly, pageable ones are
 * copied through the staging.  No counterpart in the reference (its calls
 * take one transform); a loop of reference calls over a batch becomes one
 * call. */
int zlfft_cfft_execute_batch(zlfft_plan plan, int direction, int in_layout, const void *in0, const void *in1,
                             int out_layout, void *out0, void *out1, size_t batch);
int zlfft_rfft_forward_batch(zlfft_plan plan, const void *in, int out_layout, void *out0, void *out1, size_t batch);
int zlfft_rfft_forward_sqr_mag_batch(zlfft_plan plan, const void *in, void *out, size_t batch);
int zlfft_rfft_backward_batch(zlfft_plan plan, int in_layout, const void *in0, const void *in1, void *out,
                              size_t batch);

/* ---- ragged batches ----
 * Many transforms of different orders, kinds, directions and layouts run as
 * one persistent kernel, captured once in a hipGraph: zlfft_ragged_exec is a
 * single graph launch.  Each descriptor is one call of the reference API
 * (CFFT forward/backward, RFFT forward/backward/forward_sqr_mag) on its own
 * device buffers; the set is fixed at creation (the graph holds it).
 * CFFT orders 3..14 and RFFT orders 4..15. */
#define ZLFFT_FORWARD_SQR_MAG 2

typedef struct {
    int kind;      /* ZLFFT_KIND_CFFT or ZLFFT_KIND_RFFT */
    int order;     /* size 2^order (complex for CFFT, real for RFFT) */
    int direction; /* ZLFFT_FORWARD, ZLFFT_BACKWARD; RFFT also ZLFFT_FORWARD_SQR_MAG */
    int layout;    /* ZLFFT_AOS or ZLFFT_SOA: both sides of a CFFT, the complex side of an RFFT */
    const void *in0, *in1; /* device input (in1: SoA imaginary part) */
    void *out0, *out1;     /* device output (out1: SoA imaginary part) */
} zlfft_ragged_desc;

typedef struct zlfft_ragged_s *zlfft_ragged;

/* Builds the work queue, twiddles and the captured graph (no counterpart in
 * the reference, whose calls are one transform each). */
int zlfft_ragged_create(zlfft_ragged *ragged, const zlfft_ragged_desc *descs, size_t count);
/* As zlfft_ragged_create, with the points per thread (log2: 3, 4 or 5) of each
 * CFFT order 0..14 given in logr[15] instead of the tuned table
 * (zlfft_tune.txt next to the library, or $ZLFFT_TUNE_FILE); for the
 * autotuner.  Accepted: 8 points for CFFT orders 3-7, 16 for 4-12, 32 for
 * 8-14 (an RFFT of order n uses CFFT order n - 1); ZLFFT_ERR_INVALID_ARG
 * otherwise. */
int zlfft_ragged_create_config(zlfft_ragged *ragged, const zlfft_ragged_desc *descs, size_t count,
                               const int *logr);
/* Runs every transform of the set: one hipGraphLaunch on `stream`. */
int zlfft_ragged_exec(zlfft_ragged ragged, void *stream);
int zlfft_ragged_destroy(zlfft_ragged ragged);

/* ---- STFT front-end ----
 * No counterpart in the reference: it replaces the caller's framing and
 * windowing loop in front of RFFT<F>::forward / forward_sqr_mag
 * (reference src/zldsp_fft_rfft.hpp:68-79, 104-109), as an analyzer
 * (spectrum display) would write it.  Frame f = 0..nframes-1 is
 * signal[f * hop + n], n < N = 2^order (device memory, F values), multiplied
 * by window[n] (device memory, N F values; NULL for none), and transformed:
 *   ZLFFT_STFT_BINS : N/2+1 complex bins per frame in out_layout (as forward)
 *   ZLFFT_STFT_POWER: N/2+1 squared magnitudes per frame (as forward_sqr_mag)
 *   ZLFFT_STFT_DB   : 10 log10(max(|X|^2, db_floor)) per bin (db_floor > 0)
 * LDS-resident orders run as one kernel that frames and windows at load time;
 * larger orders read the windowed frames in their first HBM pass.
 * signal must hold (nframes - 1) * hop + N values (the call cannot check it). */
#define ZLFFT_STFT_BINS 0
#define ZLFFT_STFT_POWER 1
#define ZLFFT_STFT_DB 2
int zlfft_rfft_stft_exec(zlfft_plan plan, const void *signal, size_t hop, const void *window, size_t nframes,
                         int output, int out_layout, void *out0, void *out1, float db_floor, void *stream);

/* Copies the plan's RFFT post/pre twiddles (wc, ws) for k = 1..N/4 as 2*(N/4)
 * floats to host memory (reference src/common/zldsp_fft_common_init.hpp:347-388);
 * for tests that compare them with the FP64 oracle.  FP32 plans only
 * (ZLFFT_ERR_UNSUPPORTED for FP64 plans). */
int zlfft_plan_copy_rfft_twiddles(zlfft_plan plan, float *host_out, size_t n_floats);

/* Description of the last error on the calling thread ("" if none). */
const char *zlfft_last_error(void);

/* Library version as MAJOR * 10000 + MINOR * 100 + PATCH. */
int zlfft_version(void);

#if defined(__GNUC__) || defined(__clang__)
#pragma GCC visibility pop
#endif

#ifdef __cplusplus
}
#endif

#endif /* ZLFFT_H */
