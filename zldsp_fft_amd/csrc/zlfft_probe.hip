// Device-copy reference kernels for the benchmark: the rate of a plain copy
// of the same bytes, timed the same way, is the practical ceiling of a
// memory-bound transform of that size.  Built into libzlprobe.so, which only
// bench.py and its tests load; the transform library does not use it.
//
//   structure 0: one 32 KiB chunk per workgroup (256 threads x 16 complex
//                values), as many workgroups as chunks: the fastest copy form
//                measured on MI355X.
//   structure 1: the persistent streaming kernels' form: one 256-thread
//                workgroup per resident slot (per_cu per CU), each looping
//                over chunks with the next chunk's values loaded into
//                registers while the current one is stored.
// Both use non-temporal loads and stores, as the transform kernels do.
#include <hip/hip_runtime.h>

#include <cstddef>

namespace {

typedef float zp_f2 __attribute__((ext_vector_type(2)));
constexpr int kWG = 256, kR = 16;
constexpr size_t kChunk = sizeof(zp_f2) * kWG * kR;  // bytes per chunk

__global__ void __launch_bounds__(kWG) copy_grid(const zp_f2* __restrict__ in, zp_f2* __restrict__ out) {
    const size_t base = (size_t)blockIdx.x * (kWG * kR) + threadIdx.x;
    zp_f2 v[kR];
#pragma unroll
    for (int j = 0; j < kR; ++j) v[j] = __builtin_nontemporal_load(in + base + kWG * j);
#pragma unroll
    for (int j = 0; j < kR; ++j) __builtin_nontemporal_store(v[j], out + base + kWG * j);
}

__global__ void __launch_bounds__(kWG) copy_stream(const zp_f2* __restrict__ in, zp_f2* __restrict__ out,
                                                   long long nchunks) {
    long long c = blockIdx.x;
    zp_f2 p[kR];
    // a chunk past the end is read from chunk 0 (never stored): every
    // iteration issues the same loads
    auto load = [&](long long k) {
        const zp_f2* b = in + (size_t)(k < nchunks ? k : 0) * (kWG * kR) + threadIdx.x;
#pragma unroll
        for (int j = 0; j < kR; ++j) p[j] = __builtin_nontemporal_load(b + kWG * j);
    };
    load(c);
    for (; c < nchunks; c += gridDim.x) {
        zp_f2 v[kR];
#pragma unroll
        for (int j = 0; j < kR; ++j) v[j] = p[j];
        load(c + gridDim.x);
        zp_f2* o = out + (size_t)c * (kWG * kR) + threadIdx.x;
#pragma unroll
        for (int j = 0; j < kR; ++j) __builtin_nontemporal_store(v[j], o + kWG * j);
    }
}

}  // namespace

extern "C" {

// Bytes of one chunk: a copy moves a whole number of chunks.
__attribute__((visibility("default"))) size_t zlprobe_chunk_bytes(void) { return kChunk; }

// Copies floor(bytes / chunk) chunks from in to out (device memory, 8-byte
// aligned) on `stream`.  structure 0 or 1 (above); per_cu: resident
// workgroups per CU of structure 1.  Returns 0, or -1 for invalid arguments
// or a launch error.
__attribute__((visibility("default"))) int zlprobe_copy(const void* in, void* out, size_t bytes, int structure,
                                                        int per_cu, void* stream) {
    const size_t nchunks = bytes / kChunk;
    if (!in || !out || nchunks == 0 || nchunks > 0x7fffffffull || (structure != 0 && structure != 1)) return -1;
    hipStream_t s = static_cast<hipStream_t>(stream);
    const zp_f2* src = static_cast<const zp_f2*>(in);
    zp_f2* dst = static_cast<zp_f2*>(out);
    if (structure == 0) {
        copy_grid<<<dim3((unsigned)nchunks), dim3(kWG), 0, s>>>(src, dst);
    } else {
        int dev = 0, cus = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || per_cu < 1)
            return -1;
        const size_t grid = (size_t)cus * (size_t)per_cu;
        copy_stream<<<dim3((unsigned)(grid < nchunks ? grid : nchunks)), dim3(kWG), 0, s>>>(src, dst,
                                                                                          (long long)nchunks);
    }
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // extern "C"
