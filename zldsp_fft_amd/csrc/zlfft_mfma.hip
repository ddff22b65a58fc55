// MFMA split-precision kernels for the 4096-point CFFT.
#include "zlfft_mfma.hpp"
#include "zlfft_registry.hpp"

#define ZL_MFMA_CASES(K)                                               \
    case 2: return reinterpret_cast<void*>(K<MF_F16, 1, 0>);           \
    case 3: return reinterpret_cast<void*>(K<MF_F16, 1, 1>);           \
    case 4: return reinterpret_cast<void*>(K<MF_F16, 2, 0>);           \
    case 5: return reinterpret_cast<void*>(K<MF_F16, 2, 1>);           \
    case 6: return reinterpret_cast<void*>(K<MF_BF16, 1, 0>);          \
    case 7: return reinterpret_cast<void*>(K<MF_BF16, 1, 1>);          \
    case 8: return reinterpret_cast<void*>(K<MF_BF16, 2, 0>);          \
    case 9: return reinterpret_cast<void*>(K<MF_BF16, 2, 1>);          \
    case 10: return reinterpret_cast<void*>(K<MF_BF16, 3, 0>);         \
    case 11: return reinterpret_cast<void*>(K<MF_BF16, 3, 1>);         \
    default: return nullptr;

// kernel for precision code prec (1..5, see zlfft.h) and direction: one
// transform per workgroup, or the persistent streaming form
void* zl_mfma_kernel(int prec, int backward, int streaming) {
    using namespace zl;
    if (streaming) {
        switch (prec * 2 + (backward ? 1 : 0)) { ZL_MFMA_CASES(fft4096_mfma_stream_kernel) }
    }
    switch (prec * 2 + (backward ? 1 : 0)) { ZL_MFMA_CASES(fft4096_mfma_kernel) }
}
