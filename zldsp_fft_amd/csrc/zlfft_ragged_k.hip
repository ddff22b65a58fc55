// The ragged-batch kernels (one translation unit: the kernel dispatches over
// every order and kind).
#include "zlfft_ragged.hpp"

namespace zl {

__global__ void __launch_bounds__(RWG) fft_ragged_kernel(RArgs a) {
    __shared__ float2 lds[ragged_lds_elems()];
    __shared__ int next;
    for (;;) {
        if (threadIdx.x == 0) next = (int)atomicAdd(a.counter, 1u);
        __syncthreads();
        const int i = next;
        __syncthreads();  // everyone has read `next` before it changes
        if (i >= a.nitems) break;
        const RItem it = a.items[i];
        const float2* tw = a.tw + a.tw_off[it.corder];
        const float2* rtw = a.rtw + a.rtw_off[it.corder];
        if (a.logr[it.corder] == 3)
            ragged_item<3>(a.desc, it, tw, rtw, lds);
        else
            ragged_item<4>(a.desc, it, tw, rtw, lds);
    }
}

// zeroes the queue counter (a kernel node, so the graph holds no memset node)
__global__ void ragged_reset_kernel(unsigned* counter) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *counter = 0u;
}

}  // namespace zl

void* zl_ragged_kernel() { return reinterpret_cast<void*>(zl::fft_ragged_kernel); }
void* zl_ragged_reset_kernel() { return reinterpret_cast<void*>(zl::ragged_reset_kernel); }
