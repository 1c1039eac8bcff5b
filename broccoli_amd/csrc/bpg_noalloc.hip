// bpg_noalloc.hip -- the device scratch of bpg_plan_permutation_phase_noalloc
// and the two small kernels that entry point adds (clear, quantise).
//
// The scratch is one file-scope device array per GPU, BPG_SCRATCH_BYTES
// long, held from the library's load until the process ends.  Host code
// never passes the array's name to a kernel (on the host the name is only a
// stand-in): scratch_address() takes the device address once per device,
// with that device current, when a context is created, and every use goes
// through that address.  All calls on one GPU share the array, so calls of
// the no-allocation entry point that overlap on one GPU race.
#include <hip/hip_runtime.h>

#include <climits>
#include <map>
#include <mutex>

#include "bpg_kernels.h"

static __device__ __align__(256) unsigned char scratch[bpg::BPG_SCRATCH_BYTES];

namespace bpg {

// The device address of this GPU's copy of the array, taken once per device
// (the caller makes `device` current) and kept: each GPU holds its own copy
// at its own address.
hipError_t scratch_address(int device, void** out) {
    static std::mutex mu;
    static std::map<int, void*> addr;
    std::lock_guard<std::mutex> lock(mu);
    auto it = addr.find(device);
    if (it == addr.end()) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) return e != hipSuccess ? e : hipErrorInvalidSymbol;
        it = addr.emplace(device, p).first;
    }
    *out = it->second;
    return hipSuccess;
}

__global__ __launch_bounds__(256) void k_fill_u32(unsigned* p, size_t n, unsigned v) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

hipError_t launch_fill_u32(unsigned* p, size_t n, unsigned v, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_fill_u32, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, n, v);
    return hipGetLastError();
}

// The null value the reference stores for a maximum m (KM:299, LIB:10988):
// (float)(int)(m * 10000.0f) / 10000.0f, the int conversion saturating
// (NaN: INT_MIN); an empty mask (m = -inf) gives the float of the bit
// pattern of -1e6f read as an int, / 10000 (LIB:10971).  The same
// arithmetic as the host path of bpg_plan_permutation_phase: one FP32
// product, a correctly rounded FP32 quotient.
__global__ __launch_bounds__(256) void k_quantize_max(const float* m, int K, float* out, float empty) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= K) return;
    const float t = m[k];
    if (!(t > -INFINITY)) {
        out[k] = empty;
        return;
    }
    const float x = t * 10000.0f;
    int q;
    if (x != x) q = INT_MIN;
    else if (x >= 2147483648.0f) q = INT_MAX;
    else if (x <= -2147483648.0f) q = INT_MIN;
    else q = (int)x;
    out[k] = __fdiv_rn((float)q, 10000.0f);
}

hipError_t launch_quantize_max(const float* m, int K, float* out, float empty, hipStream_t s) {
    if (K <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_quantize_max, dim3((K + 255) / 256), dim3(256), 0, s, m, K, out, empty);
    return hipGetLastError();
}

}  // namespace bpg
