// bpg_device.h -- device helpers shared by the gfx950 kernels.
#pragma once

#include <hip/hip_runtime.h>

#include "bpg.h"
#include "bpg_kernels.h"

namespace bpg {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

// ---------------------------------------------------------------------------
// helpers
// ---------------------------------------------------------------------------

// float -> bf16 bits, round to nearest even
__device__ __forceinline__ unsigned short f2bf(float f) {
    unsigned u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}
__device__ __forceinline__ float bf2f(unsigned short h) { return __uint_as_float((unsigned)h << 16); }

// order-preserving float <-> uint encoding for atomicMax
__device__ __forceinline__ unsigned enc_max(float f) {
    unsigned u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float dec_max(unsigned e) {
    return __uint_as_float((e & 0x80000000u) ? (e & 0x7FFFFFFFu) : ~e);
}


// *p, read by the whole wave from one address, as scalar-register values:
// each word goes through v_readfirstlane (read through a generic pointer, a
// structure of kernel arguments would otherwise stay in vector registers)
template <typename T>
__device__ __forceinline__ T uniform_load(const T* p) {
    static_assert(sizeof(T) % 4 == 0, "whole words");
    constexpr int W = (int)(sizeof(T) / 4);
    const unsigned* src = reinterpret_cast<const unsigned*>(p);
    unsigned w[W];
#pragma unroll
    for (int i = 0; i < W; i++) w[i] = __builtin_amdgcn_readfirstlane(src[i]);
    T v;
    __builtin_memcpy(&v, w, sizeof(T));
    return v;
}

// Blocks are dealt round-robin to the 8 XCDs.  Renumber them so that the
// blocks resident on one XCD walk consecutive logical ids, i.e. share the
// same voxel chunk (and its lines in that XCD's L2).
__device__ __forceinline__ int xcd_logical_block(int per_xcd) {
    const int b = blockIdx.x;
    return (b & 7) * per_xcd + (b >> 3);
}

}  // namespace bpg
