// bpg_kernels.hip -- gfx950 (CDNA4) kernels of the permutation-test GLM.
//
// The reference evaluates one permutation per launch, one voxel per
// work-item, re-reading the data twice per permutation
// (code/Kernels/kernelStatistics2.cpp:1562-1669).  Here the permuted fit of
// every voxel for a whole tile of permutations is one contraction
//     Z = G Y*      G: (K * regressors) x N   permutation operand
//                   Y*: N x V                  residualised data (in HBM once)
// followed, in registers, by the statistic, the per-permutation running
// maximum (max-T) and an optional supra-threshold bitmask for cluster
// inference.  Two arithmetic forms:
//   - split-BF16 MFMA: G = Ghi + Glo, Y = Yhi + Ylo (bf16 each), products
//     Ghi*Yhi + Ghi*Ylo + Glo*Yhi on v_mfma_f32_32x32x16_bf16 with FP32
//     accumulation (the Glo term is dropped when G holds exact +-1 signs);
//   - plain FP32 on the vector ALUs.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstring>

#include "bpg.h"
#include "bpg_device.h"
#include "bpg_kernels.h"

namespace bpg {


// ---------------------------------------------------------------------------
// nuisance residualisation + operand split (once per statistical map)
// ---------------------------------------------------------------------------
// y* = y - Qn (Qn' y)  (the projector I - W2 pinv(W2) of LIB:16684-16689,
// applied in FP64), yy = |y*|^2 in FP64, operand row s = dscale_s * y*_s.
// One thread per voxel.  Qn is zero-padded to QB columns and, like dscale,
// to a multiple of 16 rows, so the loops carry no guards and the scalar
// loads of a row batch; subjects are read 8 at a time (8 loads in flight
// per thread), and each group of 8 operand rows leaves as one 16-byte hi
// fragment and one 16-byte lo fragment (a wave writes 512 contiguous bytes
// per store).
// (the indices are clamped into range and the values selected afterwards,
// so the 8 loads issue together; Vc >= 1 whenever the kernel runs)
__device__ __forceinline__ void load8(const TransformArgs& a, int v, bool in, int s0, double (&y)[8]) {
    const int vc = min(v, a.Vc - 1);
    float f[8];
#pragma unroll
    for (int e = 0; e < 8; e++) f[e] = a.Yc[(size_t)min(s0 + e, a.N - 1) * a.Vc + vc];
#pragma unroll
    for (int e = 0; e < 8; e++) y[e] = (in && s0 + e < a.N) ? (double)f[e] : 0.0;
}

template <int QB>
__device__ __forceinline__ void transform_voxel(const TransformArgs& a, int v) {
    const bool in = v < a.Vc;
    double w[QB > 0 ? QB : 1];
#pragma unroll
    for (int j = 0; j < QB; j++) w[j] = 0.0;
    if (QB > 0 && in) {
        for (int s0 = 0; s0 < a.N; s0 += 8) {
            double y[8];
            load8(a, v, true, s0, y);
#pragma unroll
            for (int e = 0; e < 8; e++) {
                const double* q = a.Qn + (size_t)(s0 + e) * QB;
#pragma unroll
                for (int j = 0; j < QB; j++) w[j] += q[j] * y[e];
            }
        }
    }
    // o[e] = dscale_s * y*_s for the 8 subjects from s0 (zero beyond N)
    auto residual8 = [&](int s0, double (&o)[8], double& yy) {
        load8(a, v, in, s0, o);
#pragma unroll
        for (int e = 0; e < 8; e++) {
            const double* q = a.Qn + (size_t)(s0 + e) * QB;
            double ys = o[e];
#pragma unroll
            for (int j = 0; j < QB; j++) ys -= q[j] * w[j];
            yy += ys * ys;
            o[e] = a.dscale[s0 + e] * ys;
        }
    };
    double yy = 0.0;
    if (a.Yb) {
        const int vt = v >> 5, col = v & 31;
        uint4* Yb4 = reinterpret_cast<uint4*>(a.Yb);
        for (int kk = 0; kk < a.KK; kk++)
#pragma unroll
            for (int h = 0; h < 2; h++) {
                double o[8];
                residual8(kk * 16 + h * 8, o, yy);
                // hi, (mid,) lo: each the bf16 rounding of what the earlier
                // parts leave of o
                unsigned hi[4], mi[4], lo[4];
#pragma unroll
                for (int e = 0; e < 8; e += 2) {
                    unsigned short pc[2][3];
#pragma unroll
                    for (int d = 0; d < 2; d++) {
                        double r = o[e + d];
                        pc[d][0] = f2bf((float)r);
                        r -= (double)bf2f(pc[d][0]);
                        pc[d][1] = f2bf((float)r);
                        r -= (double)bf2f(pc[d][1]);
                        pc[d][2] = f2bf((float)r);
                    }
                    hi[e / 2] = pc[0][0] | ((unsigned)pc[1][0] << 16);
                    mi[e / 2] = pc[0][1] | ((unsigned)pc[1][1] << 16);
                    lo[e / 2] = pc[0][2] | ((unsigned)pc[1][2] << 16);
                }
                const size_t base = (((size_t)vt * a.KK + kk) * a.npy) * 64 + h * 32 + col;
                Yb4[base] = make_uint4(hi[0], hi[1], hi[2], hi[3]);
                if (a.npy == 3) {
                    Yb4[base + 64] = make_uint4(mi[0], mi[1], mi[2], mi[3]);
                    Yb4[base + 128] = make_uint4(lo[0], lo[1], lo[2], lo[3]);
                } else {
                    Yb4[base + 64] = make_uint4(mi[0], mi[1], mi[2], mi[3]);  // the two-part lo
                }
            }
    } else if (a.Ym) {
        // FP32-MFMA fragments: per group of 8 subjects, lane (h, col) holds
        // subjects 8g + h, 8g + 2 + h, 8g + 4 + h, 8g + 6 + h
        const int vt = v >> 5, col = v & 31, KG = (a.N + 7) / 8;
        float4* Ym4 = reinterpret_cast<float4*>(a.Ym);
        for (int g = 0; g < KG; g++) {
            double o[8];
            residual8(8 * g, o, yy);
            const size_t base = ((size_t)vt * KG + g) * 64 + col;
            Ym4[base] = make_float4((float)o[0], (float)o[2], (float)o[4], (float)o[6]);
            Ym4[base + 32] = make_float4((float)o[1], (float)o[3], (float)o[5], (float)o[7]);
        }
    } else {
        for (int s0 = 0; s0 < a.NS; s0 += 8) {
            double o[8];
            residual8(s0, o, yy);
#pragma unroll
            for (int e = 0; e < 8; e++)
                if (s0 + e < a.NS) {
                    if (a.Y64) a.Y64[(size_t)(s0 + e) * a.Vpad + v] = o[e];
                    else a.Y32[(size_t)(s0 + e) * a.Vpad + v] = (float)o[e];
                }
        }
    }
    // padding voxels carry a NaN norm: their statistic is NaN, which the
    // running max (fmax) ignores
    a.yy[v] = in ? yy : __builtin_nan("");
}

template <int QB>
__global__ __launch_bounds__(256) void k_transform(TransformArgs a) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= a.Vpad) return;
    transform_voxel<QB>(a, v);
}

// Several studies' transforms (one nuisance width) in one launch, so that
// the small grids of small studies fill the chip together: block b belongs
// to study j with first[j] <= b < first[j + 1] (a binary search of the
// uniform table), whose arguments it reads as scalar values.
template <int QB>
__global__ __launch_bounds__(256) void k_transform_batch(const TransformArgs* tab, const int* first, int n) {
    const int b = blockIdx.x;
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (__builtin_amdgcn_readfirstlane(first[mid]) <= b) lo = mid;
        else hi = mid - 1;
    }
    const TransformArgs a = uniform_load(tab + lo);
    const int v = (b - __builtin_amdgcn_readfirstlane(first[lo])) * 256 + (int)threadIdx.x;
    if (v >= a.Vpad) return;
    transform_voxel<QB>(a, v);
}

int transform_width(int q) {
    if (q <= 0) return 0;
    if (q <= 4) return 4;
    if (q <= 8) return 8;
    if (q <= 16) return 16;
    return q <= 24 ? 24 : -1;
}

hipError_t launch_transform(const TransformArgs& a, hipStream_t s) {
    const int blocks = (a.Vpad + 255) / 256;
    switch (transform_width(a.q)) {
        case 0: hipLaunchKernelGGL(k_transform<0>, dim3(blocks), dim3(256), 0, s, a); break;
        case 4: hipLaunchKernelGGL(k_transform<4>, dim3(blocks), dim3(256), 0, s, a); break;
        case 8: hipLaunchKernelGGL(k_transform<8>, dim3(blocks), dim3(256), 0, s, a); break;
        case 16: hipLaunchKernelGGL(k_transform<16>, dim3(blocks), dim3(256), 0, s, a); break;
        case 24: hipLaunchKernelGGL(k_transform<24>, dim3(blocks), dim3(256), 0, s, a); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_transform_batch(const TransformArgs* tab, const int* first, int n, int blocks, int QB,
                                  hipStream_t s) {
    if (n <= 0 || blocks <= 0) return hipSuccess;
    switch (QB) {
        case 0: hipLaunchKernelGGL(k_transform_batch<0>, dim3(blocks), dim3(256), 0, s, tab, first, n); break;
        case 4: hipLaunchKernelGGL(k_transform_batch<4>, dim3(blocks), dim3(256), 0, s, tab, first, n); break;
        case 8: hipLaunchKernelGGL(k_transform_batch<8>, dim3(blocks), dim3(256), 0, s, tab, first, n); break;
        case 16: hipLaunchKernelGGL(k_transform_batch<16>, dim3(blocks), dim3(256), 0, s, tab, first, n); break;
        case 24: hipLaunchKernelGGL(k_transform_batch<24>, dim3(blocks), dim3(256), 0, s, tab, first, n); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// permutation operand
// ---------------------------------------------------------------------------
// Row i of a 32-row MFMA subtile lands in lane half h = (i>>2)&1, slot
// j = 4*(i>>3) + (i&3) of the 32x32 accumulator.  Slots are grouped so one
// lane holds PP consecutive slots per permutation: permutation
// h*(16/PP) + j/PP, regressor j%PP.  Each lane then owns whole permutations
// of one voxel and the statistic needs no cross-lane traffic.
__global__ __launch_bounds__(256) void k_operand_bf16(OperandArgs a) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t total = (size_t)a.n_sub * a.KK * 64;
    if (t >= total) return;
    const int lane = (int)(t & 63);
    const size_t rest = t >> 6;
    const int kk = (int)(rest % a.KK);
    const int st = (int)(rest / a.KK);
    const int i = lane & 31, hl = lane >> 5;
    const int j = 4 * (i >> 3) + (i & 3);
    const int hh = (i >> 2) & 1;
    const int per = 16 / a.PP;  // whole permutations per lane; slots per*PP..15 stay zero
    const int k = st * (2 * per) + hh * per + j / a.PP;
    const int reg = j % a.PP;
    const bool live = j < per * a.PP && k < a.K && reg < a.pfit;
    const int s0 = kk * 16 + hl * 8;
    // gathers from clamped (in-range) indices, selected afterwards: the 8
    // loads issue together (the host checked every index against N)
    const size_t row = (size_t)min(k, a.K - 1) * a.N;
    const size_t col = (size_t)min(reg, a.pfit - 1) * a.N;
    double g[8];
    if (a.sign) {
#pragma unroll
        for (int e = 0; e < 8; e++) g[e] = (double)a.sign[row + min(s0 + e, a.N - 1)];
    } else {
        unsigned short ix[8];
#pragma unroll
        for (int e = 0; e < 8; e++) ix[e] = a.perm[row + min(s0 + e, a.N - 1)];
        if (a.bad_index) {
            bool bad = false;
#pragma unroll
            for (int e = 0; e < 8; e++)
                if (ix[e] >= a.N) {
                    bad = true;
                    ix[e] = 0;
                }
            if (bad) atomicAdd(a.bad_index, 1u);
        }
#pragma unroll
        for (int e = 0; e < 8; e++) g[e] = a.Qt[(size_t)ix[e] + col];
    }
    unsigned hi[4], lo[4];
#pragma unroll
    for (int e = 0; e < 8; e += 2) {
        double val[2];
#pragma unroll
        for (int d = 0; d < 2; d++) val[d] = (live && s0 + e + d < a.N) ? g[e + d] : 0.0;
        const unsigned short h0 = f2bf((float)val[0]), h1 = f2bf((float)val[1]);
        const unsigned short l0 = f2bf((float)(val[0] - (double)bf2f(h0)));
        const unsigned short l1 = f2bf((float)(val[1] - (double)bf2f(h1)));
        hi[e / 2] = h0 | ((unsigned)h1 << 16);
        lo[e / 2] = l0 | ((unsigned)l1 << 16);
    }
    // fragment (st, kk): 64 lanes of hi, then 64 lanes of lo
    uint4* G4 = reinterpret_cast<uint4*>(a.Gb) + (rest * 2) * 64 + lane;
    G4[0] = make_uint4(hi[0], hi[1], hi[2], hi[3]);
    G4[64] = make_uint4(lo[0], lo[1], lo[2], lo[3]);
}

hipError_t launch_operand_bf16(const OperandArgs& a, hipStream_t s) {
    const size_t total = (size_t)a.n_sub * a.KK * 64;
    const unsigned blocks = (unsigned)((total + 255) / 256);
    hipLaunchKernelGGL(k_operand_bf16, dim3(blocks), dim3(256), 0, s, a);
    return hipGetLastError();
}

// G32[(k*pfit + i)*NS + s]; rows are padded to the launch's permutation
// tile with zeros (the buffer is sized for that, K here is the padded count
// in a.n_sub).
__global__ __launch_bounds__(256) void k_operand_f32(OperandArgs a) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t total = (size_t)a.n_sub * a.pfit * a.NS;
    if (t >= total) return;
    const int s = (int)(t % a.NS);
    const size_t row = t / a.NS;
    const int i = (int)(row % a.pfit);
    const int k = (int)(row / a.pfit);
    float val = 0.0f;
    if (k < a.K && s < a.N) {
        if (a.sign) val = a.sign[(size_t)k * a.N + s];
        else {
            unsigned ix = a.perm[(size_t)k * a.N + s];
            if (a.bad_index && ix >= (unsigned)a.N) {
                atomicAdd(a.bad_index, 1u);
                ix = 0;
            }
            val = (float)a.Qt[(size_t)ix + (size_t)i * a.N];
        }
    }
    a.G32[t] = val;
}

hipError_t launch_operand_f32(const OperandArgs& a, hipStream_t s) {
    const size_t total = (size_t)a.n_sub * a.pfit * a.NS;
    const unsigned blocks = (unsigned)((total + 255) / 256);
    hipLaunchKernelGGL(k_operand_f32, dim3(blocks), dim3(256), 0, s, a);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// statistic epilogue (shared by both forms)
// ---------------------------------------------------------------------------
// SSE = |y*|^2 - |z|^2 with the subtraction in FP64 (|y*|^2 is exact in
// FP64; the difference is then rounded once).
// v_rsq_f32 / v_rcp_f32 directly: SSE is a sum of squares of the order of N,
// never denormal, so the denormal-scaling wrappers of rsqrtf / division are
// not needed.
__device__ __forceinline__ float stat_t(float z0, float sse, float sqrt_dof) {
    return z0 * sqrt_dof * __builtin_amdgcn_rsqf(sse);
}
__device__ __forceinline__ float stat_f(float num, float sse, float f_scale) {
    return num * f_scale * __builtin_amdgcn_rcpf(sse);
}

// split-BF16 MFMA contraction: bpg_contract_bf16.h, instantiated per KK in
// bpg_contract_bf16_*.hip
template <int KK>
hipError_t dispatch_bf16_pp(const ContractArgs& a, const Bf16Launch& L, int PP, bool sign, hipStream_t s, int* occ);


Bf16Launch plan_contract_bf16(int K, int PP, int KK, int Vpad, bool sign, int slots_per_xcd, int variant,
                              int chunk_pow) {
    Bf16Launch L = plan_contract_tiles(K, PP, bf16_rows_per_wave_v(KK, sign, variant), Vpad, slots_per_xcd, chunk_pow);
    L.variant = variant;
    return L;
}

Bf16Launch plan_contract_tiles(int K, int PP, int R, int Vpad, int slots_per_xcd, int chunk_pow) {
    Bf16Launch L{};
    L.n_sub = (K + 2 * (16 / PP) - 1) / (2 * (16 / PP));
    L.n_vt = Vpad / 32;
    L.n_mgroups = (L.n_sub + 4 * R - 1) / (4 * R);
    // Voxel chunks.  Blocks of one chunk length run in rounds of
    // slots_per_xcd on each XCD, and the last round of a grid that does not
    // fill it leaves most of the XCD idle for a whole block (1/17 of the
    // kernel with 4116 blocks of one block per CU).  Among the chunk counts
    // that give at least BPG_TARGET_BLOCKS/2 blocks (enough for the dispatcher
    // to even out blocks that run slower), and at most BPG_MAX_VCHUNKS partial
    // maxima per permutation, take the one with the fewest rounds x chunk
    // length, counting a block's operand loads as 2 voxel tiles.
    const int slots = slots_per_xcd > 0 ? slots_per_xcd : 1;
    const int maxc = std::max(1, std::min(BPG_MAX_VCHUNKS, L.n_vt));
    long best = -1;
    int len = std::max(1, L.n_vt);
    for (int nvc = 1; nvc <= maxc; nvc++) {
        if ((long)L.n_mgroups * nvc < BPG_TARGET_BLOCKS / 2 && nvc < maxc) continue;
        const int s = std::max(1, (L.n_vt + nvc - 1) / nvc);
        const long blocks = (long)L.n_mgroups * ((L.n_vt + s - 1) / s);
        const long rounds = ((blocks + 7) / 8 + slots - 1) / slots;
        const long cost = rounds * (s + 2);
        if (best < 0 || cost < best) {
            best = cost;
            len = s;
        }
    }
    // the autotuner's scaling of the chunk length, within the chunk limit
    if (chunk_pow > 0) len = std::min(std::max(1, L.n_vt), len << chunk_pow);
    if (chunk_pow < 0) len = std::max((L.n_vt + maxc - 1) / maxc, std::max(1, len >> -chunk_pow));
    L.vt_per_chunk = len;
    L.n_vchunks = (L.n_vt + len - 1) / len;
    L.per_xcd = (L.n_mgroups * L.n_vchunks + 7) / 8;
    return L;
}

int contract_bf16_blocks_per_cu(int KK, int PP, bool sign, bool ftest, bool out, int variant) {
    ContractArgs a{};
    a.ftest = ftest;
    int occ = out ? 1 : 0;
    Bf16Launch L{};
    L.variant = variant;
    if (launch_contract_bf16(a, L, KK, PP, sign, nullptr, &occ) != hipSuccess || occ < 1) return 1;
    return occ;
}

template <int KK>
hipError_t dispatch_persist_pp(const BatchClass& c, int PP, bool sign, int variant, bool excl, int grid,
                               hipStream_t s, int* occ);

hipError_t launch_contract_bf16_persist(const BatchClass& c, int KK, int PP, bool sign, int variant, bool excl,
                                        int grid, hipStream_t s, int* occ) {
    switch (KK) {
        case 1: return dispatch_persist_pp<1>(c, PP, sign, variant, excl, grid, s, occ);
        case 2: return dispatch_persist_pp<2>(c, PP, sign, variant, excl, grid, s, occ);
        case 3: return dispatch_persist_pp<3>(c, PP, sign, variant, excl, grid, s, occ);
        case 4: return dispatch_persist_pp<4>(c, PP, sign, variant, excl, grid, s, occ);
        case 5: return dispatch_persist_pp<5>(c, PP, sign, variant, excl, grid, s, occ);
        case 6: return dispatch_persist_pp<6>(c, PP, sign, variant, excl, grid, s, occ);
        case 7: return dispatch_persist_pp<7>(c, PP, sign, variant, excl, grid, s, occ);
        case 8: return dispatch_persist_pp<8>(c, PP, sign, variant, excl, grid, s, occ);
    }
    return hipErrorInvalidValue;
}

hipError_t launch_contract_bf16(const ContractArgs& a, const Bf16Launch& L, int KK, int PP,
                                bool sign_operand, hipStream_t s, int* occ) {
    switch (KK) {
        case 1: return dispatch_bf16_pp<1>(a, L, PP, sign_operand, s, occ);
        case 2: return dispatch_bf16_pp<2>(a, L, PP, sign_operand, s, occ);
        case 3: return dispatch_bf16_pp<3>(a, L, PP, sign_operand, s, occ);
        case 4: return dispatch_bf16_pp<4>(a, L, PP, sign_operand, s, occ);
        case 5: return dispatch_bf16_pp<5>(a, L, PP, sign_operand, s, occ);
        case 6: return dispatch_bf16_pp<6>(a, L, PP, sign_operand, s, occ);
        case 7: return dispatch_bf16_pp<7>(a, L, PP, sign_operand, s, occ);
        case 8: return dispatch_bf16_pp<8>(a, L, PP, sign_operand, s, occ);
        case 9: return dispatch_bf16_pp<9>(a, L, PP, sign_operand, s, occ);
        case 10: return dispatch_bf16_pp<10>(a, L, PP, sign_operand, s, occ);
        case 11: return dispatch_bf16_pp<11>(a, L, PP, sign_operand, s, occ);
        case 12: return dispatch_bf16_pp<12>(a, L, PP, sign_operand, s, occ);
        case 13: return dispatch_bf16_pp<13>(a, L, PP, sign_operand, s, occ);
        case 14: return dispatch_bf16_pp<14>(a, L, PP, sign_operand, s, occ);
        case 15: return dispatch_bf16_pp<15>(a, L, PP, sign_operand, s, occ);
        case 16: return dispatch_bf16_pp<16>(a, L, PP, sign_operand, s, occ);
    }
    return hipErrorInvalidValue;
}

// out[k] = max_c partial[c*K + k].  A block takes 32 permutations and splits
// the chunks into 32 slices (thread = permutation + 32 * slice, every 32nd
// chunk, 8 loads in flight); the slices meet in LDS.  With few permutations
// there are up to 1024 chunks, which one thread per permutation would walk
// one dependent load at a time.
// The per-permutation max over voxel chunks: block bx covers permutations
// 32 bx .. 32 bx + 31, 32 threads per chunk stride.
__device__ __forceinline__ void reduce_max_block(const float* __restrict__ partial, int n_part, int K,
                                                 float* __restrict__ out, int bx) {
    __shared__ float red[32][33];
    const int kl = threadIdx.x & 31, sl = threadIdx.x >> 5;
    const int k = bx * 32 + kl;
    float m = -INFINITY;
    if (k < K) {
        int c = sl;
        for (; c + 7 * 32 < n_part; c += 8 * 32) {
            float v[8];
#pragma unroll
            for (int u = 0; u < 8; u++) v[u] = partial[(size_t)(c + 32 * u) * K + k];
#pragma unroll
            for (int u = 0; u < 8; u++) m = fmaxf(m, v[u]);
        }
        for (; c < n_part; c += 32) m = fmaxf(m, partial[(size_t)c * K + k]);
    }
    red[sl][kl] = m;
    __syncthreads();
    if (sl == 0 && k < K) {
        float r = red[0][kl];
        for (int i = 1; i < 32; i++) r = fmaxf(r, red[i][kl]);
        out[k] = r;
    }
}

__global__ __launch_bounds__(1024) void k_reduce_max(const float* __restrict__ partial, int n_part, int K,
                                                     float* __restrict__ out) {
    reduce_max_block(partial, n_part, K, out, blockIdx.x);
}

hipError_t launch_reduce_max(const float* partial, int n_part, int K, float* out, hipStream_t s) {
    if (K <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_reduce_max, dim3((K + 31) / 32), dim3(1024), 0, s, partial, n_part, K, out);
    return hipGetLastError();
}

// One launch for several reductions: job blockIdx.y, its permutations
// 32 blockIdx.x .. (blocks past its K return at once).
__global__ __launch_bounds__(1024) void k_reduce_max_jobs(const ReduceJob* __restrict__ jobs) {
    const ReduceJob j = jobs[blockIdx.y];
    if ((int)blockIdx.x * 32 >= j.K) return;
    reduce_max_block(j.partial, j.n_part, j.K, j.out, blockIdx.x);
}

hipError_t launch_reduce_max_jobs(const ReduceJob* jobs, int n_jobs, int max_K, hipStream_t s) {
    if (n_jobs <= 0 || max_K <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_reduce_max_jobs, dim3((max_K + 31) / 32, n_jobs), dim3(1024), 0, s, jobs);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// FP64 recomputation of listed samples (one thread per sample).  y* = y -
// Qn (Qn' y) and the fit of y* on the permuted basis rows B[s][i] =
// Qt[perm[s] + i N] are both orthogonal projections, applied one column at
// a time (modified Gram-Schmidt: z_i = B_i' r, r -= B_i z_i), so the
// residual is formed itself, as in the FP64 form, and SSE = |r|^2.
// ---------------------------------------------------------------------------
// max(*addr, v) for floats: non-negative floats order as signed integers,
// negative ones in reverse as unsigned integers
__device__ __forceinline__ void atomic_max_f32(float* addr, float v) {
    v = v + 0.0f;  // -0 -> +0
    if (v >= 0.0f) atomicMax(reinterpret_cast<int*>(addr), __float_as_int(v));
    else atomicMin(reinterpret_cast<unsigned*>(addr), __float_as_uint(v));
}

// The block's threads share the basis in LDS (the nuisance columns read as
// broadcasts, the permuted fit columns gathered by row); MN, the subjects
// held in registers, is N rounded up to 16.
template <int MN>
__global__ __launch_bounds__(256) void k_fix_samples(SampleFixArgs a) {
    __shared__ double sQt[BPG_FIX_MAX_N * BPG_FIX_MAX_P];
    __shared__ double sQn[BPG_FIX_MAX_N * 24];
    const int N = a.N, q = a.q, pfit = a.pfit;
    for (int t = threadIdx.x; t < N * pfit; t += 256) sQt[t] = a.Qt[t];
    for (int t = threadIdx.x; t < N * q; t += 256) sQn[t] = a.Qn[(size_t)(t / q) * a.QB + t % q];
    __syncthreads();
    // block b takes segments b, b + gridDim.x, ...
    for (int g = blockIdx.x; g < BPG_SAMPLE_SEGS; g += gridDim.x) {
    const unsigned long long n = min(a.n_samples[g], a.cap);
    for (unsigned long long i = threadIdx.x; i < n; i += 256ull) {
        const unsigned long long e = a.samples[(size_t)g * a.cap + i];
        const int k = (int)(e >> 32), v = (int)(e & 0xFFFFFFFFull);
        if (k >= a.K || v >= a.Vc) continue;  // the contraction lists in-range samples only
        double r[MN];
        int row[MN];
#pragma unroll
        for (int s = 0; s < MN; s++) {
            r[s] = s < N ? (double)a.Yc[(size_t)s * a.Vc + v] : 0.0;
            row[s] = s < N ? (int)a.perm[(size_t)k * N + s] : 0;
        }
        for (int j = 0; j < q; j++) {
            double w = 0.0;
#pragma unroll
            for (int s = 0; s < MN; s++) w = fma(s < N ? sQn[s * q + j] : 0.0, r[s], w);
#pragma unroll
            for (int s = 0; s < MN; s++)
                if (s < N) r[s] = fma(-sQn[s * q + j], w, r[s]);
        }
        double z0 = 0.0;
        for (int c = 0; c < pfit; c++) {
            const double* col = sQt + c * N;
            double b[MN];
#pragma unroll
            for (int s = 0; s < MN; s++) b[s] = s < N ? col[row[s]] : 0.0;
            double z = 0.0;
#pragma unroll
            for (int s = 0; s < MN; s++) z = fma(b[s], r[s], z);
#pragma unroll
            for (int s = 0; s < MN; s++) r[s] = fma(-b[s], z, r[s]);
            if (c == 0) z0 = z;
        }
        double sse = 0.0;
#pragma unroll
        for (int s = 0; s < MN; s++) sse = fma(r[s], r[s], sse);
        const float t = (float)(z0 * sqrt(a.dof / sse));
        if (t == t) atomic_max_f32(a.maxout + k, t);
    }
    }
}

hipError_t launch_fix_samples(const SampleFixArgs& a, hipStream_t s) {
    if (a.N > BPG_FIX_MAX_N || a.pfit > BPG_FIX_MAX_P || a.q > 24 || !a.perm) return hipErrorInvalidValue;
    const unsigned blocks = BPG_SAMPLE_SEGS;
    if (a.N <= 16) hipLaunchKernelGGL(k_fix_samples<16>, dim3(blocks), dim3(256), 0, s, a);
    else if (a.N <= 32) hipLaunchKernelGGL(k_fix_samples<32>, dim3(blocks), dim3(256), 0, s, a);
    else if (a.N <= 48) hipLaunchKernelGGL(k_fix_samples<48>, dim3(blocks), dim3(256), 0, s, a);
    else hipLaunchKernelGGL(k_fix_samples<64>, dim3(blocks), dim3(256), 0, s, a);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// plain FP32 contraction (vector ALUs)
// ---------------------------------------------------------------------------
// Block = 256 threads, one voxel per thread per tile, a tile of PTT = 64/P
// permutations.  The block's operand rows sit in LDS as [s][k][i] so the
// PTT*P values of one subject are read as broadcasts; each thread streams its
// voxel's data one subject at a time and keeps PTT*P FP32 accumulators.
template <int P>
struct F32Tile {
    static constexpr int PTT = (64 / P) < 1 ? 1 : (64 / P);
    static constexpr int ROW = PTT * P;                   // floats per subject
    static constexpr int ROW4 = (ROW + 3) / 4 * 4;        // padded to float4
};

template <int P, bool FT>
__global__ __launch_bounds__(256) void k_contract_f32(ContractArgs a, F32Launch L) {
    using T = F32Tile<P>;
    constexpr int PTT = T::PTT;
    // the float4 broadcasts need a 16-byte aligned base: the dynamic array
    // is declared as float4 and the static one is a whole number of float4
    extern __shared__ float4 sG4[];  // [N][ROW4] floats
    float* sG = reinterpret_cast<float*>(sG4);
    __shared__ unsigned sMax[(PTT + 3) / 4 * 4];
    const int tid = threadIdx.x;
    const int lb = xcd_logical_block(L.per_xcd);
    if (lb >= L.n_pg * L.n_vchunks) return;
    const int pg = lb % L.n_pg;
    const int vc = lb / L.n_pg;
    const int k0 = pg * PTT;
    const int N = a.N;
    // G32 is [k][i][NS]; gather this block's rows into [s][k][i]
    for (int t = tid; t < N * T::ROW4; t += 256) {
        const int s = t / T::ROW4, ki = t % T::ROW4;
        float val = 0.0f;
        if (ki < T::ROW) {
            const int k = ki / P, i = ki % P;
            val = a.G32[((size_t)(k0 + k) * P + i) * a.NS + s];
        }
        sG[t] = val;
    }
    if (tid < PTT) sMax[tid] = 0u;
    __syncthreads();
    const int v0 = vc * L.vox_per_chunk;
    const int v1 = min(v0 + L.vox_per_chunk, a.Vpad);
    const int lane = tid & 63;
    bool bad = false;
    for (int vb = v0; vb < v1; vb += 256) {
        const int v = vb + tid;
        const bool inb = v < a.Vpad;
        const bool valid = v < a.Vc;
        float z[T::ROW];
#pragma unroll
        for (int q = 0; q < T::ROW; q++) z[q] = 0.0f;
        const float* yp = a.Y32 + (inb ? v : 0);
        for (int s = 0; s < N; s++) {
            const float ys = inb ? yp[(size_t)s * a.Vpad] : 0.0f;
            const float4* g4 = reinterpret_cast<const float4*>(sG + s * T::ROW4);
#pragma unroll
            for (int q4 = 0; q4 < T::ROW4 / 4; q4++) {
                const float4 g = g4[q4];
                if (4 * q4 + 0 < T::ROW) z[4 * q4 + 0] = fmaf(g.x, ys, z[4 * q4 + 0]);
                if (4 * q4 + 1 < T::ROW) z[4 * q4 + 1] = fmaf(g.y, ys, z[4 * q4 + 1]);
                if (4 * q4 + 2 < T::ROW) z[4 * q4 + 2] = fmaf(g.z, ys, z[4 * q4 + 2]);
                if (4 * q4 + 3 < T::ROW) z[4 * q4 + 3] = fmaf(g.w, ys, z[4 * q4 + 3]);
            }
        }
        const double yyv = inb ? a.yy[v] : 1.0;
        const float tyy = a.tau * (float)yyv;
#pragma unroll
        for (int k = 0; k < PTT; k++) {
            float ss = 0.0f, num = 0.0f;
#pragma unroll
            for (int i = 0; i < P; i++) {
                const float zi = z[k * P + i];
                ss = fmaf(zi, zi, ss);
                if (FT) {
                    if (i < a.r) num = fmaf(zi, zi, num);
                }
            }
            const float sse = (float)(yyv - (double)ss);
            if (valid && k0 + k < a.K && sse < tyy) {
                bad = true;
                a.flags[k0 + k] = 1;
            }
            float st = FT ? stat_f(num, sse, a.f_scale) : stat_t(z[k * P], sse, a.sqrt_dof);
            if (!valid) st = -INFINITY;
            const int kk = k0 + k;
            if (a.stats && valid && kk < a.K) a.stats[(size_t)kk * a.Vc + v] = st;
            if (a.bits) {
                const unsigned long long bal = __ballot(valid && st > a.cdt);
                if (lane == 0 && kk < a.K) {
                    const int w0 = (vb + (tid & ~63)) >> 5;
                    if (w0 < a.nwords) a.bits[(size_t)kk * a.nwords + w0] = (unsigned)bal;
                    if (w0 + 1 < a.nwords) a.bits[(size_t)kk * a.nwords + w0 + 1] = (unsigned)(bal >> 32);
                }
                if (a.ties && valid && kk < a.K && st > a.cdt_lo && st <= a.cdt_hi) a.ties[kk] = 1;
            }
            float m = st;
            m = fmaxf(m, __shfl_xor(m, 32, 64));
            m = fmaxf(m, __shfl_xor(m, 16, 64));
            m = fmaxf(m, __shfl_xor(m, 8, 64));
            m = fmaxf(m, __shfl_xor(m, 4, 64));
            m = fmaxf(m, __shfl_xor(m, 2, 64));
            m = fmaxf(m, __shfl_xor(m, 1, 64));
            if (lane == 0 && m > -INFINITY) atomicMax(&sMax[k], enc_max(m));
        }
    }
    if (__any(bad) && lane == 0) atomicAdd(a.unsafe, 1u);
    __syncthreads();
    if (tid < PTT && k0 + tid < a.K)
        a.partial[(size_t)vc * a.K + k0 + tid] = sMax[tid] ? dec_max(sMax[tid]) : -INFINITY;
}

namespace {

template <int P, bool FT>
hipError_t run_f32(const ContractArgs& a, const F32Launch& L, hipStream_t s) {
    using T = F32Tile<P>;
    const size_t shmem = (size_t)a.N * T::ROW4 * sizeof(float);
    hipLaunchKernelGGL((k_contract_f32<P, FT>), dim3(L.per_xcd * 8), dim3(256), shmem, s, a, L);
    return hipGetLastError();
}

template <bool FT>
hipError_t dispatch_f32(const ContractArgs& a, const F32Launch& L, hipStream_t s) {
    switch (a.pfit) {
#define BPG_P(n) case n: return run_f32<n, FT>(a, L, s);
        BPG_P(1) BPG_P(2) BPG_P(3) BPG_P(4) BPG_P(5) BPG_P(6) BPG_P(7) BPG_P(8) BPG_P(9)
        BPG_P(10) BPG_P(11) BPG_P(12) BPG_P(13) BPG_P(14) BPG_P(15) BPG_P(16) BPG_P(17)
        BPG_P(18) BPG_P(19) BPG_P(20) BPG_P(21) BPG_P(22) BPG_P(23) BPG_P(24) BPG_P(25)
#undef BPG_P
    }
    return hipErrorInvalidValue;
}

}  // namespace

// permutations per block of the FP32 form (F32Tile<P>::PTT)
int f32_perm_tile(int pfit) { return (64 / pfit) < 1 ? 1 : (64 / pfit); }

F32Launch plan_contract_f32(int K, int pfit, int Vpad) {
    F32Launch L{};
    const int PTT = f32_perm_tile(pfit);
    L.n_pg = (K + PTT - 1) / PTT;
    const int nvsteps = (Vpad + 255) / 256;
    int nvc = (2048 + L.n_pg - 1) / L.n_pg;
    if (nvc > BPG_MAX_VCHUNKS) nvc = BPG_MAX_VCHUNKS;
    if (nvc > nvsteps) nvc = nvsteps;
    if (nvc < 1) nvc = 1;
    const int steps_per = (nvsteps + nvc - 1) / nvc;
    L.vox_per_chunk = steps_per * 256;
    L.n_vchunks = (nvsteps + steps_per - 1) / steps_per;
    L.per_xcd = (L.n_pg * L.n_vchunks + 7) / 8;
    return L;
}

hipError_t launch_contract_f32(const ContractArgs& a, const F32Launch& L, hipStream_t s) {
    return a.ftest ? dispatch_f32<true>(a, L, s) : dispatch_f32<false>(a, L, s);
}

// ---------------------------------------------------------------------------
// FP64 contraction (the repeat of permutations with ill-conditioned samples)
// ---------------------------------------------------------------------------
// Block = 256 threads, one voxel per thread per step, a tile of PTT = 32/P
// permutations.  The tile's permuted basis B_k(s, i) = Qt(pi_k(s), i) (or
// sign_k(s) Qt(s, 0) for the group mean) sits in LDS as [s][k][i] doubles
// and is read as broadcasts.  Per voxel and permutation, in FP64:
//   z = B' y*,   SSE = |y* - B z|^2   (the residual itself, not
//   |y*|^2 - |z|^2, so no cancellation however small SSE is),
//   t = z_0 sqrt(dof / SSE),  F = (z_0^2 + .. + z_{r-1}^2) dof / (r SSE).
// The data pass runs twice (z, then the residual); y* comes from HBM / L2.
template <int P>
struct F64Tile {
    static constexpr int PTT = (32 / P) < 1 ? 1 : (32 / P);
    static constexpr int ROW = PTT * P;  // doubles per subject
};

template <int P, bool FT>
__global__ __launch_bounds__(256) void k_contract_f64(ContractArgs a, F32Launch L) {
    using T = F64Tile<P>;
    constexpr int PTT = T::PTT, ROW = T::ROW;
    extern __shared__ double sB[];  // [N][ROW]
    __shared__ unsigned sMax[PTT];
    const int tid = threadIdx.x;
    const int lb = xcd_logical_block(L.per_xcd);
    if (lb >= L.n_pg * L.n_vchunks) return;
    const int pg = lb % L.n_pg;
    const int vc = lb / L.n_pg;
    const int k0 = pg * PTT;
    const int N = a.N;
    for (int t = tid; t < N * ROW; t += 256) {
        const int s = t / ROW, ki = t % ROW;
        const int k = k0 + ki / P, i = ki % P;
        double val = 0.0;
        if (k < a.K) {
            if (a.sign) val = (double)a.sign[(size_t)k * N + s] * a.Qt[s + (size_t)i * N];
            else val = a.Qt[(size_t)a.perm[(size_t)k * N + s] + (size_t)i * N];
        }
        sB[t] = val;
    }
    if (tid < PTT) sMax[tid] = 0u;
    __syncthreads();
    const int v0 = vc * L.vox_per_chunk;
    const int v1 = min(v0 + L.vox_per_chunk, a.Vpad);
    const int lane = tid & 63;
    for (int vb = v0; vb < v1; vb += 256) {
        const int v = vb + tid;
        const bool inb = v < a.Vpad;
        const bool valid = v < a.Vc;
        const double* yp = a.Y64 + (inb ? v : 0);
        double z[ROW];
#pragma unroll
        for (int q = 0; q < ROW; q++) z[q] = 0.0;
        for (int s = 0; s < N; s++) {
            const double ys = inb ? yp[(size_t)s * a.Vpad] : 0.0;
            const double* row = sB + s * ROW;
#pragma unroll
            for (int q = 0; q < ROW; q++) z[q] = fma(row[q], ys, z[q]);
        }
        double sse[PTT];
#pragma unroll
        for (int k = 0; k < PTT; k++) sse[k] = 0.0;
        for (int s = 0; s < N; s++) {
            const double ys = inb ? yp[(size_t)s * a.Vpad] : 0.0;
            const double* row = sB + s * ROW;
#pragma unroll
            for (int k = 0; k < PTT; k++) {
                double e = ys;
#pragma unroll
                for (int i = 0; i < P; i++) e = fma(-row[k * P + i], z[k * P + i], e);
                sse[k] = fma(e, e, sse[k]);
            }
        }
#pragma unroll
        for (int k = 0; k < PTT; k++) {
            double stat;
            if (FT) {
                double num = 0.0;
#pragma unroll
                for (int i = 0; i < P; i++)
                    if (i < a.r) num = fma(z[k * P + i], z[k * P + i], num);
                stat = num * a.dof / ((double)a.r * sse[k]);
            } else {
                stat = z[k * P] * sqrt(a.dof / sse[k]);
            }
            float st = (float)stat;
            if (!valid) st = -INFINITY;
            const int kk = k0 + k;
            if (a.stats && valid && kk < a.K) a.stats[(size_t)kk * a.Vc + v] = st;
            if (a.bits) {
                const unsigned long long bal = __ballot(valid && st > a.cdt);
                if (lane == 0 && kk < a.K) {
                    const int w0 = (vb + (tid & ~63)) >> 5;
                    if (w0 < a.nwords) a.bits[(size_t)kk * a.nwords + w0] = (unsigned)bal;
                    if (w0 + 1 < a.nwords) a.bits[(size_t)kk * a.nwords + w0 + 1] = (unsigned)(bal >> 32);
                }
                if (a.ties && valid && kk < a.K && st > a.cdt_lo && st <= a.cdt_hi) a.ties[kk] = 1;
            }
            float m = st;
            m = fmaxf(m, __shfl_xor(m, 32, 64));
            m = fmaxf(m, __shfl_xor(m, 16, 64));
            m = fmaxf(m, __shfl_xor(m, 8, 64));
            m = fmaxf(m, __shfl_xor(m, 4, 64));
            m = fmaxf(m, __shfl_xor(m, 2, 64));
            m = fmaxf(m, __shfl_xor(m, 1, 64));
            if (lane == 0 && m > -INFINITY) atomicMax(&sMax[k], enc_max(m));
        }
    }
    __syncthreads();
    if (tid < PTT && k0 + tid < a.K)
        a.partial[(size_t)vc * a.K + k0 + tid] = sMax[tid] ? dec_max(sMax[tid]) : -INFINITY;
}

namespace {

template <int P, bool FT>
hipError_t run_f64(const ContractArgs& a, const F32Launch& L, hipStream_t s) {
    const size_t shmem = (size_t)a.N * F64Tile<P>::ROW * sizeof(double);
    hipLaunchKernelGGL((k_contract_f64<P, FT>), dim3(L.per_xcd * 8), dim3(256), shmem, s, a, L);
    return hipGetLastError();
}

template <bool FT>
hipError_t dispatch_f64(const ContractArgs& a, const F32Launch& L, hipStream_t s) {
    switch (a.pfit) {
#define BPG_P(n) case n: return run_f64<n, FT>(a, L, s);
        BPG_P(1) BPG_P(2) BPG_P(3) BPG_P(4) BPG_P(5) BPG_P(6) BPG_P(7) BPG_P(8) BPG_P(9)
        BPG_P(10) BPG_P(11) BPG_P(12) BPG_P(13) BPG_P(14) BPG_P(15) BPG_P(16) BPG_P(17)
        BPG_P(18) BPG_P(19) BPG_P(20) BPG_P(21) BPG_P(22) BPG_P(23) BPG_P(24) BPG_P(25)
#undef BPG_P
    }
    return hipErrorInvalidValue;
}

}  // namespace

int f64_perm_tile(int pfit) { return (32 / pfit) < 1 ? 1 : (32 / pfit); }

F32Launch plan_contract_f64(int K, int pfit, int Vpad) {
    F32Launch L{};
    const int PTT = f64_perm_tile(pfit);
    L.n_pg = (K + PTT - 1) / PTT;
    const int nvsteps = (Vpad + 255) / 256;
    int nvc = (2048 + L.n_pg - 1) / L.n_pg;
    if (nvc > BPG_MAX_VCHUNKS) nvc = BPG_MAX_VCHUNKS;
    if (nvc > nvsteps) nvc = nvsteps;
    if (nvc < 1) nvc = 1;
    const int steps_per = (nvsteps + nvc - 1) / nvc;
    L.vox_per_chunk = steps_per * 256;
    L.n_vchunks = (nvsteps + steps_per - 1) / steps_per;
    L.per_xcd = (L.n_pg * L.n_vchunks + 7) / 8;
    return L;
}

hipError_t launch_contract_f64(const ContractArgs& a, const F32Launch& L, hipStream_t s) {
    return a.ftest ? dispatch_f64<true>(a, L, s) : dispatch_f64<false>(a, L, s);
}

// ---------------------------------------------------------------------------
// observed maps (FP64), CalculateBetaWeightsGLM + CalculateStatisticalMaps
// GLM{T,F}Test: beta = pinv(X) y, e = y - X beta; t: sigma^2 = sum e^2 /
// (N - p); F: sigma^2 = sum (e - mean e)^2 / (N - p - 1).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_observed(ObsArgs a) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= a.Vc) return;
    const int N = a.N, p = a.p;
    double beta[BPG_MAX_REGRESSORS];
#pragma unroll
    for (int r = 0; r < BPG_MAX_REGRESSORS; r++) beta[r] = 0.0;
    for (int s = 0; s < N; s++) {
        const double ys = (double)a.Yc[(size_t)s * a.Vc + v];
#pragma unroll
        for (int r = 0; r < BPG_MAX_REGRESSORS; r++)
            if (r < p) beta[r] += a.Xp[r + (size_t)s * p] * ys;
    }
    if (a.betas)
        for (int r = 0; r < p; r++) a.betas[(size_t)r * a.Vc + v] = (float)beta[r];
    double se = 0.0, se2 = 0.0;
    for (int s = 0; s < N; s++) {
        double e = (double)a.Yc[(size_t)s * a.Vc + v];
#pragma unroll
        for (int r = 0; r < BPG_MAX_REGRESSORS; r++)
            if (r < p) e -= a.X[s + (size_t)r * N] * beta[r];
        if (a.residuals) a.residuals[(size_t)s * a.Vc + v] = (float)e;
        se += e;
        se2 += e * e;
    }
    if (!a.ftest) {
        const double var = se2 / (double)(N - p);
        if (a.resvar) a.resvar[v] = (float)var;
        for (int c = 0; c < a.nmaps; c++) {
            double cv = 0.0;
            for (int r = 0; r < p; r++) cv += a.C[r + (size_t)c * p] * beta[r];
            a.out[(size_t)c * a.Vc + v] = (float)(cv / sqrt(var * a.sc[c]));
        }
    } else {
        const double mean = se / (double)N;
        const double var = (se2 - (double)N * mean * mean) / (double)(N - p - 1);
        if (a.resvar) a.resvar[v] = (float)var;
        double cb[BPG_MAX_FROWS];
        for (int c = 0; c < a.nc; c++) {
            double acc = 0.0;
            for (int r = 0; r < p; r++) acc += a.C[r + (size_t)c * p] * beta[r];
            cb[c] = acc;
        }
        double q = 0.0;
        for (int c = 0; c < a.nc; c++)
            for (int cc = 0; cc < a.nc; cc++) q += cb[c] * a.sc[cc + (size_t)c * a.nc] * cb[cc];
        a.out[v] = (float)(q / var / (double)a.nc);
    }
}

hipError_t launch_observed(const ObsArgs& a, hipStream_t s) {
    if (a.Vc <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_observed, dim3((a.Vc + 255) / 256), dim3(256), 0, s, a);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// p-values: #{k : T > m_k} / K by binary search in the sorted null
// distribution (KC:516-652 count the same set one by one).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_pvalues(PvalArgs a) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= a.Vc) return;
    float T = a.obs[v];
    // TFCE compares the statistic itself with the TFCE null, as the
    // reference does (LIB:16824-16837)
    const bool cluster = a.inference == BPG_CLUSTER_EXTENT || a.inference == BPG_CLUSTER_MASS;
    if (cluster) {
        if (!(T > a.cdt)) {
            a.P[v] = 0.0f;
            if (a.ties) a.ties[v] = 0u;
            return;
        }
        T = (float)a.sizes[a.labels[v]];
        if (a.inference == BPG_CLUSTER_MASS) T = T / 10000.0f;
    }
    // first index whose null value is >= x (not < x)
    auto lower = [&](float x) {
        int lo = 0, hi = a.K;
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (a.sorted_max[mid] < x) lo = mid + 1;
            else hi = mid;
        }
        return lo;
    };
    const int cnt = lower(T);
    a.P[v] = (float)cnt / (float)a.K;
    if (a.ties) {
        unsigned t = 0u;
        if (cluster) {
            // every permutation whose largest cluster may differ
            t = a.cluster_ties;
        } else if (a.inference == BPG_VOXEL) {
            // null values within 1e-4 (2 + |T|) of T: a null value's statistic
            // is within 1e-4 (1 + |t|) of the exact one and its quantisation
            // may move it by one more quantum, so the comparison T > m_k can
            // differ from an exact computation's only inside this band
            const float band = 1e-4f * (2.0f + fabsf(T)) * 1.0001f;
            const int lo = lower(T - band);
            int hi = lo;
            {
                int h = a.K;
                while (hi < h) {
                    const int mid = (hi + h) >> 1;
                    if (a.sorted_max[mid] <= T + band) hi = mid + 1;
                    else h = mid;
                }
            }
            t = (unsigned)(hi - lo);
        }
        a.ties[v] = t;
    }
}

hipError_t launch_pvalues(const PvalArgs& a, hipStream_t s) {
    if (a.Vc <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_pvalues, dim3((a.Vc + 255) / 256), dim3(256), 0, s, a);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// connected components (union-find on compact indices, min-index roots)
// ---------------------------------------------------------------------------
// The reference propagates the minimum label among neighbours until nothing
// changes (KC:42-401); its fixpoint labels every component with its minimum
// linear index.  Here each supra-threshold voxel is linked to its
// lower-index neighbours with atomicMin unions, so every tree is rooted at
// its minimum index and the labels are the same.  Finds halve paths: each
// write replaces a parent by an ancestor, which concurrent finds and unions
// tolerate.
__device__ __forceinline__ unsigned uf_find(unsigned* p, unsigned x) {
    for (;;) {
        const unsigned px = p[x];
        if (px == x) return x;
        const unsigned ppx = p[px];
        if (ppx != px) p[x] = ppx;
        x = ppx;
    }
}

__device__ __forceinline__ void uf_union(unsigned* p, unsigned a, unsigned b) {
    for (;;) {
        a = uf_find(p, a);
        b = uf_find(p, b);
        if (a == b) return;
        if (a < b) {
            const unsigned t = a;
            a = b;
            b = t;
        }
        const unsigned old = atomicMin(&p[a], b);
        if (old == a) return;
        a = old;
    }
}

// CalculateTFCEValues (KC:482-510): value = sqrt(size) * h * h, added to
// the voxel's sum, every operation rounded on its own (no fused
// multiply-add); sqrt(size) comes from a table of correctly rounded values
// (v_sqrt_f32 is not correctly rounded).
__device__ __forceinline__ void tfce_add(float* acc, float sqrt_size, float h) {
#pragma clang fp contract(off)
    const float value = sqrt_size * h * h;
    *acc = *acc + value;
}

// One thread per 32-voxel word of one map.  Stages: 0 init, 1 union with the
// lower-index neighbours, 2 component sums (runs of equal roots within the
// word are added with one atomic), 3 largest sum per map.
//
// Stage 1 skips links that other links already imply (26-connectivity): the
// voxel below (z-1) touches the other 8 voxels of the plane below, and the
// voxel behind (y-1) touches its two x-neighbours, so when one of them is
// active the union with it stands for the others; their own links to it
// are made by the in-plane rules of that plane.
// The 13 neighbour entries of compact voxel c, loaded before any of them is
// tested: the loads issue together (one memory latency per voxel, where
// loads behind the tests would follow each other)
__device__ __forceinline__ void load_nbr(const int* nbr, int c, int (&nc)[BPG_NBR]) {
    const int* row = nbr + (size_t)c * BPG_NBR;
#pragma unroll
    for (int j = 0; j < BPG_NBR; j++) nc[j] = row[j];
}

__global__ __launch_bounds__(256) void k_ccl(CclArgs a, int stage) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (size_t)a.nmaps * a.nwords) return;
    const int slot = (int)(t / a.nwords);
    const int map = a.map_ids ? (int)a.map_ids[slot] : slot;
    const int w = (int)(t % a.nwords);
    // union-find scratch per slot (a map of a map_ids list, or the map)
    unsigned* parent = a.parent + (size_t)slot * a.Vc;
    unsigned* sizes = a.sizes + (size_t)slot * a.Vc;
    const unsigned* bits = a.bits + (size_t)map * a.nwords;
    unsigned word = bits[w];
    unsigned run_root = 0xFFFFFFFFu, run_sum = 0u, best = 0u;
    auto active = [&](int n) { return n >= 0 && ((bits[n >> 5] >> (n & 31)) & 1u); };
    while (word) {
        const int c = w * 32 + __ffs(word) - 1;
        word &= word - 1;
        if (c >= a.Vc) break;
        if (stage == 0) {
            parent[c] = (unsigned)c;
            sizes[c] = 0u;
        } else if (stage == 1) {
            // neighbour order: x-1, y-1, z-1, then (x-1,y-1), (x+1,y-1), then
            // the 8 others of the plane below
            int nc[BPG_NBR];
            load_nbr(a.nbr, c, nc);
            const int nx = nc[0], ny = nc[1], nz = nc[2];
            if (active(nx)) uf_union(parent, (unsigned)c, (unsigned)nx);
            const bool ay = active(ny), az = active(nz);
            if (ay) uf_union(parent, (unsigned)c, (unsigned)ny);
            if (az) uf_union(parent, (unsigned)c, (unsigned)nz);
            if (a.connectivity == 6) continue;
            if (!ay)
#pragma unroll
                for (int j = 3; j < 5; j++)
                    if (active(nc[j])) uf_union(parent, (unsigned)c, (unsigned)nc[j]);
            if (!az)
#pragma unroll
                for (int j = 5; j < BPG_NBR; j++)
                    if (active(nc[j])) uf_union(parent, (unsigned)c, (unsigned)nc[j]);
        } else if (stage == 2) {
            const unsigned root = uf_find(parent, (unsigned)c);
            // extent (KC:403-430) or mass (KC:432-458)
            const unsigned add = a.values ? (unsigned)(a.values[(size_t)map * a.Vc + c] * 10000.0f) : 1u;
            if (root != run_root) {
                if (run_sum) atomicAdd(&sizes[run_root], run_sum);
                run_root = root;
                run_sum = 0u;
            }
            run_sum += add;
            if (a.labels && map == 0) a.labels[c] = root;
        } else if (stage == 3) {
            if (parent[c] == (unsigned)c) best = max(best, sizes[c]);
        } else {
            tfce_add(a.tfce + (size_t)map * a.Vc + c, a.sqrt_size[sizes[uf_find(parent, (unsigned)c)]], a.level);
        }
    }
    if (stage == 2 && run_sum) atomicAdd(&sizes[run_root], run_sum);
    if (stage == 3 && best) atomicMax(&a.maxsize[map], best);
}

// ---- whole-map CCL in LDS ----
constexpr int CCL_BATCH = 4;  // active voxels whose neighbour rows load together
constexpr size_t CCL_LDS_BYTES = 160 * 1024;  // LDS of one CU (gfx950)
constexpr long CCL_STATIC_WORDS = 256;        // the kernels' static LDS is kept below this

// the root of x (reads only: halving the path with writes measured slower,
// 30 against 26 us per map on the F + cluster workload)
__device__ __forceinline__ unsigned lds_find(unsigned* p, unsigned x) {
    volatile unsigned* vp = p;
    for (;;) {
        const unsigned px = vp[x];
        if (px == x) return x;
        x = px;
    }
}

__device__ __forceinline__ void lds_union(unsigned* p, unsigned a, unsigned b) {
    for (;;) {
        a = lds_find(p, a);
        b = lds_find(p, b);
        if (a == b) return;
        if (a < b) {
            const unsigned t = a;
            a = b;
            b = t;
        }
        const unsigned old = atomicMin(&p[a], b);
        if (old == a) return;
        a = old;
    }
}

// LDS words of a map's bitmask (nwords), its 64-word group ranks and its
// 16-bit word ranks within the group
static long ccl_lds_fixed_words(int nwords) {
    const long ng = (nwords + 63) / 64;
    return (long)nwords + ng + ((long)nwords + 1) / 2;
}

// active voxels one block can hold when `blocks` blocks share a CU's LDS
// (lds_bytes available to one block); each takes 3 words (its voxel index,
// parent and size)
int ccl_lds_cap(int nwords, size_t lds_bytes, int blocks) {
    const long words = (long)(std::min(lds_bytes, CCL_LDS_BYTES) / 4 / blocks) - CCL_STATIC_WORDS - ccl_lds_fixed_words(nwords);
    return words > 0 ? (int)(words / 3) : 0;
}

// One map per block (map_ids[blockIdx.x], or blockIdx.x): the bitmask, the
// rank (active voxels before) of each 64-word group and, in 16 bits, of each
// word within its group in LDS; then the list of active voxels (compact
// index -> voxel, built by rank and select, so every thread takes the same
// share whatever the layout); union-find over compact indices, sizes, max.
template <int T>
__global__ __launch_bounds__(T) void k_ccl_lds(CclArgs a, int cap, const unsigned* map_ids, unsigned* overflow,
                                               unsigned* n_overflow) {
    extern __shared__ unsigned smem[];
    constexpr int NWAVES = T / 64;
    __shared__ unsigned wsum[NWAVES];
    __shared__ unsigned scan[NWAVES];  // per-wave maxima at the end
    const int map = map_ids ? (int)map_ids[blockIdx.x] : (int)blockIdx.x;
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int nw = a.nwords;
    const int ng = (nw + 63) / 64;
    unsigned* sbits = smem;                                   // [nw]
    unsigned* gbase = sbits + nw;                             // [ng] active voxels before each group
    unsigned short* srank = reinterpret_cast<unsigned short*>(gbase + ng);  // [nw] within the group
    unsigned* slist = gbase + ng + (nw + 1) / 2;              // [cap] voxel of each compact index
    unsigned* spar = slist + cap;                             // [cap]
    unsigned* ssz = spar + cap;                               // [cap]
    const unsigned* bits = a.bits + (size_t)map * nw;
    // 8 loads in flight per thread (indices clamped, stores guarded)
    for (int w0 = tid; w0 < nw; w0 += 8 * T) {
        unsigned v[8];
#pragma unroll
        for (int u = 0; u < 8; u++) v[u] = bits[min(w0 + u * T, nw - 1)];
#pragma unroll
        for (int u = 0; u < 8; u++)
            if (w0 + u * T < nw) sbits[w0 + u * T] = v[u];
    }
    __syncthreads();
    // ranks: contiguous chunks per thread; an in-wave scan of the chunk
    // counts (shuffles), then the scan of the wave totals
    const int per = (nw + T - 1) / T;
    const int w0 = min(tid * per, nw), w1 = min(w0 + per, nw);
    unsigned cnt = 0;
    for (int w = w0; w < w1; w++) cnt += __popc(sbits[w]);
    unsigned incl = cnt;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const unsigned v = (unsigned)__shfl_up((int)incl, off, 64);
        if (lane >= off) incl += v;
    }
    if (lane == 63) wsum[wave] = incl;
    __syncthreads();
    unsigned before = 0, total = 0;
#pragma unroll
    for (int i = 0; i < NWAVES; i++) {
        const unsigned v = wsum[i];
        before += i < wave ? v : 0u;
        total += v;
    }
    if (total > (unsigned)cap) {  // the whole block takes this branch
        if (tid == 0) overflow[atomicAdd(n_overflow, 1u)] = (unsigned)map;
        return;
    }
    const unsigned first = before + incl - cnt;
    unsigned base = first;
    for (int w = w0; w < w1; w++) {
        if ((w & 63) == 0) gbase[w >> 6] = base;
        base += __popc(sbits[w]);
    }
    __syncthreads();
    base = first;
    for (int w = w0; w < w1; w++) {
        srank[w] = (unsigned short)(base - gbase[w >> 6]);
        base += __popc(sbits[w]);
    }
    __syncthreads();
    auto wrank = [&](int w) { return gbase[w >> 6] + (unsigned)srank[w]; };
    // the list: compact index i is the voxel of the i-th set bit -- the last
    // group, then the last word in it, whose rank is <= i (binary searches),
    // then the (i - rank)-th set bit of that word (binary select)
    for (unsigned i = tid; i < total; i += T) {
        int lo = 0, hi = ng - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (gbase[mid] <= i) lo = mid;
            else hi = mid - 1;
        }
        const int g = lo;
        lo = g * 64;
        hi = min(nw, g * 64 + 64) - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (wrank(mid) <= i) lo = mid;
            else hi = mid - 1;
        }
        unsigned x = sbits[lo], j = i - wrank(lo);
        int b = 0;
#pragma unroll
        for (int sh = 16; sh > 0; sh >>= 1) {
            const unsigned c = __popc(x & ((1u << sh) - 1u));
            if (j >= c) {
                j -= c;
                x >>= sh;
                b += sh;
            }
        }
        slist[i] = (unsigned)(lo * 32 + b);
        spar[i] = i;
        ssz[i] = 0u;
    }
    __syncthreads();
    auto active = [&](int n) { return n >= 0 && ((sbits[n >> 5] >> (n & 31)) & 1u); };
    auto rank = [&](int n) { return wrank(n >> 5) + __popc(sbits[n >> 5] & ((1u << (n & 31)) - 1u)); };
    const int nb6 = a.connectivity == 6;
    // CCL_BATCH active voxels per thread at a time: the neighbour rows of a
    // batch are loaded together, so one memory latency serves the batch.
    // Linking by the smaller root makes the result independent of the order
    // of the unions.  (Bits past the last voxel are never set.)
    for (unsigned i0 = 0; i0 < total; i0 += CCL_BATCH * T) {
        int cs[CCL_BATCH];
        unsigned ls[CCL_BATCH];
#pragma unroll
        for (int u = 0; u < CCL_BATCH; u++) {
            const unsigned i = i0 + (unsigned)u * T + (unsigned)tid;
            ls[u] = i;
            cs[u] = i < total ? (int)slist[i] : -1;
            if (cs[u] >= a.Vc) cs[u] = -1;
        }
        int ncs[CCL_BATCH][BPG_NBR];
#pragma unroll
        for (int u = 0; u < CCL_BATCH; u++) load_nbr(a.nbr, max(cs[u], 0), ncs[u]);  // row 0 for empty slots
#pragma unroll
        for (int u = 0; u < CCL_BATCH; u++) {
            if (cs[u] < 0) continue;
            const unsigned lc = ls[u];
            const int* nc = ncs[u];
            // the same implied-link rules as the global path
            const int nx = nc[0], ny = nc[1], nz = nc[2];
            if (active(nx)) lds_union(spar, lc, rank(nx));
            const bool ay = active(ny), az = active(nz);
            if (ay) lds_union(spar, lc, rank(ny));
            if (az) lds_union(spar, lc, rank(nz));
            if (nb6) continue;
            if (!ay)
#pragma unroll
                for (int j = 3; j < 5; j++)
                    if (active(nc[j])) lds_union(spar, lc, rank(nc[j]));
            if (!az)
#pragma unroll
                for (int j = 5; j < BPG_NBR; j++)
                    if (active(nc[j])) lds_union(spar, lc, rank(nc[j]));
        }
    }
    __syncthreads();
    for (unsigned i = tid; i < total; i += T) {
        const int c = (int)slist[i];
        if (c >= a.Vc) continue;
        // extent (KC:403-430) or mass (KC:432-458)
        const unsigned add = a.values ? (unsigned)(a.values[(size_t)map * a.Vc + c] * 10000.0f) : 1u;
        atomicAdd(&ssz[lds_find(spar, i)], add);
    }
    __syncthreads();
    unsigned best = 0;
    for (unsigned i = tid; i < total; i += T)
        if (spar[i] == i) best = max(best, ssz[i]);
    for (int o = 32; o > 0; o >>= 1) best = max(best, (unsigned)__shfl_xor((int)best, o, 64));
    if (lane == 0) scan[wave] = best;
    __syncthreads();
    if (tid == 0) {
        unsigned m = 0;
        for (int i = 0; i < NWAVES; i++) m = max(m, scan[i]);
        a.maxsize[map] = m;
    }
}

bool ccl_lds_fits(int nwords, size_t lds_bytes) { return nwords > 0 && ccl_lds_cap(nwords, lds_bytes, 1) >= 4096; }

// blocks = 2: 512-thread blocks, two per CU; blocks = 1: 1024-thread blocks
// with the whole LDS (for the maps the first listed as too large)
hipError_t launch_ccl_lds(const CclArgs& a, size_t lds_bytes, int blocks, const unsigned* map_ids, int n_maps,
                          unsigned* overflow, unsigned* n_overflow, hipStream_t s) {
    if (n_maps <= 0) return hipSuccess;
    const int cap = ccl_lds_cap(a.nwords, lds_bytes, blocks);
    const size_t dyn = 4 * ((size_t)ccl_lds_fixed_words(a.nwords) + (size_t)3 * cap);
    if (blocks == 2) {
        hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(k_ccl_lds<512>),
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)dyn);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(k_ccl_lds<512>, dim3(n_maps), dim3(512), dyn, s, a, cap, map_ids, overflow, n_overflow);
    } else {
        hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(k_ccl_lds<1024>),
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)dyn);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(k_ccl_lds<1024>, dim3(n_maps), dim3(1024), dyn, s, a, cap, map_ids, overflow, n_overflow);
    }
    return hipGetLastError();
}

hipError_t launch_ccl(const CclArgs& a, hipStream_t s) {
    const size_t total = (size_t)a.nmaps * a.nwords;
    if (total == 0) return hipSuccess;
    const unsigned blocks = (unsigned)((total + 255) / 256);
    for (int stage = 0; stage < 4; stage++) {
        hipLaunchKernelGGL(k_ccl, dim3(blocks), dim3(256), 0, s, a, stage);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

// init, union and sums at one TFCE level, then the level's contributions
hipError_t launch_ccl_tfce_level(const CclArgs& a, hipStream_t s) {
    const size_t total = (size_t)a.nmaps * a.nwords;
    if (total == 0) return hipSuccess;
    const unsigned blocks = (unsigned)((total + 255) / 256);
    for (int stage : {0, 1, 2, 4}) {
        hipLaunchKernelGGL(k_ccl, dim3(blocks), dim3(256), 0, s, a, stage);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

// one thread per (map, padded voxel); Vpad is a multiple of 256, so every
// wave lies in one map and covers two whole words
__global__ __launch_bounds__(256) void k_level_bits(const float* stats, const int* nlev, int j, float level,
                                                    int nmaps, int Vc, int Vpad, unsigned* bits, int nwords) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (size_t)nmaps * Vpad) return;
    const int m = (int)(t / Vpad);
    const int v = (int)(t % Vpad);
    const bool act = v < Vc && j < nlev[m] && stats[(size_t)m * Vc + v] > level;
    const unsigned long long bal = __ballot(act);
    const int lane = threadIdx.x & 63;
    if (lane == 0) {
        const int w0 = v >> 5;
        if (w0 < nwords) bits[(size_t)m * nwords + w0] = (unsigned)bal;
        if (w0 + 1 < nwords) bits[(size_t)m * nwords + w0 + 1] = (unsigned)(bal >> 32);
    }
}

hipError_t launch_level_bits(const float* stats, const int* nlev, int j, float level, int nmaps, int Vc,
                             int Vpad, unsigned* bits, int nwords, hipStream_t s) {
    const size_t total = (size_t)nmaps * Vpad;
    if (total == 0) return hipSuccess;
    hipLaunchKernelGGL(k_level_bits, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, stats, nlev, j,
                       level, nmaps, Vc, Vpad, bits, nwords);
    return hipGetLastError();
}

// one block per map
__global__ __launch_bounds__(256) void k_rowmax(const float* x, int Vc, float* out) {
    const float* row = x + (size_t)blockIdx.x * Vc;
    float m = -INFINITY;
    for (int v = threadIdx.x; v < Vc; v += 256) m = fmaxf(m, row[v]);
    __shared__ float red[256];
    red[threadIdx.x] = m;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) red[threadIdx.x] = fmaxf(red[threadIdx.x], red[threadIdx.x + o]);
        __syncthreads();
    }
    if (threadIdx.x == 0) out[blockIdx.x] = red[0];
}

hipError_t launch_rowmax(const float* x, int nmaps, int Vc, float* out, hipStream_t s) {
    if (nmaps <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_rowmax, dim3(nmaps), dim3(256), 0, s, x, Vc, out);
    return hipGetLastError();
}

__global__ __launch_bounds__(256) void k_threshold_bits(const float* obs, int Vc, float cdt,
                                                        unsigned* bits, int nwords) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    const bool act = v < Vc && obs[v] > cdt;
    const unsigned long long bal = __ballot(act);
    const int lane = threadIdx.x & 63;
    const int w0 = (v - lane) >> 5;
    if (lane == 0) {
        if (w0 < nwords) bits[w0] = (unsigned)bal;
        if (w0 + 1 < nwords) bits[w0 + 1] = (unsigned)(bal >> 32);
    }
}

hipError_t launch_threshold_bits(const float* obs, int Vc, float cdt, unsigned* bits, int nwords,
                                 hipStream_t s) {
    const int n = nwords * 32;
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_threshold_bits, dim3((n + 255) / 256), dim3(256), 0, s, obs, Vc, cdt, bits,
                       nwords);
    return hipGetLastError();
}

}  // namespace bpg
