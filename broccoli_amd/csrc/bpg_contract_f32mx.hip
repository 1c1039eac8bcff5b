// bpg_contract_f32mx.hip -- the FP32 form of the permuted-GLM contraction on
// the FP32 matrix cores (v_mfma_f32_32x32x2_f32), with the statistic and
// max-T epilogue of the split-BF16 kernel.
//
// Z = G Y* as in bpg_contract_bf16.h, with every operand in FP32 and no
// split: one MFMA per k-step of 2 subjects.  An MFMA of this shape keeps the
// matrix pipe 64 cycles (2 x 32 x 32 x 2 flop at 64 flop per cycle per
// SIMD, the 157.3 TFLOP/s of the chip).  Each wave's R operand subtiles are
// copied once into LDS (KG KiB each; a lane reads back only what it wrote,
// so LDS is register overflow here, not a shared tile) and read per group
// of 4 k-steps (8 subjects); the data tile streams from L2 through L1, one
// float4 per lane and group, shared by the block's 4 waves, for 4 R MFMAs
// (256 R cycles of matrix work), three groups ahead.  Streaming the operand from L2 instead, as
// a first version did, fetched 57 GB per launch from HBM on the headline
// workload (the 8 MB operand exceeds an XCD's 4 MB L2) and kept the matrix
// pipe 63% busy.
//
// Fragment layouts (both 16 bytes per lane per group of 8 subjects):
//   data     Ym[(vt * KG + g) * 64 + lane]  = { Y*[8g + 2j + h][32 vt + c] }_j
//   operand  Gm[(st * KG + g) * 64 + lane]  = { G[row i][8g + 2j + h] }_j
// with c = lane % 32 (voxel column / operand row i) and h = lane / 32 (the
// subject of the k-step the lane supplies).  The accumulator layout is that
// of every 32x32 MFMA, so the operand rows are ordered exactly as for the
// split-BF16 kernel (each lane owns whole permutations of one voxel) and its
// epilogue applies unchanged.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "bpg_contract_bf16.h"
#include "bpg_kernels.h"

namespace bpg {

template <int R, int PP, bool FT, bool OUT>
__global__ __launch_bounds__(256) void k_contract_f32mx(ContractArgs a, Bf16Launch L, int KG) {
    constexpr int PER = 16 / PP;
    extern __shared__ float4 sA[];  // [wave][r][g][64 lanes]
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int lb = xcd_logical_block(L.per_xcd);
    if (lb >= L.n_mgroups * L.n_vchunks) return;
    const int mg = lb % L.n_mgroups;
    const int vc = lb / L.n_mgroups;
    const int sub0 = (mg * 4 + wave) * R;
    const float4* G = reinterpret_cast<const float4*>(a.Gm);
    const float4* Y = reinterpret_cast<const float4*>(a.Ym);
    // operand subtiles past the last one read subtile 0 (in range) and are
    // multiplied by zero-weight rows: their z is thrown away below
    int srow[R];
    bool live[R];
#pragma unroll
    for (int r = 0; r < R; r++) {
        live[r] = sub0 + r < L.n_sub;
        srow[r] = live[r] ? sub0 + r : 0;
    }
    // this wave's operand subtiles into LDS, each lane its own fragments
    float4* myA = sA + (size_t)wave * R * KG * 64 + lane;
    for (int g = 0; g < KG; g++)
#pragma unroll
        for (int r = 0; r < R; r++) myA[((size_t)r * KG + g) * 64] = G[((size_t)srow[r] * KG + g) * 64 + lane];
    float runmax[R][PER];
#pragma unroll
    for (int r = 0; r < R; r++)
#pragma unroll
        for (int P = 0; P < PER; P++) runmax[r][P] = -INFINITY;
    bool bad[R];
#pragma unroll
    for (int r = 0; r < R; r++) bad[r] = false;

    const int vt0 = vc * L.vt_per_chunk;
    const int vt1 = min(vt0 + L.vt_per_chunk, L.n_vt);
    for (int vt = vt0; vt < vt1; vt++) {
        // two accumulators per subtile (k-steps 0, 1 and 2, 3 of each group),
        // so that with one subtile per wave consecutive MFMAs do not wait on
        // each other
        f32x16 acc[R][2];
#pragma unroll
        for (int r = 0; r < R; r++)
#pragma unroll
            for (int e = 0; e < 16; e++) acc[r][0][e] = acc[r][1][e] = 0.0f;
        // a ring of 4 fragment buffers, 3 groups of prefetch: the loop is
        // unrolled over the ring so that no buffer is moved (a register move
        // of a loaded value would wait for its load); with one wave per SIMD
        // (the large-N form) 3 groups of MFMAs (768 cycles) cover an L2 read
        float4 bq[4], aq[4][R];
#pragma unroll
        for (int u = 0; u < 3; u++) {
            const int gl = min(u, KG - 1);
            bq[u] = Y[((size_t)vt * KG + gl) * 64 + lane];
#pragma unroll
            for (int r = 0; r < R; r++) aq[u][r] = myA[((size_t)r * KG + gl) * 64];
        }
        for (int g = 0; g < KG; g += 4) {
#pragma unroll
            for (int u = 0; u < 4; u++) {
                const int gl = min(g + u + 3, KG - 1);
                bq[(u + 3) & 3] = Y[((size_t)vt * KG + gl) * 64 + lane];
#pragma unroll
                for (int r = 0; r < R; r++) aq[(u + 3) & 3][r] = myA[((size_t)r * KG + gl) * 64];
                if (g + u < KG) {
                    const float4 b = bq[u];
#pragma unroll
                    for (int r = 0; r < R; r++)
                        acc[r][0] = __builtin_amdgcn_mfma_f32_32x32x2f32(aq[u][r].x, b.x, acc[r][0], 0, 0, 0);
#pragma unroll
                    for (int r = 0; r < R; r++)
                        acc[r][1] = __builtin_amdgcn_mfma_f32_32x32x2f32(aq[u][r].y, b.y, acc[r][1], 0, 0, 0);
#pragma unroll
                    for (int r = 0; r < R; r++)
                        acc[r][0] = __builtin_amdgcn_mfma_f32_32x32x2f32(aq[u][r].z, b.z, acc[r][0], 0, 0, 0);
#pragma unroll
                    for (int r = 0; r < R; r++)
                        acc[r][1] = __builtin_amdgcn_mfma_f32_32x32x2f32(aq[u][r].w, b.w, acc[r][1], 0, 0, 0);
                }
            }
        }
        const double yv = a.yy[vt * 32 + (lane & 31)];
        const float yhi = (float)yv;  // NaN for padding voxels
        const float ylo = (float)(yv - (double)yhi);
        unsigned wb_unused[16 / PP];  // bitmask words of the bits-only epilogue (not used here)
#pragma unroll
        for (int r = 0; r < R; r++)
            if (live[r])
                epi<PP, FT, OUT ? 1 : 0>(a, acc[r][0] + acc[r][1], yhi, ylo, vt, lane, sub0 + r, runmax[r], bad[r],
                                         wb_unused, TileSpan{0, 0});
    }
    // flags of the (subtile, half-wave) pairs that met an ill-conditioned
    // sample, as in the split-BF16 kernel
    bool any = false;
#pragma unroll
    for (int r = 0; r < R; r++) {
        const unsigned long long m = __ballot(bad[r]);
        any |= m != 0ull;
        if (m && (lane & 31) == 0) {
            const int hl = lane >> 5;
            if ((hl ? (m >> 32) : (m & 0xFFFFFFFFull)) != 0ull)
                for (int P = 0; P < PER; P++) {
                    const int k = (sub0 + r) * (2 * PER) + hl * PER + P;
                    if (k < a.K) a.flags[k] = 1;
                }
        }
    }
    if (any && lane == 0) atomicAdd(a.unsafe, 1u);
    const int hl = lane >> 5;
    const float scale = FT ? a.f_scale : a.sqrt_dof;
#pragma unroll
    for (int r = 0; r < R; r++)
#pragma unroll
        for (int P = 0; P < PER; P++) {
            float m = runmax[r][P];
            m = fmaxf(m, __shfl_xor(m, 16, 64));
            m = fmaxf(m, __shfl_xor(m, 8, 64));
            m = fmaxf(m, __shfl_xor(m, 4, 64));
            m = fmaxf(m, __shfl_xor(m, 2, 64));
            m = fmaxf(m, __shfl_xor(m, 1, 64));
            const int k = (sub0 + r) * (2 * PER) + hl * PER + P;
            if ((lane & 31) == 0 && k < a.K && live[r]) a.partial[(size_t)vc * a.K + k] = m * scale;
        }
}

namespace {

template <int R, int PP, bool FT>
hipError_t run_f32mx_r(const ContractArgs& a, const Bf16Launch& L, int KG, hipStream_t s, int* occ) {
    const bool out = occ ? *occ != 0 : (a.stats || a.bits);
    void (*kern)(ContractArgs, Bf16Launch, int) =
        out ? k_contract_f32mx<R, PP, FT, true> : k_contract_f32mx<R, PP, FT, false>;
    const size_t lds = f32mx_lds_bytes(KG);
    if (occ) return hipOccupancyMaxActiveBlocksPerMultiprocessor(occ, reinterpret_cast<const void*>(kern), 256, lds);
    hipLaunchKernelGGL(kern, dim3(L.per_xcd * 8), dim3(256), lds, s, a, L, KG);
    return hipGetLastError();
}

template <int PP, bool FT>
hipError_t run_f32mx(const ContractArgs& a, const Bf16Launch& L, int KG, hipStream_t s, int* occ) {
    return f32mx_rows_per_wave(KG) == 2 ? run_f32mx_r<2, PP, FT>(a, L, KG, s, occ)
                                        : run_f32mx_r<1, PP, FT>(a, L, KG, s, occ);
}

template <bool FT>
hipError_t dispatch_f32mx(const ContractArgs& a, const Bf16Launch& L, int PP, int KG, hipStream_t s, int* occ) {
    switch (PP) {
        case 1: return run_f32mx<1, FT>(a, L, KG, s, occ);
        case 2: return run_f32mx<2, FT>(a, L, KG, s, occ);
        case 3: return run_f32mx<3, FT>(a, L, KG, s, occ);
        case 4: return run_f32mx<4, FT>(a, L, KG, s, occ);
        case 5: return run_f32mx<5, FT>(a, L, KG, s, occ);
        case 8: return run_f32mx<8, FT>(a, L, KG, s, occ);
        case 16: return run_f32mx<16, FT>(a, L, KG, s, occ);
    }
    return hipErrorInvalidValue;
}

}  // namespace

hipError_t launch_contract_f32mx(const ContractArgs& a, const Bf16Launch& L, int PP, int KG, hipStream_t s,
                                 int* occ) {
    return a.ftest ? dispatch_f32mx<true>(a, L, PP, KG, s, occ) : dispatch_f32mx<false>(a, L, PP, KG, s, occ);
}

int contract_f32mx_blocks_per_cu(int PP, int KG, bool ftest, bool out) {
    ContractArgs a{};
    a.ftest = ftest;
    int occ = out ? 1 : 0;
    if (launch_contract_f32mx(a, Bf16Launch{}, PP, KG, nullptr, &occ) != hipSuccess || occ < 1) return 1;
    return occ;
}

// Operand fragments: row i of subtile st is permutation / regressor as in
// k_operand_bf16; lane (i, h) holds subjects 8g + 2j + h, j = 0..3.
__global__ __launch_bounds__(256) void k_operand_f32mx(OperandArgs a, int KG) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t total = (size_t)a.n_sub * KG * 64;
    if (t >= total) return;
    const int lane = (int)(t & 63);
    const size_t rest = t >> 6;
    const int g = (int)(rest % KG);
    const int st = (int)(rest / KG);
    const int i = lane & 31, h = lane >> 5;
    const int j = 4 * (i >> 3) + (i & 3);
    const int hh = (i >> 2) & 1;
    const int per = 16 / a.PP;
    const int k = st * (2 * per) + hh * per + j / a.PP;
    const int reg = j % a.PP;
    const bool live = j < per * a.PP && k < a.K && reg < a.pfit;
    const size_t row = (size_t)min(k, a.K - 1) * a.N;
    const size_t col = (size_t)min(reg, a.pfit - 1) * a.N;
    float v[4];
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const int s = 8 * g + 2 * q + h;
        const int sc = min(s, a.N - 1);
        float val;
        if (a.sign) {
            val = a.sign[row + sc];
        } else {
            unsigned ix = a.perm[row + sc];
            if (a.bad_index && ix >= (unsigned)a.N) {
                atomicAdd(a.bad_index, 1u);
                ix = 0;
            }
            val = (float)a.Qt[(size_t)ix + col];
        }
        v[q] = (live && s < a.N) ? val : 0.0f;
    }
    reinterpret_cast<float4*>(a.Gm)[t] = make_float4(v[0], v[1], v[2], v[3]);
}

hipError_t launch_operand_f32mx(const OperandArgs& a, int KG, hipStream_t s) {
    const size_t total = (size_t)a.n_sub * KG * 64;
    hipLaunchKernelGGL(k_operand_f32mx, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, a, KG);
    return hipGetLastError();
}

}  // namespace bpg
