// bpg_kernels.h -- host-callable launchers of the gfx950 kernels.
#pragma once

#include <hip/hip_runtime.h>

// Tuning knobs of the split-BF16 contraction: operand subtiles kept per
// wave for KK <= 4 (N <= 64) and the largest KK that keeps two, the number
// of blocks the voxel range is split over, and ping-pong tile loads.
#ifndef BPG_R_SMALL
#define BPG_R_SMALL 4
#endif
#ifndef BPG_R2_KK
#define BPG_R2_KK 8
#endif
#ifndef BPG_TARGET_BLOCKS
#define BPG_TARGET_BLOCKS 4096
#endif
#ifndef BPG_MAX_VCHUNKS
#define BPG_MAX_VCHUNKS 1024
#endif
#ifndef BPG_PINGPONG
#define BPG_PINGPONG 1
#endif

namespace bpg {

struct TransformArgs {
    const float* Yc;       // [N][Vc] compact, subject-major
    const double* Qn;      // [q][N] nuisance basis (column-major N x q)
    const double* dscale;  // [N]
    int N, q, Vc, Vpad;
    int KK;                // BF16 k-steps of 16 subjects (0: no BF16 output)
    int NS;                // FP32 operand rows (0: no FP32 output)
    double* yy;            // [Vpad] |y*|^2 (padding voxels: 1)
    unsigned* Yb;          // BF16 fragments, (Vpad/32)*KK*2*64 uint4
    float* Y32;            // [NS][Vpad]
};
hipError_t launch_transform(const TransformArgs& a, hipStream_t s);

struct OperandArgs {
    const unsigned short* perm;  // [K][N] (null for signs)
    const float* sign;           // [K][N] (null for permutations)
    const double* Qt;            // [pfit][N] rotated basis (column-major N x pfit)
    int N, pfit, K;
    int PP;                      // padded regressors per permutation (BF16 layout)
    int KK;                      // BF16 k-steps
    int n_sub;                   // 32-row subtiles
    int NS;                      // FP32 layout row length
    unsigned* Gb;                // BF16 fragments [n_sub][KK][2][64] uint4
    float* G32;                  // FP32 [K][pfit][NS]
};
hipError_t launch_operand_bf16(const OperandArgs& a, hipStream_t s);
hipError_t launch_operand_f32(const OperandArgs& a, hipStream_t s);

struct ContractArgs {
    const unsigned* Gb;    // BF16 operand fragments
    const unsigned* Yb;    // BF16 data fragments
    const float* G32;      // FP32 operand
    const float* Y32;      // FP32 data [NS][Vpad]
    const double* yy;      // [Vpad]
    float* partial;        // [n_vchunks][K] per-voxel-chunk maxima of the statistic
    float* stats;          // [K][Vc] or null
    unsigned* bits;        // [K][nwords] supra-threshold bitmask or null
    int nwords;
    int K, Vc, Vpad, N, pfit, r;
    int NS;                // FP32 operand row length
    int ftest;
    float sqrt_dof;        // t: sqrt(N - p)
    float f_scale;         // F: (N - p) / r
    float cdt;
};
// Grid of the split-BF16 contraction (see bpg_contract_bf16.h): blocks of
// 4 waves x R operand subtiles (32 permutation rows each) over chunks of
// 32-voxel tiles.
struct Bf16Launch {
    int n_mgroups, n_vchunks, vt_per_chunk, per_xcd, n_sub, n_vt;
};
int bf16_rows_per_wave(int KK);
Bf16Launch plan_contract_bf16(int K, int PP, int KK, int Vpad);
hipError_t launch_contract_bf16(const ContractArgs& a, const Bf16Launch& L, int KK, int PP,
                                bool sign_operand, hipStream_t s);

// Grid of the FP32 contraction: blocks of PTT permutations over chunks of
// 256-voxel steps.
struct F32Launch {
    int n_pg, n_vchunks, vox_per_chunk, per_xcd;
};
int f32_perm_tile(int pfit);
F32Launch plan_contract_f32(int K, int pfit, int Vpad);
hipError_t launch_contract_f32(const ContractArgs& a, const F32Launch& L, hipStream_t s);

// out[k] = max over the n_part chunk maxima partial[c*K + k] (NaN-free:
// chunks without in-mask voxels hold -inf).
hipError_t launch_reduce_max(const float* partial, int n_part, int K, float* out, hipStream_t s);

struct ObsArgs {
    const float* Yc;       // [N][Vc]
    const double* X;       // N x p column-major
    const double* Xp;      // p x N
    const double* C;       // C[r + c*p]
    const double* sc;      // t: [nc]; F: [nc*nc]
    int N, p, nc, Vc, ftest, nmaps;
    float* out;            // [nmaps][Vc]
};
hipError_t launch_observed(const ObsArgs& a, hipStream_t s);

struct PvalArgs {
    const float* obs;           // [Vc]
    const float* sorted_max;    // [K] ascending
    const unsigned* labels;     // grid labels (cluster modes) or null
    const unsigned* sizes;      // grid sizes (cluster modes) or null
    const unsigned* comp2grid;  // [Vc]
    int K, Vc, inference;
    float cdt;
    float* P;                   // [Vc]
};
hipError_t launch_pvalues(const PvalArgs& a, hipStream_t s);

// Connected components of the supra-threshold voxels of `nmaps` maps, given
// as compact-order bitmasks [nmaps][nwords].  parent/sizes are grid-sized
// per map slot ([nmaps][G]); only active voxels are touched.  sizes[root]
// counts voxels, or with `values` sums (uint)(T * 10000.0f) over the
// component (cluster mass); maxsize[map] is the largest.
struct CclArgs {
    const unsigned* bits;       // [nmaps][nwords]
    const unsigned* comp2grid;  // [Vc]
    const int* grid2comp;       // [G] (-1 outside the mask)
    unsigned* parent;           // [nmaps][G]
    unsigned* sizes;            // [nmaps][G]
    unsigned* maxsize;          // [nmaps]
    unsigned* labels;           // grid labels of map 0 (or null)
    const float* values;        // [nmaps][Vc] statistic values (cluster mass) or null
    int W, H, D, Vc, nwords, nmaps, connectivity;
};
hipError_t launch_ccl(const CclArgs& a, hipStream_t s);

// obs > cdt -> compact bitmask [nwords]
hipError_t launch_threshold_bits(const float* obs, int Vc, float cdt, unsigned* bits, int nwords,
                                 hipStream_t s);

}  // namespace bpg
