// bpg_kernels.h -- host-callable launchers of the gfx950 kernels.
#pragma once

#include <hip/hip_runtime.h>

// Tuning knobs of the split-BF16 contraction:
//  - operand subtiles kept per wave: BPG_R_SIGN for exact +-1 operands with
//    KK <= 4 (N <= 64), BPG_R_SMALL for general operands with KK <= 4, 2
//    up to KK = BPG_R2_KK, else 1 (R = 2 at KK = 13 needs 256 VGPRs and
//    about 220 AGPRs, one wave per SIMD, and halves the data-tile loads; at
//    14..16 it spills, and R = 1 ran 8% faster at KK = 15);
//  - BPG_PIPE_SMALL / _MID / _LARGE: software-pipelined epilogue for
//    general operands at KK <= 4, KK <= 8 and above;
//  - the number of blocks the voxel range is split over.
#ifndef BPG_R_SIGN
#define BPG_R_SIGN 4
#endif
#ifndef BPG_R_SMALL
#define BPG_R_SMALL 3
#endif
#ifndef BPG_R2_KK
#define BPG_R2_KK 13
#endif
// operand subtiles per wave at KK 5..8 (the rolling schedule)
#ifndef BPG_R_MID
#define BPG_R_MID 2
#endif
#ifndef BPG_PIPE_SMALL
#define BPG_PIPE_SMALL 1
#endif
#ifndef BPG_PIPE_MID
#define BPG_PIPE_MID 0
#endif
#ifndef BPG_PIPE_LARGE
#define BPG_PIPE_LARGE 1
#endif
// F statistics with the bitmask alone: thresholds compared as products
#ifndef BPG_F_BITS_PRODUCT
#define BPG_F_BITS_PRODUCT 1
#endif
#ifndef BPG_ROLL_MIN_KK
#define BPG_ROLL_MIN_KK 5
#endif
#ifndef BPG_TARGET_BLOCKS
#define BPG_TARGET_BLOCKS 4096
#endif
#ifndef BPG_MAX_VCHUNKS
#define BPG_MAX_VCHUNKS 1024
#endif

namespace bpg {

struct TransformArgs {
    const float* Yc;       // [N][Vc] compact, subject-major
    const double* Qn;      // [NR][transform_width(q)] nuisance basis, row-major, zero-padded
    const double* dscale;  // [NR], zero beyond N (NR = N rounded up to 16)
    int N, q, Vc, Vpad;
    int KK;                // BF16 k-steps of 16 subjects (with Yb)
    int npy;               // BF16 fragments per subject in Yb: 2 (hi, lo) or 3 (hi, mid, lo)
    int NS;                // FP32 / FP64 operand rows, >= N (with Y32 or Y64)
    double* yy;            // [Vpad] |y*|^2 (padding voxels: NaN)
    unsigned* Yb;          // BF16 fragments, (Vpad/32)*KK*npy*64 uint4
    float* Y32;            // [NS][Vpad]
    double* Y64;           // [NS][Vpad] (the FP64 form; dscale all ones)
    float* Ym;             // FP32-MFMA data fragments [Vpad/32][KG][64] float4 (KG = ceil(N / 8))
};
// padded width of the nuisance basis for q columns (0 for none, -1 if q > 24)
int transform_width(int q);
hipError_t launch_transform(const TransformArgs& a, hipStream_t s);
// the transforms of n studies of one nuisance width QB in one launch: study
// j takes blocks first[j] .. first[j + 1] - 1 (first[j + 1] - first[j] =
// its Vpad / 256), `blocks` in all; tab and first in device memory
hipError_t launch_transform_batch(const TransformArgs* tab, const int* first, int n, int blocks, int QB,
                                  hipStream_t s);

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
    float* Gm;                   // FP32-MFMA fragments [n_sub][KG][64] float4
    // optional: permutation indices >= N (which the host has not checked)
    // are clamped and counted here
    unsigned* bad_index;
};
hipError_t launch_operand_bf16(const OperandArgs& a, hipStream_t s);
hipError_t launch_operand_f32(const OperandArgs& a, hipStream_t s);

struct ContractArgs {
    const unsigned* Gb;    // BF16 operand fragments
    const unsigned* Yb;    // BF16 data fragments
    const float* G32;      // FP32 operand
    const float* Y32;      // FP32 data [NS][Vpad]
    const float* Gm;       // FP32-MFMA operand fragments
    const float* Ym;       // FP32-MFMA data fragments
    const unsigned* Gf8;   // split-FP8 operand fragments and scale words
    const unsigned* Gsc;
    const unsigned* Yf8;   // split-FP8 data fragments and scale words
    const unsigned* Ysc;
    const double* Y64;     // FP64 data [NS][Vpad] (unscaled y*)
    const unsigned short* perm;  // FP64 form: [K][N] permutations (or null)
    const float* sign;           // FP64 form: [K][N] signs (or null)
    const double* Qt;            // FP64 form: rotated basis, Qt[s + i*N]
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
    // conditioning check: SSE = |y*|^2 - |z|^2 loses relative accuracy as
    // SSE / |y*|^2 shrinks; a wave that meets SSE < tau * |y*|^2 adds one
    // to *unsafe
    float tau;
    float pinf;            // +infinity (a run-time value: the running max is one v_med3_f32)
    unsigned* unsafe;
    unsigned char* flags;  // [K]: 1 for a permutation with an ill-conditioned sample
    // cluster inference: a permutation whose map has an in-mask statistic in
    // (cdt_lo, cdt_hi] -- within the form's error band of the threshold --
    // gets ties[k] = 1 (its largest cluster may differ from an exact
    // computation's); cdt_lo == cdt_hi disables the check
    float cdt_lo, cdt_hi;
    unsigned char* ties;   // [K] or null
    double dof;            // FP64 form: N - p
    // split-BF16 max-T only, with a sample list (samples non-null): an
    // ill-conditioned sample is left out of the maximum and appended as
    // (k << 32 | v) for k_fix_samples, which recomputes it in FP64.  The list
    // has BPG_SAMPLE_SEGS segments of sample_cap entries, segment g at
    // samples + g * sample_cap with its own counter n_samples[g] (64-bit: it
    // never wraps), so that appends from different waves rarely meet on one
    // counter; a wave appends to segment (4 blockIdx + wave) mod
    // BPG_SAMPLE_SEGS.  A sample past its segment's capacity flags its
    // permutation and adds one to *unsafe instead, as without the list.
    unsigned long long* samples;
    unsigned long long* n_samples;
    unsigned long long sample_cap;
};
constexpr int BPG_SAMPLE_SEGS = 256;
// SSE / |y*|^2 below which a sample is counted as ill-conditioned: the
// split-BF16 z carries ~1e-5 relative error per unit of |z|^2 / SSE, the
// FP32 z ~4e-7
constexpr float BPG_TAU_BF16 = 0.125f;
constexpr float BPG_TAU_FP32 = 1.0f / 256.0f;
// the split-FP8 z carries the FP8 matrix cores' summation error, ~3e-5 of
// the sum of |products| per instruction
constexpr float BPG_TAU_FP8 = 0.5f;
// Grid of the split-BF16 contraction (see bpg_contract_bf16.h): blocks of
// 4 waves x R operand subtiles (32 permutation rows each) over chunks of
// 32-voxel tiles.
struct Bf16Launch {
    int n_mgroups, n_vchunks, vt_per_chunk, per_xcd, n_sub, n_vt;
    int variant;  // 0: the default subtiles per wave; 1: the tuner's alternative
};
// operand subtiles per wave (the planner and the kernels share this rule)
constexpr int bf16_rows_per_wave(int KK, bool sign) {
    return KK <= 4 ? (sign ? BPG_R_SIGN : BPG_R_SMALL) : (KK <= 8 ? BPG_R_MID : (KK <= BPG_R2_KK ? 2 : 1));
}
// The alternative the autotuner may choose for the max-T-only kernels up to
// 8 k-steps (N <= 128): 2 subtiles instead of 4 (signs) or 3 (N <= 64),
// 1 instead of 2 at 5..8 k-steps; elsewhere the default.
constexpr int bf16_rows_per_wave_alt(int KK, bool sign) {
    return KK <= 4 ? 2 : (KK <= 8 ? 1 : bf16_rows_per_wave(KK, sign));
}
constexpr int bf16_rows_per_wave_v(int KK, bool sign, int variant) {
    return variant ? bf16_rows_per_wave_alt(KK, sign) : bf16_rows_per_wave(KK, sign);
}
// data fragments per subject: hi + mid + lo for exact +-1 operands at
// KK <= 4 (every product exact, so the FP32 form's conditioning bound
// applies), hi + lo otherwise
constexpr int bf16_y_pieces(int KK, bool sign) { return sign && KK <= 4 ? 3 : 2; }
constexpr bool bf16_pipelined(int KK, bool sign) {
    return !sign && (KK <= 4 ? BPG_PIPE_SMALL : (KK <= 8 ? BPG_PIPE_MID : BPG_PIPE_LARGE)) != 0;
}
// slots_per_xcd: blocks of the contraction that one XCD holds at once;
// variant: subtiles per wave (bf16_rows_per_wave_v); chunk_pow: the
// voxel-chunk length scaled by 2^chunk_pow from the planner's (autotuner)
Bf16Launch plan_contract_bf16(int K, int PP, int KK, int Vpad, bool sign, int slots_per_xcd, int variant = 0,
                              int chunk_pow = 0);
// A batch of studies in one persistent contraction per class: the studies'
// launch arguments and grids in device memory, item0 the first work item of
// each study (prefix sums of n_mgroups x n_vchunks), the queue counter
// cleared before each launch.
// The work items of a batch class sit in one queue per XCD: queue x holds,
// study after study, the (permutation group, voxel chunk) items of the voxel
// chunks vc = x, x + 8, ..., so the blocks of one XCD share their chunks' data
// in that XCD's L2.
constexpr int BPG_XCDS = 8;
struct BatchStudy {
    ContractArgs a;
    Bf16Launch L;
    int item0[BPG_XCDS];  // its first item in each queue
};
struct BatchClass {
    const BatchStudy* studies;
    int n_studies;
    int n_items[BPG_XCDS];
    unsigned* counter;  // BPG_XCDS queue heads
};
// persistent launch of a class (t statistics, max-T only, KK <= 8; with
// sample lists -- excl -- permutation operands up to KK = 4); with occ set,
// stores the blocks per CU of the instantiation instead
hipError_t launch_contract_bf16_persist(const BatchClass& c, int KK, int PP, bool sign, int variant, bool excl,
                                        int grid, hipStream_t s, int* occ = nullptr);
// the largest k-step count whose max-T-only contraction keeps a sample list
constexpr int BPG_EXCL_MAX_KK = 4;
// with occ set, stores the blocks per CU of the instantiation instead of
// launching (see run_bf16)
hipError_t launch_contract_bf16(const ContractArgs& a, const Bf16Launch& L, int KK, int PP,
                                bool sign_operand, hipStream_t s, int* occ = nullptr);
// blocks of the contraction instantiation for these operands per CU
int contract_bf16_blocks_per_cu(int KK, int PP, bool sign, bool ftest, bool out, int variant = 0);

// Grid of the FP32 contraction: blocks of PTT permutations over chunks of
// 256-voxel steps.
struct F32Launch {
    int n_pg, n_vchunks, vox_per_chunk, per_xcd;
};
int f32_perm_tile(int pfit);
F32Launch plan_contract_f32(int K, int pfit, int Vpad);
hipError_t launch_contract_f32(const ContractArgs& a, const F32Launch& L, hipStream_t s);

// FP32 form on the matrix cores (bpg_contract_f32mx.hip): the split-BF16
// kernel's grid and operand row order, one v_mfma_f32_32x32x2_f32 per
// k-step of 2 subjects, operands streamed per group of 8 subjects.  For fit
// columns up to 16 (the accumulator slots of one lane); wider designs use
// the vector-ALU kernel above.
// Operand subtiles per wave: 2 up to 80 subjects (each wave's subtiles take
// 2 KG KiB of LDS, 64 KiB per block at KG = 8: two blocks per CU), 1 above.
constexpr int f32mx_rows_per_wave(int KG) { return KG <= 10 ? 2 : 1; }
constexpr size_t f32mx_lds_bytes(int KG) { return (size_t)4 * f32mx_rows_per_wave(KG) * KG * 64 * 16; }
Bf16Launch plan_contract_tiles(int K, int PP, int R, int Vpad, int slots_per_xcd, int chunk_pow = 0);
hipError_t launch_contract_f32mx(const ContractArgs& a, const Bf16Launch& L, int PP, int KG, hipStream_t s,
                                 int* occ = nullptr);
int contract_f32mx_blocks_per_cu(int PP, int KG, bool ftest, bool out);
hipError_t launch_operand_f32mx(const OperandArgs& a, int KG, hipStream_t s);

// Split-FP8 form (bpg_contract_f8.hip): BPG_F8_TERMS e4m3 terms per operand
// with an E8M0 scale per 32 subjects, products of terms i + j < T on
// v_mfma_scale_f32_32x32x64_f8f6f4, one operand term for exact +-1 signs.
// Up to 256 subjects (4 k-steps of 64) and 16 fit columns.
#ifndef BPG_F8_TERMS
#define BPG_F8_TERMS 4
#endif
constexpr int f8_rows_per_wave(int KS) { return KS <= 2 ? 2 : 1; }
constexpr size_t f8_lds_bytes(int KS, bool sign) {
    return (size_t)4 * f8_rows_per_wave(KS) * KS * ((sign ? 1 : BPG_F8_TERMS) * 128 + 64) * 16;
}
hipError_t launch_contract_f8(const ContractArgs& a, const Bf16Launch& L, int PP, int KS, bool sign, hipStream_t s,
                              int* occ = nullptr);
int contract_f8_blocks_per_cu(int PP, int KS, bool sign, bool ftest, bool out);
// data fragments from the FP32 operand rows Y32 [NS][Vpad] (k_transform's
// FP32 output): Yf8 [Vpad/32][KS][T][64][32 bytes], Ysc [Vpad/32][KS][64]
hipError_t launch_data_f8(const float* Y32, int NS, int N, int Vpad, int KS, unsigned* Yf8, unsigned* Ysc,
                          hipStream_t s);
// operand fragments Gf8 [n_sub][KS][T][64][32 bytes], Gsc [n_sub][KS][64]
hipError_t launch_operand_f8(const OperandArgs& a, int KS, bool sign, unsigned* Gf8, unsigned* Gsc, hipStream_t s);

// FP64 form (the repeat of permutations with ill-conditioned samples):
// blocks of f64_perm_tile(pfit) permutations over chunks of 256-voxel
// steps, the operand rows built in LDS from the permutations / signs and
// Qt, z and the residual sum of squares SSE = |y* - B z|^2 (B the permuted
// basis) accumulated in FP64 from the FP64 y*.
int f64_perm_tile(int pfit);
F32Launch plan_contract_f64(int K, int pfit, int Vpad);
hipError_t launch_contract_f64(const ContractArgs& a, const F32Launch& L, hipStream_t s);

// out[k] = max over the n_part chunk maxima partial[c*K + k] (NaN-free:
// chunks without in-mask voxels hold -inf).
hipError_t launch_reduce_max(const float* partial, int n_part, int K, float* out, hipStream_t s);
// several per-permutation max reductions in one launch (K <= max_K each)
struct ReduceJob {
    const float* partial;
    float* out;
    int n_part, K;
};
hipError_t launch_reduce_max_jobs(const ReduceJob* jobs, int n_jobs, int max_K, hipStream_t s);

// The samples a split-BF16 max-T-only contraction left out (ContractArgs::
// samples), recomputed in FP64 as the FP64 form computes them -- y* from
// the compact data and the nuisance basis, z = B' y* with B the permuted
// basis rows, the residual formed itself -- and merged into maxout[k] with
// an order-preserving atomic max (a NaN statistic is left out, as the
// contraction's running max leaves it out).
struct SampleFixArgs {
    const unsigned long long* samples;    // BPG_SAMPLE_SEGS segments of cap entries
    const unsigned long long* n_samples;  // per segment; entries past cap are not read
    unsigned long long cap;
    const float* Yc;              // [N][Vc]
    const double* Qn;             // [NR][QB] nuisance basis, row-major
    int QB, q;
    const double* Qt;             // N x pfit, column-major
    const unsigned short* perm;   // [K][N]
    int N, pfit, Vc, K;
    double dof;
    float* maxout;                // [K]
};
// t statistics, up to 64 subjects, 16 fit columns and 24 nuisance columns
constexpr int BPG_FIX_MAX_N = 64;
constexpr int BPG_FIX_MAX_P = 16;
hipError_t launch_fix_samples(const SampleFixArgs& a, hipStream_t s);

struct ObsArgs {
    const float* Yc;       // [N][Vc]
    const double* X;       // N x p column-major
    const double* Xp;      // p x N
    const double* C;       // C[r + c*p]
    const double* sc;      // t: [nc]; F: [nc*nc]
    int N, p, nc, Vc, ftest, nmaps;
    float* out;            // [nmaps][Vc]
    float* betas;          // [p][Vc] or null
    float* residuals;      // [N][Vc] or null
    float* resvar;         // [Vc] or null
};
hipError_t launch_observed(const ObsArgs& a, hipStream_t s);

struct PvalArgs {
    const float* obs;           // [Vc]
    const float* sorted_max;    // [K] ascending
    const unsigned* labels;     // [Vc] compact roots (cluster modes) or null
    const unsigned* sizes;      // [Vc] component sizes / masses at the roots, or null
    int K, Vc, inference;
    float cdt;
    float* P;                   // [Vc]
    // optional: per voxel, the comparisons with the null inside the error
    // band (voxel inference: null values within 1e-4 (2 + |T|) of T;
    // cluster inference: cluster_ties, the permutations whose largest
    // cluster may differ, for supra-threshold voxels; TFCE: 0)
    unsigned* ties;             // [Vc] or null
    unsigned cluster_ties;
};
hipError_t launch_pvalues(const PvalArgs& a, hipStream_t s);

// Connected components of the supra-threshold voxels of `nmaps` maps, given
// as compact-order bitmasks [nmaps][nwords].  Union-find runs on compact
// indices (in-mask voxels in linear grid order, so the minimum compact index
// of a component is its minimum grid index): parent/sizes are [nmaps][Vc],
// and only active voxels are touched.  nbr holds, per compact voxel, the
// compact indices of its 13 lower-index 26-neighbours (-1 outside the mask),
// the 3 face neighbours first, so 6-connectivity reads the first 3.
// sizes[root] counts voxels, or with `values` sums (uint)(T * 10000.0f) over
// the component (cluster mass); maxsize[map] is the largest.
constexpr int BPG_NBR = 13;
struct CclArgs {
    const unsigned* bits;       // [nmaps][nwords]
    const int* nbr;             // [Vc][BPG_NBR]
    unsigned* parent;           // [nmaps][Vc]
    unsigned* sizes;            // [nmaps][Vc]
    unsigned* maxsize;          // [nmaps]
    unsigned* labels;           // compact root of every active voxel of map 0 (or null)
    const float* values;        // [nmaps][Vc] statistic values (cluster mass) or null
    float* tfce;                // [nmaps][Vc] TFCE sums (stage 4) or null
    const float* sqrt_size;     // [Vc + 1] correctly rounded sqrt(n) (stage 4)
    float level;                // TFCE threshold of stage 4
    const unsigned* map_ids;    // [nmaps] batch map of each slot (null: slot = map)
    int Vc, nwords, nmaps, connectivity;
};
// Whole-map CCL in LDS, one block per map (map_ids[b], or b, for the n_maps
// blocks): the map's bitmask, the ranks of its words (exclusive prefix of
// popcounts) and a union-find over the map's active voxels all live in LDS.
// blocks = 2: 512-thread blocks, two per CU, each with half the LDS;
// blocks = 1: 1024-thread blocks with all of it.  Maps with more active
// voxels than the block holds are appended to overflow[*n_overflow], for
// the next tier or the global path.  Writes maxsize[map].
// lds_bytes: the device's shared memory per block (at most 160 KB is used)
bool ccl_lds_fits(int nwords, size_t lds_bytes);
hipError_t launch_ccl_lds(const CclArgs& a, size_t lds_bytes, int blocks, const unsigned* map_ids, int n_maps,
                          unsigned* overflow, unsigned* n_overflow, hipStream_t s);
// stage 4 (TFCE): tfce[map][c] += sqrt(size of c's component) * level^2
hipError_t launch_ccl_tfce_level(const CclArgs& a, hipStream_t s);

// TFCE level masks: bits[m][.] = stats[m][v] > level for maps with
// nlev[m] > j (empty otherwise)
hipError_t launch_level_bits(const float* stats, const int* nlev, int j, float level, int nmaps, int Vc,
                             int Vpad, unsigned* bits, int nwords, hipStream_t s);
// out[m] = max over v < Vc of x[m][v]
hipError_t launch_rowmax(const float* x, int nmaps, int Vc, float* out, hipStream_t s);
hipError_t launch_ccl(const CclArgs& a, hipStream_t s);

// The no-allocation permutation phase (bpg_noalloc.hip): the per-device
// scratch array (128 MiB), its device address (call with the device
// current), a fill kernel and the device-side quantisation of the maxima.
constexpr size_t BPG_SCRATCH_BYTES = (size_t)128 << 20;
hipError_t scratch_address(int device, void** out);
hipError_t launch_fill_u32(unsigned* p, size_t n, unsigned v, hipStream_t s);
hipError_t launch_quantize_max(const float* m, int K, float* out, float empty, hipStream_t s);

// obs > cdt -> compact bitmask [nwords]
hipError_t launch_threshold_bits(const float* obs, int Vc, float cdt, unsigned* bits, int nwords,
                                 hipStream_t s);

}  // namespace bpg
