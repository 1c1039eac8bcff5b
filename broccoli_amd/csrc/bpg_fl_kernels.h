// bpg_fl_kernels.h -- launchers of the first-level permutation-test kernels
// (bpg_firstlevel.hip).  Compact layout: in-mask voxels v < Vc in linear
// grid order, padded to Vpad (a multiple of 64); time-major arrays are
// [t][Vpad].  All set-up arithmetic is FP64.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>

namespace bpg {
namespace fl {

constexpr int MAXP = 25;       // regressors (the reference's beta[25])
constexpr int TILE = 64;       // voxels per tile (one wave)
constexpr int WAVES = 8;       // waves per block of the permutation kernel
constexpr int NP = 2;          // permutations per thread at once
constexpr int ROUNDS = 4;      // rounds of NP per wave
constexpr int PERMS_PER_BLOCK = WAVES * NP * ROUNDS;
constexpr int MAX_VCHUNKS = 1024;
// largest T whose whitened voxel tile ([T][64] floats) is staged in LDS
constexpr int LDS_T_MAX = 600;
// largest T of the FP32 / MFMA form ([T][64] tile in LDS next to the
// 6.25 KB of per-wave reduction space)
constexpr int MX_T_MAX = 560;
// largest T of the split-FP16 form (the tile plus the [T/8][2][NB][16][8]
// FP16 basis in LDS): one block of 16 basis columns (pq <= 16), two (pq <= 32)
constexpr int HX_T_MAX = 480;
constexpr int HX_T_MAX2 = 384;
// the split-FP16 form scales each voxel's whitened series by a power of two
// that brings its largest magnitude into [2^8, 2^9); the basis is scaled by
// HX_QSCALE = 2^4
constexpr float HX_QSCALE = 16.0f;
#define BPG_FL_MAX_FROWS 10

struct Taps {
    float f[9];
};

// e[t][v] = y[t][v] - sum_r X[t][r] beta_r, beta = Xp y.  Xp[r*T + t]
// (pinv(X)), X[t*p + r].
hipError_t launch_regress(const float* Y, const double* Xp, const double* X, int T, int p, int Vc, int Vpad,
                          double* e, hipStream_t s);
// AR(4) estimates of src[t][v] from time point `invalid` on: ar[k*Vpad + v]
hipError_t launch_ar4(const double* src, int T, int invalid, int Vc, int Vpad, double* ar, hipStream_t s);
// g[k*G + comp2grid[v]] = ar[k*Vpad + v] for k < 4 (other voxels untouched)
hipError_t launch_scatter(const double* ar, const unsigned* comp2grid, int Vc, int Vpad, size_t G, double* g,
                          hipStream_t s);
// one zero-padded 9-tap pass along axis 0 (x), 1 (y) or 2 (z) over nf
// fields of W*H*D voxels; tap j weighs the neighbour at offset 4 - j
hipError_t launch_conv(const double* in, double* out, int nf, int W, int H, int D, int axis, const Taps& taps,
                       hipStream_t s);
// the z pass at the compact voxels of 4 fields, divided by the smoothed
// mask and added to total[k*Vpad + v]
hipError_t launch_rods_accumulate(const double* g, const double* smask, const unsigned* comp2grid, int Vc,
                                  int Vpad, int W, int H, int D, const Taps& taps, double* total, hipStream_t s);
// AR(4) whitening of e with ar: wh (FP64) and, when non-null, w32 (FP32)
hipError_t launch_whiten(const double* e, const double* ar, int T, int Vc, int Vpad, double* wh, float* w32,
                         hipStream_t s);
// wscale[v] = 2^(9 - e) with max_t |w32[t][v]| = f 2^e, f in [0.5, 1)
// (1 for an all-zero series and for the padding)
hipError_t launch_wscale(const float* w32, int T, int Vc, int Vpad, float* wscale, hipStream_t s);

// The permutation kernel.  For permutation k and voxel v:
//   x_t = sum_{i<=min(t,4)} a_i x_{t-i} + w[perm_k[t]][v]      (inverse whitening)
//   z = Q' x, s = |x|^2, m = sum x                              (FP64)
//   SSE = s - |z|^2, var = (SSE - (m - u'z)^2 / T) / (T - 1)
//   (the mean term only when the constant is outside col(X): mean = 1)
//   t = (A_0 z) / sqrt(var),  F = sum_i (A_i z)^2 / (var * nrows)
// Q is an orthonormal basis of col(X) (T x pq, zero-padded columns), u =
// Q'1, A the statistic rows.
struct PermArgs {
    const float* w32;              // [T][Vpad] whitened volumes
    const double* ar;              // [4][Vpad]
    const unsigned short* perm;    // [K][Tp] (rows padded to a multiple of 8)
    const double* Q;               // [T][pq]
    const double* A;               // [nrows][pq]
    const double* u;               // [pq]
    const float* Q32;              // [Tq][16] FP32 rotated basis, rows to T rounded up to 4 zero (FP32 form)
    const float* u32;              // [16] FP32 rotated Q'1 (FP32 and split forms; [32] for the split-FP16 form)
    const unsigned* Qb;            // [Tq32/32][2][64] lanes x 8 bf16: the rotated basis, hi then lo (split-BF16 form)
    const unsigned* Qh;            // [T8][2][NB][16][8] fp16: HX_QSCALE x the rotated basis, hi then lo (split-FP16 form)
    const float* wscale;           // [Vpad] power-of-two scale of each whitened series (split-FP16 form)
    float neg_one;                 // -1 (split-FP16 form)
    unsigned char* flags;          // [K] ill-conditioned permutations (FP32 form)
    float* partial;                // [n_vchunks][K] per-chunk maxima
    float* stats;                  // [K][Vc] or null
    unsigned* bits;                // [K][nwords] stat > cdt, or null
    float cdt;
    int T, Tp, pq, nrows, ftest, mean, K, Vc, Vpad, nwords;
    int n_vtiles, n_vchunks, tiles_per_chunk, n_pgroups;
};
// pq must be one of the compiled widths (fl_width)
int fl_width(int p);
hipError_t launch_perm(PermArgs a, hipStream_t s);
// The FP32 form (pq <= 16, T <= MX_T_MAX): the recursion, |x|^2 and sum x
// in FP32, z on the FP32 matrix cores; flags[k] = 1 for a permutation with a
// sample of SSE < |x|^2 / 64 (the caller clears flags and repeats those in
// the FP64 form).
hipError_t launch_perm_mx(PermArgs a, hipStream_t s);
// The split-BF16 form (same limits): projections on the BF16 matrix cores
// from hi/lo splits of x and of the basis; flags samples with SSE < |x|^2/8.
hipError_t launch_perm_bx(PermArgs a, hipStream_t s);
// The split-FP16 form (pq <= 16 and T <= HX_T_MAX, or pq <= 32 and
// T <= HX_T_MAX2, with two 16-column blocks of the basis): each lane runs its own
// voxel's recursion and feeds 8 time points at a time, split into FP16 hi +
// lo, to v_mfma_f32_32x32x16_f16 against the scaled FP16 basis (block-
// diagonal over the two lane halves); flags samples with SSE < |x|^2/32 or
// |x|^2 >= 4e9 (a value beyond the FP16 range).
hipError_t launch_perm_hx(PermArgs a, hipStream_t s);

}  // namespace fl
}  // namespace bpg
