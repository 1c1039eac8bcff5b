/*
 * fl_oracle.h -- CPU oracle for the first-level (single-subject) permutation
 * test: regression, AR(4) whitening prior to the permutations, inverse
 * whitening of permuted whitened volumes, and the per-permutation t / F map.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library links, loads or
 * calls this code.  It is used by tests/ (to check the GPU results) and by bench.py's
 * cpu_baseline leg.
 *
 * Everything is FP64 with the reference's formulas and quirks kept; the
 * smoothing filters are the reference's float taps.  Layouts follow the
 * reference: volumes[x + y*W + z*W*H + t*W*H*D], mask[x + y*W + z*W*H]
 * (in-mask iff == 1.0f), X[t + r*T], C[r + c*p], AR fields ar[k*G + i]
 * (k = 0..3 for AR1..AR4, G = W*H*D), perm[t + k*T] (0-based).
 */
#ifndef BPG_FL_ORACLE_H
#define BPG_FL_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_FL_TAPS 9

/* Gaussian smoothing filters of `size` taps (LIB:11055-11090): taps
 * (float)exp(-u^2 / (2 sigma^2)) with sigma = fwhm / 2.354 / voxel size,
 * summed in float and divided by the sum. */
void orc_fl_filters(int size, float fwhm, float vx, float vy, float vz, float* fx, float* fy, float* fz);

/* Separable 9-tap smoothing with zero padding: along y with fy, then x with
 * fx, then z with fz (kernelConvolution.cpp:52-91, 531-569, 1201-1240).
 * The first pass multiplies the input by `certainty` (NULL: ones); the last
 * divides by `scert` (NULL: ones). */
void orc_fl_smooth(int W, int H, int D, const double* in, const float* certainty, const double* scert,
                   const float* fx, const float* fy, const float* fz, double* out);

/* Residuals of the full design (LIB:12144-12194): e = y - X pinv(X) y at
 * in-mask voxels, 0 elsewhere. */
void orc_fl_regress(int W, int H, int D, int T, int p, const double* Y, const float* mask, const double* X,
                    double* out);

/* Voxel-wise AR(4) estimates (kernelWhitening.cpp:103-235): lagged sums
 * c_k = sum_{t >= invalid + k} v_t v_{t-k} divided by T - 1 - k - invalid,
 * r_k = c_k / c0, the Toeplitz matrix of 1 and r_k + 0.001, and its inverse
 * as adjugate / (det + 0.001) applied to r.  0 outside the mask or when
 * c0 == 0. */
void orc_fl_ar4(int W, int H, int D, int T, const double* data, const float* mask, int invalid, double* ar);

/* AR(4) whitening (kernelWhitening.cpp:371-423): out_t = v_t - sum_{i<=min(t,4)}
 * a_i v_{t-i} at in-mask voxels; other voxels are copied. */
void orc_fl_whiten(int W, int H, int D, int T, const double* data, const double* ar, const float* mask,
                   double* out);

/* Inverse whitening of permuted whitened volumes
 * (kernelWhitening.cpp:480-535): x_t = sum_{i<=min(t,4)} a_i x_{t-i} +
 * w_{perm[t]} at in-mask voxels, 0 elsewhere. */
void orc_fl_color(int W, int H, int D, int T, const double* white, const double* ar, const float* mask,
                  const unsigned short* perm, double* out);

/* Whitening prior to the permutations (LIB:16363-16452, after the
 * regression of LIB:16469-16473): three rounds of AR(4) estimation (invalid
 * time points 0, then 4), normalised smoothing of the estimates with the AR
 * smoothing filter, accumulation into the total estimates and whitening of
 * the residuals with the totals.  Returns the whitened volumes and the total
 * estimates times the mask. */
void orc_fl_prepare(int W, int H, int D, int T, int p, const float* Y, const float* mask, const float* X,
                    float fwhm, float vx, float vy, float vz, double* whitened, double* ar);

/* First-level permutation statistic of volumes x (kernelStatistics3.cpp:
 * 2107-2190 for t, kernelStatistics5.cpp:1943-2060 for F): beta = pinv(X) x,
 * residual variance with Welford's running mean over T values divided by
 * T - 1, t = c'beta / sqrt(var c'(X'X)^-1 c), F = (C beta)' (C (X'X)^-1 C')^-1
 * (C beta) / (var nc).  test 0: t of contrast `contrast`; 1: F over all nc
 * contrasts.  NaN-free only where var > 0; out-of-mask voxels get 0. */
void orc_fl_stat(int test, int W, int H, int D, int T, int p, int nc, const double* x, const float* mask,
                 const double* X, const double* C, int contrast, double* out);

/* Random time-point permutations (LIB:16877-16922): cumulative
 * random_shuffle of 0..T-1 from the oracle's rand() stream, each rejected
 * until distinct from the identity and every earlier row; row k is the
 * (k+1)-th accepted shuffle. */
int orc_gen_perm_first_level(int T, long K, unsigned short* perm);

/* The null distribution of one statistical map: for each permutation k,
 * inverse whitening with perm k, the statistic, and the quantised max over
 * the mask (inference 0), the largest cluster extent (1) or mass (2) above
 * cdt.  stats (optional, K*G) receives the statistic maps. */
int orc_fl_null(int test, int inference, int W, int H, int D, int T, int p, int nc, const double* white,
                const double* ar, const float* mask, const float* X, const float* C, int contrast, long K,
                const unsigned short* perm, float cdt, int connectivity, float* maxdist, double* stats);

#ifdef __cplusplus
}
#endif
#endif
