/*
 * bpg_oracle.h -- CPU oracle for the second-level permutation-test GLM.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library links, loads or
 * calls this code.  It is used by tests/ (to check the GPU results), by
 * __graft_entry__.smoke() (to check the GPU results) and by bench.py's cpu_baseline leg
 * (as the timed CPU restatement of the reference algorithm).
 *
 * Two precisions restate the reference path:
 *   - "ref32": the reference's OpenCL kernels restated op-for-op in FP32
 *     (same loop order, same formulas, same quirks), vectorised across voxels
 *     the way a CPU OpenCL runtime vectorises work-items;
 *   - "f64":   the same semantics in FP64 with the model (pseudo-inverse,
 *     contrast scalars, nuisance projector) rebuilt in FP64 from X and C by
 *     Householder QR.  This is the accuracy yardstick.
 *
 * Layouts follow the reference (code/BROCCOLI_LIB/broccoli_lib.h):
 *   Y[x + y*W + z*W*H + s*W*H*D], mask[x + y*W + z*W*H] (in-mask iff == 1.0f),
 *   X[s + r*N], xtxxt[s + r*N] = pinv(X)(r,s), C[r + c*p],
 *   ctxtxc[c] (t) or [c + cc*nc] (F), perm[s + k*N] (0-based), sign[s + k*N].
 */
#ifndef BPG_ORACLE_H
#define BPG_ORACLE_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* codes, as in code/BROCCOLI_LIB/broccoli_constants.h:51-54, 78-83 */
#define ORC_TTEST 0
#define ORC_FTEST 1
#define ORC_GROUP_MEAN 2
#define ORC_VOXEL 0
#define ORC_CLUSTER_EXTENT 1
#define ORC_CLUSTER_MASS 2
#define ORC_TFCE 3
#define ORC_TWOSAMPLE 0
#define ORC_CORRELATION 1
#define ORC_MEAN 2

#define ORC_PREC_REF32 0
#define ORC_PREC_F64 1

/* ---------------- generators (LIB:16925-17122) ---------------- */
void orc_srand(unsigned int seed);
int orc_rand(void);
void orc_random_shuffle_u16(unsigned short* a, int n);
void orc_random_shuffle_int(int* a, int n);
int orc_gen_sign_matrix(int N, long K, float* sign);
int orc_gen_perm_two_sample(int N, int n1, int n2, long K, unsigned short* perm);
int orc_gen_perm_correlation(int N, long K, unsigned short* perm);

/* ---------------- CLI model logic (RGL:38-42, 908-1194) ---------------- */
double orc_stirling_factorial(unsigned long n);
int orc_detect_designs(int N, int p, int nc, const float* X, const float* C,
                       int ftest, int* designs, int* n1, int* n2);
long orc_max_permutations(int test, int design, int N, int n2);
int orc_model_precompute(int N, int p, int nc, int ftest, const float* X,
                         const float* C, float* xtxxt, float* ctxtxc);

/* ---------------- FP64 linear algebra ---------------- */
int orc_qr64(int m, int n, const double* A, double* Q, double* R);
int orc_pinv64(int N, int p, const double* X, double* Xpinv);
int orc_nuisance_projector64(int N, int p, const double* X, int r,
                             const double* Crows, double* R);

/* ---------------- per-map statistics ---------------- */
void orc_transform32(int W, int H, int D, int N, const float* Rf, const float* mask, float* Y);
void orc_perm_map32(int test, int W, int H, int D, int N, int p, int nc,
                    const float* Y, const float* mask, const float* X,
                    const float* xtxxt, const float* C, const float* ctxtxc,
                    int contrast, const unsigned short* perm, const float* sign,
                    float* out);
void orc_obs_maps32(int ftest, int W, int H, int D, int N, int p, int nc,
                    const float* Y, const float* mask, const float* X,
                    const float* xtxxt, const float* C, const float* ctxtxc,
                    float* out);
float orc_max_quantized(int W, int H, int D, const float* map, const float* mask);
int orc_quantize(float t);

/* FP64 maps: Y is FP64, grid layout.  Model in FP64 (Xd[s + r*N], Cd[r + c*p]). */
void orc_perm_map64(int test, int W, int H, int D, int N, int p, int nc,
                    const double* Y, const float* mask, const double* Xd,
                    const double* Cd, int contrast, const unsigned short* perm,
                    const float* sign, double* out);
void orc_obs_maps64(int ftest, int W, int H, int D, int N, int p, int nc,
                    const double* Y, const float* mask, const double* Xd,
                    const double* Cd, double* out);

/* ---------------- clusters (KC:42-480, LIB:18163-18297) ---------------- */
unsigned orc_ccl(int W, int H, int D, const float* map, const float* mask,
                 float thr, int connectivity, unsigned* labels, unsigned* sizes);
unsigned orc_cluster_mass(int W, int H, int D, const float* map, const float* mask,
                          float thr, int connectivity, const unsigned* labels,
                          unsigned* masses);

/* Threshold-free cluster enhancement of a permuted map (LIB:16756-16762,
 * LIB:18464-18508, KC:482-510): thresholds h = 0, 0.2846, ... (float
 * accumulation) up to the map's quantised max; at each h the components of
 * {T > h} (connectivity as orc_ccl) add sqrt(size) * h * h to their voxels;
 * returns the quantised max of the sums over the mask.  tfce (optional,
 * may be NULL) receives the per-voxel sums. */
float orc_tfce_max(int W, int H, int D, const float* map, const float* mask, int connectivity,
                   float* tfce);

/* ---------------- p-values (KC:516-652) ---------------- */
void orc_pvalues(int inference, int W, int H, int D, const float* map,
                 const float* mask, const unsigned* labels, const unsigned* sizes,
                 float thr, long K, const float* maxdist, float* P);

/* ---------------- the whole path (LIB:13144-13376, 16589-16873) ------- */
typedef struct {
    int test;             /* ORC_TTEST / ORC_FTEST / ORC_GROUP_MEAN */
    int inference;        /* ORC_VOXEL / ORC_CLUSTER_EXTENT / ORC_CLUSTER_MASS / ORC_TFCE */
    int W, H, D;
    int N, p, nc;
    const float* Y;
    const float* mask;
    const float* X;
    const float* xtxxt;
    const float* C;
    const float* ctxtxc;
    const long* K;                        /* per statistical map */
    const unsigned short* const* perm;    /* per map, [s + k*N] (t / F) */
    const float* sign;                    /* [s + k*N] (group mean) */
    float cdt;
    int connectivity;                     /* 26 (reference) or 6 */
    int nthreads;                         /* OpenMP threads, <=0: default */
} orc_study;

/* outputs: maps[c*WHD], pvalues[c*WHD], maxdist[c][K[c]] (float**) */
int orc_randomise(const orc_study* st, int precision, float* maps,
                  float* pvalues, float** maxdist);

/* statistic maps of the first K permutations of map c, out[k*W*H*D + i] */
int orc_perm_stats(const orc_study* st, int c, int precision, long K, double* out);

/* FP64 max over the in-mask voxels of the statistic of the first K
 * permutations of map c (not quantised), maxv[K] */
int orc_perm_max64(const orc_study* st, int c, long K, double* maxv);

/* the timed CPU restatement: permutation phase only, one contrast, voxel
 * inference, ref32 precision.  Returns max-T values in maxdist[K]. */
int orc_permutation_phase32(const orc_study* st, int contrast, long K,
                            float* maxdist);

#ifdef __cplusplus
}
#endif
#endif
