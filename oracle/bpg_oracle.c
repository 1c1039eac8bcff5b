/*
 * bpg_oracle.c -- CPU oracle for the second-level permutation-test GLM.
 *
 * TEST INFRASTRUCTURE ONLY (see bpg_oracle.h).  The product library never
 * links or calls this file; tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py use it to check the GPU results or as the
 * timed CPU restatement of the reference algorithm.
 *
 * Every function names the reference lines it restates.  Reference paths are
 * relative to the BROCCOLI repository:
 *   LIB = code/BROCCOLI_LIB/broccoli_lib.cpp
 *   RGL = code/Bash_Wrapper/RandomiseGroupLevel.cpp
 *   K1/K2/K4 = code/Kernels/kernelStatistics{1,2,4}.cpp
 *   KC = code/Kernels/kernelClusterize.cpp, KM = code/Kernels/kernelMisc.cpp
 *
 * Built with -ffp-contract=off so every FP32 operation rounds exactly as
 * written (no silent FMA contraction), which makes the ref32 results
 * reproducible bit-for-bit across compilers and vector widths.
 */
#include "bpg_oracle.h"

#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define VB 8 /* voxel lanes processed together (vectorised across x) */
#define MAXP 25
#define MAXC 10

static inline size_t idx3(int x, int y, int z, int W, int H) {
    return (size_t)x + (size_t)y * W + (size_t)z * W * H;
}

/* ===================================================================== */
/* Generators                                                            */
/* ===================================================================== */

void orc_srand(unsigned int seed) { srand(seed); }
int orc_rand(void) { return rand(); }

/* libstdc++ std::random_shuffle(first, last) (bits/stl_algo.h, GCC 11):
 * for i = 1 .. n-1: j = rand() % (i + 1); if (i != j) swap(a[i], a[j]).
 * Called by LIB:16978 (two-sample), LIB:17042 (correlation). */
void orc_random_shuffle_u16(unsigned short* a, int n) {
    for (int i = 1; i < n; i++) {
        int j = rand() % (i + 1);
        if (i != j) {
            unsigned short t = a[i];
            a[i] = a[j];
            a[j] = t;
        }
    }
}

void orc_random_shuffle_int(int* a, int n) {
    for (int i = 1; i < n; i++) {
        int j = rand() % (i + 1);
        if (i != j) {
            int t = a[i];
            a[i] = a[j];
            a[j] = t;
        }
    }
}

/* GenerateSignMatrixSecondLevel, LIB:17069-17122.  Row 0 is all +1; each
 * later row multiplies the running flips by 2*(rand()%2)-1 until the vector
 * differs from every earlier row. */
int orc_gen_sign_matrix(int N, long K, float* sign) {
    int* hist = (int*)malloc(sizeof(int) * (size_t)N * (size_t)(K > 0 ? K : 1));
    int* flips = (int*)malloc(sizeof(int) * (size_t)N);
    if (!hist || !flips) { free(hist); free(flips); return -1; }
    for (int i = 0; i < N; i++) flips[i] = 1;
    memcpy(hist, flips, sizeof(int) * N);
    for (int i = 0; i < N; i++) sign[i] = (float)flips[i];
    for (long p = 1; p < K; p++) {
        for (;;) {
            for (int i = 0; i < N; i++) flips[i] *= (2 * (int)(rand() % 2) - 1);
            int unique = 1;
            for (long r = 0; r < p; r++) {
                if (memcmp(hist + (size_t)r * N, flips, sizeof(int) * N) == 0) { unique = 0; break; }
            }
            if (unique) { memcpy(hist + (size_t)p * N, flips, sizeof(int) * N); break; }
        }
        for (int i = 0; i < N; i++) sign[i + (size_t)p * N] = (float)flips[i];
    }
    free(hist);
    free(flips);
    return 0;
}

/* Builds the subject order from the group labels, LIB:16946-16962 and
 * LIB:16994-17009: position i takes the next unused subject of group g[i],
 * ascending within each group; group 1 is subjects 0..n1-1. */
static void two_sample_fill(const int* g, int N, int n1, unsigned short* out) {
    int a = 0, b = 0;
    for (int i = 0; i < N; i++) {
        if (g[i] == 1) out[i] = (unsigned short)(a++);
        else if (g[i] == -1) out[i] = (unsigned short)(n1 + b++);
    }
}

/* GeneratePermutationMatrixSecondLevelTwoSample, LIB:16925-17011. */
int orc_gen_perm_two_sample(int N, int n1, int n2, long K, unsigned short* perm) {
    (void)n2;
    int* hist = (int*)malloc(sizeof(int) * (size_t)N * (size_t)(K > 0 ? K : 1));
    int* g = (int*)malloc(sizeof(int) * (size_t)N);
    if (!hist || !g) { free(hist); free(g); return -1; }
    for (int i = 0; i < N; i++) g[i] = (i < n1) ? 1 : -1;
    memcpy(hist, g, sizeof(int) * N);
    long nh = 1;
    two_sample_fill(g, N, n1, perm);
    for (long p = 1; p < K; p++) {
        for (;;) {
            orc_random_shuffle_int(g, N);
            int unique = 1;
            for (long r = 0; r < nh; r++) {
                if (memcmp(hist + (size_t)r * N, g, sizeof(int) * N) == 0) { unique = 0; break; }
            }
            if (unique) { memcpy(hist + (size_t)nh * N, g, sizeof(int) * N); nh++; break; }
        }
        two_sample_fill(g, N, n1, perm + (size_t)p * N);
    }
    free(hist);
    free(g);
    return 0;
}

/* GeneratePermutationMatrixSecondLevelCorrelation, LIB:17014-17066. */
int orc_gen_perm_correlation(int N, long K, unsigned short* perm) {
    unsigned short* hist = (unsigned short*)malloc(sizeof(unsigned short) * (size_t)N * (size_t)(K > 0 ? K : 1));
    unsigned short* cur = (unsigned short*)malloc(sizeof(unsigned short) * (size_t)N);
    if (!hist || !cur) { free(hist); free(cur); return -1; }
    for (int i = 0; i < N; i++) cur[i] = (unsigned short)i;
    memcpy(hist, cur, sizeof(unsigned short) * N);
    memcpy(perm, cur, sizeof(unsigned short) * N);
    for (long p = 1; p < K; p++) {
        for (;;) {
            orc_random_shuffle_u16(cur, N);
            int unique = 1;
            for (long r = 0; r < p; r++) {
                if (memcmp(hist + (size_t)r * N, cur, sizeof(unsigned short) * N) == 0) { unique = 0; break; }
            }
            if (unique) { memcpy(hist + (size_t)p * N, cur, sizeof(unsigned short) * N); break; }
        }
        memcpy(perm + (size_t)p * N, cur, sizeof(unsigned short) * N);
    }
    free(hist);
    free(cur);
    return 0;
}

/* ===================================================================== */
/* CLI model logic                                                       */
/* ===================================================================== */

/* RGL:38-42: Stirling approximation with pi ~ 3.14 and e ~ 2.7183. */
double orc_stirling_factorial(unsigned long n) {
    return (n == 1 || n == 0) ? 1.0
                              : round(sqrt(2.0 * 3.14 * (double)n) * pow(((double)n / 2.7183), (double)n));
}

/* Design-type detection, RGL:908-1044.  t: per contrast, the distinct
 * values of X*c (compared as float); F: distinct values over all of X*C'. */
int orc_detect_designs(int N, int p, int nc, const float* X, const float* C,
                       int ftest, int* designs, int* n1, int* n2) {
    float* uniq = (float*)malloc(sizeof(float) * (size_t)N * (size_t)(nc > 0 ? nc : 1));
    float* vals = (float*)malloc(sizeof(float) * (size_t)N * (size_t)(nc > 0 ? nc : 1));
    if (!uniq || !vals) { free(uniq); free(vals); return -1; }
    int ngroups = ftest ? 1 : nc;
    for (int c = 0; c < nc; c++) {
        for (int s = 0; s < N; s++) {
            double acc = 0.0;
            for (int r = 0; r < p; r++) acc += (double)X[s + (size_t)r * N] * (double)C[r + (size_t)c * p];
            vals[s + (size_t)c * N] = (float)acc;
        }
    }
    for (int g = 0; g < ngroups; g++) {
        int nu = 0;
        for (int s = 0; s < N; s++) {
            int c0 = ftest ? 0 : g, c1 = ftest ? nc : g + 1;
            for (int c = c0; c < c1; c++) {
                float v = vals[s + (size_t)c * N];
                int un = 1;
                for (int i = 0; i < nu; i++) if (v == uniq[i]) { un = 0; break; }
                if (un) uniq[nu++] = v;
            }
        }
        designs[g] = (nu == 2) ? ORC_TWOSAMPLE : (nu > 2) ? ORC_CORRELATION : ORC_MEAN;
        n1[g] = 0;
        n2[g] = 0;
        if (nu == 2) {
            int col = ftest ? 0 : g;
            for (int s = 0; s < N; s++) if (vals[s + (size_t)col * N] == uniq[0]) n1[g]++;
            n2[g] = N - n1[g];
        }
    }
    free(uniq);
    free(vals);
    return 0;
}

/* Largest permutation count for a design, RGL:1048-1134 (as a double, the
 * reference compares in double): 2^N sign flips, C(N, n2) two-sample label
 * shuffles, Stirling N! for correlation, 1 for a mean design. */
long orc_max_permutations(int test, int design, int N, int n2) {
    double m;
    if (test == ORC_GROUP_MEAN) m = pow(2.0, (double)N);
    else if (design == ORC_TWOSAMPLE)
        m = round(exp(lgamma(N + 1) - lgamma(N - n2 + 1) - lgamma(n2 + 1)));
    else if (design == ORC_CORRELATION) m = orc_stirling_factorial((unsigned long)N);
    else m = 1.0;
    if (m > 9.0e18) return (long)9000000000000000000L;
    return (long)m;
}

/* ===================================================================== */
/* FP64 linear algebra (textbook Householder QR)                          */
/* ===================================================================== */

/* A: m x n column-major (m >= n).  Q: m x m column-major (full), R: m x n.
 * Either output may be NULL. */
int orc_qr64(int m, int n, const double* A, double* Q, double* R) {
    double* a = (double*)malloc(sizeof(double) * (size_t)m * n);
    double* q = (double*)calloc((size_t)m * m, sizeof(double));
    double* v = (double*)malloc(sizeof(double) * (size_t)m);
    if (!a || !q || !v) { free(a); free(q); free(v); return -1; }
    memcpy(a, A, sizeof(double) * (size_t)m * n);
    for (int i = 0; i < m; i++) q[i + (size_t)i * m] = 1.0;
    int steps = (n < m - 1) ? n : m - 1;
    for (int k = 0; k < steps; k++) {
        double norm = 0.0;
        for (int i = k; i < m; i++) norm += a[i + (size_t)k * m] * a[i + (size_t)k * m];
        norm = sqrt(norm);
        if (norm == 0.0) continue;
        double alpha = (a[k + (size_t)k * m] > 0) ? -norm : norm;
        for (int i = 0; i < m; i++) v[i] = 0.0;
        for (int i = k; i < m; i++) v[i] = a[i + (size_t)k * m];
        v[k] -= alpha;
        double vv = 0.0;
        for (int i = k; i < m; i++) vv += v[i] * v[i];
        if (vv == 0.0) continue;
        /* a <- (I - 2 v v'/vv) a */
        for (int j = k; j < n; j++) {
            double d = 0.0;
            for (int i = k; i < m; i++) d += v[i] * a[i + (size_t)j * m];
            d = 2.0 * d / vv;
            for (int i = k; i < m; i++) a[i + (size_t)j * m] -= d * v[i];
        }
        /* q <- q (I - 2 v v'/vv) */
        for (int i = 0; i < m; i++) {
            double d = 0.0;
            for (int l = k; l < m; l++) d += q[i + (size_t)l * m] * v[l];
            d = 2.0 * d / vv;
            for (int l = k; l < m; l++) q[i + (size_t)l * m] -= d * v[l];
        }
    }
    if (Q) memcpy(Q, q, sizeof(double) * (size_t)m * m);
    if (R) {
        for (int j = 0; j < n; j++)
            for (int i = 0; i < m; i++) R[i + (size_t)j * m] = (i <= j) ? a[i + (size_t)j * m] : 0.0;
    }
    free(a);
    free(q);
    free(v);
    return 0;
}

/* pinv(X) = R^-1 Q1' from the thin QR; Xpinv[s + r*N] = pinv(X)(r,s)
 * (the layout of the reference's xtxxt, RGL:1148). */
int orc_pinv64(int N, int p, const double* X, double* Xpinv) {
    double* Q = (double*)malloc(sizeof(double) * (size_t)N * N);
    double* R = (double*)malloc(sizeof(double) * (size_t)N * p);
    if (!Q || !R) { free(Q); free(R); return -1; }
    if (orc_qr64(N, p, X, Q, R)) { free(Q); free(R); return -1; }
    /* solve R b = Q1(s,:)' for each s */
    for (int s = 0; s < N; s++) {
        double b[MAXP];
        for (int r = 0; r < p; r++) b[r] = Q[s + (size_t)r * N];
        for (int r = p - 1; r >= 0; r--) {
            double acc = b[r];
            for (int j = r + 1; j < p; j++) acc -= R[r + (size_t)j * N] * b[j];
            b[r] = acc / R[r + (size_t)r * N];
        }
        for (int r = 0; r < p; r++) Xpinv[s + (size_t)r * N] = b[r];
    }
    free(Q);
    free(R);
    return 0;
}

/* Nuisance projector of LIB:16654-16727: R = I - W2 pinv(W2) with W2 the
 * nuisance part of X C_full^-1.  Its column space is X * null(C), so R is
 * built here from an orthonormal basis of null(C) (Householder QR of C')
 * and an orthonormal basis of X*null(C).  Crows[j + i*p] = C(i,j), i < r.
 * R is N x N column-major. */
int orc_nuisance_projector64(int N, int p, const double* X, int r,
                             const double* Crows, double* R) {
    int q = p - r;
    for (size_t i = 0; i < (size_t)N * N; i++) R[i] = 0.0;
    for (int i = 0; i < N; i++) R[i + (size_t)i * N] = 1.0;
    if (q <= 0) return 0;
    double* Qc = (double*)malloc(sizeof(double) * (size_t)p * p);
    double* W2 = (double*)malloc(sizeof(double) * (size_t)N * q);
    double* Qw = (double*)malloc(sizeof(double) * (size_t)N * N);
    if (!Qc || !W2 || !Qw) { free(Qc); free(W2); free(Qw); return -1; }
    /* C' is p x r column-major: element (j,i) = C(i,j) = Crows[j + i*p] */
    if (orc_qr64(p, r, Crows, Qc, NULL)) { free(Qc); free(W2); free(Qw); return -1; }
    for (int j = 0; j < q; j++)
        for (int s = 0; s < N; s++) {
            double acc = 0.0;
            for (int k = 0; k < p; k++) acc += X[s + (size_t)k * N] * Qc[k + (size_t)(r + j) * p];
            W2[s + (size_t)j * N] = acc;
        }
    if (orc_qr64(N, q, W2, Qw, NULL)) { free(Qc); free(W2); free(Qw); return -1; }
    for (int j = 0; j < N; j++)
        for (int i = 0; i < N; i++) {
            double acc = 0.0;
            for (int k = 0; k < q; k++) acc += Qw[i + (size_t)k * N] * Qw[j + (size_t)k * N];
            R[i + (size_t)j * N] -= acc;
        }
    free(Qc);
    free(W2);
    free(Qw);
    return 0;
}

/* Gauss-Jordan inverse with partial pivoting, n x n column-major, in place. */
static int inv64(int n, double* A) {
    double* Ainv = (double*)calloc((size_t)n * n, sizeof(double));
    if (!Ainv) return -1;
    for (int i = 0; i < n; i++) Ainv[i + (size_t)i * n] = 1.0;
    for (int k = 0; k < n; k++) {
        int piv = k;
        for (int i = k + 1; i < n; i++)
            if (fabs(A[i + (size_t)k * n]) > fabs(A[piv + (size_t)k * n])) piv = i;
        if (A[piv + (size_t)k * n] == 0.0) { free(Ainv); return -1; }
        if (piv != k)
            for (int j = 0; j < n; j++) {
                double t = A[k + (size_t)j * n]; A[k + (size_t)j * n] = A[piv + (size_t)j * n]; A[piv + (size_t)j * n] = t;
                t = Ainv[k + (size_t)j * n]; Ainv[k + (size_t)j * n] = Ainv[piv + (size_t)j * n]; Ainv[piv + (size_t)j * n] = t;
            }
        double d = A[k + (size_t)k * n];
        for (int j = 0; j < n; j++) { A[k + (size_t)j * n] /= d; Ainv[k + (size_t)j * n] /= d; }
        for (int i = 0; i < n; i++) {
            if (i == k) continue;
            double f = A[i + (size_t)k * n];
            if (f == 0.0) continue;
            for (int j = 0; j < n; j++) {
                A[i + (size_t)j * n] -= f * A[k + (size_t)j * n];
                Ainv[i + (size_t)j * n] -= f * Ainv[k + (size_t)j * n];
            }
        }
    }
    memcpy(A, Ainv, sizeof(double) * (size_t)n * n);
    free(Ainv);
    return 0;
}

/* FP64 contrast scalars: t: c (X'X)^-1 c'; F: (C (X'X)^-1 C')^-1 stored
 * [c + cc*nc] (RGL:1166-1192). */
static int contrast_scalars64(int N, int p, int nc, int ftest, const double* X,
                              const double* Cd, double* out) {
    double xtx[MAXP * MAXP];
    for (int i = 0; i < p; i++)
        for (int j = 0; j < p; j++) {
            double acc = 0.0;
            for (int s = 0; s < N; s++) acc += X[s + (size_t)i * N] * X[s + (size_t)j * N];
            xtx[i + j * p] = acc;
        }
    if (inv64(p, xtx)) return -1;
    if (!ftest) {
        for (int c = 0; c < nc; c++) {
            double acc = 0.0;
            for (int i = 0; i < p; i++)
                for (int j = 0; j < p; j++) acc += Cd[i + c * p] * xtx[i + j * p] * Cd[j + c * p];
            out[c] = acc;
        }
    } else {
        double m[MAXC * MAXC];
        for (int a = 0; a < nc; a++)
            for (int b = 0; b < nc; b++) {
                double acc = 0.0;
                for (int i = 0; i < p; i++)
                    for (int j = 0; j < p; j++) acc += Cd[i + a * p] * xtx[i + j * p] * Cd[j + b * p];
                m[a + b * nc] = acc;
            }
        if (inv64(nc, m)) return -1;
        for (int a = 0; a < nc; a++)
            for (int b = 0; b < nc; b++) out[a + b * nc] = m[a + b * nc];
    }
    return 0;
}

/* Model precompute of RGL:1137-1194 in FP64, cast to float. */
int orc_model_precompute(int N, int p, int nc, int ftest, const float* X,
                         const float* C, float* xtxxt, float* ctxtxc) {
    double* Xd = (double*)malloc(sizeof(double) * (size_t)N * p);
    double* Xp = (double*)malloc(sizeof(double) * (size_t)N * p);
    double Cd[MAXP * MAXC], sc[MAXC * MAXC];
    if (!Xd || !Xp) { free(Xd); free(Xp); return -1; }
    for (size_t i = 0; i < (size_t)N * p; i++) Xd[i] = (double)X[i];
    for (int i = 0; i < p * nc; i++) Cd[i] = (double)C[i];
    int rc = orc_pinv64(N, p, Xd, Xp);
    if (!rc) rc = contrast_scalars64(N, p, nc, ftest, Xd, Cd, sc);
    if (!rc) {
        for (size_t i = 0; i < (size_t)N * p; i++) xtxxt[i] = (float)Xp[i];
        int ns = ftest ? nc * nc : nc;
        for (int i = 0; i < ns; i++) ctxtxc[i] = (float)sc[i];
    }
    free(Xd);
    free(Xp);
    return rc;
}

/* ===================================================================== */
/* FP32 restatement of the OpenCL kernels                                 */
/* ===================================================================== */

/* TransformData, K1:1684-1719: y*_v = sum_vv R[vv + v*N] * y_vv, in place. */
void orc_transform32(int W, int H, int D, int N, const float* Rf, const float* mask, float* Y) {
    size_t V = (size_t)W * H * D;
#pragma omp parallel
    {
        float* tmp = (float*)malloc(sizeof(float) * (size_t)N);
#pragma omp for schedule(static)
        for (long i = 0; i < (long)V; i++) {
            if (mask[i] != 1.0f) continue;
            for (int v = 0; v < N; v++) {
                float acc = 0.0f;
                for (int vv = 0; vv < N; vv++) acc += Rf[vv + (size_t)v * N] * Y[i + (size_t)vv * V];
                tmp[v] = acc;
            }
            for (int v = 0; v < N; v++) Y[i + (size_t)v * V] = tmp[v];
        }
        free(tmp);
    }
}

static inline float rsqrt32(float x) { return 1.0f / sqrtf(x); }

/* CalculateStatisticalMaps{GLMTTest,GLMFTest,Mean}SecondLevelPermutation,
 * K2:1562-1669, K4:1943-2060, K2:1458-1559, restated per voxel in FP32 with
 * the kernels' loop order (beta loop over subjects, then the residual loop).
 * VB consecutive x-voxels are processed as independent lanes. */
void orc_perm_map32(int test, int W, int H, int D, int N, int p, int nc,
                    const float* Y, const float* mask, const float* X,
                    const float* xtxxt, const float* C, const float* ctxtxc,
                    int contrast, const unsigned short* perm, const float* sign,
                    float* out) {
    size_t V = (size_t)W * H * D;
    int P = (test == ORC_GROUP_MEAN) ? 1 : p;
#pragma omp parallel for collapse(2) schedule(dynamic, 8)
    for (int z = 0; z < D; z++)
        for (int y = 0; y < H; y++) {
            for (int x0 = 0; x0 < W; x0 += VB) {
                int nl = (W - x0 < VB) ? W - x0 : VB;
                size_t base = idx3(x0, y, z, W, H);
                int any = 0;
                for (int l = 0; l < nl; l++) any |= (mask[base + l] == 1.0f);
                if (!any) continue;
                float beta[MAXP][VB];
                float vareps[VB];
                for (int r = 0; r < P; r++)
                    for (int l = 0; l < VB; l++) beta[r][l] = 0.0f;
                for (int s = 0; s < N; s++) {
                    int ps = (test == ORC_GROUP_MEAN) ? s : perm[s];
                    float sg = (test == ORC_GROUP_MEAN) ? sign[s] : 1.0f;
                    const float* ys = Y + base + (size_t)s * V;
                    for (int r = 0; r < P; r++) {
                        float w = xtxxt[(size_t)N * r + ps];
                        for (int l = 0; l < VB; l++) {
                            float value = (l < nl) ? ys[l] : 0.0f;
                            if (test == ORC_GROUP_MEAN) value = value * sg;
                            beta[r][l] += value * w;
                        }
                    }
                }
                for (int l = 0; l < VB; l++) vareps[l] = 0.0f;
                for (int s = 0; s < N; s++) {
                    int ps = (test == ORC_GROUP_MEAN) ? s : perm[s];
                    float sg = (test == ORC_GROUP_MEAN) ? sign[s] : 1.0f;
                    const float* ys = Y + base + (size_t)s * V;
                    float eps[VB];
                    for (int l = 0; l < VB; l++) {
                        eps[l] = (l < nl) ? ys[l] : 0.0f;
                        if (test == ORC_GROUP_MEAN) eps[l] = eps[l] * sg;
                    }
                    for (int r = 0; r < P; r++) {
                        float xv = X[(size_t)N * r + ps];
                        for (int l = 0; l < VB; l++) eps[l] -= xv * beta[r][l];
                    }
                    for (int l = 0; l < VB; l++) vareps[l] += eps[l] * eps[l];
                }
                for (int l = 0; l < nl; l++) {
                    if (mask[base + l] != 1.0f) continue;
                    float res;
                    if (test == ORC_GROUP_MEAN) {
                        float ve = vareps[l] / ((float)N - 1.0f);
                        res = beta[0][l] * rsqrt32(ve * ctxtxc[0]);
                    } else if (test == ORC_TTEST) {
                        float ve = vareps[l] / ((float)N - p);
                        float cv = 0.0f;
                        for (int r = 0; r < p; r++) cv += C[(size_t)p * contrast + r] * beta[r][l];
                        res = cv * rsqrt32(ve * ctxtxc[contrast]);
                    } else {
                        float ve = vareps[l] / ((float)N - p);
                        float cb[MAXC], tb[MAXC];
                        for (int c = 0; c < nc; c++) {
                            cb[c] = 0.0f;
                            for (int r = 0; r < p; r++) cb[c] += C[(size_t)p * c + r] * beta[r][l];
                        }
                        for (int c = 0; c < nc; c++) {
                            tb[c] = 0.0f;
                            for (int cc = 0; cc < nc; cc++)
                                tb[c] += 1.0f / ve * ctxtxc[cc + c * nc] * cb[cc];
                        }
                        float scalar = 0.0f;
                        for (int c = 0; c < nc; c++) scalar += cb[c] * tb[c];
                        res = scalar / (float)nc;
                    }
                    out[base + l] = res;
                }
            }
        }
}

/* Observed maps on the original data and design: CalculateBetaWeightsGLM
 * (K1:51-128) followed by CalculateStatisticalMapsGLMTTest (K1:1456-1556) or
 * ...GLMFTest (K1:1560-1673).  The F variance is sum (e - mean e)^2 /
 * (N - p - 1), unlike the permuted F.  out: t [v + c*V], F [v]. */
void orc_obs_maps32(int ftest, int W, int H, int D, int N, int p, int nc,
                    const float* Y, const float* mask, const float* X,
                    const float* xtxxt, const float* C, const float* ctxtxc,
                    float* out) {
    size_t V = (size_t)W * H * D;
    int nmaps = ftest ? 1 : nc;
#pragma omp parallel for schedule(static)
    for (long i = 0; i < (long)V; i++) {
        if (mask[i] != 1.0f) {
            for (int c = 0; c < nmaps; c++) out[i + (size_t)c * V] = 0.0f;
            continue;
        }
        float beta[MAXP];
        for (int r = 0; r < p; r++) beta[r] = 0.0f;
        for (int s = 0; s < N; s++) {
            float temp = Y[i + (size_t)s * V] * 1.0f;
            for (int r = 0; r < p; r++) beta[r] += temp * xtxxt[(size_t)N * r + s];
        }
        float meaneps = 0.0f;
        for (int s = 0; s < N; s++) {
            float eps = Y[i + (size_t)s * V];
            for (int r = 0; r < p; r++) eps -= X[(size_t)N * r + s] * beta[r];
            meaneps += eps;
        }
        meaneps /= (float)N;
        float vareps = 0.0f;
        for (int s = 0; s < N; s++) {
            float eps = Y[i + (size_t)s * V];
            for (int r = 0; r < p; r++) eps -= X[(size_t)N * r + s] * beta[r];
            if (ftest) vareps += (eps - meaneps) * (eps - meaneps);
            else vareps += eps * eps;
        }
        if (!ftest) {
            vareps /= ((float)N - p);
            for (int c = 0; c < nc; c++) {
                float cv = 0.0f;
                for (int r = 0; r < p; r++) cv += C[(size_t)p * c + r] * beta[r];
                out[i + (size_t)c * V] = cv * rsqrt32(vareps * ctxtxc[c]);
            }
        } else {
            vareps /= ((float)N - (float)p - 1.0f);
            float cb[MAXC], tb[MAXC];
            for (int c = 0; c < nc; c++) {
                cb[c] = 0.0f;
                for (int r = 0; r < p; r++) cb[c] += C[(size_t)p * c + r] * beta[r];
            }
            for (int c = 0; c < nc; c++) {
                tb[c] = 0.0f;
                for (int cc = 0; cc < nc; cc++) tb[c] += 1.0f / vareps * ctxtxc[cc + c * nc] * cb[cc];
            }
            float scalar = 0.0f;
            for (int c = 0; c < nc; c++) scalar += cb[c] * tb[c];
            out[i] = scalar / (float)nc;
        }
    }
}

/* (int)(T * 10000.0f), KM:299. */
int orc_quantize(float t) { return (int)(t * 10000.0f); }

/* CalculateMaxAtomic, KM:282-300 + LIB:10965-10989: the int max starts from
 * the bit pattern of -1e6f (SetMemory writes a float into the int buffer,
 * LIB:10971) and is returned as (float)max / 10000.0f. */
float orc_max_quantized(int W, int H, int D, const float* map, const float* mask) {
    size_t V = (size_t)W * H * D;
    float init = -1000000.0f;
    int m0;
    memcpy(&m0, &init, sizeof(int));
    int m = m0;
#pragma omp parallel for reduction(max : m) schedule(static)
    for (long i = 0; i < (long)V; i++) {
        if (mask[i] != 1.0f) continue;
        int q = (int)(map[i] * 10000.0f);
        if (q > m) m = q;
    }
    return (float)((float)m / 10000.0f);
}

/* ===================================================================== */
/* FP64 statistics                                                       */
/* ===================================================================== */

typedef struct {
    int N, p, nc, test;
    double Xp[MAXP * 512];
    double sc[MAXC * MAXC];
} model64;

static int model64_init(model64* m, int test, int N, int p, int nc, const double* Xd, const double* Cd) {
    if (N > 512 || p > MAXP || nc > MAXC) return -1;
    m->N = N; m->p = p; m->nc = nc; m->test = test;
    if (orc_pinv64(N, p, Xd, m->Xp)) return -1;
    return contrast_scalars64(N, p, nc, test == ORC_FTEST, Xd, Cd, m->sc);
}

/* Same semantics as orc_perm_map32 (K2:1458/1562, K4:1943) in FP64, with
 * pinv(X) and the contrast scalars rebuilt in FP64 (model64_init). */
void orc_perm_map64(int test, int W, int H, int D, int N, int p, int nc,
                    const double* Y, const float* mask, const double* Xd,
                    const double* Cd, int contrast, const unsigned short* perm,
                    const float* sign, double* out) {
    size_t V = (size_t)W * H * D;
    model64* m = (model64*)malloc(sizeof(model64));
    int P = (test == ORC_GROUP_MEAN) ? 1 : p;
    const double* Xm = Xd;
    if (!m || model64_init(m, test == ORC_GROUP_MEAN ? ORC_TTEST : test, N, P,
                           test == ORC_GROUP_MEAN ? 1 : nc, Xd, Cd)) {
        free(m);
        return;
    }
#pragma omp parallel for schedule(static)
    for (long i = 0; i < (long)V; i++) {
        if (mask[i] != 1.0f) continue;
        double beta[MAXP];
        double yv[512];
        for (int s = 0; s < N; s++) {
            yv[s] = Y[i + (size_t)s * V];
            if (test == ORC_GROUP_MEAN) yv[s] *= (double)sign[s];
        }
        for (int r = 0; r < P; r++) beta[r] = 0.0;
        for (int s = 0; s < N; s++) {
            int ps = (test == ORC_GROUP_MEAN) ? s : perm[s];
            for (int r = 0; r < P; r++) beta[r] += yv[s] * m->Xp[ps + (size_t)r * N];
        }
        double sse = 0.0;
        for (int s = 0; s < N; s++) {
            int ps = (test == ORC_GROUP_MEAN) ? s : perm[s];
            double e = yv[s];
            for (int r = 0; r < P; r++) e -= Xm[ps + (size_t)r * N] * beta[r];
            sse += e * e;
        }
        double var = sse / (double)(N - P);
        if (test == ORC_GROUP_MEAN) {
            out[i] = beta[0] / sqrt(var * m->sc[0]);
        } else if (test == ORC_TTEST) {
            double cv = 0.0;
            for (int r = 0; r < p; r++) cv += Cd[r + (size_t)p * contrast] * beta[r];
            out[i] = cv / sqrt(var * m->sc[contrast]);
        } else {
            double cb[MAXC];
            for (int c = 0; c < nc; c++) {
                cb[c] = 0.0;
                for (int r = 0; r < p; r++) cb[c] += Cd[r + (size_t)p * c] * beta[r];
            }
            double q = 0.0;
            for (int a = 0; a < nc; a++)
                for (int b = 0; b < nc; b++) q += cb[a] * m->sc[b + a * nc] * cb[b];
            out[i] = q / var / (double)nc;
        }
    }
    free(m);
}

/* Observed maps in FP64 with the K1 formulas (F variance over N - p - 1 of
 * the mean-centred residuals). */
void orc_obs_maps64(int ftest, int W, int H, int D, int N, int p, int nc,
                    const double* Y, const float* mask, const double* Xd,
                    const double* Cd, double* out) {
    size_t V = (size_t)W * H * D;
    int nmaps = ftest ? 1 : nc;
    model64* m = (model64*)malloc(sizeof(model64));
    if (!m || model64_init(m, ftest ? ORC_FTEST : ORC_TTEST, N, p, nc, Xd, Cd)) { free(m); return; }
#pragma omp parallel for schedule(static)
    for (long i = 0; i < (long)V; i++) {
        if (mask[i] != 1.0f) {
            for (int c = 0; c < nmaps; c++) out[i + (size_t)c * V] = 0.0;
            continue;
        }
        double beta[MAXP], e[512];
        for (int r = 0; r < p; r++) beta[r] = 0.0;
        for (int s = 0; s < N; s++)
            for (int r = 0; r < p; r++) beta[r] += Y[i + (size_t)s * V] * m->Xp[s + (size_t)r * N];
        double mean = 0.0;
        for (int s = 0; s < N; s++) {
            double v = Y[i + (size_t)s * V];
            for (int r = 0; r < p; r++) v -= Xd[s + (size_t)r * N] * beta[r];
            e[s] = v;
            mean += v;
        }
        mean /= (double)N;
        double sse = 0.0;
        for (int s = 0; s < N; s++) sse += ftest ? (e[s] - mean) * (e[s] - mean) : e[s] * e[s];
        if (!ftest) {
            double var = sse / (double)(N - p);
            for (int c = 0; c < nc; c++) {
                double cv = 0.0;
                for (int r = 0; r < p; r++) cv += Cd[r + (size_t)p * c] * beta[r];
                out[i + (size_t)c * V] = cv / sqrt(var * m->sc[c]);
            }
        } else {
            double var = sse / (double)(N - p - 1);
            double cb[MAXC];
            for (int c = 0; c < nc; c++) {
                cb[c] = 0.0;
                for (int r = 0; r < p; r++) cb[c] += Cd[r + (size_t)p * c] * beta[r];
            }
            double q = 0.0;
            for (int a = 0; a < nc; a++)
                for (int b = 0; b < nc; b++) q += cb[a] * m->sc[b + a * nc] * cb[b];
            out[i] = q / var / (double)nc;
        }
    }
    free(m);
}

/* ===================================================================== */
/* Clusters                                                              */
/* ===================================================================== */

/* Connected components of {mask == 1 && map > thr} with 26- (KC:93-369,
 * HDR:76-104) or 6-connectivity.  Labels follow the fixpoint of the
 * reference's label propagation (KC:42-401): every voxel of a component
 * carries the smallest linear index in it; other voxels carry 3*W*H*D.
 * The scan is the breadth-first search of LIB:18163-18297: the first voxel
 * met in raster order opens a component, and it is the component's minimum
 * index.  sizes[label] = voxel count (KC:403-430).  Returns the largest size
 * (KC:460-480), 0 if there is none. */
unsigned orc_ccl(int W, int H, int D, const float* map, const float* mask,
                 float thr, int connectivity, unsigned* labels, unsigned* sizes) {
    size_t V = (size_t)W * H * D;
    unsigned none = (unsigned)(V * 3);
    unsigned* queue = (unsigned*)malloc(sizeof(unsigned) * V);
    for (size_t i = 0; i < V; i++) {
        labels[i] = none;
        if (sizes) sizes[i] = 0;
    }
    int off[26][3];
    int no = 0;
    for (int dz = -1; dz <= 1; dz++)
        for (int dy = -1; dy <= 1; dy++)
            for (int dx = -1; dx <= 1; dx++) {
                int nz = (dx != 0) + (dy != 0) + (dz != 0);
                if (nz == 0) continue;
                if (connectivity == 6 && nz != 1) continue;
                off[no][0] = dx; off[no][1] = dy; off[no][2] = dz; no++;
            }
    unsigned best = 0;
    for (int z = 0; z < D; z++)
        for (int y = 0; y < H; y++)
            for (int x = 0; x < W; x++) {
                size_t i = idx3(x, y, z, W, H);
                if (labels[i] != none) continue;
                if (!(mask[i] == 1.0f && map[i] > thr)) continue;
                unsigned lab = (unsigned)i;
                size_t head = 0, tail = 0;
                queue[tail++] = (unsigned)i;
                labels[i] = lab;
                while (head < tail) {
                    unsigned c = queue[head++];
                    int cx = (int)(c % W), cy = (int)((c / W) % H), cz = (int)(c / ((size_t)W * H));
                    for (int k = 0; k < no; k++) {
                        int xx = cx + off[k][0], yy = cy + off[k][1], zz = cz + off[k][2];
                        if (xx < 0 || yy < 0 || zz < 0 || xx >= W || yy >= H || zz >= D) continue;
                        size_t j = idx3(xx, yy, zz, W, H);
                        if (labels[j] != none) continue;
                        if (!(mask[j] == 1.0f && map[j] > thr)) continue;
                        labels[j] = lab;
                        queue[tail++] = (unsigned)j;
                    }
                }
                if (sizes) sizes[lab] = (unsigned)tail;
                if ((unsigned)tail > best) best = (unsigned)tail;
            }
    free(queue);
    return best;
}

/* CalculateClusterMasses, KC:432-458: mass[label] += (uint)(T * 10000.0f)
 * over the component; returns the largest (KC:460-480). */
unsigned orc_cluster_mass(int W, int H, int D, const float* map, const float* mask,
                          float thr, int connectivity, const unsigned* labels,
                          unsigned* masses) {
    (void)connectivity;
    size_t V = (size_t)W * H * D;
    for (size_t i = 0; i < V; i++) masses[i] = 0;
    for (size_t i = 0; i < V; i++)
        if (mask[i] == 1.0f && map[i] > thr) masses[labels[i]] += (unsigned)(map[i] * 10000.0f);
    unsigned best = 0;
    for (size_t i = 0; i < V; i++) if (masses[i] > best) best = masses[i];
    return best;
}

/* CalculatePermutationPValues{VoxelLevel,ClusterExtent,ClusterMass}
 * Inference, KC:516-652: P = #{k : T > max_k} / K, counted as a float sum
 * of 1.0f as in the kernels. */
void orc_pvalues(int inference, int W, int H, int D, const float* map,
                 const float* mask, const unsigned* labels, const unsigned* sizes,
                 float thr, long K, const float* maxdist, float* P) {
    size_t V = (size_t)W * H * D;
#pragma omp parallel for schedule(static)
    for (long i = 0; i < (long)V; i++) {
        if (mask[i] != 1.0f) { P[i] = 0.0f; continue; }
        float tv;
        /* TFCE compares the statistic itself with the TFCE null (LIB:16824) */
        if (inference == ORC_VOXEL || inference == ORC_TFCE) tv = map[i];
        else {
            if (!(map[i] > thr)) { P[i] = 0.0f; continue; }
            tv = (float)sizes[labels[i]];
            if (inference == ORC_CLUSTER_MASS) tv = tv / 10000.0f;
        }
        float sum = 0.0f;
        for (long k = 0; k < K; k++)
            if (tv > maxdist[k]) sum += 1.0f;
        P[i] = sum / (float)K;
    }
}

/* ===================================================================== */
/* The whole path                                                        */
/* ===================================================================== */

static int build_R(const orc_study* st, int c, double* R) {
    int N = st->N, p = st->p;
    double* Xd = (double*)malloc(sizeof(double) * (size_t)N * p);
    double Cr[MAXP * MAXC];
    int r;
    if (!Xd) return -1;
    for (size_t i = 0; i < (size_t)N * p; i++) Xd[i] = (double)st->X[i];
    if (st->test == ORC_TTEST) {
        r = 1;
        for (int j = 0; j < p; j++) Cr[j] = (double)st->C[j + (size_t)c * p];
    } else {
        r = st->nc;
        for (int i = 0; i < p * st->nc; i++) Cr[i] = (double)st->C[i];
    }
    int rc = orc_nuisance_projector64(N, p, Xd, r, Cr, R);
    free(Xd);
    return rc;
}

float orc_tfce_max(int W, int H, int D, const float* map, const float* mask, int connectivity,
                   float* tfce) {
    size_t V = (size_t)W * H * D;
    unsigned* lab = (unsigned*)malloc(sizeof(unsigned) * V);
    unsigned* sz = (unsigned*)malloc(sizeof(unsigned) * V);
    float* acc = tfce ? tfce : (float*)malloc(sizeof(float) * V);
    for (size_t i = 0; i < V; i++) acc[i] = 0.0f;
    const float top = orc_max_quantized(W, H, D, map, mask);
    const float delta = 0.2846f;
    for (float h = 0.0f; h <= top; h += delta) {
        orc_ccl(W, H, D, map, mask, h, connectivity, lab, sz);
        for (size_t i = 0; i < V; i++) {
            if (mask[i] != 1.0f || lab[i] >= V) continue;
            const float value = sqrtf((float)sz[lab[i]]) * h * h;
            acc[i] += value;
        }
    }
    const float best = orc_max_quantized(W, H, D, acc, mask);
    free(lab);
    free(sz);
    if (!tfce) free(acc);
    return best;
}

/* Statistic of one permutation map -> max statistic (A.5): quantised max
 * (voxel), largest cluster size, or largest cluster mass / 1e4. */
static float null_statistic(const orc_study* st, const float* map, unsigned* labels,
                            unsigned* aux) {
    if (st->inference == ORC_VOXEL)
        return orc_max_quantized(st->W, st->H, st->D, map, st->mask);
    if (st->inference == ORC_TFCE)
        return orc_tfce_max(st->W, st->H, st->D, map, st->mask, st->connectivity, NULL);
    unsigned best = orc_ccl(st->W, st->H, st->D, map, st->mask, st->cdt,
                            st->connectivity, labels, aux);
    if (st->inference == ORC_CLUSTER_EXTENT) return (float)best;
    unsigned bm = orc_cluster_mass(st->W, st->H, st->D, map, st->mask, st->cdt,
                                   st->connectivity, labels, aux);
    return (float)bm / 10000.0f;
}

/* Perform{Mean,GLMTTest,GLMFTest}SecondLevelPermutationWrapper
 * (LIB:13144-13376) -> ApplyPermutationTestSecondLevel (LIB:16589-16790),
 * observed maps (LIB:15858-15950), CalculatePermutationPValues
 * (LIB:16794-16873). */
int orc_randomise(const orc_study* st, int precision, float* maps,
                  float* pvalues, float** maxdist) {
    int W = st->W, H = st->H, D = st->D, N = st->N, p = st->p, nc = st->nc;
    size_t V = (size_t)W * H * D;
    int nmaps = (st->test == ORC_TTEST) ? nc : 1;
#ifdef _OPENMP
    if (st->nthreads > 0) omp_set_num_threads(st->nthreads);
#endif
    float* work = (float*)malloc(sizeof(float) * V * N);
    float* map = (float*)malloc(sizeof(float) * V);
    unsigned* labels = (unsigned*)malloc(sizeof(unsigned) * V);
    unsigned* aux = (unsigned*)malloc(sizeof(unsigned) * V);
    double* R = (double*)malloc(sizeof(double) * (size_t)N * N);
    double* work64 = NULL;
    double* map64 = NULL;
    double* Xd = (double*)malloc(sizeof(double) * (size_t)N * p);
    double Cd[MAXP * MAXC];
    if (!work || !map || !labels || !aux || !R || !Xd) goto fail;
    for (size_t i = 0; i < (size_t)N * p; i++) Xd[i] = (double)st->X[i];
    for (int i = 0; i < p * nc; i++) Cd[i] = (double)st->C[i];
    if (precision == ORC_PREC_F64) {
        work64 = (double*)malloc(sizeof(double) * V * N);
        map64 = (double*)malloc(sizeof(double) * V);
        if (!work64 || !map64) goto fail;
    }

    for (int c = 0; c < nmaps; c++) {
        int transform = (st->test == ORC_TTEST && p > 1) || (st->test == ORC_FTEST && p > nc);
        if (transform && build_R(st, c, R)) goto fail;
        if (precision == ORC_PREC_REF32) {
            memcpy(work, st->Y, sizeof(float) * V * N);
            if (transform) {
                float* Rf = (float*)malloc(sizeof(float) * (size_t)N * N);
                for (size_t i = 0; i < (size_t)N * N; i++) Rf[i] = (float)R[i];
                orc_transform32(W, H, D, N, Rf, st->mask, work);
                free(Rf);
            }
        } else {
            for (size_t i = 0; i < V * N; i++) work64[i] = (double)st->Y[i];
            if (transform) {
#pragma omp parallel for schedule(static)
                for (long i = 0; i < (long)V; i++) {
                    if (st->mask[i] != 1.0f) continue;
                    double tmp[512];
                    for (int v = 0; v < N; v++) {
                        double acc = 0.0;
                        for (int vv = 0; vv < N; vv++) acc += R[vv + (size_t)v * N] * work64[i + (size_t)vv * V];
                        tmp[v] = acc;
                    }
                    for (int v = 0; v < N; v++) work64[i + (size_t)v * V] = tmp[v];
                }
            }
        }
        for (long k = 0; k < st->K[c]; k++) {
            const unsigned short* pk = (st->test == ORC_GROUP_MEAN) ? NULL : st->perm[c] + (size_t)k * N;
            const float* sk = (st->test == ORC_GROUP_MEAN) ? st->sign + (size_t)k * N : NULL;
            memset(map, 0, sizeof(float) * V);
            if (precision == ORC_PREC_REF32) {
                orc_perm_map32(st->test, W, H, D, N, p, nc, work, st->mask, st->X, st->xtxxt,
                               st->C, st->ctxtxc, c, pk, sk, map);
                maxdist[c][k] = null_statistic(st, map, labels, aux);
            } else {
                memset(map64, 0, sizeof(double) * V);
                orc_perm_map64(st->test, W, H, D, N, p, nc, work64, st->mask, Xd, Cd, c, pk, sk, map64);
                if (st->inference == ORC_VOXEL) {
                    float init = -1000000.0f;
                    int m;
                    memcpy(&m, &init, sizeof(int));
                    for (size_t i = 0; i < V; i++)
                        if (st->mask[i] == 1.0f) {
                            double q = trunc(map64[i] * 10000.0);
                            if ((int)q > m) m = (int)q;
                        }
                    maxdist[c][k] = (float)((float)m / 10000.0f);
                } else if (st->inference == ORC_TFCE) {
                    for (size_t i = 0; i < V; i++) map[i] = (float)map64[i];
                    maxdist[c][k] = null_statistic(st, map, labels, aux);
                } else {
                    /* threshold in FP64, label on the FP64 map */
                    for (size_t i = 0; i < V; i++)
                        map[i] = (map64[i] > (double)st->cdt) ? 1.0f : 0.0f;
                    orc_study s2 = *st;
                    s2.cdt = 0.5f;
                    float mm = null_statistic(&s2, map, labels, aux);
                    if (st->inference == ORC_CLUSTER_MASS) {
                        /* mass needs the statistic values: recompute */
                        for (size_t i = 0; i < V; i++) map[i] = (float)map64[i];
                        mm = null_statistic(st, map, labels, aux);
                    }
                    maxdist[c][k] = mm;
                }
            }
        }
    }
    /* observed maps on the original data */
    {
        int pe = (st->test == ORC_GROUP_MEAN) ? 1 : p;
        int ne = (st->test == ORC_GROUP_MEAN) ? 1 : nc;
        if (precision == ORC_PREC_REF32) {
            orc_obs_maps32(st->test == ORC_FTEST, W, H, D, N, pe, ne, st->Y, st->mask, st->X,
                           st->xtxxt, st->C, st->ctxtxc, maps);
        } else {
            double* obs = (double*)malloc(sizeof(double) * V * nmaps);
            if (!obs) goto fail;
            for (size_t i = 0; i < V * N; i++) work64[i] = (double)st->Y[i];
            orc_obs_maps64(st->test == ORC_FTEST, W, H, D, N, pe, ne, work64, st->mask, Xd, Cd, obs);
            for (size_t i = 0; i < V * nmaps; i++) maps[i] = (float)obs[i];
            free(obs);
        }
    }
    for (int c = 0; c < nmaps; c++) {
        float* mc = maps + (size_t)c * V;
        if (st->inference == ORC_CLUSTER_EXTENT || st->inference == ORC_CLUSTER_MASS) {
            orc_ccl(W, H, D, mc, st->mask, st->cdt, st->connectivity, labels, aux);
            if (st->inference == ORC_CLUSTER_MASS)
                orc_cluster_mass(W, H, D, mc, st->mask, st->cdt, st->connectivity, labels, aux);
        }
        orc_pvalues(st->inference, W, H, D, mc, st->mask, labels, aux, st->cdt, st->K[c],
                    maxdist[c], pvalues + (size_t)c * V);
    }
    free(work); free(map); free(labels); free(aux); free(R); free(work64); free(map64); free(Xd);
    return 0;
fail:
    free(work); free(map); free(labels); free(aux); free(R); free(work64); free(map64); free(Xd);
    return -1;
}

/* The reference's K loop for one contrast in voxel mode (LIB:16733-16763):
 * one full pass over the (already transformed) data per permutation, then
 * the masked quantised max.  st->Y must hold the transformed data. */
int orc_permutation_phase32(const orc_study* st, int contrast, long K, float* maxdist) {
    size_t V = (size_t)st->W * st->H * st->D;
#ifdef _OPENMP
    if (st->nthreads > 0) omp_set_num_threads(st->nthreads);
#endif
    float* map = (float*)calloc(V, sizeof(float));
    if (!map) return -1;
    for (long k = 0; k < K; k++) {
        const unsigned short* pk = (st->test == ORC_GROUP_MEAN) ? NULL : st->perm[contrast] + (size_t)k * st->N;
        const float* sk = (st->test == ORC_GROUP_MEAN) ? st->sign + (size_t)k * st->N : NULL;
        orc_perm_map32(st->test, st->W, st->H, st->D, st->N, st->p, st->nc, st->Y, st->mask, st->X,
                       st->xtxxt, st->C, st->ctxtxc, contrast, pk, sk, map);
        maxdist[k] = orc_max_quantized(st->W, st->H, st->D, map, st->mask);
    }
    free(map);
    return 0;
}

/* Transformed data of map c (LIB:16651-16691): ref32 restates the FP32
 * TransformData kernel with R cast to float; f64 applies R in FP64. */
static int transformed(const orc_study* st, int c, int precision, float* w32, double* w64) {
    int N = st->N, p = st->p, nc = st->nc;
    size_t V = (size_t)st->W * st->H * st->D;
    int transform = (st->test == ORC_TTEST && p > 1) || (st->test == ORC_FTEST && p > nc);
    double* R = NULL;
    if (transform) {
        R = (double*)malloc(sizeof(double) * (size_t)N * N);
        if (!R || build_R(st, c, R)) { free(R); return -1; }
    }
    if (precision == ORC_PREC_REF32) {
        memcpy(w32, st->Y, sizeof(float) * V * N);
        if (transform) {
            float* Rf = (float*)malloc(sizeof(float) * (size_t)N * N);
            for (size_t i = 0; i < (size_t)N * N; i++) Rf[i] = (float)R[i];
            orc_transform32(st->W, st->H, st->D, N, Rf, st->mask, w32);
            free(Rf);
        }
    } else {
        for (size_t i = 0; i < V * N; i++) w64[i] = (double)st->Y[i];
        if (transform) {
#pragma omp parallel for schedule(static)
            for (long i = 0; i < (long)V; i++) {
                if (st->mask[i] != 1.0f) continue;
                double tmp[512];
                for (int v = 0; v < N; v++) {
                    double acc = 0.0;
                    for (int vv = 0; vv < N; vv++) acc += R[vv + (size_t)v * N] * w64[i + (size_t)vv * V];
                    tmp[v] = acc;
                }
                for (int v = 0; v < N; v++) w64[i + (size_t)v * V] = tmp[v];
            }
        }
    }
    free(R);
    return 0;
}

/* Statistic maps of the first K permutations of map c: out[k*W*H*D + i]
 * (0 outside the mask), as double.  For parity checks. */
int orc_perm_stats(const orc_study* st, int c, int precision, long K, double* out) {
    int W = st->W, H = st->H, D = st->D, N = st->N, p = st->p, nc = st->nc;
    size_t V = (size_t)W * H * D;
#ifdef _OPENMP
    if (st->nthreads > 0) omp_set_num_threads(st->nthreads);
#endif
    float* w32 = NULL;
    double* w64 = NULL;
    double Cd[MAXP * MAXC];
    double* Xd = (double*)malloc(sizeof(double) * (size_t)N * p);
    float* map = (float*)malloc(sizeof(float) * V);
    if (precision == ORC_PREC_REF32) w32 = (float*)malloc(sizeof(float) * V * N);
    else w64 = (double*)malloc(sizeof(double) * V * N);
    if (!Xd || !map || (!w32 && !w64)) { free(Xd); free(map); free(w32); free(w64); return -1; }
    for (size_t i = 0; i < (size_t)N * p; i++) Xd[i] = (double)st->X[i];
    for (int i = 0; i < p * nc; i++) Cd[i] = (double)st->C[i];
    if (transformed(st, c, precision, w32, w64)) { free(Xd); free(map); free(w32); free(w64); return -1; }
    for (long k = 0; k < K; k++) {
        const unsigned short* pk = (st->test == ORC_GROUP_MEAN) ? NULL : st->perm[c] + (size_t)k * N;
        const float* sk = (st->test == ORC_GROUP_MEAN) ? st->sign + (size_t)k * N : NULL;
        double* o = out + (size_t)k * V;
        if (precision == ORC_PREC_REF32) {
            memset(map, 0, sizeof(float) * V);
            orc_perm_map32(st->test, W, H, D, N, p, nc, w32, st->mask, st->X, st->xtxxt, st->C,
                           st->ctxtxc, c, pk, sk, map);
            for (size_t i = 0; i < V; i++) o[i] = (double)map[i];
        } else {
            memset(o, 0, sizeof(double) * V);
            orc_perm_map64(st->test, W, H, D, N, p, nc, w64, st->mask, Xd, Cd, c, pk, sk, o);
        }
    }
    free(Xd); free(map); free(w32); free(w64);
    return 0;
}


/* FP64 maximum of the statistic over the in-mask voxels for the first K
 * permutations of map c, not quantised: maxv[k] = max_v t_k(v), with
 * orc_perm_map64's semantics (the reference's beta loop then residual loop,
 * K2:1562-1669, K4:1943-2060, K2:1458-1559, in FP64 on the FP64-transformed
 * data).  For the full-size accuracy check: the in-mask voxels are
 * compacted, taken in blocks of OB that stay in cache while every
 * permutation runs over them, OL voxels at a time as independent lanes;
 * each thread keeps its own maxima, merged at the end.  NaN statistics
 * (all-zero residualised data) are left out of the max. */
#define OB 256
#define OL 8
/* beta and the residual sum of squares of OL voxels (lanes) for one
 * permutation, the reference's two loops (K2:1624-1663) in FP64; PP is the
 * compile-time regressor count (0: use P) so the accumulators stay in
 * registers.  Contracted multiply-adds are allowed: the result is FP64. */
#define ORC_LANES(PP)                                                              \
    static void __attribute__((optimize("fp-contract=fast")))                     \
    lanes##PP(int N, int P_, size_t Vp, const double* yc, const double* Xp,         \
              const double* Xd, const unsigned short* pk, const float* sk,          \
              double beta[MAXP][OL], double sse[OL]) {                             \
        const int P = PP ? PP : P_;                                                \
        double b[PP ? PP : MAXP][OL];                                              \
        for (int r = 0; r < P; r++)                                                \
            for (int l = 0; l < OL; l++) b[r][l] = 0.0;                            \
        for (int s = 0; s < N; s++) {                                              \
            const int ps = pk ? pk[s] : s;                                         \
            const double sg = sk ? (double)sk[s] : 1.0;                            \
            const double* ys = yc + (size_t)s * Vp;                                \
            for (int r = 0; r < P; r++) {                                          \
                const double w = Xp[ps + (size_t)r * N] * sg;                      \
                for (int l = 0; l < OL; l++) b[r][l] += ys[l] * w;                 \
            }                                                                      \
        }                                                                          \
        double q[OL];                                                              \
        for (int l = 0; l < OL; l++) q[l] = 0.0;                                   \
        for (int s = 0; s < N; s++) {                                              \
            const int ps = pk ? pk[s] : s;                                         \
            const double sg = sk ? (double)sk[s] : 1.0;                            \
            const double* ys = yc + (size_t)s * Vp;                                \
            double e[OL];                                                          \
            for (int l = 0; l < OL; l++) e[l] = ys[l] * sg;                        \
            for (int r = 0; r < P; r++) {                                          \
                const double xv = Xd[ps + (size_t)r * N];                          \
                for (int l = 0; l < OL; l++) e[l] -= xv * b[r][l];                 \
            }                                                                      \
            for (int l = 0; l < OL; l++) q[l] += e[l] * e[l];                      \
        }                                                                          \
        for (int r = 0; r < P; r++)                                                \
            for (int l = 0; l < OL; l++) beta[r][l] = b[r][l];                     \
        for (int l = 0; l < OL; l++) sse[l] = q[l];                                \
    }
ORC_LANES(0)
ORC_LANES(1)
ORC_LANES(2)
ORC_LANES(3)
ORC_LANES(4)
ORC_LANES(5)
ORC_LANES(6)
ORC_LANES(7)
ORC_LANES(8)
typedef void (*orc_lanes_fn)(int, int, size_t, const double*, const double*, const double*,
                             const unsigned short*, const float*, double[MAXP][OL], double[OL]);

int orc_perm_max64(const orc_study* st, int c, long K, double* maxv) {
    int N = st->N, p = st->p, nc = st->nc, test = st->test;
    size_t V = (size_t)st->W * st->H * st->D;
    int P = (test == ORC_GROUP_MEAN) ? 1 : p;
#ifdef _OPENMP
    if (st->nthreads > 0) omp_set_num_threads(st->nthreads);
#endif
    double* w64 = (double*)malloc(sizeof(double) * V * N);
    double* Xd = (double*)malloc(sizeof(double) * (size_t)N * p);
    model64* m = (model64*)malloc(sizeof(model64));
    double Cd[MAXP * MAXC];
    if (!w64 || !Xd || !m || N > 512) { free(w64); free(Xd); free(m); return -1; }
    for (size_t i = 0; i < (size_t)N * p; i++) Xd[i] = (double)st->X[i];
    for (int i = 0; i < p * nc; i++) Cd[i] = (double)st->C[i];
    if (transformed(st, c, ORC_PREC_F64, NULL, w64) ||
        model64_init(m, test == ORC_GROUP_MEAN ? ORC_TTEST : test, N, P, test == ORC_GROUP_MEAN ? 1 : nc, Xd, Cd)) {
        free(w64); free(Xd); free(m);
        return -1;
    }
    size_t Vc = 0;
    for (size_t i = 0; i < V; i++) Vc += st->mask[i] == 1.0f;
    size_t Vp = (Vc + OB - 1) / OB * OB;
    double* yc = (double*)calloc((size_t)N * (Vp ? Vp : 1), sizeof(double));
    if (!yc) { free(w64); free(Xd); free(m); return -1; }
    {
        size_t j = 0;
        for (size_t i = 0; i < V; i++)
            if (st->mask[i] == 1.0f) {
                for (int s = 0; s < N; s++) yc[(size_t)s * Vp + j] = w64[i + (size_t)s * V];
                j++;
            }
    }
    free(w64);
    for (long k = 0; k < K; k++) maxv[k] = -INFINITY;
    long nblk = (long)(Vp / OB);
    const unsigned short* perm0 = (test == ORC_GROUP_MEAN) ? NULL : st->perm[c];
    double cvec[MAXP];
    for (int r = 0; r < P; r++) cvec[r] = (test == ORC_GROUP_MEAN) ? 1.0 : Cd[r + (size_t)p * c];
    static const orc_lanes_fn lanes_tab[9] = {lanes0, lanes1, lanes2, lanes3, lanes4,
                                              lanes5, lanes6, lanes7, lanes8};
    const orc_lanes_fn lanes = P <= 8 ? lanes_tab[P] : lanes0;
#pragma omp parallel
    {
        double* lm = (double*)malloc(sizeof(double) * (size_t)K);
        for (long k = 0; k < K; k++) lm[k] = -INFINITY;
#pragma omp for schedule(dynamic, 1)
        for (long b = 0; b < nblk; b++) {
            size_t v0 = (size_t)b * OB;
            for (long k = 0; k < K; k++) {
                const unsigned short* pk = perm0 ? perm0 + (size_t)k * N : NULL;
                const float* sk = (test == ORC_GROUP_MEAN) ? st->sign + (size_t)k * N : NULL;
                double best = -INFINITY;
                for (size_t l0 = v0; l0 < v0 + OB && l0 < Vc; l0 += OL) {
                    double beta[MAXP][OL], sse[OL];
                    lanes(N, P, Vp, yc + l0, m->Xp, Xd, pk, sk, beta, sse);
                    for (int l = 0; l < OL; l++) {
                        if (l0 + l >= Vc) break;
                        double var = sse[l] / (double)(N - P), t;
                        if (test == ORC_FTEST) {
                            double cb[MAXC], q = 0.0;
                            for (int a = 0; a < nc; a++) {
                                cb[a] = 0.0;
                                for (int r = 0; r < p; r++) cb[a] += Cd[r + (size_t)p * a] * beta[r][l];
                            }
                            for (int a = 0; a < nc; a++)
                                for (int bb = 0; bb < nc; bb++) q += cb[a] * m->sc[bb + a * nc] * cb[bb];
                            t = q / var / (double)nc;
                        } else {
                            double cv = 0.0;
                            for (int r = 0; r < P; r++) cv += cvec[r] * beta[r][l];
                            t = cv / sqrt(var * m->sc[test == ORC_GROUP_MEAN ? 0 : c]);
                        }
                        if (t > best) best = t;
                    }
                }
                if (best > lm[k]) lm[k] = best;
            }
        }
#pragma omp critical
        for (long k = 0; k < K; k++)
            if (lm[k] > maxv[k]) maxv[k] = lm[k];
        free(lm);
    }
    free(yc); free(Xd); free(m);
    return 0;
}
