/*
 * fl_oracle.c -- CPU oracle for the first-level permutation test.
 *
 * TEST INFRASTRUCTURE ONLY (see fl_oracle.h).  FP64 restatement of
 * BROCCOLI_LIB::ApplyPermutationTestFirstLevel (LIB:16454-16558) and the
 * kernels it launches; each function names the reference lines it follows.
 */
#include "fl_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "bpg_oracle.h"

#define IDX3(x, y, z) ((size_t)(x) + (size_t)W * ((size_t)(y) + (size_t)H * (size_t)(z)))

/* LIB:11055-11090 */
void orc_fl_filters(int size, float fwhm, float vx, float vy, float vz, float* fx, float* fy, float* fz) {
    const int half = (size - 1) / 2;
    const double sx = (double)fwhm / 2.354 / (double)vx;
    const double sy = (double)fwhm / 2.354 / (double)vy;
    const double sz = (double)fwhm / 2.354 / (double)vz;
    float sumx = 0.0f, sumy = 0.0f, sumz = 0.0f;
    for (int i = 0; i < size; i++) {
        const double u = (double)(i - half);
        fx[i] = (float)exp(-(u * u) / (2.0 * sx * sx));
        fy[i] = (float)exp(-(u * u) / (2.0 * sy * sy));
        fz[i] = (float)exp(-(u * u) / (2.0 * sz * sz));
        sumx += fx[i];
        sumy += fy[i];
        sumz += fz[i];
    }
    for (int i = 0; i < size; i++) {
        fx[i] /= sumx;
        fy[i] /= sumy;
        fz[i] /= sumz;
    }
}

/* kernelConvolution.cpp:52-91 (y, filter Y, input times certainty),
 * 531-569 (x, filter X), 1201-1240 (z, filter Z, divided by the smoothed
 * certainty).  Tap j of a 9-tap filter weighs the neighbour at offset 4 - j
 * (the filters are symmetric). */
void orc_fl_smooth(int W, int H, int D, const double* in, const float* certainty, const double* scert,
                   const float* fx, const float* fy, const float* fz, double* out) {
    const size_t G = (size_t)W * H * D;
    double* a = (double*)malloc(sizeof(double) * G);
    double* b = (double*)malloc(sizeof(double) * G);
    if (!a || !b) { free(a); free(b); return; }
#pragma omp parallel for collapse(2) schedule(static)
    for (int z = 0; z < D; z++)
        for (int y = 0; y < H; y++)
            for (int x = 0; x < W; x++) {
                double s = 0.0;
                for (int o = -4; o <= 4; o++) {
                    const int yy = y + o;
                    if (yy < 0 || yy >= H) continue;
                    const size_t j = IDX3(x, yy, z);
                    const double c = certainty ? (double)certainty[j] : 1.0;
                    s += in[j] * c * (double)fy[4 - o];
                }
                a[IDX3(x, y, z)] = s;
            }
#pragma omp parallel for collapse(2) schedule(static)
    for (int z = 0; z < D; z++)
        for (int y = 0; y < H; y++)
            for (int x = 0; x < W; x++) {
                double s = 0.0;
                for (int o = -4; o <= 4; o++) {
                    const int xx = x + o;
                    if (xx < 0 || xx >= W) continue;
                    s += a[IDX3(xx, y, z)] * (double)fx[4 - o];
                }
                b[IDX3(x, y, z)] = s;
            }
#pragma omp parallel for collapse(2) schedule(static)
    for (int z = 0; z < D; z++)
        for (int y = 0; y < H; y++)
            for (int x = 0; x < W; x++) {
                double s = 0.0;
                for (int o = -4; o <= 4; o++) {
                    const int zz = z + o;
                    if (zz < 0 || zz >= D) continue;
                    s += b[IDX3(x, y, zz)] * (double)fz[4 - o];
                }
                const size_t i = IDX3(x, y, z);
                out[i] = scert ? s / scert[i] : s;
            }
    free(a);
    free(b);
}

/* LIB:12144-12194: CalculateBetaWeightsGLM with all time points kept, then
 * RemoveLinearFit (kernelStatistics1.cpp:1727), in FP64 with pinv(X). */
void orc_fl_regress(int W, int H, int D, int T, int p, const double* Y, const float* mask, const double* X,
                    double* out) {
    const size_t G = (size_t)W * H * D;
    double* Xp = (double*)malloc(sizeof(double) * (size_t)T * p);
    if (!Xp) return;
    orc_pinv64(T, p, X, Xp); /* Xp[t + r*T] = pinv(X)(r, t) */
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < G; i++) {
        if (mask[i] != 1.0f) {
            for (int t = 0; t < T; t++) out[i + (size_t)t * G] = 0.0;
            continue;
        }
        double beta[64];
        for (int r = 0; r < p; r++) {
            double s = 0.0;
            for (int t = 0; t < T; t++) s += Xp[t + (size_t)r * T] * Y[i + (size_t)t * G];
            beta[r] = s;
        }
        for (int t = 0; t < T; t++) {
            double e = Y[i + (size_t)t * G];
            for (int r = 0; r < p; r++) e -= X[t + (size_t)r * T] * beta[r];
            out[i + (size_t)t * G] = e;
        }
    }
    free(Xp);
}

/* 4x4 determinant and solve by Gaussian elimination with partial pivoting */
static double det_solve4(const double M[4][4], const double* rhs, double* sol) {
    double A[4][5];
    for (int i = 0; i < 4; i++) {
        for (int j = 0; j < 4; j++) A[i][j] = M[i][j];
        A[i][4] = rhs[i];
    }
    double det = 1.0;
    for (int c = 0; c < 4; c++) {
        int piv = c;
        for (int r = c + 1; r < 4; r++)
            if (fabs(A[r][c]) > fabs(A[piv][c])) piv = r;
        if (A[piv][c] == 0.0) {
            for (int i = 0; i < 4; i++) sol[i] = 0.0;
            return 0.0;
        }
        if (piv != c) {
            for (int j = 0; j < 5; j++) {
                const double t = A[c][j];
                A[c][j] = A[piv][j];
                A[piv][j] = t;
            }
            det = -det;
        }
        det *= A[c][c];
        for (int r = c + 1; r < 4; r++) {
            const double f = A[r][c] / A[c][c];
            for (int j = c; j < 5; j++) A[r][j] -= f * A[c][j];
        }
    }
    for (int r = 3; r >= 0; r--) {
        double s = A[r][4];
        for (int j = r + 1; j < 4; j++) s -= A[r][j] * sol[j];
        sol[r] = s / A[r][r];
    }
    return det;
}

/* kernelWhitening.cpp:103-235 (EstimateAR4Models) and the Invert_4x4 of
 * kernelWhitening.cpp:59-98: adjugate / (det + 0.001), i.e. the solution of
 * M a = r scaled by det / (det + 0.001). */
void orc_fl_ar4(int W, int H, int D, int T, const double* data, const float* mask, int invalid, double* ar) {
    const size_t G = (size_t)W * H * D;
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < G; i++) {
        for (int k = 0; k < 4; k++) ar[(size_t)k * G + i] = 0.0;
        if (mask[i] != 1.0f) continue;
        double c[5] = {0, 0, 0, 0, 0};
        for (int t = invalid; t < T; t++) {
            const double v = data[i + (size_t)t * G];
            for (int k = 0; k < 5; k++)
                if (t - k >= invalid) c[k] += v * data[i + (size_t)(t - k) * G];
        }
        for (int k = 0; k < 5; k++) c[k] /= ((double)T - 1.0 - (double)k - (double)invalid);
        if (c[0] == 0.0) continue;
        const double r1 = c[1] / c[0], r2 = c[2] / c[0], r3 = c[3] / c[0], r4 = c[4] / c[0];
        const double a = r1 + 0.001, b = r2 + 0.001, d = r3 + 0.001;
        const double M[4][4] = {{1.0, a, b, d}, {a, 1.0, a, b}, {b, a, 1.0, a}, {d, b, a, 1.0}};
        const double rhs[4] = {r1, r2, r3, r4};
        double sol[4];
        const double det = det_solve4(M, rhs, sol);
        const double scale = det / (det + 0.001);
        for (int k = 0; k < 4; k++) ar[(size_t)k * G + i] = sol[k] * scale;
    }
}

/* kernelWhitening.cpp:371-423 (ApplyWhiteningAR4) */
void orc_fl_whiten(int W, int H, int D, int T, const double* data, const double* ar, const float* mask,
                   double* out) {
    const size_t G = (size_t)W * H * D;
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < G; i++) {
        if (mask[i] != 1.0f) {
            for (int t = 0; t < T; t++) out[i + (size_t)t * G] = data[i + (size_t)t * G];
            continue;
        }
        const double a[4] = {ar[i], ar[G + i], ar[2 * G + i], ar[3 * G + i]};
        for (int t = 0; t < T; t++) {
            double v = data[i + (size_t)t * G];
            for (int k = 1; k <= 4 && k <= t; k++) v -= a[k - 1] * data[i + (size_t)(t - k) * G];
            out[i + (size_t)t * G] = v;
        }
    }
}

/* kernelWhitening.cpp:480-535 (GeneratePermutedVolumesFirstLevel) */
void orc_fl_color(int W, int H, int D, int T, const double* white, const double* ar, const float* mask,
                  const unsigned short* perm, double* out) {
    const size_t G = (size_t)W * H * D;
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < G; i++) {
        if (mask[i] != 1.0f) {
            for (int t = 0; t < T; t++) out[i + (size_t)t * G] = 0.0;
            continue;
        }
        const double a[4] = {ar[i], ar[G + i], ar[2 * G + i], ar[3 * G + i]};
        for (int t = 0; t < T; t++) {
            double v = white[i + (size_t)perm[t] * G];
            for (int k = 1; k <= 4 && k <= t; k++) v += a[k - 1] * out[i + (size_t)(t - k) * G];
            out[i + (size_t)t * G] = v;
        }
    }
}

/* LIB:16363-16452 (PerformWhiteningPriorPermutations) after the regression
 * of LIB:16469-16473 (PerformRegression) */
void orc_fl_prepare(int W, int H, int D, int T, int p, const float* Y, const float* mask, const float* X,
                    float fwhm, float vx, float vy, float vz, double* whitened, double* ar) {
    const size_t G = (size_t)W * H * D;
    double* Yd = (double*)malloc(sizeof(double) * G * T);
    double* e = (double*)malloc(sizeof(double) * G * T);
    double* Xd = (double*)malloc(sizeof(double) * (size_t)T * p);
    double* cur = (double*)malloc(sizeof(double) * 4 * G);
    double* sm = (double*)malloc(sizeof(double) * G);
    double* ones = (double*)malloc(sizeof(double) * G);
    double* smask = (double*)malloc(sizeof(double) * G);
    if (!Yd || !e || !Xd || !cur || !sm || !ones || !smask) goto done;
    for (size_t i = 0; i < G * (size_t)T; i++) Yd[i] = (double)Y[i];
    for (size_t i = 0; i < (size_t)T * p; i++) Xd[i] = (double)X[i];
    orc_fl_regress(W, H, D, T, p, Yd, mask, Xd, e);
    {
        float fx[ORC_FL_TAPS], fy[ORC_FL_TAPS], fz[ORC_FL_TAPS];
        orc_fl_filters(ORC_FL_TAPS, fwhm, vx, vy, vz, fx, fy, fz);
        /* the smoothed mask (PerformSmoothing, certainty 1) */
        for (size_t i = 0; i < G; i++) ones[i] = (double)mask[i];
        orc_fl_smooth(W, H, D, ones, NULL, NULL, fx, fy, fz, smask);
        for (size_t i = 0; i < 4 * G; i++) ar[i] = 0.0;
        memcpy(whitened, e, sizeof(double) * G * T);
        for (int it = 0; it < 3; it++) {
            orc_fl_ar4(W, H, D, T, whitened, mask, it == 0 ? 0 : 4, cur);
            for (int k = 0; k < 4; k++) {
                /* PerformSmoothingNormalized in place (LIB:11545-11620): the
                 * certainty-weighted smoothing over the smoothed certainty,
                 * times the certainty */
                orc_fl_smooth(W, H, D, cur + (size_t)k * G, mask, smask, fx, fy, fz, sm);
                for (size_t i = 0; i < G; i++) {
                    const double v = mask[i] == 0.0f ? 0.0 : sm[i] * (double)mask[i];
                    ar[(size_t)k * G + i] += v;
                }
            }
            orc_fl_whiten(W, H, D, T, e, ar, mask, whitened);
        }
        /* LIB:16442-16445: the totals times the mask */
        for (int k = 0; k < 4; k++)
            for (size_t i = 0; i < G; i++) ar[(size_t)k * G + i] *= (double)mask[i];
    }
done:
    free(Yd);
    free(e);
    free(Xd);
    free(cur);
    free(sm);
    free(ones);
    free(smask);
}

/* kernelStatistics3.cpp:2107-2190 and kernelStatistics5.cpp:1943-2060 */
void orc_fl_stat(int test, int W, int H, int D, int T, int p, int nc, const double* x, const float* mask,
                 const double* X, const double* C, int contrast, double* out) {
    const size_t G = (size_t)W * H * D;
    double* Xp = (double*)malloc(sizeof(double) * (size_t)T * p);
    double* XtXi = (double*)malloc(sizeof(double) * (size_t)p * p);
    double* M = (double*)malloc(sizeof(double) * (size_t)nc * nc);
    if (!Xp || !XtXi || !M) { free(Xp); free(XtXi); free(M); return; }
    orc_pinv64(T, p, X, Xp);
    /* (X'X)^-1 = pinv(X) pinv(X)' */
    for (int a = 0; a < p; a++)
        for (int b = 0; b < p; b++) {
            double s = 0.0;
            for (int t = 0; t < T; t++) s += Xp[t + (size_t)a * T] * Xp[t + (size_t)b * T];
            XtXi[a + (size_t)b * p] = s;
        }
    /* c'(X'X)^-1 c for t; the inverse of C (X'X)^-1 C' for F */
    double ctc = 0.0;
    if (test == 0) {
        for (int a = 0; a < p; a++)
            for (int b = 0; b < p; b++)
                ctc += C[a + (size_t)contrast * p] * XtXi[a + (size_t)b * p] * C[b + (size_t)contrast * p];
    } else {
        double* S = (double*)malloc(sizeof(double) * (size_t)nc * nc);
        for (int i = 0; i < nc; i++)
            for (int j = 0; j < nc; j++) {
                double s = 0.0;
                for (int a = 0; a < p; a++)
                    for (int b = 0; b < p; b++)
                        s += C[a + (size_t)i * p] * XtXi[a + (size_t)b * p] * C[b + (size_t)j * p];
                S[i + (size_t)j * nc] = s;
            }
        /* Gauss-Jordan inverse */
        for (int i = 0; i < nc * nc; i++) M[i] = 0.0;
        for (int i = 0; i < nc; i++) M[i + (size_t)i * nc] = 1.0;
        for (int c = 0; c < nc; c++) {
            int piv = c;
            for (int r = c + 1; r < nc; r++)
                if (fabs(S[r + (size_t)c * nc]) > fabs(S[piv + (size_t)c * nc])) piv = r;
            for (int j = 0; j < nc; j++) {
                double t = S[c + (size_t)j * nc];
                S[c + (size_t)j * nc] = S[piv + (size_t)j * nc];
                S[piv + (size_t)j * nc] = t;
                t = M[c + (size_t)j * nc];
                M[c + (size_t)j * nc] = M[piv + (size_t)j * nc];
                M[piv + (size_t)j * nc] = t;
            }
            const double d = S[c + (size_t)c * nc];
            for (int j = 0; j < nc; j++) {
                S[c + (size_t)j * nc] /= d;
                M[c + (size_t)j * nc] /= d;
            }
            for (int r = 0; r < nc; r++) {
                if (r == c) continue;
                const double f = S[r + (size_t)c * nc];
                for (int j = 0; j < nc; j++) {
                    S[r + (size_t)j * nc] -= f * S[c + (size_t)j * nc];
                    M[r + (size_t)j * nc] -= f * M[c + (size_t)j * nc];
                }
            }
        }
        free(S);
    }
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < G; i++) {
        if (mask[i] != 1.0f) {
            out[i] = 0.0;
            continue;
        }
        double beta[64];
        for (int r = 0; r < p; r++) {
            double s = 0.0;
            for (int t = 0; t < T; t++) s += x[i + (size_t)t * G] * Xp[t + (size_t)r * T];
            beta[r] = s;
        }
        double mean = 0.0, var = 0.0, n = 0.0;
        for (int t = 0; t < T; t++) {
            double e = x[i + (size_t)t * G];
            for (int r = 0; r < p; r++) e -= X[t + (size_t)r * T] * beta[r];
            n += 1.0;
            const double delta = e - mean;
            mean += delta / n;
            var += delta * (e - mean);
        }
        var /= (n - 1.0);
        if (test == 0) {
            double cb = 0.0;
            for (int r = 0; r < p; r++) cb += C[r + (size_t)contrast * p] * beta[r];
            out[i] = cb / sqrt(var * ctc);
        } else {
            double cbeta[64];
            for (int c = 0; c < nc; c++) {
                double s = 0.0;
                for (int r = 0; r < p; r++) s += C[r + (size_t)c * p] * beta[r];
                cbeta[c] = s;
            }
            double f = 0.0;
            for (int a = 0; a < nc; a++)
                for (int b = 0; b < nc; b++) f += cbeta[a] * M[a + (size_t)b * nc] * cbeta[b];
            out[i] = f / var / (double)nc;
        }
    }
    free(Xp);
    free(XtXi);
    free(M);
}

/* LIB:16877-16922 */
int orc_gen_perm_first_level(int T, long K, unsigned short* perm) {
    unsigned short* hist = (unsigned short*)malloc(sizeof(unsigned short) * (size_t)T * (size_t)(K + 1));
    unsigned short* cur = (unsigned short*)malloc(sizeof(unsigned short) * (size_t)T);
    if (!hist || !cur) { free(hist); free(cur); return -1; }
    for (int i = 0; i < T; i++) cur[i] = (unsigned short)i;
    memcpy(hist, cur, sizeof(unsigned short) * T);
    for (long p = 0; p < K; p++) {
        for (;;) {
            orc_random_shuffle_u16(cur, T);
            int unique = 1;
            for (long r = 0; r < p + 1; r++)
                if (memcmp(hist + (size_t)r * T, cur, sizeof(unsigned short) * T) == 0) { unique = 0; break; }
            if (unique) { memcpy(hist + (size_t)(p + 1) * T, cur, sizeof(unsigned short) * T); break; }
        }
        memcpy(perm + (size_t)p * T, cur, sizeof(unsigned short) * T);
    }
    free(hist);
    free(cur);
    return 0;
}

/* LIB:16486-16540: the permutation loop of one contrast */
int orc_fl_null(int test, int inference, int W, int H, int D, int T, int p, int nc, const double* white,
                const double* ar, const float* mask, const float* X, const float* C, int contrast, long K,
                const unsigned short* perm, float cdt, int connectivity, float* maxdist, double* stats) {
    const size_t G = (size_t)W * H * D;
    double* x = (double*)malloc(sizeof(double) * G * T);
    double* map = (double*)malloc(sizeof(double) * G);
    float* mapf = (float*)malloc(sizeof(float) * G);
    unsigned* labels = (unsigned*)malloc(sizeof(unsigned) * G);
    unsigned* aux = (unsigned*)malloc(sizeof(unsigned) * G);
    double* Xd = (double*)malloc(sizeof(double) * (size_t)T * p);
    double* Cd = (double*)malloc(sizeof(double) * (size_t)p * nc);
    int rc = 0;
    if (!x || !map || !mapf || !labels || !aux || !Xd || !Cd) { rc = -1; goto done; }
    for (size_t i = 0; i < (size_t)T * p; i++) Xd[i] = (double)X[i];
    for (size_t i = 0; i < (size_t)p * nc; i++) Cd[i] = (double)C[i];
    for (long k = 0; k < K; k++) {
        orc_fl_color(W, H, D, T, white, ar, mask, perm + (size_t)k * T, x);
        orc_fl_stat(test, W, H, D, T, p, nc, x, mask, Xd, Cd, contrast, map);
        if (stats) memcpy(stats + (size_t)k * G, map, sizeof(double) * G);
        if (inference == ORC_VOXEL) {
            const float init = -1000000.0f;
            int m;
            memcpy(&m, &init, sizeof(int));
            for (size_t i = 0; i < G; i++)
                if (mask[i] == 1.0f) {
                    const double q = trunc(map[i] * 10000.0);
                    if ((int)q > m) m = (int)q;
                }
            maxdist[k] = (float)((float)m / 10000.0f);
        } else {
            for (size_t i = 0; i < G; i++) mapf[i] = (float)map[i];
            const unsigned best = orc_ccl(W, H, D, mapf, mask, cdt, connectivity, labels, aux);
            if (inference == ORC_CLUSTER_EXTENT) {
                maxdist[k] = (float)best;
            } else {
                const unsigned bm = orc_cluster_mass(W, H, D, mapf, mask, cdt, connectivity, labels, aux);
                maxdist[k] = (float)bm / 10000.0f;
            }
        }
    }
done:
    free(x);
    free(map);
    free(mapf);
    free(labels);
    free(aux);
    free(Xd);
    free(Cd);
    return rc;
}
