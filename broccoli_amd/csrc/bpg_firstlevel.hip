// bpg_firstlevel.hip -- gfx950 kernels of the first-level permutation test
// (BROCCOLI_LIB::ApplyPermutationTestFirstLevel, LIB:16454-16558).
//
// Set-up, once per study, in FP64: the regression of the full design
// (LIB:12144-12194), three rounds of AR(4) estimation, normalised smoothing
// of the estimates and whitening (LIB:16363-16452).  Then one fused kernel
// per batch of permutations: inverse whitening of the permuted whitened
// volumes (kernelWhitening.cpp:480-535), the per-voxel GLM statistic
// (kernelStatistics3.cpp:2107-2190, kernelStatistics5.cpp:1943-2060) and the
// per-permutation max / supra-threshold bitmask, without writing the
// permuted volumes anywhere.
#include <hip/hip_runtime.h>

#include <cfloat>
#include <cmath>

#include "bpg_device.h"
#include "bpg_fl_kernels.h"

namespace bpg {
namespace fl {

// Scalar-memory views: wave-uniform addresses in the constant address space
// are read with scalar loads into SGPRs.
typedef const __attribute__((address_space(4))) double* cdouble_p;
typedef const __attribute__((address_space(4))) unsigned* cuint_p;

// ---------------------------------------------------------------------------
// set-up
// ---------------------------------------------------------------------------

__global__ __launch_bounds__(256) void k_regress(const float* __restrict__ Y, const double* __restrict__ Xp,
                                                 const double* __restrict__ X, int T, int p, int Vc, int Vpad,
                                                 double* __restrict__ e) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= Vpad) return;
    if (v >= Vc) {
        for (int t = 0; t < T; t++) e[(size_t)t * Vpad + v] = 0.0;
        return;
    }
    double beta[MAXP];
#pragma unroll
    for (int r = 0; r < MAXP; r++) beta[r] = 0.0;
    for (int t = 0; t < T; t++) {
        const double y = (double)Y[(size_t)t * Vpad + v];
#pragma unroll
        for (int r = 0; r < MAXP; r++)
            if (r < p) beta[r] = fma(Xp[(size_t)r * T + t], y, beta[r]);
    }
    for (int t = 0; t < T; t++) {
        double acc = (double)Y[(size_t)t * Vpad + v];
#pragma unroll
        for (int r = 0; r < MAXP; r++)
            if (r < p) acc = fma(-X[(size_t)t * p + r], beta[r], acc);
        e[(size_t)t * Vpad + v] = acc;
    }
}

hipError_t launch_regress(const float* Y, const double* Xp, const double* X, int T, int p, int Vc, int Vpad,
                          double* e, hipStream_t s) {
    hipLaunchKernelGGL(k_regress, dim3((Vpad + 255) / 256), dim3(256), 0, s, Y, Xp, X, T, p, Vc, Vpad, e);
    return hipGetLastError();
}

// Solves M a = r for the 4x4 system by elimination with pivoting by
// compare-and-swap (all indices compile-time, so everything stays in
// registers); returns det(M).
__device__ __forceinline__ double solve4(double (&A)[4][5], double (&sol)[4]) {
    double det = 1.0;
#pragma unroll
    for (int c = 0; c < 4; c++) {
#pragma unroll
        for (int r = c + 1; r < 4; r++) {
            const bool sw = fabs(A[r][c]) > fabs(A[c][c]);
#pragma unroll
            for (int j = 0; j < 5; j++) {
                const double lo = A[c][j], hi = A[r][j];
                A[c][j] = sw ? hi : lo;
                A[r][j] = sw ? lo : hi;
            }
            det = sw ? -det : det;
        }
        det *= A[c][c];
        const double piv = A[c][c];
#pragma unroll
        for (int r = c + 1; r < 4; r++) {
            const double f = piv != 0.0 ? A[r][c] / piv : 0.0;
#pragma unroll
            for (int j = 0; j < 5; j++) A[r][j] = fma(-f, A[c][j], A[r][j]);
        }
    }
#pragma unroll
    for (int r = 3; r >= 0; r--) {
        double acc = A[r][4];
#pragma unroll
        for (int j = r + 1; j < 4; j++) acc = fma(-A[r][j], sol[j], acc);
        sol[r] = A[r][r] != 0.0 ? acc / A[r][r] : 0.0;
    }
    return det;
}

// kernelWhitening.cpp:103-235: lag sums c_k over t >= invalid + k, divided
// by T - 1 - k - invalid; r_k = c_k / c0; the Toeplitz system of 1 and
// r_k + 0.001 solved and scaled by det / (det + 0.001) (the adjugate over
// det + 0.001 of Invert_4x4, kernelWhitening.cpp:59-98).
__global__ __launch_bounds__(256) void k_ar4(const double* __restrict__ src, int T, int invalid, int Vc, int Vpad,
                                             double* __restrict__ ar) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= Vpad) return;
    double out[4] = {0.0, 0.0, 0.0, 0.0};
    if (v < Vc) {
        double c[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
        double h0 = 0.0, h1 = 0.0, h2 = 0.0, h3 = 0.0;  // v_{t-1}, .., v_{t-4}
        for (int t = invalid; t < T; t++) {
            const double x = src[(size_t)t * Vpad + v];
            const int n = t - invalid;  // earlier samples available
            c[0] = fma(x, x, c[0]);
            if (n >= 1) c[1] = fma(x, h0, c[1]);
            if (n >= 2) c[2] = fma(x, h1, c[2]);
            if (n >= 3) c[3] = fma(x, h2, c[3]);
            if (n >= 4) c[4] = fma(x, h3, c[4]);
            h3 = h2;
            h2 = h1;
            h1 = h0;
            h0 = x;
        }
#pragma unroll
        for (int k = 0; k < 5; k++) c[k] /= ((double)T - 1.0 - (double)k - (double)invalid);
        if (c[0] != 0.0) {
            const double r1 = c[1] / c[0], r2 = c[2] / c[0], r3 = c[3] / c[0], r4 = c[4] / c[0];
            const double a = r1 + 0.001, b = r2 + 0.001, d = r3 + 0.001;
            double A[4][5] = {{1.0, a, b, d, r1}, {a, 1.0, a, b, r2}, {b, a, 1.0, a, r3}, {d, b, a, 1.0, r4}};
            double sol[4];
            const double det = solve4(A, sol);
            const double scale = det / (det + 0.001);
#pragma unroll
            for (int k = 0; k < 4; k++) out[k] = sol[k] * scale;
        }
    }
#pragma unroll
    for (int k = 0; k < 4; k++) ar[(size_t)k * Vpad + v] = out[k];
}

hipError_t launch_ar4(const double* src, int T, int invalid, int Vc, int Vpad, double* ar, hipStream_t s) {
    hipLaunchKernelGGL(k_ar4, dim3((Vpad + 255) / 256), dim3(256), 0, s, src, T, invalid, Vc, Vpad, ar);
    return hipGetLastError();
}

__global__ __launch_bounds__(256) void k_scatter(const double* __restrict__ ar, const unsigned* __restrict__ c2g,
                                                 int Vc, int Vpad, size_t G, double* __restrict__ g) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= Vc) return;
    const size_t i = c2g[v];
#pragma unroll
    for (int k = 0; k < 4; k++) g[(size_t)k * G + i] = ar[(size_t)k * Vpad + v];
}

hipError_t launch_scatter(const double* ar, const unsigned* comp2grid, int Vc, int Vpad, size_t G, double* g,
                          hipStream_t s) {
    if (Vc <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_scatter, dim3((Vc + 255) / 256), dim3(256), 0, s, ar, comp2grid, Vc, Vpad, G, g);
    return hipGetLastError();
}

// kernelConvolution.cpp:52-91 / 531-569 / 1201-1240 without the
// certainty factors (applied by the caller's input and the rods pass)
__global__ __launch_bounds__(256) void k_conv(const double* __restrict__ in, double* __restrict__ out, int nf,
                                              int W, int H, int D, int axis, Taps taps) {
    const size_t G = (size_t)W * H * D;
    const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (size_t)nf * G) return;
    const size_t f = idx / G, gi = idx - f * G;
    const int x = (int)(gi % W), y = (int)((gi / W) % H), z = (int)(gi / ((size_t)W * H));
    const int pos = axis == 0 ? x : axis == 1 ? y : z;
    const int len = axis == 0 ? W : axis == 1 ? H : D;
    const size_t stride = axis == 0 ? 1 : axis == 1 ? (size_t)W : (size_t)W * H;
    const double* base = in + f * G;
    double acc = 0.0;
#pragma unroll
    for (int o = -4; o <= 4; o++) {
        const int q = pos + o;
        if (q >= 0 && q < len) acc = fma(base[gi + (ptrdiff_t)o * (ptrdiff_t)stride], (double)taps.f[4 - o], acc);
    }
    out[idx] = acc;
}

hipError_t launch_conv(const double* in, double* out, int nf, int W, int H, int D, int axis, const Taps& taps,
                       hipStream_t s) {
    const size_t total = (size_t)nf * W * H * D;
    if (total == 0) return hipSuccess;
    hipLaunchKernelGGL(k_conv, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, in, out, nf, W, H, D, axis,
                       taps);
    return hipGetLastError();
}

__global__ __launch_bounds__(256) void k_rods_accumulate(const double* __restrict__ g,
                                                         const double* __restrict__ smask,
                                                         const unsigned* __restrict__ c2g, int Vc, int Vpad, int W,
                                                         int H, int D, Taps taps, double* __restrict__ total) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= 4 * Vc) return;
    const int k = idx / Vc, v = idx - k * Vc;
    const size_t G = (size_t)W * H * D;
    const size_t gi = c2g[v];
    const int z = (int)(gi / ((size_t)W * H));
    const size_t plane = (size_t)W * H;
    const double* base = g + (size_t)k * G;
    double acc = 0.0;
#pragma unroll
    for (int o = -4; o <= 4; o++) {
        const int q = z + o;
        if (q >= 0 && q < D) acc = fma(base[gi + (ptrdiff_t)o * (ptrdiff_t)plane], (double)taps.f[4 - o], acc);
    }
    // in-mask voxels: the final multiplication by the mask is by 1
    total[(size_t)k * Vpad + v] += acc / smask[gi];
}

hipError_t launch_rods_accumulate(const double* g, const double* smask, const unsigned* comp2grid, int Vc,
                                  int Vpad, int W, int H, int D, const Taps& taps, double* total, hipStream_t s) {
    if (Vc <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_rods_accumulate, dim3((4 * Vc + 255) / 256), dim3(256), 0, s, g, smask, comp2grid, Vc,
                       Vpad, W, H, D, taps, total);
    return hipGetLastError();
}

// kernelWhitening.cpp:371-423
__global__ __launch_bounds__(256) void k_whiten(const double* __restrict__ e, const double* __restrict__ ar, int T,
                                                int Vc, int Vpad, double* __restrict__ wh, float* __restrict__ w32) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= Vpad) return;
    const double a1 = ar[v], a2 = ar[(size_t)Vpad + v], a3 = ar[2 * (size_t)Vpad + v], a4 = ar[3 * (size_t)Vpad + v];
    double h0 = 0.0, h1 = 0.0, h2 = 0.0, h3 = 0.0;
    for (int t = 0; t < T; t++) {
        const double x = v < Vc ? e[(size_t)t * Vpad + v] : 0.0;
        double o = x;
        o = fma(-a1, h0, o);
        o = fma(-a2, h1, o);
        o = fma(-a3, h2, o);
        o = fma(-a4, h3, o);
        h3 = h2;
        h2 = h1;
        h1 = h0;
        h0 = x;
        wh[(size_t)t * Vpad + v] = o;
        if (w32) w32[(size_t)t * Vpad + v] = (float)o;
    }
}

hipError_t launch_whiten(const double* e, const double* ar, int T, int Vc, int Vpad, double* wh, float* w32,
                         hipStream_t s) {
    hipLaunchKernelGGL(k_whiten, dim3((Vpad + 255) / 256), dim3(256), 0, s, e, ar, T, Vc, Vpad, wh, w32);
    return hipGetLastError();
}

__global__ __launch_bounds__(256) void k_wscale(const float* __restrict__ w32, int T, int Vc, int Vpad,
                                                float* __restrict__ wscale) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= Vpad) return;
    float mx = 0.0f;
    if (v < Vc)
        for (int t = 0; t < T; t++) mx = fmaxf(mx, fabsf(w32[(size_t)t * Vpad + v]));
    float sc = 1.0f;
    if (mx > 0.0f && mx <= FLT_MAX) {
        int e;
        (void)frexpf(mx, &e);
        sc = ldexpf(1.0f, 9 - max(e, -100));
    }
    wscale[v] = sc;
}

hipError_t launch_wscale(const float* w32, int T, int Vc, int Vpad, float* wscale, hipStream_t s) {
    hipLaunchKernelGGL(k_wscale, dim3((Vpad + 255) / 256), dim3(256), 0, s, w32, T, Vc, Vpad, wscale);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// the permutation kernel
// ---------------------------------------------------------------------------
// Block: 8 waves over one 64-voxel tile at a time (lane = voxel) and a
// group of 64 permutations (each wave 4 rounds of 2); blocks of one voxel
// chunk walk its tiles.  The tile's whitened volumes ([T][64] floats) sit
// in LDS when T <= LDS_T_MAX; permutation indices and the design are
// wave-uniform scalar loads.  Each thread keeps, per permutation, 4 AR
// history values, pq projections, |x|^2 and sum x in FP64.

__device__ __forceinline__ float wave_max(float x) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) x = fmaxf(x, __shfl_xor(x, o));
    return x;
}

template <int PQ, bool USE_LDS, bool MEAN>
__global__ __launch_bounds__(512) void k_fl_perm(PermArgs a) {
    extern __shared__ float tile[];
    __shared__ float smax[WAVES * ROUNDS * NP];
    const int n_blocks = a.n_vchunks * a.n_pgroups;
    const int lid = xcd_logical_block((int)(gridDim.x >> 3));
    if (lid >= n_blocks) return;
    const int vchunk = lid / a.n_pgroups;
    const int pgroup = lid - vchunk * a.n_pgroups;
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int tile0 = vchunk * a.tiles_per_chunk;
    const int tile1 = min(tile0 + a.tiles_per_chunk, a.n_vtiles);
    const cdouble_p Q = (cdouble_p)a.Q;
    const cdouble_p A = (cdouble_p)a.A;
    const cdouble_p U = (cdouble_p)a.u;
    const double invT = 1.0 / (double)a.T, invTm1 = 1.0 / ((double)a.T - 1.0);
    float* my = smax + wave * (ROUNDS * NP);
    if (lane == 0)
        for (int i = 0; i < ROUNDS * NP; i++) my[i] = -INFINITY;

    for (int tl = tile0; tl < tile1; tl++) {
        const int v = tl * TILE + lane;
        if (USE_LDS) {
            __syncthreads();
            const int n = a.T * TILE;
            for (int i = threadIdx.x; i < n; i += WAVES * 64)
                tile[i] = a.w32[(size_t)(i >> 6) * a.Vpad + (size_t)tl * TILE + (i & 63)];
            __syncthreads();
        }
        const double a1 = a.ar[v], a2 = a.ar[(size_t)a.Vpad + v];
        const double a3 = a.ar[2 * (size_t)a.Vpad + v], a4 = a.ar[3 * (size_t)a.Vpad + v];
        const bool in_mask = v < a.Vc;
        for (int r = 0; r < ROUNDS; r++) {
            const int kbase = pgroup * PERMS_PER_BLOCK + r * (WAVES * NP) + wave * NP;
            if (kbase >= a.K) break;
            cuint_p prow[NP];
#pragma unroll
            for (int j = 0; j < NP; j++) {
                const int k = min(kbase + j, a.K - 1);
                prow[j] = (cuint_p)(a.perm + (size_t)k * a.Tp);
            }
            double z[NP][PQ], h[NP][4], s[NP], m[NP];
#pragma unroll
            for (int j = 0; j < NP; j++) {
#pragma unroll
                for (int q = 0; q < PQ; q++) z[j][q] = 0.0;
#pragma unroll
                for (int q = 0; q < 4; q++) h[j][q] = 0.0;
                s[j] = 0.0;
                m[j] = 0.0;
            }
            // one time step of every permutation of the thread
            auto step = [&](int t, const unsigned (&idx)[NP]) {
                double qv[PQ];
#pragma unroll
                for (int q = 0; q < PQ; q++) qv[q] = Q[(size_t)t * PQ + q];
#pragma unroll
                for (int j = 0; j < NP; j++) {
                    const float w = USE_LDS ? tile[idx[j] * TILE + lane] : a.w32[(size_t)idx[j] * a.Vpad + v];
                    double x = (double)w;
                    x = fma(a4, h[j][3], x);
                    x = fma(a3, h[j][2], x);
                    x = fma(a2, h[j][1], x);
                    x = fma(a1, h[j][0], x);
                    h[j][3] = h[j][2];
                    h[j][2] = h[j][1];
                    h[j][1] = h[j][0];
                    h[j][0] = x;
                    s[j] = fma(x, x, s[j]);
                    if (MEAN) m[j] += x;
#pragma unroll
                    for (int q = 0; q < PQ; q++) z[j][q] = fma(qv[q], x, z[j][q]);
                }
            };
            // whole groups of 8 time points: no exits inside, so the AR
            // history rotates by register renaming (8 is a multiple of 4)
            const int Tfull = a.T & ~7;
            for (int t8 = 0; t8 < Tfull; t8 += 8) {
                unsigned pk[NP][4];  // 8 indices of each permutation
#pragma unroll
                for (int j = 0; j < NP; j++)
#pragma unroll
                    for (int q = 0; q < 4; q++) pk[j][q] = prow[j][(t8 >> 1) + q];
#pragma unroll
                for (int tt = 0; tt < 8; tt++) {
                    unsigned idx[NP];
#pragma unroll
                    for (int j = 0; j < NP; j++) idx[j] = (pk[j][tt >> 1] >> ((tt & 1) * 16)) & 0xFFFFu;
                    step(t8 + tt, idx);
                }
            }
            for (int t = Tfull; t < a.T; t++) {
                unsigned idx[NP];
#pragma unroll
                for (int j = 0; j < NP; j++) idx[j] = (prow[j][t >> 1] >> ((t & 1) * 16)) & 0xFFFFu;
                step(t, idx);
            }
#pragma unroll
            for (int j = 0; j < NP; j++) {
                const int k = kbase + j;
                const bool valid = k < a.K;
                double ss = 0.0, uz = 0.0;
#pragma unroll
                for (int q = 0; q < PQ; q++) {
                    ss = fma(z[j][q], z[j][q], ss);
                    uz = fma(U[q], z[j][q], uz);
                }
                const double sse = s[j] - ss;
                // without the constant in col(X) the residuals need not sum to 0
                const double me = MEAN ? m[j] - uz : 0.0;
                const double var = (sse - me * me * invT) * invTm1;
                double stat;
                if (!a.ftest) {
                    double d = 0.0;
#pragma unroll
                    for (int q = 0; q < PQ; q++) d = fma(A[q], z[j][q], d);
                    stat = d / sqrt(var);
                } else {
                    double acc = 0.0;
                    for (int i = 0; i < a.nrows; i++) {
                        double d = 0.0;
#pragma unroll
                        for (int q = 0; q < PQ; q++) d = fma(A[(size_t)i * PQ + q], z[j][q], d);
                        acc = fma(d, d, acc);
                    }
                    stat = acc / (var * (double)a.nrows);
                }
                const float st = (float)stat;
                const bool act = in_mask && valid;
                if (a.stats && act) a.stats[(size_t)k * a.Vc + v] = st;
                if (a.bits) {
                    const unsigned long long bal = __ballot(act && st > a.cdt);
                    if (lane == 0 && valid) {
                        a.bits[(size_t)k * a.nwords + 2 * (size_t)tl] = (unsigned)bal;
                        a.bits[(size_t)k * a.nwords + 2 * (size_t)tl + 1] = (unsigned)(bal >> 32);
                    }
                }
                const float mx = wave_max(act ? st : -INFINITY);
                if (lane == 0) my[r * NP + j] = fmaxf(my[r * NP + j], mx);
            }
        }
    }
    if (lane == 0) {
        for (int r = 0; r < ROUNDS; r++)
            for (int j = 0; j < NP; j++) {
                const int k = pgroup * PERMS_PER_BLOCK + r * (WAVES * NP) + wave * NP + j;
                if (k < a.K) a.partial[(size_t)vchunk * a.K + k] = my[r * NP + j];
            }
    }
}

// ---------------------------------------------------------------------------
// the FP32 form: the recursion, |x|^2 and sum x in FP32 on the vector ALUs
// (the reference's arithmetic), the projections z = Q'x on the FP32 matrix
// cores.  Per group of 4 time points and permutation, each lane holds x for
// its voxel at t..t+3 in 4 registers; two permlane32 and two permlane16
// swaps turn them into the A operands of 4 MFMA 16x16x4 (16 voxels x 4 time
// points each), B = Q[t..t+3][0..15] (4 rows, one coalesced load), so the
// 64 voxels x 16 columns of z accumulate in 4 x 4 registers.  The basis is
// rotated so that its first column is the t contrast direction (first r
// columns span the F rows): the numerator is column 0 (or the sum of
// squares of the first r columns).  Samples with SSE < |x|^2 / 64 flag
// their permutation for an FP64 repeat.
// ---------------------------------------------------------------------------

typedef float f32x4 __attribute__((ext_vector_type(4)));

// sum over the 16 lanes of a row, in every lane of the row (DPP row
// rotations, no LDS)
__device__ __forceinline__ float sum16(float x) {
    x += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), 0x128, 0xF, 0xF, false));  // row_ror:8
    x += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), 0x124, 0xF, 0xF, false));  // row_ror:4
    x += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), 0x122, 0xF, 0xF, false));  // row_ror:2
    x += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), 0x121, 0xF, 0xF, false));  // row_ror:1
    return x;
}

__device__ __forceinline__ void transpose4(float (&X)[4]) {
    // [X_k] row blocks of 16 lanes -> X_g[16k + i] = X_k[16g + i] (the
    // permlane32 swap exchanges the upper half of its first operand with
    // the lower half of its second; permlane16 odd rows of the first with
    // even rows of the second)
    auto p02 = __builtin_amdgcn_permlane32_swap(__float_as_uint(X[0]), __float_as_uint(X[2]), false, false);
    auto p13 = __builtin_amdgcn_permlane32_swap(__float_as_uint(X[1]), __float_as_uint(X[3]), false, false);
    auto q01 = __builtin_amdgcn_permlane16_swap(p02[0], p13[0], false, false);
    auto q23 = __builtin_amdgcn_permlane16_swap(p02[1], p13[1], false, false);
    X[0] = __uint_as_float(q01[0]);
    X[1] = __uint_as_float(q01[1]);
    X[2] = __uint_as_float(q23[0]);
    X[3] = __uint_as_float(q23[1]);
}

template <bool MEAN>
__global__ __launch_bounds__(512) void k_fl_perm_mx(PermArgs a) {
    extern __shared__ float tile[];  // [T][64] whitened tile
    __shared__ float smax[WAVES * ROUNDS * NP];
    __shared__ float red[WAVES][3][TILE];
    const int n_blocks = a.n_vchunks * a.n_pgroups;
    const int lid = xcd_logical_block((int)(gridDim.x >> 3));
    if (lid >= n_blocks) return;
    const int vchunk = lid / a.n_pgroups;
    const int pgroup = lid - vchunk * a.n_pgroups;
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int tile0 = vchunk * a.tiles_per_chunk;
    const int tile1 = min(tile0 + a.tiles_per_chunk, a.n_vtiles);
    // this lane's column c = lane % 16 of the rotated basis
    const int c = lane & 15;
    const float uc = a.u32[c];
    const float invT = 1.0f / (float)a.T, invTm1 = 1.0f / ((float)a.T - 1.0f);
    float* my = smax + wave * (ROUNDS * NP);
    if (lane == 0)
        for (int i = 0; i < ROUNDS * NP; i++) my[i] = -INFINITY;

    for (int tl = tile0; tl < tile1; tl++) {
        const int v = tl * TILE + lane;
        __syncthreads();
        {
            const int n = a.T * TILE;
            for (int i = threadIdx.x; i < n; i += WAVES * 64)
                tile[i] = a.w32[(size_t)(i >> 6) * a.Vpad + (size_t)tl * TILE + (i & 63)];
        }
        __syncthreads();
        const float a1 = (float)a.ar[v], a2 = (float)a.ar[(size_t)a.Vpad + v];
        const float a3 = (float)a.ar[2 * (size_t)a.Vpad + v], a4 = (float)a.ar[3 * (size_t)a.Vpad + v];
        const bool in_mask = v < a.Vc;
        for (int r = 0; r < ROUNDS; r++) {
            const int kbase = pgroup * PERMS_PER_BLOCK + r * (WAVES * NP) + wave * NP;
            if (kbase >= a.K) break;
            cuint_p prow[NP];
#pragma unroll
            for (int j = 0; j < NP; j++) {
                const int k = min(kbase + j, a.K - 1);
                prow[j] = (cuint_p)(a.perm + (size_t)k * a.Tp);
            }
            f32x4 acc[NP][4];
            float h[NP][4], s[NP], m[NP];
#pragma unroll
            for (int j = 0; j < NP; j++) {
#pragma unroll
                for (int g = 0; g < 4; g++) acc[j][g] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
                for (int q = 0; q < 4; q++) h[j][q] = 0.0f;
                s[j] = 0.0f;
                m[j] = 0.0f;
            }
            // 4 time points from t0 (n valid): recursion, swaps, 4 MFMAs per
            // permutation
            auto block4 = [&](int t0, const unsigned (&idx)[NP][4], int n) {
                const float b = a.Q32[(size_t)t0 * 16 + lane];  // B[k][col] = Q[t0 + lane/16][lane%16]
#pragma unroll
                for (int j = 0; j < NP; j++) {
                    float X[4];
#pragma unroll
                    for (int k = 0; k < 4; k++) {
                        if (k < n) {
                            float x = tile[idx[j][k] * TILE + lane];
                            x = fmaf(a4, h[j][3], x);
                            x = fmaf(a3, h[j][2], x);
                            x = fmaf(a2, h[j][1], x);
                            x = fmaf(a1, h[j][0], x);
                            h[j][3] = h[j][2];
                            h[j][2] = h[j][1];
                            h[j][1] = h[j][0];
                            h[j][0] = x;
                            s[j] = fmaf(x, x, s[j]);
                            if (MEAN) m[j] += x;
                            X[k] = x;
                        } else {
                            X[k] = 0.0f;
                        }
                    }
                    transpose4(X);
#pragma unroll
                    for (int g = 0; g < 4; g++) acc[j][g] = __builtin_amdgcn_mfma_f32_16x16x4f32(X[g], b, acc[j][g], 0, 0, 0);
                }
            };
            const int Tfull = a.T & ~7;
            for (int t8 = 0; t8 < Tfull; t8 += 8) {
                unsigned pk[NP][4];
#pragma unroll
                for (int j = 0; j < NP; j++)
#pragma unroll
                    for (int q = 0; q < 4; q++) pk[j][q] = prow[j][(t8 >> 1) + q];
#pragma unroll
                for (int hb = 0; hb < 2; hb++) {
                    unsigned idx[NP][4];
#pragma unroll
                    for (int j = 0; j < NP; j++)
#pragma unroll
                        for (int k = 0; k < 4; k++) {
                            const int tt = hb * 4 + k;
                            idx[j][k] = (pk[j][tt >> 1] >> ((tt & 1) * 16)) & 0xFFFFu;
                        }
                    block4(t8 + hb * 4, idx, 4);
                }
            }
            for (int t0 = Tfull; t0 < a.T; t0 += 4) {
                const int n = min(4, a.T - t0);
                unsigned idx[NP][4];
#pragma unroll
                for (int j = 0; j < NP; j++)
#pragma unroll
                    for (int k = 0; k < 4; k++) {
                        const int t = min(t0 + k, a.T - 1);
                        idx[j][k] = (prow[j][t >> 1] >> ((t & 1) * 16)) & 0xFFFFu;
                    }
                block4(t0, idx, n);
            }
            // epilogue: per voxel |z|^2, the statistic rows and u'z from the
            // 16 lanes that hold its columns, back to lane = voxel via LDS
#pragma unroll
            for (int j = 0; j < NP; j++) {
                const int k = kbase + j;
                const bool valid = k < a.K;
#pragma unroll
                for (int g = 0; g < 4; g++)
#pragma unroll
                    for (int rr = 0; rr < 4; rr++) {
                        const float zc = acc[j][g][rr];
                        const float z2 = zc * zc;
                        const float ss = sum16(z2);
                        const float uz = MEAN ? sum16(uc * zc) : 0.0f;
                        // t: column 0 (held by the writing lane); F: the
                        // first nrows columns
                        const float num = a.ftest ? sum16(c < a.nrows ? z2 : 0.0f) : zc;
                        const int vv = 16 * g + 4 * (lane >> 4) + rr;  // voxel of this row
                        if (c == 0) {
                            red[wave][0][vv] = ss;
                            red[wave][1][vv] = num;
                            red[wave][2][vv] = uz;
                        }
                    }
                const float ss = red[wave][0][lane], num = red[wave][1][lane], uz = red[wave][2][lane];
                const float sse = s[j] - ss;
                const float me = MEAN ? m[j] - uz : 0.0f;
                const float var = (sse - me * me * invT) * invTm1;
                const float st = a.ftest ? num / (var * (float)a.nrows) : num / sqrtf(var);
                const bool act = in_mask && valid;
                // conditioning: a sample whose residual sum of squares is
                // below 1/64 of |x|^2 needs the FP64 form
                const bool bad = act && !(sse >= s[j] * (1.0f / 64.0f));
                if (__ballot(bad) && lane == 0) a.flags[k] = 1;
                if (a.stats && act) a.stats[(size_t)k * a.Vc + v] = st;
                if (a.bits) {
                    const unsigned long long bal = __ballot(act && st > a.cdt);
                    if (lane == 0 && valid) {
                        a.bits[(size_t)k * a.nwords + 2 * (size_t)tl] = (unsigned)bal;
                        a.bits[(size_t)k * a.nwords + 2 * (size_t)tl + 1] = (unsigned)(bal >> 32);
                    }
                }
                const float mx = wave_max(act ? st : -INFINITY);
                if (lane == 0) my[r * NP + j] = fmaxf(my[r * NP + j], mx);
            }
        }
    }
    if (lane == 0) {
        for (int r = 0; r < ROUNDS; r++)
            for (int j = 0; j < NP; j++) {
                const int k = pgroup * PERMS_PER_BLOCK + r * (WAVES * NP) + wave * NP + j;
                if (k < a.K) a.partial[(size_t)vchunk * a.K + k] = my[r * NP + j];
            }
    }
}

// ---------------------------------------------------------------------------
// the split-BF16 form: as the FP32 form, but the projections run on the BF16
// matrix cores (v_mfma_f32_16x16x32_bf16), which overlap the vector work.
// Per 32 time points each lane splits its voxel's x into bf16 hi + lo (8
// per 4 registers, |x - hi - lo| <= 2^-18 |x|); the permlane transposes
// turn each register slice into the A operands of the 4 row groups, and 3
// products (hi.hi, hi.lo, lo.hi) per group accumulate z.  B is the rotated
// basis split the same way, laid out per lane on the host.  Samples with
// SSE < |x|^2 / 8 flag their permutation for the FP64 form.
// ---------------------------------------------------------------------------

__device__ __forceinline__ void transpose4u(unsigned (&X)[4]) {
    auto p02 = __builtin_amdgcn_permlane32_swap(X[0], X[2], false, false);
    auto p13 = __builtin_amdgcn_permlane32_swap(X[1], X[3], false, false);
    auto q01 = __builtin_amdgcn_permlane16_swap(p02[0], p13[0], false, false);
    auto q23 = __builtin_amdgcn_permlane16_swap(p02[1], p13[1], false, false);
    X[0] = q01[0];
    X[1] = q01[1];
    X[2] = q23[0];
    X[3] = q23[1];
}

typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));

// (x0, x1) -> packed bf16 hi parts and bf16 remainders, round to nearest
// even: one v_cvt_pk_bf16_f32 per pair for each
__device__ __forceinline__ void split_pair(float x0, float x1, unsigned& hi, unsigned& lo) {
    const bf16x2 h = __builtin_convertvector(f32x2{x0, x1}, bf16x2);
    hi = __builtin_bit_cast(unsigned, h);
    const float h0 = __uint_as_float(hi << 16), h1 = __uint_as_float(hi & 0xFFFF0000u);
    lo = __builtin_bit_cast(unsigned, __builtin_convertvector(f32x2{x0 - h0, x1 - h1}, bf16x2));
}

template <bool MEAN>
__global__ __launch_bounds__(512) void k_fl_perm_bx(PermArgs a) {
    extern __shared__ float tile[];  // [T][64] whitened tile
    __shared__ float smax[WAVES * ROUNDS * NP];
    __shared__ float red[WAVES][3][TILE];
    const int n_blocks = a.n_vchunks * a.n_pgroups;
    const int lid = xcd_logical_block((int)(gridDim.x >> 3));
    if (lid >= n_blocks) return;
    const int vchunk = lid / a.n_pgroups;
    const int pgroup = lid - vchunk * a.n_pgroups;
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int tile0 = vchunk * a.tiles_per_chunk;
    const int tile1 = min(tile0 + a.tiles_per_chunk, a.n_vtiles);
    const bf16x8* QB = reinterpret_cast<const bf16x8*>(a.Qb);
    const int c = lane & 15;
    const float uc = a.u32[c];
    const float invT = 1.0f / (float)a.T, invTm1 = 1.0f / ((float)a.T - 1.0f);
    float* my = smax + wave * (ROUNDS * NP);
    if (lane == 0)
        for (int i = 0; i < ROUNDS * NP; i++) my[i] = -INFINITY;

    for (int tl = tile0; tl < tile1; tl++) {
        const int v = tl * TILE + lane;
        __syncthreads();
        {
            const int n = a.T * TILE;
            for (int i = threadIdx.x; i < n; i += WAVES * 64)
                tile[i] = a.w32[(size_t)(i >> 6) * a.Vpad + (size_t)tl * TILE + (i & 63)];
        }
        __syncthreads();
        const float a1 = (float)a.ar[v], a2 = (float)a.ar[(size_t)a.Vpad + v];
        const float a3 = (float)a.ar[2 * (size_t)a.Vpad + v], a4 = (float)a.ar[3 * (size_t)a.Vpad + v];
        const bool in_mask = v < a.Vc;
        for (int r = 0; r < ROUNDS; r++) {
            const int kbase = pgroup * PERMS_PER_BLOCK + r * (WAVES * NP) + wave * NP;
            if (kbase >= a.K) break;
            cuint_p prow[NP];
#pragma unroll
            for (int j = 0; j < NP; j++) {
                const int k = min(kbase + j, a.K - 1);
                prow[j] = (cuint_p)(a.perm + (size_t)k * a.Tp);
            }
            f32x4 acc[NP][4];
            float h[NP][4], s[NP], m[NP];
#pragma unroll
            for (int j = 0; j < NP; j++) {
#pragma unroll
                for (int g = 0; g < 4; g++) acc[j][g] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
                for (int q = 0; q < 4; q++) h[j][q] = 0.0f;
                s[j] = 0.0f;
                m[j] = 0.0f;
            }
            // 32 time points from t0, n of them valid
            auto block32 = [&](int t0, int n) {
                const bf16x8 bh = QB[((size_t)(t0 >> 5) * 2) * 64 + lane];
                const bf16x8 bl = QB[((size_t)(t0 >> 5) * 2 + 1) * 64 + lane];
                // one permutation at a time keeps 32 split registers live
#pragma unroll
                for (int j = 0; j < NP; j++) {
                    unsigned hi[16], lo[16];  // [chunk * 4 + w]: bf16 pair of times 8 chunk + 2w, +1
#pragma unroll
                    for (int ch = 0; ch < 4; ch++) {
                        unsigned pk[4];
#pragma unroll
                        for (int q = 0; q < 4; q++) {
                            const int t = min(t0 + 8 * ch + 2 * q, a.Tp - 2);
                            pk[q] = prow[j][t >> 1];
                        }
                        float xs[8];
#pragma unroll
                        for (int e = 0; e < 8; e++) {
                            if (8 * ch + e < n) {
                                const unsigned idx = (pk[e >> 1] >> ((e & 1) * 16)) & 0xFFFFu;
                                float x = tile[idx * TILE + lane];
                                x = fmaf(a4, h[j][3], x);
                                x = fmaf(a3, h[j][2], x);
                                x = fmaf(a2, h[j][1], x);
                                x = fmaf(a1, h[j][0], x);
                                h[j][3] = h[j][2];
                                h[j][2] = h[j][1];
                                h[j][1] = h[j][0];
                                h[j][0] = x;
                                s[j] = fmaf(x, x, s[j]);
                                if (MEAN) m[j] += x;
                                xs[e] = x;
                            } else {
                                xs[e] = 0.0f;
                            }
                        }
#pragma unroll
                        for (int w = 0; w < 4; w++) split_pair(xs[2 * w], xs[2 * w + 1], hi[ch * 4 + w], lo[ch * 4 + w]);
                    }
                    // per register slice w: chunk-major -> row-group-major
                    unsigned ah[4][4], al[4][4];  // [group][w]
#pragma unroll
                    for (int w = 0; w < 4; w++) {
                        unsigned xh[4] = {hi[w], hi[4 + w], hi[8 + w], hi[12 + w]};
                        unsigned xl[4] = {lo[w], lo[4 + w], lo[8 + w], lo[12 + w]};
                        transpose4u(xh);
                        transpose4u(xl);
#pragma unroll
                        for (int g = 0; g < 4; g++) {
                            ah[g][w] = xh[g];
                            al[g][w] = xl[g];
                        }
                    }
#pragma unroll
                    for (int g = 0; g < 4; g++) {
                        typedef unsigned u32x4 __attribute__((ext_vector_type(4)));
                        const bf16x8 Ah = __builtin_bit_cast(bf16x8, u32x4{ah[g][0], ah[g][1], ah[g][2], ah[g][3]});
                        const bf16x8 Al = __builtin_bit_cast(bf16x8, u32x4{al[g][0], al[g][1], al[g][2], al[g][3]});
                        acc[j][g] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(Ah, bh, acc[j][g], 0, 0, 0);
                        acc[j][g] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(Ah, bl, acc[j][g], 0, 0, 0);
                        acc[j][g] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(Al, bh, acc[j][g], 0, 0, 0);
                    }
                }
            };
            const int Tfull = a.T & ~31;
            for (int t0 = 0; t0 < Tfull; t0 += 32) block32(t0, 32);
            if (Tfull < a.T) block32(Tfull, a.T - Tfull);
#pragma unroll
            for (int j = 0; j < NP; j++) {
                const int k = kbase + j;
                const bool valid = k < a.K;
#pragma unroll
                for (int g = 0; g < 4; g++)
#pragma unroll
                    for (int rr = 0; rr < 4; rr++) {
                        const float zc = acc[j][g][rr];
                        const float z2 = zc * zc;
                        const float ss = sum16(z2);
                        const float uz = MEAN ? sum16(uc * zc) : 0.0f;
                        const float num = a.ftest ? sum16(c < a.nrows ? z2 : 0.0f) : zc;
                        const int vv = 16 * g + 4 * (lane >> 4) + rr;
                        if (c == 0) {
                            red[wave][0][vv] = ss;
                            red[wave][1][vv] = num;
                            red[wave][2][vv] = uz;
                        }
                    }
                const float ss = red[wave][0][lane], num = red[wave][1][lane], uz = red[wave][2][lane];
                const float sse = s[j] - ss;
                const float me = MEAN ? m[j] - uz : 0.0f;
                const float var = (sse - me * me * invT) * invTm1;
                const float st = a.ftest ? num / (var * (float)a.nrows) : num / sqrtf(var);
                const bool act = in_mask && valid;
                // the split-BF16 z is good to about 1e-5 of |x|: samples with
                // SSE below 1/8 of |x|^2 need the FP64 form
                const bool bad = act && !(sse >= s[j] * 0.125f);
                if (__ballot(bad) && lane == 0) a.flags[k] = 1;
                if (a.stats && act) a.stats[(size_t)k * a.Vc + v] = st;
                if (a.bits) {
                    const unsigned long long bal = __ballot(act && st > a.cdt);
                    if (lane == 0 && valid) {
                        a.bits[(size_t)k * a.nwords + 2 * (size_t)tl] = (unsigned)bal;
                        a.bits[(size_t)k * a.nwords + 2 * (size_t)tl + 1] = (unsigned)(bal >> 32);
                    }
                }
                const float mx = wave_max(act ? st : -INFINITY);
                if (lane == 0) my[r * NP + j] = fmaxf(my[r * NP + j], mx);
            }
        }
    }
    if (lane == 0) {
        for (int r = 0; r < ROUNDS; r++)
            for (int j = 0; j < NP; j++) {
                const int k = pgroup * PERMS_PER_BLOCK + r * (WAVES * NP) + wave * NP + j;
                if (k < a.K) a.partial[(size_t)vchunk * a.K + k] = my[r * NP + j];
            }
    }
}

// ---------------------------------------------------------------------------
// the split-FP16 form: the FP32 recursion of the other forms, the
// projections on v_mfma_f32_32x32x16_f16 from FP16 hi/lo splits (two
// instructions per value), 8 time points per product.  Each voxel's series
// is scaled by a power of two (exact, and the statistic is scale-free) so
// that its values sit well inside the FP16 range; a sample with
// |x|^2 >= 4e9 may have left it and is flagged, as is one with
// SSE < |x|^2 / 32.
// ---------------------------------------------------------------------------

// Lane l supplies B column l % 32 for k-half l / 32 (its own x at 8 time
// points), so a column mixes the series of lanes j and j + 32.  A is
// block-diagonal: rows 0..15 (the basis columns for lane j's series) read
// k 0..7 only, rows 16..31 (for lane j + 32's) k 8..15 only, so D rows
// 0..15 are z of lane j's voxel and rows 16..31 z of lane j + 32's.  Row i
// of D lands in lane half (i / 4) % 2, register i % 4 + 4 (i / 8) (lane
// layouts probed on the device).  The A operand of a lane is 8 basis values
// (its column i % 16) or zeros: lanes with i < 16 in the upper half, or
// i >= 16 in the lower, read a zero block and do not advance.  Half of
// each product is zeros; 3 products per 8 time points and permutation
// keep the matrix pipe at a fraction of the vector work.
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

// x -> FP16 hi (rounded toward zero) and the remainder, rounded toward zero:
// |x - hi - lo| < 2^-20 |x| in the FP16 normal range.  m1 is -1 held in a
// register, so that x - hi is one v_fma_mix_f32 reading the FP16 half.
__device__ __forceinline__ void split8_f16(const float (&x)[8], float m1, f16x8& hi, f16x8& lo) {
    unsigned h[4], l[4];
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const auto hp = __builtin_amdgcn_cvt_pkrtz(x[2 * q], x[2 * q + 1]);
        const auto lp = __builtin_amdgcn_cvt_pkrtz(__builtin_fmaf((float)hp[0], m1, x[2 * q]),
                                                   __builtin_fmaf((float)hp[1], m1, x[2 * q + 1]));
        h[q] = __builtin_bit_cast(unsigned, hp);
        l[q] = __builtin_bit_cast(unsigned, lp);
    }
    typedef unsigned u32x4 __attribute__((ext_vector_type(4)));
    hi = __builtin_bit_cast(f16x8, u32x4{h[0], h[1], h[2], h[3]});
    lo = __builtin_bit_cast(f16x8, u32x4{l[0], l[1], l[2], l[3]});
}

template <bool MEAN, int NB>
__global__ __launch_bounds__(512) void k_fl_perm_hx(PermArgs a) {
    // [T][64] scaled whitened tile, then NB * 512 zero bytes, then the basis:
    // [T8][hi, lo][NB blocks of 16 columns][8 time points] FP16
    extern __shared__ float tile[];
    __shared__ float smax[WAVES * ROUNDS * NP];
    const int n_blocks = a.n_vchunks * a.n_pgroups;
    const int lid = xcd_logical_block((int)(gridDim.x >> 3));
    if (lid >= n_blocks) return;
    const int vchunk = lid / a.n_pgroups;
    const int pgroup = lid - vchunk * a.n_pgroups;
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int tile0 = vchunk * a.tiles_per_chunk;
    const int tile1 = min(tile0 + a.tiles_per_chunk, a.n_vtiles);
    const int T8 = (a.T + 7) >> 3;
    char* qbase = reinterpret_cast<char*>(tile + (size_t)a.T * TILE);
    {
        // zero block + basis, once per block (the first barrier below orders it)
        unsigned* dst = reinterpret_cast<unsigned*>(qbase);
        for (int i = threadIdx.x; i < NB * 128; i += WAVES * 64) dst[i] = 0u;
        const int n = T8 * NB * 128;
        for (int i = threadIdx.x; i < n; i += WAVES * 64) dst[NB * 128 + i] = a.Qh[i];
    }
    // this lane's A operand: column (lane & 15) of each column block, or
    // zeros (the zero block covers every offset a zero lane reads)
    const bool live_a = ((lane & 31) < 16) == (lane < 32);
    const f16x8* qa = reinterpret_cast<const f16x8*>(qbase + (live_a ? NB * 512 + (lane & 15) * 16 : 0));
    const int qstep = live_a ? NB * 32 : 0;  // f16x8 per block of 8 time points
    const int h = lane >> 5;
    const float m1 = a.neg_one;
    const float invT = 1.0f / (float)a.T, invTm1 = 1.0f / ((float)a.T - 1.0f);
    constexpr float iq = 1.0f / HX_QSCALE;
    float* my = smax + wave * (ROUNDS * NP);
    if (lane == 0)
        for (int i = 0; i < ROUNDS * NP; i++) my[i] = -INFINITY;

    for (int tl = tile0; tl < tile1; tl++) {
        const int v = tl * TILE + lane;
        __syncthreads();
        {
            const float sc = a.wscale[(size_t)tl * TILE + (threadIdx.x & 63)];
            const int n = a.T * TILE;
            for (int i = threadIdx.x; i < n; i += WAVES * 64)
                tile[i] = a.w32[(size_t)(i >> 6) * a.Vpad + (size_t)tl * TILE + (i & 63)] * sc;
        }
        __syncthreads();
        const float a1 = (float)a.ar[v], a2 = (float)a.ar[(size_t)a.Vpad + v];
        const float a3 = (float)a.ar[2 * (size_t)a.Vpad + v], a4 = (float)a.ar[3 * (size_t)a.Vpad + v];
        const bool in_mask = v < a.Vc;
        for (int r = 0; r < ROUNDS; r++) {
            const int kbase = pgroup * PERMS_PER_BLOCK + r * (WAVES * NP) + wave * NP;
            if (kbase >= a.K) break;
            cuint_p prow[NP];
#pragma unroll
            for (int j = 0; j < NP; j++) {
                const int k = min(kbase + j, a.K - 1);
                prow[j] = (cuint_p)(a.perm + (size_t)k * a.Tp);
            }
            f32x16 acc[NP][NB];
            float hs[NP][4], s[NP], m[NP];
#pragma unroll
            for (int j = 0; j < NP; j++) {
#pragma unroll
                for (int nb = 0; nb < NB; nb++)
#pragma unroll
                    for (int e = 0; e < 16; e++) acc[j][nb][e] = 0.0f;
#pragma unroll
                for (int q = 0; q < 4; q++) hs[j][q] = 0.0f;
                s[j] = 0.0f;
                m[j] = 0.0f;
            }
            // time points 8c .. 8c + n - 1
            auto block8 = [&](int c, int n) {
                f16x8 qh[NB], ql[NB];
#pragma unroll
                for (int nb = 0; nb < NB; nb++) {
                    qh[nb] = qa[(size_t)c * qstep + nb * 16];
                    ql[nb] = qa[(size_t)c * qstep + (NB + nb) * 16];
                }
#pragma unroll
                for (int j = 0; j < NP; j++) {
                    const unsigned pk[4] = {prow[j][4 * c], prow[j][4 * c + 1], prow[j][4 * c + 2],
                                            prow[j][4 * c + 3]};
                    float xs[8];
#pragma unroll
                    for (int e = 0; e < 8; e++) {
                        if (e < n) {
                            const unsigned idx = (pk[e >> 1] >> ((e & 1) * 16)) & 0xFFFFu;
                            float x = tile[idx * TILE + lane];
                            x = fmaf(a4, hs[j][3], x);
                            x = fmaf(a3, hs[j][2], x);
                            x = fmaf(a2, hs[j][1], x);
                            x = fmaf(a1, hs[j][0], x);
                            hs[j][3] = hs[j][2];
                            hs[j][2] = hs[j][1];
                            hs[j][1] = hs[j][0];
                            hs[j][0] = x;
                            s[j] = fmaf(x, x, s[j]);
                            if (MEAN) m[j] += x;
                            xs[e] = x;
                        } else {
                            xs[e] = 0.0f;
                        }
                    }
                    f16x8 xh, xl;
                    split8_f16(xs, m1, xh, xl);
#pragma unroll
                    for (int nb = 0; nb < NB; nb++) {
                        acc[j][nb] = __builtin_amdgcn_mfma_f32_32x32x16_f16(qh[nb], xh, acc[j][nb], 0, 0, 0);
                        acc[j][nb] = __builtin_amdgcn_mfma_f32_32x32x16_f16(qh[nb], xl, acc[j][nb], 0, 0, 0);
                        acc[j][nb] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ql[nb], xh, acc[j][nb], 0, 0, 0);
                    }
                }
            };
            const int T8f = a.T >> 3;
            for (int c = 0; c < T8f; c++) block8(c, 8);
            if (T8f < T8) block8(T8f, a.T - 8 * T8f);
#pragma unroll
            for (int j = 0; j < NP; j++) {
                const int k = kbase + j;
                const bool valid = k < a.K;
                // Register r of lane half h holds column (r % 4) + 8 ((r % 8) / 4)
                // + 4 h of voxel (lane % 32) for r < 8 and of voxel
                // (lane % 32) + 32 for r >= 8.  Each lane sums what it holds
                // for both voxels; one exchange with lane ^ 32 per sum
                // completes its own voxel's.
                float sv[2] = {0.0f, 0.0f}, fv[2] = {0.0f, 0.0f}, uv[2] = {0.0f, 0.0f};
#pragma unroll
                for (int nb = 0; nb < NB; nb++)
#pragma unroll
                    for (int rr = 0; rr < 16; rr++) {
                        const float z = acc[j][nb][rr] * iq;
                        const int vv = rr >> 3;
                        // the column in lane half 0 (block nb)
                        const int c0 = 16 * nb + (rr & 3) + 8 * ((rr & 7) >> 2);
                        const int col = c0 + 4 * h;
                        sv[vv] = fmaf(z, z, sv[vv]);
                        if (col < a.nrows) fv[vv] = fmaf(z, z, fv[vv]);
                        if (MEAN) uv[vv] = fmaf(h ? a.u32[c0 + 4] : a.u32[c0], z, uv[vv]);
                    }
                const float ss = (h ? sv[1] : sv[0]) + __shfl_xor(h ? sv[0] : sv[1], 32);
                const float fn = (h ? fv[1] : fv[0]) + __shfl_xor(h ? fv[0] : fv[1], 32);
                const float uz = MEAN ? (h ? uv[1] : uv[0]) + __shfl_xor(h ? uv[0] : uv[1], 32) : 0.0f;
                // column 0 of voxel (lane % 32) + 32 sits in register 8 of lane half 0
                const float z8 = __shfl_xor(acc[j][0][8] * iq, 32);
                const float z0 = h ? z8 : acc[j][0][0] * iq;
                const float sse = s[j] - ss;
                const float me = MEAN ? m[j] - uz : 0.0f;
                const float var = (sse - me * me * invT) * invTm1;
                const float st = a.ftest ? fn / (var * (float)a.nrows) : z0 / sqrtf(var);
                const bool act = in_mask && valid;
                const bool bad = act && !(sse >= s[j] * 0.03125f && s[j] < 4.0e9f);
                if (__ballot(bad) && lane == 0) a.flags[k] = 1;
                if (a.stats && act) a.stats[(size_t)k * a.Vc + v] = st;
                if (a.bits) {
                    const unsigned long long bal = __ballot(act && st > a.cdt);
                    if (lane == 0 && valid) {
                        a.bits[(size_t)k * a.nwords + 2 * (size_t)tl] = (unsigned)bal;
                        a.bits[(size_t)k * a.nwords + 2 * (size_t)tl + 1] = (unsigned)(bal >> 32);
                    }
                }
                const float mx = wave_max(act ? st : -INFINITY);
                if (lane == 0) my[r * NP + j] = fmaxf(my[r * NP + j], mx);
            }
        }
    }
    if (lane == 0) {
        for (int r = 0; r < ROUNDS; r++)
            for (int j = 0; j < NP; j++) {
                const int k = pgroup * PERMS_PER_BLOCK + r * (WAVES * NP) + wave * NP + j;
                if (k < a.K) a.partial[(size_t)vchunk * a.K + k] = my[r * NP + j];
            }
    }
}

// Launch with `lds` bytes of dynamic LDS; the limit is raised on every call
// (it is a property of the kernel on the current device).
template <typename F>
hipError_t launch_lds(F* kern, unsigned grid, size_t lds, const PermArgs& a, hipStream_t s) {
    const hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                             hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(kern, dim3(grid), dim3(WAVES * 64), lds, s, a);
    return hipGetLastError();
}

hipError_t launch_perm_hx(PermArgs a, hipStream_t s) {
    if (a.K <= 0 || a.n_vchunks <= 0) return hipSuccess;
    const int nb = a.pq <= 16 ? 1 : 2;
    if (a.T > (nb == 1 ? HX_T_MAX : HX_T_MAX2) || a.pq > 32 || a.nrows > BPG_FL_MAX_FROWS || a.Tp < a.T)
        return hipErrorInvalidValue;
    const int n_blocks = a.n_vchunks * a.n_pgroups;
    const unsigned grid = (unsigned)((n_blocks + 7) / 8 * 8);
    const size_t lds = (size_t)a.T * TILE * sizeof(float) + (size_t)nb * 512 + (size_t)((a.T + 7) / 8) * nb * 512;
    if (nb == 1)
        return a.mean ? launch_lds(&k_fl_perm_hx<true, 1>, grid, lds, a, s)
                      : launch_lds(&k_fl_perm_hx<false, 1>, grid, lds, a, s);
    return a.mean ? launch_lds(&k_fl_perm_hx<true, 2>, grid, lds, a, s)
                  : launch_lds(&k_fl_perm_hx<false, 2>, grid, lds, a, s);
}

hipError_t launch_perm_bx(PermArgs a, hipStream_t s) {
    if (a.K <= 0 || a.n_vchunks <= 0) return hipSuccess;
    if (a.T > MX_T_MAX || a.pq > 16 || a.nrows > BPG_FL_MAX_FROWS) return hipErrorInvalidValue;
    const int n_blocks = a.n_vchunks * a.n_pgroups;
    const unsigned grid = (unsigned)((n_blocks + 7) / 8 * 8);
    const size_t lds = (size_t)a.T * TILE * sizeof(float);
    return a.mean ? launch_lds(&k_fl_perm_bx<true>, grid, lds, a, s) : launch_lds(&k_fl_perm_bx<false>, grid, lds, a, s);
}

hipError_t launch_perm_mx(PermArgs a, hipStream_t s) {
    if (a.K <= 0 || a.n_vchunks <= 0) return hipSuccess;
    if (a.T > MX_T_MAX || a.pq > 16 || a.nrows > BPG_FL_MAX_FROWS) return hipErrorInvalidValue;
    const int n_blocks = a.n_vchunks * a.n_pgroups;
    const unsigned grid = (unsigned)((n_blocks + 7) / 8 * 8);
    const size_t lds = (size_t)a.T * TILE * sizeof(float);
    return a.mean ? launch_lds(&k_fl_perm_mx<true>, grid, lds, a, s) : launch_lds(&k_fl_perm_mx<false>, grid, lds, a, s);
}

int fl_width(int p) {
    static const int widths[] = {1, 2, 3, 4, 5, 6, 7, 8, 10, 12, 14, 16, 20, 25};
    for (int w : widths)
        if (p <= w) return w;
    return -1;
}

template <int PQ, bool MEAN>
hipError_t launch_perm_m(const PermArgs& a, hipStream_t s) {
    const int n_blocks = a.n_vchunks * a.n_pgroups;
    const unsigned grid = (unsigned)((n_blocks + 7) / 8 * 8);
    if (a.T <= LDS_T_MAX) {
        return launch_lds(&k_fl_perm<PQ, true, MEAN>, grid, (size_t)a.T * TILE * sizeof(float), a, s);
    } else {
        hipLaunchKernelGGL((k_fl_perm<PQ, false, MEAN>), dim3(grid), dim3(WAVES * 64), 0, s, a);
    }
    return hipGetLastError();
}

template <int PQ>
hipError_t launch_perm_w(const PermArgs& a, hipStream_t s) {
    return a.mean ? launch_perm_m<PQ, true>(a, s) : launch_perm_m<PQ, false>(a, s);
}

hipError_t launch_perm(PermArgs a, hipStream_t s) {
    if (a.K <= 0 || a.n_vchunks <= 0) return hipSuccess;
    switch (a.pq) {
        case 1: return launch_perm_w<1>(a, s);
        case 2: return launch_perm_w<2>(a, s);
        case 3: return launch_perm_w<3>(a, s);
        case 4: return launch_perm_w<4>(a, s);
        case 5: return launch_perm_w<5>(a, s);
        case 6: return launch_perm_w<6>(a, s);
        case 7: return launch_perm_w<7>(a, s);
        case 8: return launch_perm_w<8>(a, s);
        case 10: return launch_perm_w<10>(a, s);
        case 12: return launch_perm_w<12>(a, s);
        case 14: return launch_perm_w<14>(a, s);
        case 16: return launch_perm_w<16>(a, s);
        case 20: return launch_perm_w<20>(a, s);
        case 25: return launch_perm_w<25>(a, s);
    }
    return hipErrorInvalidValue;
}

}  // namespace fl
}  // namespace bpg
