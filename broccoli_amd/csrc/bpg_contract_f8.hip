// bpg_contract_f8.hip -- the split-FP8 form of the permuted-GLM contraction on
// gfx950's block-scaled FP8 matrix cores (v_mfma_scale_f32_32x32x64_f8f6f4,
// both operands OCP FP8 E4M3, one E8M0 scale per 32 subjects).
//
// Every operand element x is the sum of T e4m3 terms, each on its own
// power-of-two scale per block of 32 subjects:
//   x = sum_t 2^{e_t} c_t,  c_t = e4m3(r_t / 2^{e_t}),  r_{t+1} = r_t - 2^{e_t} c_t,
// with e_t the smallest exponent that brings the block's largest |r_t| to
// at most 448 (the e4m3 maximum).  Each term holds 4 significant bits, so
// the residual after T terms is below 2^-4T of the block maximum.  The
// contraction takes the products of operand term i and data term j with
// i + j < T (the dropped ones are below 2^-4T), T(T+1)/2 MFMAs per k-step of
// 64 subjects; exact +-1 operands (sign flips) need one operand term, so T
// products.
//
// Lane layout of the MFMA, probed on the device: for a 32x32x64 step, lane l
// supplies 32 bytes of row (A) or column (B) l % 32; bytes 0-15 hold
// subjects 16 (l / 32) + 0..15 of the first block of 32, bytes 16-31 the
// same of the second block; the scale of (row or column c, block h) comes
// from lane c + 32 h.  The accumulator layout is that of every 32x32 MFMA,
// so the operand rows are ordered as for the split-BF16 kernel and its
// epilogue applies unchanged.
//
// The same probe showed that the FP8 dot-product hardware does not sum
// exactly: on random e4m3 operands with unit scales one instruction's
// result was off by up to 2^-12.6 of its largest product (3e-5 of the sum of
// |products|), far above FP32 rounding.  That error floor, not the number
// of terms, bounds this form's accuracy.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>

#include "bpg_contract_bf16.h"
#include "bpg_kernels.h"

namespace bpg {

typedef int i32x8 __attribute__((ext_vector_type(8)));

// OCP FP8 E4M3 (bias 7, no infinities, 0x7F / 0xFF NaN, largest 448):
// round to nearest even, saturating at +-448.  x is an FP64 value.
__device__ __forceinline__ unsigned e4m3_encode(double x) {
    const unsigned sign = x < 0.0 ? 0x80u : 0u;
    double a = fabs(x);
    if (!(a < 448.0)) return sign | 0x7Eu;
    if (a < 0.015625) {  // below 2^-6: subnormal, steps of 2^-9
        const double m = rint(a * 512.0);
        return sign | (unsigned)m;  // m == 8 is the smallest normal (0x08)
    }
    int e;
    const double f = frexp(a, &e);  // a = f 2^e, f in [0.5, 1)
    double m = rint((f * 2.0 - 1.0) * 8.0);
    int ex = e - 1;
    if (m >= 8.0) {
        m = 0.0;
        ex += 1;
    }
    unsigned code = ((unsigned)(ex + 7) << 3) | (unsigned)m;
    if (code > 0x7Eu) code = 0x7Eu;
    return sign | code;
}

__device__ __forceinline__ double e4m3_decode(unsigned c) {
    const double s = (c & 0x80u) ? -1.0 : 1.0;
    const unsigned e = (c >> 3) & 15u, m = c & 7u;
    return s * (e == 0 ? (double)m * 0.001953125 : (1.0 + (double)m * 0.125) * ldexp(1.0, (int)e - 7));
}

// Splits 32 values (one block) into T terms.  codes[t][0..31], scale bytes
// (E8M0: exponent + 127) per term.  An all-zero residual gets scale 2^0
// and zero codes.
template <int T>
__device__ __forceinline__ void split_block(double (&r)[32], unsigned char (&codes)[T][32], unsigned (&sc)[T]) {
#pragma unroll
    for (int t = 0; t < T; t++) {
        double m = 0.0;
#pragma unroll
        for (int s = 0; s < 32; s++) m = fmax(m, fabs(r[s]));
        int e = 0;
        if (m > 0.0) {
            int ex;
            frexp(m / 448.0, &ex);  // m / 448 = f 2^ex, f in [0.5, 1): 2^ex >= m / 448
            e = max(-127, min(127, ex));
        }
        sc[t] = (unsigned)(e + 127);
        const double inv = ldexp(1.0, -e), scl = ldexp(1.0, e);
#pragma unroll
        for (int s = 0; s < 32; s++) {
            const unsigned c = e4m3_encode(r[s] * inv);
            codes[t][s] = (unsigned char)c;
            r[s] -= e4m3_decode(c) * scl;
        }
    }
}

// Writes one block's terms: subjects 0-15 of the block go to lane c, 16-31
// to lane 32 + c, at byte offset 16 h of each lane's 32 bytes; the scale
// bytes to lane c + 32 h's scale word, byte t.
template <int T>
__device__ __forceinline__ void store_block(uint4* frag, unsigned* scw, size_t term_stride, int c, int h,
                                            const unsigned char (&codes)[T][32], const unsigned (&sc)[T]) {
#pragma unroll
    for (int t = 0; t < T; t++) {
#pragma unroll
        for (int half = 0; half < 2; half++) {
            unsigned w[4];
#pragma unroll
            for (int q = 0; q < 4; q++)
                w[q] = (unsigned)codes[t][16 * half + 4 * q] | ((unsigned)codes[t][16 * half + 4 * q + 1] << 8) |
                       ((unsigned)codes[t][16 * half + 4 * q + 2] << 16) |
                       ((unsigned)codes[t][16 * half + 4 * q + 3] << 24);
            frag[(size_t)t * term_stride + (size_t)(32 * half + c) * 2 + h] = make_uint4(w[0], w[1], w[2], w[3]);
        }
    }
    unsigned word = 0;
#pragma unroll
    for (int t = 0; t < T; t++) word |= sc[t] << (8 * t);
    scw[32 * h + c] = word;
}

// Data fragments: thread = (voxel, k-step, block).  The residualised values
// y*_s dscale_s come from the FP32 operand rows Y32 [NS][Vpad] that
// k_transform wrote for this form (exact in FP64 arithmetic from there).
template <int T>
__global__ __launch_bounds__(256) void k_data_f8(const float* Y32, int NS, int N, int Vpad, int KS, unsigned* Yf8,
                                                 unsigned* Ysc) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t total = (size_t)Vpad * KS * 2;
    if (t >= total) return;
    const int v = (int)(t % Vpad);
    const int h = (int)((t / Vpad) % 2);
    const int ks = (int)(t / Vpad / 2);
    double r[32];
#pragma unroll
    for (int s = 0; s < 32; s++) {
        const int sub = 64 * ks + 32 * h + s;
        r[s] = sub < N ? (double)Y32[(size_t)min(sub, NS - 1) * Vpad + v] : 0.0;
    }
    unsigned char codes[T][32];
    unsigned sc[T];
    split_block<T>(r, codes, sc);
    const int vt = v >> 5, c = v & 31;
    uint4* frag = reinterpret_cast<uint4*>(Yf8) + ((size_t)(vt * KS + ks) * T) * 128;
    store_block<T>(frag, Ysc + (size_t)(vt * KS + ks) * 64, 128, c, h, codes, sc);
}

// Operand fragments: thread = (subtile, k-step, row, block); rows ordered as
// in k_operand_bf16.
template <int T>
__global__ __launch_bounds__(256) void k_operand_f8(OperandArgs a, int KS, int TG, unsigned* Gf8, unsigned* Gsc) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t total = (size_t)a.n_sub * KS * 64;
    if (t >= total) return;
    const int i = (int)(t & 31);
    const int h = (int)((t >> 5) & 1);
    const size_t rest = t >> 6;
    const int ks = (int)(rest % KS);
    const int st = (int)(rest / KS);
    const int j = 4 * (i >> 3) + (i & 3);
    const int hh = (i >> 2) & 1;
    const int per = 16 / a.PP;
    const int k = st * (2 * per) + hh * per + j / a.PP;
    const int reg = j % a.PP;
    const bool live = j < per * a.PP && k < a.K && reg < a.pfit;
    const size_t row = (size_t)min(k, a.K - 1) * a.N;
    const size_t col = (size_t)min(reg, a.pfit - 1) * a.N;
    double r[32];
#pragma unroll
    for (int s = 0; s < 32; s++) {
        const int sub = 64 * ks + 32 * h + s;
        const int sc = min(sub, a.N - 1);
        double val = a.sign ? (double)a.sign[row + sc] : a.Qt[(size_t)a.perm[row + sc] + col];
        r[s] = (live && sub < a.N) ? val : 0.0;
    }
    unsigned char codes[T][32];
    unsigned sc[T];
    split_block<T>(r, codes, sc);
    // operand terms beyond TG (exact operands) are zero by construction
    (void)TG;
    uint4* frag = reinterpret_cast<uint4*>(Gf8) + ((size_t)(st * KS + ks) * T) * 128;
    store_block<T>(frag, Gsc + (size_t)(st * KS + ks) * 64, 128, i, h, codes, sc);
}

// acc[(i + j) & 1] += (operand term i) x (data term j) for i + j < T: two
// accumulators, so consecutive MFMAs do not wait on each other.  The term
// indices are the scale bytes' op_sel, which must be constants.
template <int GI, int YJ, int T, int TG>
__device__ __forceinline__ void f8_products(f32x16 (&acc)[2], const i32x8 (&at)[TG], const i32x8 (&bt)[T], int as,
                                            int bs) {
    if constexpr (GI < TG) {
        if constexpr (YJ < T) {
            if constexpr (GI + YJ < T)
                acc[(GI + YJ) & 1] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(at[GI], bt[YJ], acc[(GI + YJ) & 1],
                                                                                   0, 0, GI, as, YJ, bs);
            f8_products<GI, YJ + 1, T, TG>(acc, at, bt, as, bs);
        } else {
            f8_products<GI + 1, 0, T, TG>(acc, at, bt, as, bs);
        }
    }
}

// Contraction.  Block = 4 waves x R operand subtiles; each wave's subtiles
// (T terms x KS k-steps x 32 bytes per lane, plus a scale word per k-step)
// are staged in LDS (a per-lane area), the data fragments stream from L2
// through L1 (shared by the block's waves), the next k-step's prefetched.
template <int T, int TG, int R, int PP, bool FT, bool OUT>
__global__ __launch_bounds__(256) void k_contract_f8(ContractArgs a, Bf16Launch L, int KS) {
    constexpr int PER = 16 / PP;
    extern __shared__ uint4 sA8[];
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int lb = xcd_logical_block(L.per_xcd);
    if (lb >= L.n_mgroups * L.n_vchunks) return;
    const int mg = lb % L.n_mgroups;
    const int vc = lb / L.n_mgroups;
    const int sub0 = (mg * 4 + wave) * R;
    const uint4* G = reinterpret_cast<const uint4*>(a.Gf8);
    const uint4* Y = reinterpret_cast<const uint4*>(a.Yf8);
    int srow[R];
    bool live[R];
#pragma unroll
    for (int r = 0; r < R; r++) {
        live[r] = sub0 + r < L.n_sub;
        srow[r] = live[r] ? sub0 + r : 0;
    }
    // per wave: [r][ks][TG terms][2 uint4] fragments, then [r][ks] scale words
    uint4* myA = sA8 + (size_t)wave * R * KS * (TG * 128 + 64);
    unsigned* myS = reinterpret_cast<unsigned*>(myA + (size_t)R * KS * TG * 2 * 64);
    for (int ks = 0; ks < KS; ks++)
#pragma unroll
        for (int r = 0; r < R; r++) {
#pragma unroll
            for (int t = 0; t < TG; t++)
#pragma unroll
                for (int q = 0; q < 2; q++)
                    myA[(((size_t)r * KS + ks) * TG + t) * 128 + lane * 2 + q] =
                        G[((size_t)(srow[r] * KS + ks) * T + t) * 128 + lane * 2 + q];
            myS[((size_t)r * KS + ks) * 64 + lane] = a.Gsc[(size_t)(srow[r] * KS + ks) * 64 + lane];
        }
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
        f32x16 acc[R][2];
#pragma unroll
        for (int r = 0; r < R; r++)
#pragma unroll
            for (int e = 0; e < 16; e++) acc[r][0][e] = acc[r][1][e] = 0.0f;
        for (int ks = 0; ks < KS; ks++) {
            i32x8 bt[T];
            const size_t yb = ((size_t)(vt * KS + ks) * T) * 128 + lane * 2;
#pragma unroll
            for (int t = 0; t < T; t++) {
                const uint4 lo = Y[yb + (size_t)t * 128], hi = Y[yb + (size_t)t * 128 + 1];
                bt[t] = i32x8{(int)lo.x, (int)lo.y, (int)lo.z, (int)lo.w, (int)hi.x, (int)hi.y, (int)hi.z, (int)hi.w};
            }
            const int bs = (int)a.Ysc[(size_t)(vt * KS + ks) * 64 + lane];
#pragma unroll
            for (int r = 0; r < R; r++) {
                i32x8 at[TG];
#pragma unroll
                for (int t = 0; t < TG; t++) {
                    const uint4 lo = myA[(((size_t)r * KS + ks) * TG + t) * 128 + lane * 2];
                    const uint4 hi = myA[(((size_t)r * KS + ks) * TG + t) * 128 + lane * 2 + 1];
                    at[t] = i32x8{(int)lo.x, (int)lo.y, (int)lo.z, (int)lo.w, (int)hi.x, (int)hi.y, (int)hi.z, (int)hi.w};
                }
                const int as = (int)myS[((size_t)r * KS + ks) * 64 + lane];
                f8_products<0, 0, T, TG>(acc[r], at, bt, as, bs);
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

constexpr int F8_TERMS = BPG_F8_TERMS;

template <int TG, int R, int PP, bool FT>
hipError_t run_f8(const ContractArgs& a, const Bf16Launch& L, int KS, hipStream_t s, int* occ) {
    const bool out = occ ? *occ != 0 : (a.stats || a.bits);
    void (*kern)(ContractArgs, Bf16Launch, int) =
        out ? k_contract_f8<F8_TERMS, TG, R, PP, FT, true> : k_contract_f8<F8_TERMS, TG, R, PP, FT, false>;
    const size_t lds = f8_lds_bytes(KS, TG == 1);
    if (occ) return hipOccupancyMaxActiveBlocksPerMultiprocessor(occ, reinterpret_cast<const void*>(kern), 256, lds);
    hipLaunchKernelGGL(kern, dim3(L.per_xcd * 8), dim3(256), lds, s, a, L, KS);
    return hipGetLastError();
}

template <int PP, bool FT>
hipError_t run_f8_pp(const ContractArgs& a, const Bf16Launch& L, int KS, bool sign, hipStream_t s, int* occ) {
    const int R = f8_rows_per_wave(KS);
    if (sign) {
        if constexpr (PP == 1 && !FT)
            return R == 2 ? run_f8<1, 2, 1, false>(a, L, KS, s, occ) : run_f8<1, 1, 1, false>(a, L, KS, s, occ);
        return hipErrorInvalidValue;
    }
    return R == 2 ? run_f8<F8_TERMS, 2, PP, FT>(a, L, KS, s, occ) : run_f8<F8_TERMS, 1, PP, FT>(a, L, KS, s, occ);
}

template <bool FT>
hipError_t dispatch_f8(const ContractArgs& a, const Bf16Launch& L, int PP, int KS, bool sign, hipStream_t s, int* occ) {
    switch (PP) {
        case 1: return run_f8_pp<1, FT>(a, L, KS, sign, s, occ);
        case 2: return run_f8_pp<2, FT>(a, L, KS, sign, s, occ);
        case 3: return run_f8_pp<3, FT>(a, L, KS, sign, s, occ);
        case 4: return run_f8_pp<4, FT>(a, L, KS, sign, s, occ);
        case 5: return run_f8_pp<5, FT>(a, L, KS, sign, s, occ);
        case 8: return run_f8_pp<8, FT>(a, L, KS, sign, s, occ);
        case 16: return run_f8_pp<16, FT>(a, L, KS, sign, s, occ);
    }
    return hipErrorInvalidValue;
}

}  // namespace

hipError_t launch_contract_f8(const ContractArgs& a, const Bf16Launch& L, int PP, int KS, bool sign, hipStream_t s,
                              int* occ) {
    return a.ftest ? dispatch_f8<true>(a, L, PP, KS, sign, s, occ) : dispatch_f8<false>(a, L, PP, KS, sign, s, occ);
}

int contract_f8_blocks_per_cu(int PP, int KS, bool sign, bool ftest, bool out) {
    ContractArgs a{};
    a.ftest = ftest;
    int occ = out ? 1 : 0;
    if (launch_contract_f8(a, Bf16Launch{}, PP, KS, sign, nullptr, &occ) != hipSuccess || occ < 1) return 1;
    return occ;
}

hipError_t launch_data_f8(const float* Y32, int NS, int N, int Vpad, int KS, unsigned* Yf8, unsigned* Ysc,
                          hipStream_t s) {
    const size_t total = (size_t)Vpad * KS * 2;
    hipLaunchKernelGGL(k_data_f8<F8_TERMS>, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, Y32, NS, N,
                       Vpad, KS, Yf8, Ysc);
    return hipGetLastError();
}

hipError_t launch_operand_f8(const OperandArgs& a, int KS, bool sign, unsigned* Gf8, unsigned* Gsc, hipStream_t s) {
    const size_t total = (size_t)a.n_sub * KS * 64;
    hipLaunchKernelGGL(k_operand_f8<F8_TERMS>, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, a, KS,
                       sign ? 1 : F8_TERMS, Gf8, Gsc);
    return hipGetLastError();
}

}  // namespace bpg
