// bpg_contract_bf16.h -- split-BF16 MFMA contraction with the fused
// statistic and max-T epilogue (instantiated per KK in
// bpg_contract_bf16_*.hip so the instantiations compile in parallel).
//
// Operand tile (A, per wave, resident): R subtiles of 32 permutation rows x
// KK k-steps of 16 subjects, hi and lo bf16 parts, in the fragment layout of
// v_mfma_f32_32x32x16_bf16 (lane l: row l%32, subjects 16*kk + 8*(l/32) +
// 0..7).  Data tile (B, streamed): 32 voxels x KK k-steps, hi and lo, same
// layout with voxels as columns.  Accumulator: lane l holds column l%32 and
// rows 8*b + 4*(l/32) + a at slot 4*b + a; the operand rows are ordered so
// slots j = P*PP .. P*PP+PP-1 of a lane are the PP regressors of one
// permutation, so each lane finishes PER = 16/PP whole permutations of one
// voxel (PP = 3 or 5 leave slot 15 as a zero row).
//
// Epilogue per (permutation, voxel), with z = Qt' P' y* (orthonormal basis,
// first axes spanning the contrast):
//   SSE = |y*|^2 - |z|^2           (|y*|^2 held as an FP64-accurate float pair)
//   t: u = z_0 / sqrt(SSE)          F: u = (z_0^2 + .. + z_{r-1}^2) / SSE
// The running max keeps u; the positive constant sqrt(N-p) (t) or
// (N-p)/r (F) is applied once after the max, which is exact: rounding is
// monotone, so max(fl(c*u)) = fl(c*max(u)).
//
// Two schedules: sequential (each subtile's MFMA chain, then its epilogue)
// and software-pipelined (the chain of subtile i issued together with the
// epilogue of subtile i-1, which may belong to the previous voxel tile,
// interleaved one MFMA to a few VALU instructions by sched_group_barrier).
#pragma once

#include "bpg_device.h"

namespace bpg {

// Views of the fragment and norm arrays in the global address space: the
// kernels read them through pointers that may come from memory (the
// persistent form's study table), which the compiler would otherwise treat
// as generic and access with flat instructions.
typedef const bf16x8 __attribute__((address_space(1))) gbf16x8;
typedef const double __attribute__((address_space(1))) gdouble;
typedef float __attribute__((address_space(1))) gfloat;
typedef unsigned char __attribute__((address_space(1))) gbyte;
typedef unsigned long long __attribute__((address_space(1))) gu64;


// v_writelane_b32: old with lane `lane` replaced by the uniform `value` (the
// LLVM intrinsic; this compiler exposes no builtin for it)
__device__ int lane_write(int value, int lane, int old) __asm("llvm.amdgcn.writelane.i32");

// The data tile of voxel tile vt: NPY fragments per k-step (hi, lo; or hi,
// mid, lo for exact +-1 operands at small N), bm unused with NPY = 2.
template <int KK, int NPY>
__device__ __forceinline__ void load_tile(const gbf16x8* Y, const gdouble* yy, int vt, int lane,
                                          bf16x8 (&bh)[KK], bf16x8 (&bm)[NPY == 3 ? KK : 1], bf16x8 (&bl)[KK],
                                          double& yv) {
#pragma unroll
    for (int kk = 0; kk < KK; kk++) {
        const size_t base = (((size_t)vt * KK + kk) * NPY) * 64 + lane;
        bh[kk] = Y[base];
        if constexpr (NPY == 3) bm[kk] = Y[base + 64];
        bl[kk] = Y[base + (NPY - 1) * 64];
    }
    yv = yy[vt * 32 + (lane & 31)];
}

// z for one subtile and one voxel tile: lo*hi + hi*lo + hi*hi; for exact
// +-1 operands the lo*hi product vanishes and, with NPY = 3, hi*mid joins
// (every product then exact: the error is the FP32 accumulation's)
template <int KK, bool SIGN, int NPY>
__device__ __forceinline__ f32x16 chain(const bf16x8 (&ah)[KK], const bf16x8 (&al)[KK], const bf16x8 (&bh)[KK],
                                        const bf16x8 (&bm)[NPY == 3 ? KK : 1], const bf16x8 (&bl)[KK]) {
    f32x16 acc;
#pragma unroll
    for (int e = 0; e < 16; e++) acc[e] = 0.0f;
#pragma unroll
    for (int kk = 0; kk < KK; kk++) {
        if (!SIGN) acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al[kk], bh[kk], acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[kk], bl[kk], acc, 0, 0, 0);
        if constexpr (NPY == 3) acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[kk], bm[kk], acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[kk], bh[kk], acc, 0, 0, 0);
    }
    return acc;
}

// The voxel tiles of one work item, first..last.
struct TileSpan {
    int first, last;
};

// The rolling pipelined schedule: pipelined general operands (two data
// fragments) at 5..8 k-steps.  Above 8 the ping-pong pipelined schedule
// runs, whose accumulators need the AGPRs of a single wave per SIMD.
constexpr bool rolling_pipelined(int KK, bool sign, bool pipe) {
    return pipe && bf16_y_pieces(KK, sign) == 2 && KK >= BPG_ROLL_MIN_KK && KK <= 8;
}

// Statistic and running max of global subtile st over voxel tile vt; vt < 0
// marks the empty item that opens the pipelined schedule (its NaN norm is
// dropped by fmaxf and it writes nothing).  OUT = 0: the max only, no
// branches.  OUT = 1: also the statistic of every voxel (stats) and/or the
// supra-threshold bitmask (bits), one word per (permutation, tile).
// OUT = 2 (cluster extent, bits only, no max): lane (vt mod 32) of each half
// keeps the bitmask word of its permutation in wb, and every 32 tiles (and
// at the item's last tile) the half-wave stores its 32 words as one 128-byte
// run of the permutation's row.
// OUT = 3 (max only, with a sample list): an ill-conditioned sample's
// statistic becomes NaN, which the running max drops, and the sample is
// appended to the list (append_samples) for its FP64 recomputation.
//
// Appends the samples of this tile whose bit P is set in excl: one atomic
// add per wave reserves the entries of all its samples in the wave's
// segment, and each lane writes its own in (slot P, lane) order.  Past the
// segment's capacity, the permutation is flagged for a whole repeat instead.
// (a call, not inlined: the rare path then holds no registers in the
// contraction loop around it)
template <int PER>
__device__ __attribute__((noinline)) void append_samples(unsigned long long* samples, unsigned long long* n_samples,
                                                         unsigned long long cap, unsigned char* flags,
                                                         unsigned* unsafe, int K, unsigned excl, int st, int vt,
                                                         int lane) {
    const int hl = lane >> 5;
    const unsigned v = (unsigned)(vt * 32 + (lane & 31));
    const int seg = (int)((blockIdx.x * 4u + (threadIdx.x >> 6)) & (unsigned)(BPG_SAMPLE_SEGS - 1));
    unsigned long long m[PER];
    unsigned total = 0u;
#pragma unroll
    for (int P = 0; P < PER; P++) {
        m[P] = __ballot(((excl >> P) & 1u) != 0u);
        total += (unsigned)__popcll(m[P]);
    }
    const int leader = __builtin_ctzll(__ballot(excl != 0u));
    unsigned long long base = 0ull;
    if (lane == leader) base = atomicAdd(n_samples + seg, (unsigned long long)total);
    base = __shfl(base, leader, 64);
    gu64* out = (gu64*)(samples + (size_t)seg * cap);
    unsigned long long off = 0ull;
#pragma unroll
    for (int P = 0; P < PER; P++) {
        const unsigned rank =
            __builtin_amdgcn_mbcnt_hi((unsigned)(m[P] >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m[P], 0u));
        const int k = st * (2 * PER) + hl * PER + P;
        if (((excl >> P) & 1u) != 0u && k < K) {
            const unsigned long long idx = base + off + rank;
            if (idx < cap) {
                out[idx] = ((unsigned long long)k << 32) | v;
            } else {
                ((gbyte*)flags)[k] = 1;
                atomicAdd(unsafe, 1u);
            }
        }
        off += (unsigned long long)__popcll(m[P]);
    }
}

template <int PP, bool FT, int OUT>
__device__ __forceinline__ void epi(const ContractArgs& a, const f32x16& acc, float yhi, float ylo, int vt,
                                    int lane, int st, float (&rm)[16 / PP], bool& bad, unsigned (&wb)[16 / PP],
                                    const TileSpan& sp) {
    constexpr int PER = 16 / PP;
    const float tyy = a.tau * yhi;  // NaN on padding: never flagged
    float u[PER];
    // F, bits only (OUT = 2): the statistic's numerator num f_scale and SSE,
    // compared with the thresholds as products (F > c <=> num f_scale > c SSE
    // for SSE > 0; an SSE <= 0 is ill-conditioned and its permutation
    // repeated), which spares the reciprocal
    constexpr bool FPROD = FT && OUT == 2 && BPG_F_BITS_PRODUCT;
    float fnum[PER], fsse[PER];
    unsigned excl = 0u;  // OUT = 3: bit P for an ill-conditioned sample P
#pragma unroll
    for (int P = 0; P < PER; P++) {
        const float z0 = acc[P * PP];
        float sse;
        float num = 0.0f;
        if (FT) {
            float ss = z0 * z0;
            num = ss;
#pragma unroll
            for (int g = 1; g < PP; g++) {
                const float z = acc[P * PP + g];
                ss = fmaf(z, z, ss);
                num = (g < a.r) ? ss : num;
            }
            sse = (yhi - ss) + ylo;
        } else {
            // |y*|^2 - |z|^2 as one FMA chain from |y*|^2 (one rounding per
            // step, one instruction fewer than squaring and subtracting)
            float d = fmaf(-z0, z0, yhi);
#pragma unroll
            for (int g = 1; g < PP; g++) {
                const float z = acc[P * PP + g];
                d = fmaf(-z, z, d);
            }
            sse = d + ylo;
        }
        const bool ill = sse < tyy;
        if constexpr (FPROD) {
            fnum[P] = num * a.f_scale;
            fsse[P] = sse;
        } else {
            u[P] = FT ? num * __builtin_amdgcn_rcpf(sse) : z0 * __builtin_amdgcn_rsqf(sse);
        }
        if constexpr (OUT == 3) {
            u[P] = ill ? __builtin_nanf("") : u[P];
            excl |= (unsigned)ill << P;
        } else {
            bad |= ill;
        }
        // max(u, rm) as med3(u, rm, +inf): one v_med3_f32, where fmaxf costs
        // a second instruction to quiet the loop-carried rm; a NaN u (padding
        // voxel) yields rm, as fmaxf does (checked on the device)
        if constexpr (OUT != 2) rm[P] = __builtin_amdgcn_fmed3f(u[P], rm[P], a.pinf);
    }
    if constexpr (OUT == 3) {
        // rare: a uniform branch around the list appends
        if (__ballot(excl != 0u) != 0ull)
            append_samples<PER>(a.samples, a.n_samples, a.sample_cap, a.flags, a.unsafe, a.K, excl, st, vt, lane);
    }
    if constexpr (OUT == 1) {
        const int hl = lane >> 5;
        const int v = vt * 32 + (lane & 31);
        const bool valid = vt >= 0 && v < a.Vc;
        const float scale = FT ? a.f_scale : a.sqrt_dof;
        const int kbase = st * (2 * PER) + hl * PER;
#pragma unroll
        for (int P = 0; P < PER; P++) {
            const float stv = u[P] * scale;
            const int k = kbase + P;
            if (a.stats && valid && k < a.K) a.stats[(size_t)k * a.Vc + v] = stv;
            if (a.bits) {
                const unsigned long long bal = __ballot(valid && stv > a.cdt);
                if ((lane & 31) == 0 && k < a.K && vt >= 0)
                    a.bits[(size_t)k * a.nwords + vt] = hl ? (unsigned)(bal >> 32) : (unsigned)bal;
                // a statistic inside the error band of the threshold (rare):
                // the map's largest cluster may differ from an exact one's
                if (a.ties && valid && k < a.K && stv > a.cdt_lo && stv <= a.cdt_hi) a.ties[k] = 1;
            }
        }
    }
    if constexpr (OUT == 2) {
        const int v = vt * 32 + (lane & 31);
        const unsigned long long vmask = __ballot(vt >= 0 && v < a.Vc);
        const float scale = FT ? a.f_scale : a.sqrt_dof;
        const int slot = vt & 31;
        unsigned long long tmask = 0ull;
        float stv[PER];
        bool tie[PER];
#pragma unroll
        for (int P = 0; P < PER; P++) {
            bool over;
            if constexpr (FPROD) {
                over = fnum[P] > a.cdt * fsse[P];
                tie[P] = fnum[P] > a.cdt_lo * fsse[P] && fnum[P] <= a.cdt_hi * fsse[P];
            } else {
                stv[P] = u[P] * scale;
                over = stv[P] > a.cdt;
                tie[P] = stv[P] > a.cdt_lo && stv[P] <= a.cdt_hi;
            }
            // the two words of the ballot into lanes slot and 32 + slot
            const unsigned long long bal = __ballot(over) & vmask;
            wb[P] = (unsigned)lane_write((int)(unsigned)bal, slot, (int)wb[P]);
            wb[P] = (unsigned)lane_write((int)(unsigned)(bal >> 32), 32 + slot, (int)wb[P]);
            tmask |= __ballot(tie[P]);
        }
        if (tmask & vmask) {
            // a statistic inside the error band of the threshold (rare):
            // the map's largest cluster may differ from an exact one's
            const int hl = lane >> 5;
            const int kbase = st * (2 * PER) + hl * PER;
#pragma unroll
            for (int P = 0; P < PER; P++) {
                const unsigned long long tb = __ballot(tie[P]) & vmask;
                const int k = kbase + P;
                if ((lane & 31) == 0 && k < a.K && (hl ? (tb >> 32) : (tb & 0xFFFFFFFFull)) != 0ull) a.ties[k] = 1;
            }
        }
        if (vt >= 0 && ((vt & 31) == 31 || vt == sp.last)) {
            const int t = (vt & ~31) + (lane & 31);
            const int kbase = st * (2 * PER) + (lane >> 5) * PER;
            if (t >= sp.first && t <= vt) {
#pragma unroll
                for (int P = 0; P < PER; P++) {
                    const int k = kbase + P;
                    if (k < a.K) a.bits[(size_t)k * a.nwords + t] = wb[P];
                }
            }
        }
    }
}

// pipeline state: the accumulator awaiting its epilogue and its tile's norms
struct Pending {
    f32x16 acc;
    float yhi, ylo;
    int vt;
};

template <int KK, int R, int PP, bool SIGN, bool FT, int OUT, bool PIPE, int NPY>
__device__ __forceinline__ void contract_tile(const ContractArgs& a, const bf16x8 (&ahi)[R][KK],
                                              const bf16x8 (&alo)[R][KK], const bf16x8 (&bh)[KK],
                                              const bf16x8 (&bm)[NPY == 3 ? KK : 1], const bf16x8 (&bl)[KK],
                                              double yv, int vt, int lane, int sub0,
                                              float (&runmax)[R][16 / PP], Pending& pd,
                                              bool (&bad)[PIPE ? 1 : R], unsigned (&wbuf)[R][16 / PP],
                                              const TileSpan& sp) {
    const float yhi = (float)yv;  // NaN for padding voxels
    const float ylo = (float)(yv - (double)yhi);
    if constexpr (!PIPE) {
#pragma unroll
        for (int r = 0; r < R; r++) {
            const f32x16 acc = chain<KK, SIGN, NPY>(ahi[r], alo[r], bh, bm, bl);
            epi<PP, FT, OUT>(a, acc, yhi, ylo, vt, lane, sub0 + r, runmax[r], bad[PIPE ? 0 : r], wbuf[r], sp);
        }
    } else {
        constexpr int NM = KK * (SIGN ? NPY : 3);
        constexpr int PER = 16 / PP;
        constexpr int NV = (PER * (PP + 5) + NM - 1) / NM;
#pragma unroll
        for (int r = 0; r < R; r++) {
            const f32x16 acc = chain<KK, SIGN, NPY>(ahi[r], alo[r], bh, bm, bl);
            if (r == 0)
                epi<PP, FT, OUT>(a, pd.acc, pd.yhi, pd.ylo, pd.vt, lane, sub0 + R - 1, runmax[R - 1], bad[0],
                                 wbuf[R - 1], sp);
            else
                epi<PP, FT, OUT>(a, pd.acc, yhi, ylo, vt, lane, sub0 + r - 1, runmax[r - 1], bad[0], wbuf[r - 1], sp);
#pragma unroll
            for (int i = 0; i < NM; i++) {
                __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);   // one MFMA
                __builtin_amdgcn_sched_group_barrier(0x002, NV, 0);  // NV VALU
            }
            pd.acc = acc;
        }
        pd.yhi = yhi;
        pd.ylo = ylo;
        pd.vt = vt;
    }
}

// Block = 4 waves, each with R resident operand subtiles; the block walks a
// chunk of voxel tiles, loading tile t+1 (ping-pong registers, straight from
// L2; the 4 waves share the lines through the CU's L1) while tile t is
// multiplied.  At the end, a 32-lane butterfly per permutation and one
// partial maximum per (voxel chunk, permutation).
// One work item: permutation group mg (4 waves x R subtiles) over voxel
// chunk vc.
template <int KK, int R, int PP, bool SIGN, bool FT, int OUT, bool PIPE>
__device__ __forceinline__ void contract_item(const ContractArgs& a, const Bf16Launch& L, int mg, int vc) {
    constexpr int PER = 16 / PP;
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int sub0 = (mg * 4 + wave) * R;

    const gbf16x8* G = (const gbf16x8*)a.Gb;
    const gbf16x8* Y = (const gbf16x8*)a.Yb;
    const gdouble* YY = (const gdouble*)a.yy;

    bf16x8 ahi[R][KK], alo[R][KK];
#pragma unroll
    for (int r = 0; r < R; r++) {
        const int st = sub0 + r;
#pragma unroll
        for (int kk = 0; kk < KK; kk++) {
            bf16x8 h = {}, l = {};
            if (st < L.n_sub) {
                const size_t base = (((size_t)st * KK + kk) * 2) * 64 + lane;
                h = G[base];
                if (!SIGN) l = G[base + 64];
            }
            ahi[r][kk] = h;
            alo[r][kk] = l;
        }
    }
    float runmax[R][PER];
#pragma unroll
    for (int r = 0; r < R; r++)
#pragma unroll
        for (int P = 0; P < PER; P++) runmax[r][P] = -INFINITY;

    Pending pd;
#pragma unroll
    for (int e = 0; e < 16; e++) pd.acc[e] = 0.0f;
    pd.yhi = __builtin_nanf("");
    pd.ylo = 0.0f;
    pd.vt = -1;

    const int vt0 = vc * L.vt_per_chunk;
    const int vt1 = min(vt0 + L.vt_per_chunk, L.n_vt);
    const TileSpan sp{vt0, vt1 - 1};
    unsigned wbuf[R][PER];  // OUT = 2: the bitmask words of the lane's tile slot
#pragma unroll
    for (int r = 0; r < R; r++)
#pragma unroll
        for (int P = 0; P < PER; P++) wbuf[r][P] = 0u;
    constexpr int NPY = bf16_y_pieces(KK, SIGN);
    constexpr int NBM = NPY == 3 ? KK : 1;
    bf16x8 b0h[KK], b0l[KK], b1h[KK], b1l[KK], b0m[NBM], b1m[NBM];
    double y0 = 0.0, y1 = 0.0;
    // this lane met an ill-conditioned sample, per resident subtile (one
    // flag for all of them in the pipelined schedule, which has no registers
    // to spare); the steady state pays one compare per sample
    constexpr int NB = PIPE ? 1 : R;
    bool bad[NB];
#pragma unroll
    for (int r = 0; r < NB; r++) bad[r] = false;
    // Rolling schedule (two data fragments, KK >= 5, no software pipelining):
    // one data buffer; the k-step loop is outermost, and once k-step kk has
    // fed every resident subtile its registers take the next tile's k-step
    // kk (prefetched a whole tile ahead, as with two buffers, for half the
    // registers); the R accumulation chains interleave.
    constexpr bool ROLL = !PIPE && NPY == 2 && KK >= BPG_ROLL_MIN_KK;
    if constexpr (ROLL) {
        if (vt0 < vt1) load_tile<KK, NPY>(Y, YY, vt0, lane, b0h, b0m, b0l, y0);
        for (int vt = vt0; vt < vt1; vt++) {
            const int vn = min(vt + 1, vt1 - 1);  // the last tile reloads itself
            f32x16 acc[R];
#pragma unroll
            for (int r = 0; r < R; r++)
#pragma unroll
                for (int e = 0; e < 16; e++) acc[r][e] = 0.0f;
#pragma unroll
            for (int kk = 0; kk < KK; kk++) {
#pragma unroll
                for (int r = 0; r < R; r++) {
                    if (!SIGN) acc[r] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(alo[r][kk], b0h[kk], acc[r], 0, 0, 0);
                    acc[r] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ahi[r][kk], b0l[kk], acc[r], 0, 0, 0);
                    acc[r] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ahi[r][kk], b0h[kk], acc[r], 0, 0, 0);
                }
                const size_t base = (((size_t)vn * KK + kk) * 2) * 64 + lane;
                b0h[kk] = Y[base];
                b0l[kk] = Y[base + 64];
            }
            const double yn = YY[vn * 32 + (lane & 31)];
            const float yhi = (float)y0;  // NaN for padding voxels
            const float ylo = (float)(y0 - (double)yhi);
#pragma unroll
            for (int r = 0; r < R; r++)
                epi<PP, FT, OUT>(a, acc[r], yhi, ylo, vt, lane, sub0 + r, runmax[r], bad[r], wbuf[r], sp);
            y0 = yn;
        }
    }
    // Rolling pipelined schedule (two data fragments, KK 5..8, pipelined):
    // one data buffer; the subtiles' chains run one after another, each
    // interleaved with the epilogue of the accumulator before it (the last
    // subtile of the previous tile for the first), and the last chain hands
    // each k-step's registers to the next tile's data once it has used them
    // (every load then has KK - 1 k-steps of MFMAs to arrive).
    constexpr bool RPIPE = rolling_pipelined(KK, SIGN, PIPE);
    if constexpr (RPIPE) {
        constexpr int NM = KK * (SIGN ? 2 : 3);
        constexpr int NV = (PER * (PP + 5) + NM - 1) / NM;
        if (vt0 < vt1) load_tile<KK, NPY>(Y, YY, vt0, lane, b0h, b0m, b0l, y0);
        for (int vt = vt0; vt < vt1; vt++) {
            const int vn = min(vt + 1, vt1 - 1);  // the last tile reloads itself
            const double yn = YY[vn * 32 + (lane & 31)];
            const float yhi = (float)y0;  // NaN for padding voxels
            const float ylo = (float)(y0 - (double)yhi);
#pragma unroll
            for (int r = 0; r < R; r++) {
                f32x16 acc;
#pragma unroll
                for (int e = 0; e < 16; e++) acc[e] = 0.0f;
#pragma unroll
                for (int kk = 0; kk < KK; kk++) {
                    if (!SIGN) acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(alo[r][kk], b0h[kk], acc, 0, 0, 0);
                    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ahi[r][kk], b0l[kk], acc, 0, 0, 0);
                    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ahi[r][kk], b0h[kk], acc, 0, 0, 0);
                    if (r == R - 1) {
                        const size_t base = (((size_t)vn * KK + kk) * 2) * 64 + lane;
                        b0h[kk] = Y[base];
                        b0l[kk] = Y[base + 64];
                    }
                }
                if (r == 0)
                    epi<PP, FT, OUT>(a, pd.acc, pd.yhi, pd.ylo, pd.vt, lane, sub0 + R - 1, runmax[R - 1], bad[0],
                                     wbuf[R - 1], sp);
                else
                    epi<PP, FT, OUT>(a, pd.acc, yhi, ylo, vt, lane, sub0 + r - 1, runmax[r - 1], bad[0],
                                     wbuf[r - 1], sp);
#pragma unroll
                for (int i = 0; i < NM; i++) {
                    __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);   // one MFMA
                    __builtin_amdgcn_sched_group_barrier(0x002, NV, 0);  // NV VALU
                }
                pd.acc = acc;
            }
            pd.yhi = yhi;
            pd.ylo = ylo;
            pd.vt = vt;
            y0 = yn;
        }
    }
    if (!ROLL && !RPIPE && vt0 < vt1) load_tile<KK, NPY>(Y, YY, vt0, lane, b0h, b0m, b0l, y0);
    for (int vt = vt0; vt < vt1 && !ROLL && !RPIPE; vt += 2) {
        if (vt + 1 < vt1) load_tile<KK, NPY>(Y, YY, vt + 1, lane, b1h, b1m, b1l, y1);
        contract_tile<KK, R, PP, SIGN, FT, OUT, PIPE, NPY>(a, ahi, alo, b0h, b0m, b0l, y0, vt, lane, sub0, runmax,
                                                           pd, bad, wbuf, sp);
        if (vt + 1 >= vt1) break;
        if (vt + 2 < vt1) load_tile<KK, NPY>(Y, YY, vt + 2, lane, b0h, b0m, b0l, y0);
        contract_tile<KK, R, PP, SIGN, FT, OUT, PIPE, NPY>(a, ahi, alo, b1h, b1m, b1l, y1, vt + 1, lane, sub0,
                                                           runmax, pd, bad, wbuf, sp);
    }
    if constexpr (PIPE)
        epi<PP, FT, OUT>(a, pd.acc, pd.yhi, pd.ylo, pd.vt, lane, sub0 + R - 1, runmax[R - 1], bad[0], wbuf[R - 1], sp);
    // rows past K (zero operand rows) have z = 0 and SSE = |y*|^2: never
    // flagged.  Rare: a wave that met an ill-conditioned sample adds one to
    // the counter and flags the permutations of each (subtile, half-wave)
    // where it happened (of every resident subtile in the pipelined form).
    // With a sample list (OUT = 3) the epilogue has listed them instead.
    if constexpr (OUT != 3) {
        bool any = false;
#pragma unroll
        for (int b = 0; b < NB; b++) {
            const unsigned long long m = __ballot(bad[b]);
            any |= m != 0ull;
            if (m && (lane & 31) == 0) {
                const int hl = lane >> 5;
                if ((hl ? (m >> 32) : (m & 0xFFFFFFFFull)) != 0ull)
                    for (int r = (PIPE ? 0 : b); r < (PIPE ? R : b + 1); r++)
                        for (int P = 0; P < PER; P++) {
                            const int k = (sub0 + r) * (2 * PER) + hl * PER + P;
                            if (k < a.K) ((gbyte*)a.flags)[k] = 1;
                        }
            }
        }
        if (any && lane == 0) atomicAdd(a.unsafe, 1u);
    }

    if constexpr (OUT == 2) return;  // cluster extent: no maxima
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
            // one writer per (voxel chunk, permutation): no atomics
            if ((lane & 31) == 0 && k < a.K) ((gfloat*)a.partial)[(size_t)vc * a.K + k] = m * scale;
        }
}

template <int KK, int R, int PP, bool SIGN, bool FT, int OUT, bool PIPE>
__device__ __forceinline__ void contract_body(const ContractArgs& a, const Bf16Launch& L) {
    const int lb = xcd_logical_block(L.per_xcd);
    if (lb >= L.n_mgroups * L.n_vchunks) return;
    contract_item<KK, R, PP, SIGN, FT, OUT, PIPE>(a, L, lb % L.n_mgroups, lb / L.n_mgroups);
}

// (the rolling pipelined schedule held to 256 registers, two waves per SIMD;
// the ping-pong schedule above 8 k-steps keeps its 256 VGPRs and AGPRs at
// one wave per SIMD)
template <int KK, int R, int PP, bool SIGN, bool FT, int OUT, bool PIPE>
__global__ __launch_bounds__(256, (rolling_pipelined(KK, SIGN, PIPE) ? 2 : 1))
void k_contract_bf16(ContractArgs a, Bf16Launch L) {
    contract_body<KK, R, PP, SIGN, FT, OUT, PIPE>(a, L);
}

// Persistent form over the studies of one batch class (same KK, PP, operand
// kind and subtile count).  A block starts on the queue of the XCD it runs
// on (its XCC_ID hardware register): item i of study j in queue x is
// (mg, vc) = (i % n_mgroups, x + 8 * (i / n_mgroups)) after the items of
// studies 0..j-1.  A block whose queue is empty takes items from the other
// XCDs' queues in turn, and leaves when all eight are empty (every block
// reaches that exit).
// (the rolling schedule held to 256 registers as in k_contract_bf16_roll:
// the accumulators then stay in VGPRs, and two waves share each SIMD)
template <int KK, int R, bool SIGN, bool PIPE>
constexpr int persist_min_blocks() {
    return (!PIPE && bf16_y_pieces(KK, SIGN) == 2 && KK >= BPG_ROLL_MIN_KK) ? (R >= 3 ? 1 : 2) : 1;
}

template <int KK, int R, int PP, bool SIGN, bool PIPE, int OUT>
__global__ __launch_bounds__(256, (persist_min_blocks<KK, R, SIGN, PIPE>())) void k_contract_bf16_persist(BatchClass c) {
    __shared__ int s_item;
    // s_getreg_b32 hwreg(HW_REG_XCC_ID, 0, 4): the XCD of this block
    const int home = __builtin_amdgcn_s_getreg((3 << 11) | (0 << 6) | 20) & (BPG_XCDS - 1);
    for (int q = 0; q < BPG_XCDS; q++) {
        const int x = (home + q) & (BPG_XCDS - 1);
        for (;;) {
            if (threadIdx.x == 0) s_item = (int)atomicAdd(c.counter + x, 1u);
            __syncthreads();
            // one value for the whole block: scalar registers from here on
            const int item = __builtin_amdgcn_readfirstlane(s_item);
            __syncthreads();
            if (item >= c.n_items[x]) break;
            int j = 0;
            while (j + 1 < c.n_studies && c.studies[j + 1].item0[x] <= item) j++;
            // the study's arguments as values, not references into memory
            // the item's stores might alias, and in scalar registers
            const ContractArgs a = uniform_load(&c.studies[j].a);
            const Bf16Launch L = uniform_load(&c.studies[j].L);
            const int local = item - c.studies[j].item0[x];
            contract_item<KK, R, PP, SIGN, false, OUT, PIPE>(a, L, local % L.n_mgroups,
                                                             x + BPG_XCDS * (local / L.n_mgroups));
        }
    }
}

// The rolling schedule (two data fragments, KK 5..8), held to 256 registers
// so that two waves share each SIMD
template <int KK, int R, int PP, bool SIGN, bool FT, int OUT, bool PIPE>
__global__ __launch_bounds__(256, R >= 3 ? 1 : 2) void k_contract_bf16_roll(ContractArgs a, Bf16Launch L) {
    contract_body<KK, R, PP, SIGN, FT, OUT, PIPE>(a, L);
}

// Launches the instantiation for these operands, or, with occ set, stores
// in *occ the blocks of that instantiation one CU holds at once (on entry
// *occ != 0 asks for the one that writes statistics or bitmasks).
template <int KK, int R, int PP, bool SIGN, bool FT, bool PIPE>
hipError_t run_bf16(const ContractArgs& a, const Bf16Launch& L, hipStream_t s, int* occ) {
    constexpr bool ROLL = !PIPE && bf16_y_pieces(KK, SIGN) == 2 && KK >= BPG_ROLL_MIN_KK;
    // 0: maxima only; 1: statistics and / or bitmask words; 2: the bitmask
    // (and ties) alone, in 128-byte runs; 3: maxima only, with a sample list
    // (t statistics, permutation operands, KK <= BPG_EXCL_MAX_KK)
    const int mode = occ ? *occ
                         : (a.bits && a.ties && !a.stats) ? 2
                         : (a.stats || a.bits)            ? 1
                         : a.samples                      ? 3
                                                          : 0;
    constexpr bool EXCL = !SIGN && !FT && KK <= BPG_EXCL_MAX_KK && !ROLL;
    if (mode == 3 && !EXCL) return hipErrorInvalidValue;
    void (*kern)(ContractArgs, Bf16Launch) = nullptr;
    if constexpr (ROLL)
        kern = mode == 2   ? k_contract_bf16_roll<KK, R, PP, SIGN, FT, 2, PIPE>
               : mode == 1 ? k_contract_bf16_roll<KK, R, PP, SIGN, FT, 1, PIPE>
                           : k_contract_bf16_roll<KK, R, PP, SIGN, FT, 0, PIPE>;
    else
        kern = mode == 2   ? k_contract_bf16<KK, R, PP, SIGN, FT, 2, PIPE>
               : mode == 1 ? k_contract_bf16<KK, R, PP, SIGN, FT, 1, PIPE>
                           : k_contract_bf16<KK, R, PP, SIGN, FT, 0, PIPE>;
    if constexpr (EXCL)
        if (mode == 3) kern = k_contract_bf16<KK, R, PP, SIGN, FT, 3, PIPE>;
    if (occ) return hipOccupancyMaxActiveBlocksPerMultiprocessor(occ, reinterpret_cast<const void*>(kern), 256, 0);
    hipLaunchKernelGGL(kern, dim3(L.per_xcd * 8), dim3(256), 0, s, a, L);
    return hipGetLastError();
}

template <int KK, int PP>
hipError_t dispatch_bf16_kind(const ContractArgs& a, const Bf16Launch& L, bool sign, hipStream_t s, int* occ) {
    constexpr int RS = bf16_rows_per_wave(KK, true);
    constexpr int R = bf16_rows_per_wave(KK, false);
    constexpr bool PIPE = bf16_pipelined(KK, false);
    // the autotuner's alternative subtile count, compiled for the max-T-only
    // kernels up to 8 k-steps
    constexpr bool ALT = KK <= 8;
    constexpr int RSA = bf16_rows_per_wave_alt(KK, true);
    constexpr int RA = bf16_rows_per_wave_alt(KK, false);
    const bool out = occ ? *occ != 0 : (a.stats || a.bits || a.samples);
    const bool alt = ALT && L.variant == 1 && !out;
    if (sign) {
        if (a.ftest || PP != 1) return hipErrorInvalidValue;
        if constexpr (PP == 1) {
            if constexpr (ALT)
                if (alt) return run_bf16<KK, RSA, 1, true, false, bf16_pipelined(KK, true)>(a, L, s, occ);
            return run_bf16<KK, RS, 1, true, false, bf16_pipelined(KK, true)>(a, L, s, occ);
        }
    }
    if constexpr (ALT)
        if (alt)
            return a.ftest ? run_bf16<KK, RA, PP, false, true, PIPE>(a, L, s, occ)
                           : run_bf16<KK, RA, PP, false, false, PIPE>(a, L, s, occ);
    return a.ftest ? run_bf16<KK, R, PP, false, true, PIPE>(a, L, s, occ)
                   : run_bf16<KK, R, PP, false, false, PIPE>(a, L, s, occ);
}

template <int KK, int R, int PP, bool SIGN>
hipError_t run_bf16_persist(const BatchClass& c, bool excl, int grid, hipStream_t s, int* occ) {
    constexpr bool PIPE = bf16_pipelined(KK, SIGN);
    constexpr bool ROLL = !PIPE && bf16_y_pieces(KK, SIGN) == 2 && KK >= BPG_ROLL_MIN_KK;
    constexpr bool EXCL = !SIGN && KK <= BPG_EXCL_MAX_KK && !ROLL;
    if (excl && !EXCL) return hipErrorInvalidValue;
    void (*kern)(BatchClass) = k_contract_bf16_persist<KK, R, PP, SIGN, PIPE, 0>;
    if constexpr (EXCL)
        if (excl) kern = k_contract_bf16_persist<KK, R, PP, SIGN, PIPE, 3>;
    if (occ) return hipOccupancyMaxActiveBlocksPerMultiprocessor(occ, reinterpret_cast<const void*>(kern), 256, 0);
    hipLaunchKernelGGL(kern, dim3(grid), dim3(256), 0, s, c);
    return hipGetLastError();
}

template <int KK, int PP>
hipError_t dispatch_persist_kind(const BatchClass& c, bool sign, int variant, bool excl, int grid, hipStream_t s,
                                 int* occ) {
    constexpr int RS = bf16_rows_per_wave(KK, true), RSA = bf16_rows_per_wave_alt(KK, true);
    constexpr int R = bf16_rows_per_wave(KK, false), RA = bf16_rows_per_wave_alt(KK, false);
    if (sign) {
        if (excl) return hipErrorInvalidValue;
        if constexpr (PP == 1)
            return variant ? run_bf16_persist<KK, RSA, 1, true>(c, false, grid, s, occ)
                           : run_bf16_persist<KK, RS, 1, true>(c, false, grid, s, occ);
        return hipErrorInvalidValue;
    }
    // a class with sample lists runs the default subtile count
    return variant && !excl ? run_bf16_persist<KK, RA, PP, false>(c, false, grid, s, occ)
                            : run_bf16_persist<KK, R, PP, false>(c, excl, grid, s, occ);
}

// The persistent contraction of a batch class (t statistics, max-T only),
// compiled up to 8 k-steps (N <= 128).
template <int KK>
hipError_t dispatch_persist_pp(const BatchClass& c, int PP, bool sign, int variant, bool excl, int grid,
                               hipStream_t s, int* occ) {
    if constexpr (KK > 8) {
        return hipErrorInvalidValue;
    } else {
        switch (PP) {
            case 1: return dispatch_persist_kind<KK, 1>(c, sign, variant, excl, grid, s, occ);
            case 2: return dispatch_persist_kind<KK, 2>(c, sign, variant, excl, grid, s, occ);
            case 3: return dispatch_persist_kind<KK, 3>(c, sign, variant, excl, grid, s, occ);
            case 4: return dispatch_persist_kind<KK, 4>(c, sign, variant, excl, grid, s, occ);
            case 5: return dispatch_persist_kind<KK, 5>(c, sign, variant, excl, grid, s, occ);
            case 8: return dispatch_persist_kind<KK, 8>(c, sign, variant, excl, grid, s, occ);
            case 16: return dispatch_persist_kind<KK, 16>(c, sign, variant, excl, grid, s, occ);
        }
        return hipErrorInvalidValue;
    }
}

template <int KK>
hipError_t dispatch_bf16_pp(const ContractArgs& a, const Bf16Launch& L, int PP, bool sign, hipStream_t s,
                            int* occ) {
    switch (PP) {
        case 1: return dispatch_bf16_kind<KK, 1>(a, L, sign, s, occ);
        case 2: return dispatch_bf16_kind<KK, 2>(a, L, sign, s, occ);
        case 3: return dispatch_bf16_kind<KK, 3>(a, L, sign, s, occ);
        case 4: return dispatch_bf16_kind<KK, 4>(a, L, sign, s, occ);
        case 5: return dispatch_bf16_kind<KK, 5>(a, L, sign, s, occ);
        case 8: return dispatch_bf16_kind<KK, 8>(a, L, sign, s, occ);
        case 16: return dispatch_bf16_kind<KK, 16>(a, L, sign, s, occ);
    }
    return hipErrorInvalidValue;
}

}  // namespace bpg
