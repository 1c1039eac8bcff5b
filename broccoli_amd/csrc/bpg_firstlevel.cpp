// bpg_firstlevel.cpp -- host side of the first-level permutation test
// (include/bpg.h): validation, the FP64 model, the device set-up sequence
// (regression, AR(4) whitening) and the permutation phases.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <new>
#include <string>
#include <unordered_set>
#include <vector>

#include "bpg_fl_kernels.h"
#include "bpg_internal.h"
#include "bpg_kernels.h"
#include "bpg_runtime.h"

using bpg::DBuf;
using bpg::HBuf;

struct bpg_fl_plan {
    bpg_ctx* ctx = nullptr;
    bpg_fl_desc d{};
    bpg::FlModel model;
    int Vc = 0, Vpad = 0, nwords = 0, Tp = 0;
    size_t G = 0;
    std::vector<unsigned> comp2grid;
    hipStream_t stream = nullptr;
    hipEvent_t ev[4] = {};
    bpg_timing timing{};
    double setup_ms = 0.0;

    DBuf<float> w32;           // [T][Vpad] whitened volumes
    DBuf<double> ar;           // [4][Vpad] total AR estimates
    DBuf<double> Q, A, u;      // model
    DBuf<float> Q32, u32;      // rotated model of the FP32 form ([Tq][16], [16])
    DBuf<unsigned short> Qb;   // rotated model of the split-BF16 form ([Tq32/32][2][64][8])
    DBuf<unsigned short> Qh;   // rotated model of the split-FP16 form ([T8][2][16][8])
    DBuf<float> wscale;        // [Vpad] power-of-two scale of each whitened series
    DBuf<unsigned char> flags; // [K] permutations to repeat in FP64
    HBuf<unsigned char> h_flags;
    DBuf<int> d_nbr;
    DBuf<unsigned short> perm; // [K][Tp]
    DBuf<float> partial, maxout, stats;
    DBuf<unsigned> bits, maxsize;
    bpg::ClusterScratch cl;
    HBuf<unsigned short> h_perm;
    HBuf<float> h_max;

    void release_all() {
        w32.release(); ar.release(); Q.release(); A.release(); u.release(); d_nbr.release(); perm.release();
        partial.release(); maxout.release(); stats.release(); bits.release(); maxsize.release(); cl.release();
        h_perm.release(); h_max.release();
        Q32.release(); u32.release(); Qb.release(); Qh.release(); wscale.release(); flags.release(); h_flags.release();
    }
};

namespace {

bpg::fl::Taps make_taps(const float* f) {
    bpg::fl::Taps t;
    for (int i = 0; i < 9; i++) t.f[i] = f[i];
    return t;
}

// CreateSmoothingFilters (LIB:11055-11090): float taps exp(-u^2 / (2 s^2))
// with s = fwhm / 2.354 / voxel size, divided by their float sum.
void smoothing_filters(float fwhm, const float* vox, float (*f)[9]) {
    for (int a = 0; a < 3; a++) {
        const double sigma = (double)fwhm / 2.354 / (double)vox[a];
        float sum = 0.0f;
        for (int i = 0; i < 9; i++) {
            const double u = (double)(i - 4);
            f[a][i] = (float)std::exp(-(u * u) / (2.0 * sigma * sigma));
            sum += f[a][i];
        }
        for (int i = 0; i < 9; i++) f[a][i] /= sum;
    }
}

// The basis of the FP32 and split-BF16 forms for statistical map `map`:
// Q M with M orthogonal (p x p) whose first r columns are the statistic rows
// A_i (orthonormal), so z' = M'z has z'_i = A_i z for i < r and |z'| = |z|;
// completed by Gram-Schmidt on the unit vectors.  rot[t*W + j] (FP64, W = 16
// or 32 columns, rows
// to Tq, zeros beyond T and p), u32 = M'u.
void rotated_basis(const bpg::FlModel& m, int map, int Tq, int W, std::vector<double>& rot, std::vector<float>& u32) {
    const int p = m.p, r = m.nrows[map];
    const std::vector<double>& A = m.A[map];
    std::vector<std::vector<double>> cols;
    for (int i = 0; i < r; i++) {
        std::vector<double> v(p);
        for (int j = 0; j < p; j++) v[j] = A[(size_t)i * m.pq + j];
        cols.push_back(v);
    }
    for (int e = 0; e < p && (int)cols.size() < p; e++) {
        std::vector<double> v(p, 0.0);
        v[e] = 1.0;
        for (int pass = 0; pass < 2; pass++)
            for (const auto& q : cols) {
                double d = 0.0;
                for (int j = 0; j < p; j++) d += q[j] * v[j];
                for (int j = 0; j < p; j++) v[j] -= d * q[j];
            }
        double n2 = 0.0;
        for (double x : v) n2 += x * x;
        if (n2 < 1e-6) continue;
        for (double& x : v) x /= std::sqrt(n2);
        cols.push_back(v);
    }
    rot.assign((size_t)Tq * W, 0.0);
    u32.assign(W, 0.0f);
    for (int i = 0; i < (int)cols.size(); i++) {
        double uu = 0.0;
        for (int j = 0; j < p; j++) uu += m.u[j] * cols[i][j];
        u32[i] = (float)uu;
        for (int t = 0; t < m.T; t++) {
            double acc = 0.0;
            for (int j = 0; j < p; j++) acc += m.Q[(size_t)t * m.pq + j] * cols[i][j];
            rot[(size_t)t * W + i] = acc;
        }
    }
}

// bf16 bits of f, round to nearest even
unsigned short bf16_bits(float f) {
    unsigned u;
    std::memcpy(&u, &f, 4);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}
float bf16_value(unsigned short b) {
    const unsigned u = (unsigned)b << 16;
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}

// The B operands of the split-BF16 form: per block of 32 time points, hi then
// lo, lane l holding rot[t0 + 8 (l / 16) + e][l % 16] for e = 0..7.
void split_basis(const std::vector<double>& rot, int Tq32, std::vector<unsigned short>& qb) {
    const int nb = Tq32 / 32;
    qb.assign((size_t)nb * 2 * 64 * 8, 0);
    for (int b = 0; b < nb; b++)
        for (int l = 0; l < 64; l++)
            for (int e = 0; e < 8; e++) {
                const int t = 32 * b + 8 * (l / 16) + e;
                const double q = (size_t)t * 16 + (l % 16) < rot.size() ? rot[(size_t)t * 16 + (l % 16)] : 0.0;
                const unsigned short hb = bf16_bits((float)q);
                const unsigned short lb = bf16_bits((float)(q - (double)bf16_value(hb)));
                qb[(((size_t)b * 2 + 0) * 64 + l) * 8 + e] = hb;
                qb[(((size_t)b * 2 + 1) * 64 + l) * 8 + e] = lb;
            }
}

// The A operands of the split-FP16 form (NB blocks of 16 basis columns,
// rot of width 16 NB): for time block c (8 points),
// qh[(((c * 2 + hl) * NB + nb) * 16 + col) * 8 + k] is the FP16 hi part
// (hl = 0) or remainder (hl = 1), rounded to nearest even, of
// HX_QSCALE * rot[(8c + k) * 16 NB + 16 nb + col] (zero beyond T).
void split_basis_f16(const std::vector<double>& rot, int T, int NB, std::vector<unsigned short>& qh) {
    const int T8 = (T + 7) / 8, W = 16 * NB;
    qh.assign((size_t)T8 * 2 * NB * 16 * 8, 0);
    for (int c = 0; c < T8; c++)
        for (int nb = 0; nb < NB; nb++)
            for (int col = 0; col < 16; col++)
                for (int k = 0; k < 8; k++) {
                    const int t = 8 * c + k;
                    const double q =
                        t < T ? (double)bpg::fl::HX_QSCALE * rot[(size_t)t * W + 16 * nb + col] : 0.0;
                    const _Float16 h = (_Float16)q;
                    const _Float16 l = (_Float16)(q - (double)h);
                    const size_t o = (((size_t)c * 2 * NB + nb) * 16 + col) * 8 + k;
                    std::memcpy(&qh[o], &h, 2);
                    std::memcpy(&qh[o + (size_t)NB * 16 * 8], &l, 2);
                }
}

// The device set-up: regression, then three rounds of AR(4) estimation,
// normalised smoothing, accumulation and whitening (LIB:16363-16452).
int run_setup(bpg_fl_plan* pl, const float* yc) {
    bpg_ctx* ctx = pl->ctx;
    const bpg_fl_desc& d = pl->d;
    const int T = d.T, p = d.p, Vc = pl->Vc, Vpad = pl->Vpad;
    const size_t G = pl->G;
    hipStream_t st = pl->stream;
    DBuf<float> Y;
    DBuf<double> Xp, X, e, wh, cur, g0, g1, smask, m0;
    DBuf<unsigned> c2g;
    auto cleanup = [&] {
        Y.release(); Xp.release(); X.release(); e.release(); wh.release(); cur.release();
        g0.release(); g1.release(); smask.release(); m0.release(); c2g.release();
    };
    int rc = BPG_OK;
    [&] {
        auto H = [&](hipError_t e_, const char* what) {
            if (e_ != hipSuccess && rc == BPG_OK) {
                ctx->err = std::string(what) + ": " + hipGetErrorString(e_);
                rc = e_ == hipErrorOutOfMemory ? BPG_ERR_NOMEM : BPG_ERR_HIP;
            }
            return rc == BPG_OK;
        };
        if (!H(Y.ensure((size_t)T * Vpad), "allocation") || !H(Xp.ensure((size_t)p * T), "allocation") ||
            !H(X.ensure((size_t)T * p), "allocation") || !H(e.ensure((size_t)T * Vpad), "allocation") ||
            !H(wh.ensure((size_t)T * Vpad), "allocation") || !H(cur.ensure(4 * (size_t)Vpad), "allocation") ||
            !H(g0.ensure(4 * G), "allocation") || !H(g1.ensure(4 * G), "allocation") ||
            !H(smask.ensure(G), "allocation") || !H(m0.ensure(G), "allocation") ||
            !H(c2g.ensure(std::max(Vc, 1)), "allocation") || !H(pl->w32.ensure((size_t)T * Vpad), "allocation") ||
            !H(pl->ar.ensure(4 * (size_t)Vpad), "allocation") || !H(pl->wscale.ensure((size_t)Vpad), "allocation"))
            return;
        std::vector<double> maskd(G);
        for (size_t i = 0; i < G; i++) maskd[i] = (double)d.mask[i];
        if (!H(hipMemcpyAsync(Y.p, yc, (size_t)T * Vpad * sizeof(float), hipMemcpyHostToDevice, st), "upload") ||
            !H(hipMemcpyAsync(Xp.p, pl->model.Xp.data(), (size_t)p * T * sizeof(double), hipMemcpyHostToDevice, st),
               "upload") ||
            !H(hipMemcpyAsync(X.p, pl->model.X.data(), (size_t)T * p * sizeof(double), hipMemcpyHostToDevice, st),
               "upload") ||
            !H(hipMemcpyAsync(m0.p, maskd.data(), G * sizeof(double), hipMemcpyHostToDevice, st), "upload") ||
            (Vc > 0 && !H(hipMemcpyAsync(c2g.p, pl->comp2grid.data(), (size_t)Vc * sizeof(unsigned),
                                         hipMemcpyHostToDevice, st), "upload")))
            return;
        float f[3][9];
        smoothing_filters(d.ar_smoothing_fwhm, d.voxel_size, f);
        const bpg::fl::Taps fx = make_taps(f[0]), fy = make_taps(f[1]), fz = make_taps(f[2]);
        if (!H(hipEventRecord(pl->ev[0], st), "event")) return;
        if (!H(bpg::fl::launch_regress(Y.p, Xp.p, X.p, T, p, Vc, Vpad, e.p, st), "regression")) return;
        // the smoothed mask (PerformSmoothing with certainty 1, LIB:11274)
        if (!H(bpg::fl::launch_conv(m0.p, g0.p, 1, d.W, d.H, d.D, 1, fy, st), "smoothing") ||
            !H(bpg::fl::launch_conv(g0.p, g1.p, 1, d.W, d.H, d.D, 0, fx, st), "smoothing") ||
            !H(bpg::fl::launch_conv(g1.p, smask.p, 1, d.W, d.H, d.D, 2, fz, st), "smoothing"))
            return;
        if (!H(hipMemsetAsync(pl->ar.p, 0, 4 * (size_t)Vpad * sizeof(double), st), "memset")) return;
        for (int it = 0; it < 3; it++) {
            // AR(4) of the current whitened volumes (the residuals at first)
            if (!H(bpg::fl::launch_ar4(it == 0 ? e.p : wh.p, T, it == 0 ? 0 : 4, Vc, Vpad, cur.p, st), "AR(4)"))
                return;
            // normalised smoothing (LIB:11545-11620): S(a * mask) / S(mask),
            // times the mask; the estimates are 0 outside the mask
            if (!H(hipMemsetAsync(g0.p, 0, 4 * G * sizeof(double), st), "memset") ||
                !H(bpg::fl::launch_scatter(cur.p, c2g.p, Vc, Vpad, G, g0.p, st), "smoothing") ||
                !H(bpg::fl::launch_conv(g0.p, g1.p, 4, d.W, d.H, d.D, 1, fy, st), "smoothing") ||
                !H(bpg::fl::launch_conv(g1.p, g0.p, 4, d.W, d.H, d.D, 0, fx, st), "smoothing") ||
                !H(bpg::fl::launch_rods_accumulate(g0.p, smask.p, c2g.p, Vc, Vpad, d.W, d.H, d.D, fz, pl->ar.p,
                                                   st), "smoothing"))
                return;
            // whitening of the residuals with the running totals
            if (!H(bpg::fl::launch_whiten(e.p, pl->ar.p, T, Vc, Vpad, wh.p, it == 2 ? pl->w32.p : nullptr, st),
                   "whitening"))
                return;
        }
        if (!H(bpg::fl::launch_wscale(pl->w32.p, T, Vc, Vpad, pl->wscale.p, st), "series scale")) return;
        if (!H(hipEventRecord(pl->ev[1], st), "event")) return;
        H(hipStreamSynchronize(st), "set-up");
    }();
    (void)hipStreamSynchronize(st);  // nothing in flight before the buffers go
    if (rc == BPG_OK) pl->setup_ms = bpg::elapsed(pl->ev[0], pl->ev[1]);
    cleanup();
    return rc;
}

}  // namespace

extern "C" int bpg_fl_plan_create(bpg_ctx* ctx, const bpg_fl_desc* desc, bpg_fl_plan** out) {
    if (!ctx || !desc || !out) return BPG_ERR_ARG;
    *out = nullptr;
    const bpg_fl_desc& d = *desc;
    if (d.test != BPG_TTEST && d.test != BPG_FTEST) BPG_FAIL(ctx, BPG_ERR_ARG, "test must be t or F");
    if (d.precision != BPG_PREC_FP32 && d.precision != BPG_PREC_FP64 && d.precision != BPG_PREC_SPLIT_BF16 &&
        d.precision != BPG_PREC_SPLIT_F16)
        BPG_FAIL(ctx, BPG_ERR_ARG,
                 "precision must be BPG_PREC_SPLIT_F16, BPG_PREC_SPLIT_BF16, BPG_PREC_FP32 or BPG_PREC_FP64");
    if (d.inference == BPG_TFCE)
        BPG_FAIL(ctx, BPG_ERR_UNSUPPORTED,
                 "TFCE is not available for the first-level permutation test (the reference leaves its "
                 "distribution unset, LIB:16524-16530)");
    if (d.inference != BPG_VOXEL && d.inference != BPG_CLUSTER_EXTENT && d.inference != BPG_CLUSTER_MASS)
        BPG_FAIL(ctx, BPG_ERR_ARG, "unknown inference mode");
    if (d.inference != BPG_VOXEL && d.connectivity != 26 && d.connectivity != 6)
        BPG_FAIL(ctx, BPG_ERR_ARG, "connectivity must be 26 or 6");
    if (d.W <= 0 || d.H <= 0 || d.D <= 0) BPG_FAIL(ctx, BPG_ERR_ARG, "empty grid");
    if (d.T < 10 || d.T > 65535) BPG_FAIL(ctx, BPG_ERR_ARG, "the number of time points must be in 10..65535");
    if (d.p <= 0 || d.p > BPG_MAX_REGRESSORS || d.p >= d.T)
        BPG_FAIL(ctx, BPG_ERR_ARG, "the number of regressors must be in 1..25 and below T");
    if (d.nc <= 0 || (d.test == BPG_FTEST && d.nc > BPG_MAX_FROWS) || d.nc > BPG_MAX_REGRESSORS)
        BPG_FAIL(ctx, BPG_ERR_ARG, "number of contrasts out of range");
    if (!d.volumes || !d.mask || !d.X || !d.C) BPG_FAIL(ctx, BPG_ERR_ARG, "null input");
    if (!(d.voxel_size[0] > 0) || !(d.voxel_size[1] > 0) || !(d.voxel_size[2] > 0) || !(d.ar_smoothing_fwhm > 0))
        BPG_FAIL(ctx, BPG_ERR_ARG, "voxel sizes and the AR smoothing FWHM must be positive");
    const int pq = bpg::fl::fl_width(d.p);
    auto* pl = new (std::nothrow) bpg_fl_plan;
    if (!pl) return BPG_ERR_NOMEM;
    pl->ctx = ctx;
    pl->d = d;
    {
        std::string err;
        const int rc = bpg::build_fl_model(d.T, d.p, d.nc, d.test == BPG_FTEST, d.X, d.C, pq, pl->model, err);
        if (rc != BPG_OK) {
            delete pl;
            BPG_FAIL(ctx, BPG_ERR_DESIGN, err);
        }
    }
    if (hipSetDevice(ctx->device) != hipSuccess) {
        delete pl;
        BPG_FAIL(ctx, BPG_ERR_HIP, "hipSetDevice failed");
    }
    pl->G = (size_t)d.W * d.H * d.D;
    std::vector<int> grid2comp;
    bpg::compact_mask(d.mask, pl->G, pl->comp2grid, grid2comp);
    pl->Vc = (int)pl->comp2grid.size();
    pl->Vpad = std::max(bpg::fl::TILE, (pl->Vc + bpg::fl::TILE - 1) / bpg::fl::TILE * bpg::fl::TILE);
    pl->nwords = pl->Vpad / 32;
    pl->Tp = (d.T + 7) / 8 * 8;
    auto fail = [&](int rc) {
        pl->release_all();
        for (auto& e : pl->ev)
            if (e) (void)hipEventDestroy(e);
        if (pl->stream) (void)hipStreamDestroy(pl->stream);
        delete pl;
        return rc;
    };
    if (hipStreamCreate(&pl->stream) != hipSuccess) { ctx->err = "hipStreamCreate failed"; return fail(BPG_ERR_HIP); }
    for (auto& e : pl->ev)
        if (hipEventCreate(&e) != hipSuccess) { ctx->err = "hipEventCreate failed"; return fail(BPG_ERR_HIP); }
    // compact, time-major volumes (zero padding beyond Vc)
    std::vector<float> yc((size_t)d.T * pl->Vpad, 0.0f);
    for (int t = 0; t < d.T; t++) {
        const float* src = d.volumes + (size_t)t * pl->G;
        float* dst = yc.data() + (size_t)t * pl->Vpad;
        for (int v = 0; v < pl->Vc; v++) dst[v] = src[pl->comp2grid[v]];
    }
    if (pl->Vc > 0) {
        const int rc = run_setup(pl, yc.data());
        if (rc != BPG_OK) return fail(rc);
    }
    const bpg::FlModel& m = pl->model;
    size_t na = 0;
    for (const auto& a : m.A) na = std::max(na, a.size());
    if (pl->Q.ensure(m.Q.size()) || pl->u.ensure(m.u.size()) || pl->A.ensure(na)) {
        ctx->err = "device allocation failed";
        return fail(BPG_ERR_NOMEM);
    }
    if (hipMemcpy(pl->Q.p, m.Q.data(), m.Q.size() * sizeof(double), hipMemcpyHostToDevice) ||
        hipMemcpy(pl->u.p, m.u.data(), m.u.size() * sizeof(double), hipMemcpyHostToDevice)) {
        ctx->err = "upload of the model failed";
        return fail(BPG_ERR_HIP);
    }
    {
        // the fast forms' rotated bases (the split-FP16 one up to 32 columns)
        const int Tq = (d.T + 3) & ~3, Tq32 = (d.T + 31) & ~31;
        if (pl->u32.ensure(32) || pl->Qh.ensure((size_t)(Tq32 / 8) * 2 * 2 * 16 * 8) ||
            (m.pq <= 16 && (pl->Q32.ensure((size_t)Tq * 16) || pl->Qb.ensure((size_t)Tq32 / 32 * 2 * 64 * 8)))) {
            ctx->err = "device allocation failed";
            return fail(BPG_ERR_NOMEM);
        }
    }
    if (d.inference != BPG_VOXEL) {
        std::vector<int> nbr;
        bpg::build_neighbours(d.W, d.H, d.D, pl->comp2grid, grid2comp, nbr);
        if (pl->d_nbr.ensure(nbr.size()) ||
            hipMemcpy(pl->d_nbr.p, nbr.data(), nbr.size() * sizeof(int), hipMemcpyHostToDevice)) {
            ctx->err = "upload of the neighbour table failed";
            return fail(BPG_ERR_HIP);
        }
    }
    pl->timing.transform_ms = pl->setup_ms;
    *out = pl;
    return BPG_OK;
}

extern "C" void bpg_fl_plan_destroy(bpg_fl_plan* pl) {
    if (!pl) return;
    (void)hipSetDevice(pl->ctx->device);
    if (pl->stream) (void)hipStreamSynchronize(pl->stream);
    pl->release_all();
    for (auto& e : pl->ev)
        if (e) (void)hipEventDestroy(e);
    if (pl->stream) (void)hipStreamDestroy(pl->stream);
    delete pl;
}

extern "C" long bpg_fl_plan_num_voxels(const bpg_fl_plan* pl) { return pl ? pl->Vc : -1; }

extern "C" int bpg_fl_plan_ar_estimates(bpg_fl_plan* pl, float* ar) {
    if (!pl || !ar) return BPG_ERR_ARG;
    bpg_ctx* ctx = pl->ctx;
    if (pl->Vc == 0) {
        std::fill(ar, ar + 4 * pl->G, 0.0f);
        return BPG_OK;
    }
    BPG_HIP(ctx, hipSetDevice(ctx->device));
    std::vector<double> h(4 * (size_t)pl->Vpad);
    BPG_HIP(ctx, hipMemcpy(h.data(), pl->ar.p, h.size() * sizeof(double), hipMemcpyDeviceToHost));
    std::fill(ar, ar + 4 * pl->G, 0.0f);
    for (int k = 0; k < 4; k++)
        for (int v = 0; v < pl->Vc; v++) ar[(size_t)k * pl->G + pl->comp2grid[v]] = (float)h[(size_t)k * pl->Vpad + v];
    return BPG_OK;
}

extern "C" int bpg_fl_plan_whitened(bpg_fl_plan* pl, float* volumes) {
    if (!pl || !volumes) return BPG_ERR_ARG;
    bpg_ctx* ctx = pl->ctx;
    const int T = pl->d.T;
    if (pl->Vc == 0) {
        std::fill(volumes, volumes + (size_t)T * pl->G, 0.0f);
        return BPG_OK;
    }
    BPG_HIP(ctx, hipSetDevice(ctx->device));
    std::vector<float> h((size_t)T * pl->Vpad);
    BPG_HIP(ctx, hipMemcpy(h.data(), pl->w32.p, h.size() * sizeof(float), hipMemcpyDeviceToHost));
    std::fill(volumes, volumes + (size_t)T * pl->G, 0.0f);
    for (int t = 0; t < T; t++)
        for (int v = 0; v < pl->Vc; v++)
            volumes[(size_t)t * pl->G + pl->comp2grid[v]] = h[(size_t)t * pl->Vpad + v];
    return BPG_OK;
}

namespace {

// One permutation phase in the requested fast form (FP32 or split-BF16
// projections; when allowed and the study fits it) or the FP64 form.
// *fast_used tells which ran.
int fl_phase(bpg_fl_plan* pl, int map, const unsigned short* perm, long K, float* maxdist, float* stats,
             bool allow_fast, bool* fast_used) {
    bpg_ctx* ctx = pl->ctx;
    const bpg_fl_desc& d = pl->d;
    const bpg::FlModel& m = pl->model;
    if (!perm || !maxdist) BPG_FAIL(ctx, BPG_ERR_ARG, "null permutation matrix or output");
    if (map < 0 || map >= m.nmaps) BPG_FAIL(ctx, BPG_ERR_ARG, "statistical map index out of range");
    if (K <= 0 || K > (1L << 24)) BPG_FAIL(ctx, BPG_ERR_ARG, "number of permutations out of range");
    BPG_HIP(ctx, hipSetDevice(ctx->device));
    const auto t0 = std::chrono::steady_clock::now();
    hipStream_t st = pl->stream;
    const int T = d.T, Tp = pl->Tp, Vc = pl->Vc;
    const bool cluster = d.inference != BPG_VOXEL;
    const bool mass = d.inference == BPG_CLUSTER_MASS;
    const bool want_stats = stats != nullptr || mass;

    // staging: rows padded to Tp, indices checked
    BPG_HIP(ctx, pl->h_perm.ensure((size_t)K * Tp));
    unsigned short worst = 0;
    for (long k = 0; k < K; k++) {
        const unsigned short* src = perm + (size_t)k * T;
        unsigned short* dst = pl->h_perm.p + (size_t)k * Tp;
        for (int t = 0; t < T; t++) {
            dst[t] = src[t];
            worst = std::max(worst, src[t]);
        }
        for (int t = T; t < Tp; t++) dst[t] = 0;
    }
    if (worst >= T) BPG_FAIL(ctx, BPG_ERR_ARG, "permutation index out of range");
    if (Vc == 0) {
        // empty mask: nothing to launch (the reference's empty max, or an
        // empty largest cluster)
        for (long k = 0; k < K; k++) maxdist[k] = cluster ? 0.0f : bpg::empty_max();
        pl->timing = bpg_timing{};
        pl->timing.precision_used = BPG_PREC_FP64;
        *fast_used = false;
        return BPG_OK;
    }

    const int n_vtiles = pl->Vpad / bpg::fl::TILE;
    const int tpc = (n_vtiles + bpg::fl::MAX_VCHUNKS - 1) / bpg::fl::MAX_VCHUNKS;
    const int n_vchunks = (n_vtiles + tpc - 1) / tpc;
    const int n_pgroups = (int)((K + bpg::fl::PERMS_PER_BLOCK - 1) / bpg::fl::PERMS_PER_BLOCK);
    BPG_HIP(ctx, pl->perm.ensure((size_t)K * Tp));
    BPG_HIP(ctx, pl->partial.ensure((size_t)n_vchunks * K));
    BPG_HIP(ctx, pl->maxout.ensure(K));
    BPG_HIP(ctx, pl->h_max.ensure(K));
    if (want_stats) BPG_HIP(ctx, pl->stats.ensure((size_t)K * std::max(Vc, 1)));
    if (cluster) {
        BPG_HIP(ctx, pl->bits.ensure((size_t)K * pl->nwords));
        BPG_HIP(ctx, pl->maxsize.ensure(K));
    }
    // the fast forms and their limits (the split-FP16 form up to 32 basis
    // columns, the others up to 16); the FP64 form otherwise and for the
    // repeats
    const bool frows = allow_fast && d.precision != BPG_PREC_FP64 && m.nrows[map] <= BPG_FL_MAX_FROWS;
    const int nb = m.pq <= 16 ? 1 : 2;
    const bool hx = frows && d.precision == BPG_PREC_SPLIT_F16 && m.pq <= 32 &&
                    T <= (nb == 1 ? bpg::fl::HX_T_MAX : bpg::fl::HX_T_MAX2);
    const bool fast16 = frows && !hx && m.pq <= 16 && T <= bpg::fl::MX_T_MAX;
    const bool bx = fast16 && (d.precision == BPG_PREC_SPLIT_BF16 || d.precision == BPG_PREC_SPLIT_F16);
    const bool mx = fast16 && !bx;
    const bool fast = hx || fast16;
    const int W = hx ? 16 * nb : 16;
    *fast_used = fast;
    const std::vector<double>& Arows = m.A[map];
    BPG_HIP(ctx, hipMemcpyAsync(pl->A.p, Arows.data(), Arows.size() * sizeof(double), hipMemcpyHostToDevice, st));
    std::vector<double> rot;
    std::vector<float> q32, u32;
    std::vector<unsigned short> qb;
    if (fast) {
        const int Tq = (T + 3) & ~3, Tq32 = (T + 31) & ~31;
        rotated_basis(m, map, Tq32, W, rot, u32);
        if (hx) {
            split_basis_f16(rot, T, nb, qb);
            BPG_HIP(ctx, hipMemcpyAsync(pl->Qh.p, qb.data(), qb.size() * sizeof(unsigned short), hipMemcpyHostToDevice,
                                        st));
        } else if (bx) {
            split_basis(rot, Tq32, qb);
            BPG_HIP(ctx, hipMemcpyAsync(pl->Qb.p, qb.data(), qb.size() * sizeof(unsigned short), hipMemcpyHostToDevice,
                                        st));
        } else {
            q32.assign((size_t)Tq * 16, 0.0f);
            for (size_t i = 0; i < q32.size(); i++) q32[i] = (float)rot[i];
            BPG_HIP(ctx, hipMemcpyAsync(pl->Q32.p, q32.data(), q32.size() * sizeof(float), hipMemcpyHostToDevice, st));
        }
        BPG_HIP(ctx, hipMemcpyAsync(pl->u32.p, u32.data(), W * sizeof(float), hipMemcpyHostToDevice, st));
        BPG_HIP(ctx, pl->flags.ensure(K));
        BPG_HIP(ctx, pl->h_flags.ensure(K));
        BPG_HIP(ctx, hipMemsetAsync(pl->flags.p, 0, K, st));
    }
    BPG_HIP(ctx, hipMemcpyAsync(pl->perm.p, pl->h_perm.p, (size_t)K * Tp * sizeof(unsigned short),
                                hipMemcpyHostToDevice, st));
    bpg::fl::PermArgs a{};
    a.w32 = pl->w32.p;
    a.ar = pl->ar.p;
    a.perm = pl->perm.p;
    a.Q = pl->Q.p;
    a.A = pl->A.p;
    a.u = pl->u.p;
    a.partial = pl->partial.p;
    a.stats = want_stats ? pl->stats.p : nullptr;
    a.bits = cluster ? pl->bits.p : nullptr;
    a.cdt = d.cdt;
    a.T = T; a.Tp = Tp; a.pq = m.pq; a.nrows = m.nrows[map]; a.ftest = m.ftest ? 1 : 0;
    a.mean = m.const_in_span ? 0 : 1;
    a.K = (int)K; a.Vc = Vc; a.Vpad = pl->Vpad; a.nwords = pl->nwords;
    a.n_vtiles = n_vtiles; a.n_vchunks = n_vchunks; a.tiles_per_chunk = tpc; a.n_pgroups = n_pgroups;
    if (fast) {
        a.Q32 = pl->Q32.p;
        a.Qb = reinterpret_cast<const unsigned*>(pl->Qb.p);
        a.Qh = reinterpret_cast<const unsigned*>(pl->Qh.p);
        a.wscale = pl->wscale.p;
        a.neg_one = -1.0f;
        a.u32 = pl->u32.p;
        a.flags = pl->flags.p;
    }
    BPG_HIP(ctx, hipEventRecord(pl->ev[0], st));
    BPG_HIP(ctx, hx   ? bpg::fl::launch_perm_hx(a, st)
                 : bx ? bpg::fl::launch_perm_bx(a, st)
                 : mx ? bpg::fl::launch_perm_mx(a, st)
                      : bpg::fl::launch_perm(a, st));
    BPG_HIP(ctx, hipEventRecord(pl->ev[1], st));
    if (fast) BPG_HIP(ctx, hipMemcpyAsync(pl->h_flags.p, pl->flags.p, K, hipMemcpyDeviceToHost, st));
    if (cluster) {
        bpg::CclArgs cc{};
        cc.bits = pl->bits.p;
        cc.nbr = pl->d_nbr.p;
        cc.values = mass ? pl->stats.p : nullptr;
        cc.Vc = Vc;
        cc.nwords = pl->nwords;
        cc.connectivity = d.connectivity;
        const int rc = bpg::cluster_max(ctx, pl->cl, cc, K, pl->maxsize.p, st);
        if (rc != BPG_OK) return rc;
        BPG_HIP(ctx, hipEventRecord(pl->ev[2], st));
        BPG_HIP(ctx, hipMemcpyAsync(pl->h_max.p, pl->maxsize.p, K * sizeof(unsigned), hipMemcpyDeviceToHost, st));
    } else {
        BPG_HIP(ctx, bpg::launch_reduce_max(pl->partial.p, n_vchunks, (int)K, pl->maxout.p, st));
        BPG_HIP(ctx, hipEventRecord(pl->ev[2], st));
        BPG_HIP(ctx, hipMemcpyAsync(pl->h_max.p, pl->maxout.p, K * sizeof(float), hipMemcpyDeviceToHost, st));
    }
    if (stats) BPG_HIP(ctx, hipMemcpyAsync(stats, pl->stats.p, (size_t)K * Vc * sizeof(float), hipMemcpyDeviceToHost, st));
    BPG_HIP(ctx, hipStreamSynchronize(st));
    const unsigned* host = reinterpret_cast<const unsigned*>(pl->h_max.p);
    for (long k = 0; k < K; k++) {
        if (cluster) {
            // CalculateLargestCluster, LIB:18453-18460: extent, or mass / 1e4
            maxdist[k] = mass ? (float)host[k] / 10000.0f : (float)host[k];
        } else if (Vc == 0 || !(pl->h_max.p[k] > -INFINITY)) {
            maxdist[k] = bpg::empty_max();
        } else {
            maxdist[k] = (float)((float)bpg::quantize(pl->h_max.p[k]) / 10000.0f);
        }
    }
    const auto t1 = std::chrono::steady_clock::now();
    bpg_timing& tm = pl->timing;
    tm = bpg_timing{};
    tm.phase_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
    tm.transform_ms = pl->setup_ms;
    tm.contract_ms = bpg::elapsed(pl->ev[0], pl->ev[1]);
    tm.cluster_ms = cluster ? bpg::elapsed(pl->ev[1], pl->ev[2]) : 0.0;
    tm.contract_launches = 1;
    tm.perm_voxels = K * (long)Vc;
    tm.fit_columns = d.p;
    tm.slots_per_perm = fast ? W : m.pq;
    tm.precision_used = hx ? BPG_PREC_SPLIT_F16 : bx ? BPG_PREC_SPLIT_BF16 : mx ? BPG_PREC_FP32 : BPG_PREC_FP64;
    tm.fp32_permutations = fast ? K : 0;
    return BPG_OK;
}

}  // namespace

extern "C" int bpg_fl_plan_permutation_phase(bpg_fl_plan* pl, int map, const unsigned short* perm, long K,
                                             float* maxdist, float* stats) {
    if (!pl) return BPG_ERR_ARG;
    const auto t0 = std::chrono::steady_clock::now();
    bool fast = false;
    int rc = fl_phase(pl, map, perm, K, maxdist, stats, true, &fast);
    if (rc != BPG_OK || !fast) return rc;
    // permutations with an ill-conditioned sample: repeated in FP64, their
    // results replace the FP32 ones
    std::vector<long> S;
    for (long k = 0; k < K; k++)
        if (pl->h_flags.p[k]) S.push_back(k);
    if (S.empty()) return BPG_OK;
    const bpg_timing keep = pl->timing;
    const int T = pl->d.T, Vc = pl->Vc;
    const long Ks = (long)S.size();
    std::vector<unsigned short> sp((size_t)Ks * T);
    for (long i = 0; i < Ks; i++) std::copy(perm + S[i] * T, perm + (S[i] + 1) * T, sp.begin() + i * T);
    std::vector<float> smd(Ks), sst(stats ? (size_t)Ks * Vc : 0);
    bool unused = false;
    rc = fl_phase(pl, map, sp.data(), Ks, smd.data(), stats ? sst.data() : nullptr, false, &unused);
    if (rc != BPG_OK) return rc;
    for (long i = 0; i < Ks; i++) {
        maxdist[S[i]] = smd[i];
        if (stats) std::copy(sst.begin() + i * Vc, sst.begin() + (i + 1) * Vc, stats + S[i] * Vc);
    }
    pl->timing = keep;
    pl->timing.unsafe_waves = (int)Ks;
    pl->timing.fp32_permutations = K - Ks;
    pl->timing.phase_ms =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return BPG_OK;
}

extern "C" int bpg_fl_plan_last_timing(const bpg_fl_plan* pl, bpg_timing* out) {
    if (!pl || !out) return BPG_ERR_ARG;
    *out = pl->timing;
    return BPG_OK;
}
