// bpg_capi.cpp -- context, plan and permutation-phase entry points of the
// C ABI (include/bpg.h).  Host code: validates, builds the FP64 models,
// owns device buffers and one HIP stream per plan, and enqueues the
// kernels of bpg_kernels.hip.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <climits>
#include <cmath>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "bpg_internal.h"
#include "bpg_kernels.h"
#include "bpg_runtime.h"

// the least chance of an ill-conditioned sample per permutation for which a
// max-T-only phase keeps a sample list
#ifndef BPG_LIST_MIN_CHANCE
#define BPG_LIST_MIN_CHANCE 0.01
#endif

namespace {

using bpg::DBuf;
using bpg::HBuf;
using bpg::quantize;
using bpg::empty_max;
using bpg::elapsed;

int pow2ceil(int x) {
    int r = 1;
    while (r < x) r <<= 1;
    return r;
}

// Accumulator slots per permutation in the split-BF16 layout: a lane holds 16
// slots, i.e. 16/PP whole permutations.  3 and 5 regressors fit 5 and 3
// permutations per lane unpadded (15 of 16 slots); other counts round up to
// a power of two, which fits as many permutations as any value between.
int bf16_slots(int pfit) { return (pfit == 3 || pfit == 5) ? pfit : pow2ceil(pfit); }

// FP32 operand row count for N subjects: multiples of 4 up to 64, then of 16.
int f32_rows(int N) { return N <= 64 ? (N + 3) / 4 * 4 : (N + 15) / 16 * 16; }

// The autotuner's bucket of a split-BF16 contraction.
bpg::TuneKey tune_key(int KK, int PP, bool sign, bool ftest, bool out, int Vpad) {
    int vb = 0;
    while ((1 << vb) < Vpad) vb++;
    return bpg::TuneKey(KK, PP, sign ? 1 : 0, ftest ? 1 : 0, out ? 1 : 0, vb);
}

// The split-BF16 grid: the candidate bpg_plan_tune has in force, else the
// tuned entry of this bucket, else the planner's default.  The alternative
// subtile count exists for the max-T-only kernels only.
// (a contraction with a sample list runs the default subtile count: excl)
bpg::Bf16Launch bf16_launch(const bpg_ctx* ctx, long K, int PP, int KK, int Vpad, bool sign, bool ftest, bool out,
                            bool excl = false) {
    bpg::TuneEntry e;
    if (ctx->tune_forced) {
        e = ctx->tune_force;
    } else {
        const auto it = ctx->tune.find(tune_key(KK, PP, sign, ftest, out, Vpad));
        if (it != ctx->tune.end()) e = it->second;
    }
    const int variant = out || excl ? 0 : e.variant;
    const int slots = ctx->cus / 8 * bpg::contract_bf16_blocks_per_cu(KK, PP, sign, ftest, out, variant);
    bpg::Bf16Launch L = bpg::plan_contract_bf16((int)K, PP, KK, Vpad, sign, slots, variant, e.chunk_pow);
    if (out && L.n_vt > 32) {
        // chunks of whole 32-tile groups: the bitmask runs a block stores
        // then never share a 128-byte run with another block's
        L.vt_per_chunk = std::min(L.n_vt, (L.vt_per_chunk + 31) / 32 * 32);
        L.n_vchunks = (L.n_vt + L.vt_per_chunk - 1) / L.vt_per_chunk;
        L.per_xcd = (L.n_mgroups * L.n_vchunks + 7) / 8;
    }
    return L;
}

}  // namespace

struct bpg_plan {
    bpg_ctx* ctx = nullptr;
    bpg_study_desc d{};
    int nmaps = 1;
    std::vector<bpg::MapModel> models;
    bpg::ObsModel obs;
    int Vc = 0, Vpad = 0, nwords = 0;
    size_t G = 0;
    std::vector<unsigned> comp2grid;
    hipStream_t stream = nullptr;
    hipEvent_t ev[6] = {};
    // the permutations' upload runs on a stream of its own, beside the
    // nuisance transform; the operand build waits for it
    hipStream_t copy = nullptr;
    hipEvent_t copied = nullptr;
    bpg_timing timing{};
    bool have_obs = false;
    // cluster inference: per map, the tied maps of its last permutation phase
    // (cdt_tie_maps), and the per-permutation tie flags of the last phase
    std::vector<long> map_ties, last_K;
    std::vector<unsigned char> last_ties;
    // the phase enqueued on `stream` and not yet finished
    struct Pending {
        bool active = false;
        long K = 0;
        bool cluster = false, mass = false, tfce = false;
        int form = BPG_PREC_FP32;  // arithmetic of this run (BPG_PREC_*)
        bool mx = false;           // FP32 on the matrix cores
        bool exact = false;        // split-BF16 with every product exact (three data parts)
        int map = 0;
        long fp32_perms = 0;  // permutations whose results come from the FP32 form
        long fp64_perms = 0;  // ... from the FP64 form
        bool have_times = false;  // kernel times kept from before a repeat
        double t_transform = 0, t_operand = 0, t_contract = 0, t_cluster = 0;
        const unsigned short* perm = nullptr;
        const float* sign = nullptr;
        float* stats = nullptr;
        int pfit = 0, PP = 0, launches = 0;
        unsigned long long sample_cap = 0;  // the sample list's per-segment capacity (0: no list)
        long fp64_samples = 0;              // samples recomputed from the list
        std::chrono::steady_clock::time_point t0;
    } pending;

    DBuf<float> Yc;
    DBuf<int> d_nbr;
    DBuf<double> yy, Qt, Qn, dscale;
    DBuf<unsigned> Yb, Gb;
    DBuf<float> Y32, G32;
    DBuf<float> Ym, Gm;
    DBuf<unsigned> Yf8, Ysc, Gf8, Gsc;
    DBuf<double> Y64;
    DBuf<unsigned char> ties;
    HBuf<unsigned char> h_ties;
    DBuf<unsigned short> perm;
    DBuf<float> sign;
    DBuf<float> partial, maxout;
    DBuf<float> stats;
    DBuf<unsigned> bits;
    DBuf<unsigned> maxsize, labels;
    bpg::ClusterScratch cl;
    DBuf<float> tfce, sqrt_size;
    DBuf<int> nlev;
    DBuf<unsigned> unsafe;
    DBuf<unsigned char> flags;
    DBuf<unsigned long long> samples, n_samples;
    HBuf<unsigned long long> h_nsamples;
    DBuf<double> oX, oXp, oC, osc;
    DBuf<float> d_obs, sorted, P, glm_out;
    DBuf<unsigned> d_pties;
    // per-map models on the device, [Qt | Qn (NR x QB, row-major) | dscale]
    // at mdev_off[map] (uploaded once at creation; the no-allocation phase
    // reads them)
    DBuf<double> mdev;
    std::vector<size_t> mdev_off;
    HBuf<double> h_model;
    HBuf<unsigned short> h_perm;
    HBuf<float> h_sign, h_max;
    HBuf<int> h_nlev;
    HBuf<unsigned> h_unsafe;
    HBuf<unsigned char> h_flags;

    void release_all() {
        h_model.release(); h_perm.release(); h_sign.release(); h_max.release(); h_nlev.release(); h_unsafe.release(); h_flags.release();
        h_ties.release(); Y64.release(); ties.release(); Ym.release(); Gm.release();
        Yf8.release(); Ysc.release(); Gf8.release(); Gsc.release();
        Yc.release(); d_nbr.release();
        yy.release(); Qt.release(); Qn.release(); dscale.release();
        Yb.release(); Gb.release(); Y32.release(); G32.release();
        perm.release(); sign.release(); partial.release(); maxout.release(); stats.release(); bits.release();
        cl.release(); maxsize.release(); labels.release(); tfce.release(); sqrt_size.release(); nlev.release(); unsafe.release(); flags.release();
        samples.release(); n_samples.release(); h_nsamples.release();
        oX.release(); oXp.release(); oC.release(); osc.release();
        d_obs.release(); sorted.release(); P.release(); glm_out.release(); d_pties.release(); mdev.release();
    }
};

extern "C" int bpg_create(int device, bpg_ctx** out) {
    if (!out) return BPG_ERR_ARG;
    *out = nullptr;
    bpg_ctx* c = new (std::nothrow) bpg_ctx();
    if (!c) return BPG_ERR_NOMEM;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) {
        delete c;
        return BPG_ERR_NODEVICE;
    }
    if (device < 0 || device >= count) {
        delete c;
        return BPG_ERR_NODEVICE;
    }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) {
        delete c;
        return BPG_ERR_HIP;
    }
    c->device = device;
    c->arch = prop.gcnArchName;
    c->lds_per_block = prop.sharedMemPerBlockOptin ? prop.sharedMemPerBlockOptin : prop.sharedMemPerBlock;
    if (prop.multiProcessorCount > 0) c->cus = prop.multiProcessorCount;
    c->name = prop.name[0] ? std::string(prop.name) : "AMD GPU " + c->arch;
    if (c->arch.compare(0, 6, "gfx950") != 0) {
        delete c;
        return BPG_ERR_UNSUPPORTED;
    }
    // the no-allocation entry point's scratch: its address on this device,
    // taken with the device current
    {
        hipError_t e = hipSetDevice(device);
        if (e == hipSuccess) e = bpg::scratch_address(device, &c->scratch);
        if (e != hipSuccess) {
            c->scratch = nullptr;
            c->scratch_err = std::string("device scratch unavailable: ") + hipGetErrorString(e);
        }
    }
    // tuned shapes cached for this device (bpg_tune_save)
    if (const char* tc = std::getenv("BPG_TUNE_CACHE")) (void)bpg_tune_load(c, tc);
    *out = c;
    return BPG_OK;
}

extern "C" void bpg_destroy(bpg_ctx* ctx) { delete ctx; }

extern "C" const char* bpg_last_error(const bpg_ctx* ctx) { return ctx ? ctx->err.c_str() : "no context"; }

extern "C" const char* bpg_device_name(const bpg_ctx* ctx) { return ctx ? ctx->name.c_str() : ""; }

namespace {

// The argument checks of bpg_plan_create (everything but the data).
int check_study_args(const bpg_study_desc& d, std::string& err) {
    auto fail = [&](const char* msg) { err = msg; return BPG_ERR_ARG; };
    if (d.test != BPG_TTEST && d.test != BPG_FTEST && d.test != BPG_GROUP_MEAN) return fail("unknown statistical test");
    if (d.inference != BPG_VOXEL && d.inference != BPG_CLUSTER_EXTENT && d.inference != BPG_CLUSTER_MASS &&
        d.inference != BPG_TFCE)
        return fail("unknown inference mode");
    if (d.precision != BPG_PREC_FP32 && d.precision != BPG_PREC_SPLIT_BF16 && d.precision != BPG_PREC_FP64 &&
        d.precision != BPG_PREC_FP32_VALU && d.precision != BPG_PREC_SPLIT_FP8)
        return fail("unknown precision");
    if (d.connectivity != 26 && d.connectivity != 6) return fail("connectivity must be 26 or 6");
    if (d.W <= 0 || d.H <= 0 || d.D <= 0) return fail("grid dimensions must be positive");
    if ((double)d.W * d.H * d.D * 3.0 >= 4294967295.0) return fail("grid too large for 32-bit labels");
    if (!d.Y || !d.mask || !d.X || !d.C) return fail("null input pointer");
    const int p = d.test == BPG_GROUP_MEAN ? 1 : d.p;
    if (p < 1 || p > BPG_MAX_REGRESSORS) return fail("number of regressors must be 1..25");
    if (d.nc < 1) return fail("number of contrasts must be positive");
    if (d.test == BPG_FTEST && (d.nc > BPG_MAX_FROWS || d.nc > p)) return fail("F contrast must have 1..min(10, p) rows");
    if (d.N <= p || d.N > BPG_MAX_SUBJECTS) return fail("need p < N <= 256 subjects");
    if (d.test == BPG_FTEST && d.N <= p + 1) return fail("F test needs N > p + 1");
    return BPG_OK;
}

// A NaN or infinite value inside the mask makes that voxel's statistic NaN
// for every permutation.  The reference then converts NaN with
// (int)(T * 1e4) (KM:299), whose result depends on the device (x86:
// INT_MIN, which never wins the max; AMD GPUs: 0), so there is no result to
// reproduce: such data are refused.  (Finite data whose residualised values
// are all zero, e.g. a voxel constant over subjects with the intercept in
// the nuisance space, give 0/0 = NaN as well; the max leaves those samples
// out and they never exceed a threshold, which is what comparisons with
// NaN give in the reference.)
const char* kNonFinite = "the data hold a NaN or infinite value inside the mask";

}  // namespace

extern "C" int bpg_validate_study(const bpg_study_desc* desc, char* err, int err_len) {
    std::string e;
    int rc = BPG_ERR_ARG;
    if (!desc) {
        e = "null study";
    } else if ((rc = check_study_args(*desc, e)) == BPG_OK) {
        const bpg_study_desc& d = *desc;
        const size_t G = (size_t)d.W * d.H * d.D;
        std::vector<unsigned> c2g;
        std::vector<int> g2c;
        bpg::compact_mask(d.mask, G, c2g, g2c);
        for (int s = 0; s < d.N && rc == BPG_OK; s++)
            for (unsigned g : c2g)
                if (!std::isfinite(d.Y[(size_t)s * G + g])) {
                    e = kNonFinite;
                    rc = BPG_ERR_ARG;
                    break;
                }
        bpg_study_desc dm = d;
        if (d.test == BPG_GROUP_MEAN) dm.p = dm.nc = 1;
        const int nmaps = d.test == BPG_TTEST ? d.nc : 1;
        for (int m = 0; m < nmaps && rc == BPG_OK; m++) {
            bpg::MapModel mm;
            rc = bpg::build_map_model(dm, m, mm, e);
        }
    }
    if (err && err_len > 0) {
        std::strncpy(err, e.c_str(), (size_t)err_len - 1);
        err[err_len - 1] = '\0';
    }
    return rc;
}

extern "C" int bpg_plan_create(bpg_ctx* ctx, const bpg_study_desc* desc, bpg_plan** out) {
    if (!ctx || !desc || !out) return BPG_ERR_ARG;
    *out = nullptr;
    const bpg_study_desc& d = *desc;
    {
        const int rc = check_study_args(d, ctx->err);
        if (rc != BPG_OK) return rc;
    }
    const int p = d.test == BPG_GROUP_MEAN ? 1 : d.p;

    BPG_HIP(ctx, hipSetDevice(ctx->device));
    bpg_plan* pl = new (std::nothrow) bpg_plan();
    if (!pl) return BPG_ERR_NOMEM;
    pl->ctx = ctx;
    pl->d = d;
    pl->d.p = p;
    if (d.test == BPG_GROUP_MEAN) pl->d.nc = 1;
    pl->d.Y = pl->d.mask = pl->d.X = pl->d.C = nullptr;  // not retained
    pl->nmaps = d.test == BPG_TTEST ? d.nc : 1;
    // models (FP64, host)
    bpg_study_desc dm = d;
    dm.p = p;
    if (d.test == BPG_GROUP_MEAN) dm.nc = 1;
    pl->models.resize(pl->nmaps);
    for (int m = 0; m < pl->nmaps; m++) {
        const int rc = bpg::build_map_model(dm, m, pl->models[m], ctx->err);
        if (rc) { delete pl; return rc; }
    }
    {
        const int rc = bpg::build_obs_model(dm, pl->obs, ctx->err);
        if (rc) { delete pl; return rc; }
    }
    // the split-BF16 layout holds at most 16 fit columns per permutation
    // (one lane's accumulator slots); wider designs run the FP32 form
    int max_pfit = 0;
    for (const auto& m : pl->models) max_pfit = std::max(max_pfit, m.pfit);
    if ((d.precision == BPG_PREC_SPLIT_BF16 || d.precision == BPG_PREC_SPLIT_FP8) && max_pfit > 16)
        pl->d.precision = BPG_PREC_FP32;
    // compaction: in-mask voxels in linear order (mask == 1.0f, K2:1586)
    pl->G = (size_t)d.W * d.H * d.D;
    std::vector<int> grid2comp;
    bpg::compact_mask(d.mask, pl->G, pl->comp2grid, grid2comp);
    pl->Vc = (int)pl->comp2grid.size();
    pl->Vpad = std::max(32, (pl->Vc + 255) / 256 * 256);
    pl->nwords = pl->Vpad / 32;
    std::vector<float> yc((size_t)d.N * std::max(pl->Vc, 1));
    bool finite = true;
    for (int s = 0; s < d.N; s++) {
        const float* src = d.Y + (size_t)s * pl->G;
        float* dst = yc.data() + (size_t)s * pl->Vc;
        for (int v = 0; v < pl->Vc; v++) {
            dst[v] = src[pl->comp2grid[v]];
            finite &= std::isfinite(dst[v]);
        }
    }
    if (!finite) {
        ctx->err = kNonFinite;
        delete pl;
        return BPG_ERR_ARG;
    }
    auto fail = [&](int rc) {
        pl->release_all();
        if (pl->stream) (void)hipStreamDestroy(pl->stream);
        if (pl->copy) (void)hipStreamDestroy(pl->copy);
        if (pl->copied) (void)hipEventDestroy(pl->copied);
        delete pl;
        return rc;
    };
    if (hipStreamCreate(&pl->stream) != hipSuccess || hipStreamCreate(&pl->copy) != hipSuccess) {
        ctx->err = "hipStreamCreate failed";
        return fail(BPG_ERR_HIP);
    }
    if (hipEventCreateWithFlags(&pl->copied, hipEventDisableTiming) != hipSuccess) {
        ctx->err = "hipEventCreate failed";
        return fail(BPG_ERR_HIP);
    }
    for (auto& e : pl->ev)
        if (hipEventCreate(&e) != hipSuccess) { ctx->err = "hipEventCreate failed"; return fail(BPG_ERR_HIP); }
    // cluster inference: compact indices of the 13 lower-index 26-neighbours
    // of every in-mask voxel (face neighbours first), -1 outside the mask
    std::vector<int> nbr;
    if (d.inference != BPG_VOXEL) bpg::build_neighbours(d.W, d.H, d.D, pl->comp2grid, grid2comp, nbr);
    if (pl->Yc.ensure(yc.size()) || (!nbr.empty() && pl->d_nbr.ensure(nbr.size()))) {
        ctx->err = "device allocation of the study failed";
        return fail(BPG_ERR_NOMEM);
    }
    if (hipMemcpy(pl->Yc.p, yc.data(), yc.size() * sizeof(float), hipMemcpyHostToDevice) ||
        (!nbr.empty() && hipMemcpy(pl->d_nbr.p, nbr.data(), nbr.size() * sizeof(int), hipMemcpyHostToDevice))) {
        ctx->err = "upload of the study failed";
        return fail(BPG_ERR_HIP);
    }
    if (pl->Vc > 0) {
        // every map's model on the device, and the data-side buffers of the
        // plan's form at their largest size: a phase then allocates only
        // what depends on K (and the no-allocation phase nothing)
        const int N = d.N, NR = (N + 15) / 16 * 16, KK = (N + 15) / 16, KG = (N + 7) / 8;
        std::vector<double> hm;
        size_t yb = 0;
        int max_fit = 0;
        for (const auto& m : pl->models) max_fit = std::max(max_fit, m.pfit);
        for (const auto& m : pl->models) {
            const int QB = std::max(bpg::transform_width(m.q), 1);
            pl->mdev_off.push_back(hm.size());
            hm.insert(hm.end(), m.Qt.begin(), m.Qt.end());
            const size_t q0 = hm.size();
            hm.resize(q0 + (size_t)NR * QB + NR, 0.0);
            for (int j = 0; j < m.q; j++)
                for (int s = 0; s < N; s++) hm[q0 + (size_t)s * QB + j] = m.Qn[(size_t)j * N + s];
            std::copy(m.dscale.begin(), m.dscale.end(), hm.begin() + q0 + (size_t)NR * QB);
            yb = std::max(yb, (size_t)(pl->Vpad / 32) * KK * bpg::bf16_y_pieces(KK, m.sign_operand) * 64 * 4);
        }
        if (pl->mdev.ensure(hm.size()) || pl->yy.ensure(pl->Vpad) ||
            (pl->d.precision == BPG_PREC_SPLIT_BF16 && pl->Yb.ensure(yb)) ||
            (pl->d.precision == BPG_PREC_FP32 && max_fit <= 16 &&
             pl->Ym.ensure((size_t)(pl->Vpad / 32) * KG * 64 * 4)) ||
            ((pl->d.precision == BPG_PREC_FP32_VALU || pl->d.precision == BPG_PREC_SPLIT_FP8 ||
              (pl->d.precision == BPG_PREC_FP32 && max_fit > 16)) &&
             pl->Y32.ensure((size_t)f32_rows(N) * pl->Vpad))) {
            ctx->err = "device allocation of the study failed";
            return fail(BPG_ERR_NOMEM);
        }
        if (hipMemcpy(pl->mdev.p, hm.data(), hm.size() * sizeof(double), hipMemcpyHostToDevice)) {
            ctx->err = "upload of the model failed";
            return fail(BPG_ERR_HIP);
        }
    }
    *out = pl;
    return BPG_OK;
}

extern "C" void bpg_plan_destroy(bpg_plan* pl) {
    if (!pl) return;
    (void)hipSetDevice(pl->ctx->device);
    if (pl->stream) (void)hipStreamSynchronize(pl->stream);
    if (pl->copy) (void)hipStreamSynchronize(pl->copy);
    pl->release_all();
    for (auto& e : pl->ev)
        if (e) (void)hipEventDestroy(e);
    if (pl->copied) (void)hipEventDestroy(pl->copied);
    if (pl->stream) (void)hipStreamDestroy(pl->stream);
    if (pl->copy) (void)hipStreamDestroy(pl->copy);
    delete pl;
}

extern "C" long bpg_plan_num_voxels(const bpg_plan* pl) { return pl ? pl->Vc : -1; }

namespace {

// The regularised incomplete beta function I_x(a, b) from its continued
// fraction (Abramowitz & Stegun 26.5.8), evaluated by the modified Lentz
// method; on the side x > (a+1)/(a+b+2) through I_x(a,b) = 1 - I_{1-x}(b,a).
double inc_beta(double x, double a, double b) {
    if (!(x > 0.0)) return 0.0;
    if (!(x < 1.0)) return 1.0;
    const bool flip = x > (a + 1.0) / (a + b + 2.0);
    if (flip) {
        std::swap(a, b);
        x = 1.0 - x;
    }
    const double front =
        std::exp(std::lgamma(a + b) - std::lgamma(a) - std::lgamma(b) + a * std::log(x) + b * std::log1p(-x)) / a;
    const double tiny = 1e-300;
    auto clamp = [&](double v) { return std::fabs(v) < tiny ? tiny : v; };
    double c = 1.0, d = 1.0 / clamp(1.0 - (a + b) * x / (a + 1.0)), h = d;
    for (int m = 1; m <= 400; m++) {
        const double m2 = 2.0 * m;
        double t = m * (b - m) * x / ((a + m2 - 1.0) * (a + m2));
        d = 1.0 / clamp(1.0 + t * d);
        c = clamp(1.0 + t / c);
        h *= d * c;
        t = -(a + m) * (a + b + m) * x / ((a + m2) * (a + m2 + 1.0));
        d = 1.0 / clamp(1.0 + t * d);
        c = clamp(1.0 + t / c);
        const double step = d * c;
        h *= step;
        if (std::fabs(step - 1.0) < 1e-13) break;
    }
    const double r = front * h;
    return flip ? 1.0 - r : r;
}

// The chance that a permutation of a study meets an ill-conditioned sample
// (SSE < tau |y*|^2) at one of its V voxels, were the data Gaussian noise:
// SSE / |y*|^2 is then Beta(dof / 2, pfit / 2) at every voxel.  Orders the
// batch's work and chooses the sample list; the results meet the accuracy
// gate whatever it says.
double ill_conditioned_chance(double dof, int pfit, long V, double tau) {
    if (dof <= 0.0 || pfit <= 0 || V <= 0) return 0.0;
    const double pv = inc_beta(tau, 0.5 * dof, 0.5 * pfit);
    if (pv >= 1.0) return 1.0;
    return -std::expm1((double)V * std::log1p(-pv));
}


// The per-segment capacity of the sample list of a split-BF16 max-T-only
// phase (its ill-conditioned samples recomputed one by one in FP64), or 0
// for none: a list pays when a permutation is likely to meet such a sample
// (a chance of 1% or more; below that the few flagged permutations are
// repeated whole) and the expected count of them, K V pv for Gaussian noise,
// is modest.  The BPG_SAMPLE_SEGS segments hold 8 times that count, at
// least 256 and at most 32 Ki entries each (64 MiB in all); samples past
// their segment's end flag their permutations for a whole repeat.
// Environment: BPG_SAMPLE_LIST=off|on forces the list off or on (on: where
// the kernels support it), BPG_SAMPLE_CAP sets the per-segment capacity.
unsigned long long sample_list_cap(double dof, int pfit, long V, long K, double tau) {
    const char* mode = std::getenv("BPG_SAMPLE_LIST");
    const char* capenv = std::getenv("BPG_SAMPLE_CAP");
    const unsigned long long forced = capenv ? std::strtoull(capenv, nullptr, 10) : 0ull;
    if (mode && std::strcmp(mode, "off") == 0) return 0ull;
    const bool on = mode && std::strcmp(mode, "on") == 0;
    if (dof <= 0.0 || pfit <= 0 || V <= 0 || K <= 0) return 0ull;
    const double pv = inc_beta(tau, 0.5 * dof, 0.5 * pfit);
    const double expected = pv * (double)V * (double)K;
    const bool pays = ill_conditioned_chance(dof, pfit, V, tau) >= BPG_LIST_MIN_CHANCE && expected <= (double)(1 << 20);
    if (!on && !pays) return 0ull;
    if (forced) return forced;
    const double want = std::min(8.0 * expected / bpg::BPG_SAMPLE_SEGS, 32768.0);
    return std::max<unsigned long long>(256ull, (unsigned long long)want);
}

}  // namespace

namespace {

// Everything of one permutation phase up to the copies of its results into
// pinned host memory, enqueued on the plan's stream; phase_finish waits and
// quantises.  Between the two the host is free (to enqueue other plans).
// form: the arithmetic (BPG_PREC_SPLIT_BF16, _FP32, _FP64), or -1 for the
// plan's own.
int phase_enqueue(bpg_plan* pl, int map, const unsigned short* perm, const float* sign, long K, float* stats,
                  int form = -1) {
    bpg_ctx* ctx = pl->ctx;
    const bpg_study_desc& d = pl->d;
    if (map < 0 || map >= pl->nmaps) BPG_FAIL(ctx, BPG_ERR_ARG, "statistical map index out of range");
    if (K <= 0 || K > (1L << 26)) BPG_FAIL(ctx, BPG_ERR_ARG, "number of permutations out of range");
    const bpg::MapModel& m = pl->models[map];
    if (m.sign_operand ? !sign : !perm) BPG_FAIL(ctx, BPG_ERR_ARG, "missing permutation / sign matrix");
    const int N = d.N;
    if (pl->Vc == 0) {
        // empty mask: nothing to launch; the null is the reference's empty
        // max (phase_finish) or an empty largest cluster
        if (!m.sign_operand)
            for (size_t i = 0; i < (size_t)K * N; i++)
                if (perm[i] >= N) BPG_FAIL(ctx, BPG_ERR_ARG, "permutation index out of range");
        BPG_HIP(ctx, pl->h_max.ensure(K));
        BPG_HIP(ctx, pl->h_unsafe.ensure(1));
        const bool clu = d.inference == BPG_CLUSTER_EXTENT || d.inference == BPG_CLUSTER_MASS;
        for (long k = 0; k < K; k++) {
            if (clu) reinterpret_cast<unsigned*>(pl->h_max.p)[k] = 0u;
            else pl->h_max.p[k] = -INFINITY;
        }
        *pl->h_unsafe.p = 0;
        auto& ps = pl->pending;
        ps = bpg_plan::Pending{};
        ps.active = true;
        ps.map = map;
        ps.K = K;
        ps.cluster = clu;
        ps.mass = d.inference == BPG_CLUSTER_MASS;
        ps.tfce = d.inference == BPG_TFCE;
        ps.form = d.precision;
        ps.have_times = true;
        ps.pfit = m.pfit;
        ps.perm = perm;
        ps.sign = sign;
        ps.t0 = std::chrono::steady_clock::now();
        return BPG_OK;
    }
    BPG_HIP(ctx, hipSetDevice(ctx->device));
    const auto t0 = std::chrono::steady_clock::now();
    hipStream_t st = pl->stream;
    if (form < 0) form = d.precision;
    const bool bf = form == BPG_PREC_SPLIT_BF16;
    const bool f64 = form == BPG_PREC_FP64;
    const bool cluster = d.inference == BPG_CLUSTER_EXTENT || d.inference == BPG_CLUSTER_MASS;
    // cluster mass and TFCE need the statistic of every voxel: the
    // contraction then also writes the statistic values
    const bool mass = d.inference == BPG_CLUSTER_MASS;
    const bool tfce = d.inference == BPG_TFCE;
    const bool want_stats = stats != nullptr || mass || tfce;
    const int pfit = m.pfit;
    const int KK = (N + 15) / 16;
    const int npy = bpg::bf16_y_pieces(KK, m.sign_operand);
    const int NS = f32_rows(N);
    const int PP = bf16_slots(pfit);
    const int Vc = pl->Vc, Vpad = pl->Vpad;
    // split-BF16 max-T of a t test whose permutations are likely to meet
    // ill-conditioned samples: those samples go to a list for FP64
    const unsigned long long scap =
        bf && !want_stats && !cluster && !tfce && !m.sign_operand && d.test == BPG_TTEST && npy == 2 &&
                KK <= bpg::BPG_EXCL_MAX_KK && pfit <= bpg::BPG_FIX_MAX_P && m.q <= 24
            ? sample_list_cap((double)m.dof, pfit, Vc, K, bpg::BPG_TAU_BF16)
            : 0ull;
    const bpg::Bf16Launch Lb =
        bf ? bf16_launch(ctx, K, PP, KK, Vpad, m.sign_operand, d.test == BPG_FTEST, want_stats || cluster, scap > 0)
           : bpg::Bf16Launch{};
    // FP32 on the matrix cores up to 16 fit columns, else on the vector ALUs
    const bool mx = form == BPG_PREC_FP32 && pfit <= 16;
    const int KG = (N + 7) / 8;
    // split-FP8 (the data split from the FP32 rows of the vector-ALU form)
    const bool f8 = form == BPG_PREC_SPLIT_FP8;
    const int KS = (N + 63) / 64;
    const bpg::Bf16Launch L8 =
        f8 ? bpg::plan_contract_tiles((int)K, PP, bpg::f8_rows_per_wave(KS), Vpad,
                                      ctx->cus / 8 *
                                          bpg::contract_f8_blocks_per_cu(PP, KS, m.sign_operand, d.test == BPG_FTEST,
                                                                         want_stats || cluster))
           : bpg::Bf16Launch{};
    const bpg::Bf16Launch Lm =
        mx ? bpg::plan_contract_tiles((int)K, PP, bpg::f32mx_rows_per_wave(KG), Vpad,
                                      ctx->cus / 8 *
                                          bpg::contract_f32mx_blocks_per_cu(PP, KG, d.test == BPG_FTEST,
                                                                            want_stats || cluster))
           : bpg::Bf16Launch{};
    const bpg::F32Launch Lf = bf || mx || f8 ? bpg::F32Launch{}
                              : f64 ? bpg::plan_contract_f64((int)K, pfit, Vpad)
                                    : bpg::plan_contract_f32((int)K, pfit, Vpad);
    const int n_sub = bf ? Lb.n_sub : f8 ? L8.n_sub : Lm.n_sub;
    const int n_part = bf ? Lb.n_vchunks : mx ? Lm.n_vchunks : f8 ? L8.n_vchunks : Lf.n_vchunks;
    const int PT = bf || f64 || mx || f8 ? 1 : bpg::f32_perm_tile(pfit);
    const long Kpad = (K + PT - 1) / PT * PT;
    const int NS64 = (N + 7) / 8 * 8;  // FP64 data rows (zero beyond N)

    // buffers
    BPG_HIP(ctx, pl->yy.ensure(Vpad));
    // nuisance basis row-major, zero-padded to QB columns and NR rows
    const int QB = bpg::transform_width(m.q);
    const int NR = (N + 15) / 16 * 16;
    if (QB < 0) BPG_FAIL(ctx, BPG_ERR_DESIGN, "more than 24 nuisance regressors");
    if (f64) {
        BPG_HIP(ctx, pl->Qt.ensure((size_t)N * pfit));
        BPG_HIP(ctx, pl->Qn.ensure((size_t)NR * std::max(QB, 1)));
        BPG_HIP(ctx, pl->dscale.ensure(NR));
    }
    if (bf) {
        BPG_HIP(ctx, pl->Yb.ensure((size_t)(Vpad / 32) * KK * npy * 64 * 4));
        BPG_HIP(ctx, pl->Gb.ensure((size_t)n_sub * KK * 2 * 64 * 4));
    } else if (f64) {
        BPG_HIP(ctx, pl->Y64.ensure((size_t)NS64 * Vpad));
    } else if (mx) {
        BPG_HIP(ctx, pl->Ym.ensure((size_t)(Vpad / 32) * KG * 64 * 4));
        BPG_HIP(ctx, pl->Gm.ensure((size_t)n_sub * KG * 64 * 4));
    } else if (f8) {
        BPG_HIP(ctx, pl->Y32.ensure((size_t)NS * Vpad));
        BPG_HIP(ctx, pl->Yf8.ensure((size_t)(Vpad / 32) * KS * BPG_F8_TERMS * 64 * 8));
        BPG_HIP(ctx, pl->Ysc.ensure((size_t)(Vpad / 32) * KS * 64));
        BPG_HIP(ctx, pl->Gf8.ensure((size_t)n_sub * KS * BPG_F8_TERMS * 64 * 8));
        BPG_HIP(ctx, pl->Gsc.ensure((size_t)n_sub * KS * 64));
    } else {
        BPG_HIP(ctx, pl->Y32.ensure((size_t)NS * Vpad));
        BPG_HIP(ctx, pl->G32.ensure((size_t)Kpad * pfit * NS));
    }
    if (m.sign_operand) BPG_HIP(ctx, pl->sign.ensure((size_t)K * N));
    else BPG_HIP(ctx, pl->perm.ensure((size_t)K * N));
    BPG_HIP(ctx, pl->partial.ensure((size_t)n_part * K));
    BPG_HIP(ctx, pl->maxout.ensure(K));
    BPG_HIP(ctx, pl->unsafe.ensure(1));
    BPG_HIP(ctx, pl->h_unsafe.ensure(1));
    BPG_HIP(ctx, pl->flags.ensure(K));
    BPG_HIP(ctx, pl->h_flags.ensure(K));
    if (cluster) {
        BPG_HIP(ctx, pl->ties.ensure(K));
        BPG_HIP(ctx, pl->h_ties.ensure(K));
    }
    if (scap) {
        BPG_HIP(ctx, pl->samples.ensure(scap * bpg::BPG_SAMPLE_SEGS));
        BPG_HIP(ctx, pl->n_samples.ensure(bpg::BPG_SAMPLE_SEGS));
        BPG_HIP(ctx, pl->h_nsamples.ensure(bpg::BPG_SAMPLE_SEGS));
    }
    if (want_stats) BPG_HIP(ctx, pl->stats.ensure((size_t)K * std::max(Vc, 1)));
    if (cluster || tfce) BPG_HIP(ctx, pl->bits.ensure((size_t)K * pl->nwords));

    // pinned staging (free at entry: every phase ends with a stream sync)
    BPG_HIP(ctx, pl->h_max.ensure(K));
    if (m.sign_operand) BPG_HIP(ctx, pl->h_sign.ensure((size_t)K * N));
    else BPG_HIP(ctx, pl->h_perm.ensure((size_t)K * N));
    // the map's model [Qt | Qn | dscale]: the copy uploaded at creation; the
    // FP64 form reads y* itself (its operand carries the scale), so it takes
    // one with unit scales, uploaded here
    const size_t nQt = m.Qt.size();
    const double* dQt = pl->mdev.p + pl->mdev_off[map];
    const double* dQn = dQt + nQt;
    const double* dds = dQn + (size_t)NR * std::max(QB, 1);
    BPG_HIP(ctx, hipEventRecord(pl->ev[0], st));
    // the contraction's counters and flags, cleared ahead of the transform
    // (beside the permutations' staging rather than before the contraction)
    BPG_HIP(ctx, hipMemsetAsync(pl->unsafe.p, 0, sizeof(unsigned), st));
    BPG_HIP(ctx, hipMemsetAsync(pl->flags.p, 0, K, st));
    if (cluster) BPG_HIP(ctx, hipMemsetAsync(pl->ties.p, 0, K, st));
    if (scap)
        BPG_HIP(ctx, hipMemsetAsync(pl->n_samples.p, 0, bpg::BPG_SAMPLE_SEGS * sizeof(unsigned long long), st));
    if (f64) {
        const size_t nQn = (size_t)NR * QB;
        BPG_HIP(ctx, pl->h_model.ensure(nQt + nQn + NR));
        double* hm = pl->h_model.p;
        std::memcpy(hm, m.Qt.data(), nQt * sizeof(double));
        std::fill(hm + nQt, hm + nQt + nQn + NR, 0.0);
        for (int j = 0; j < m.q; j++)
            for (int s = 0; s < N; s++) hm[nQt + (size_t)s * QB + j] = m.Qn[(size_t)j * N + s];
        std::fill(hm + nQt + nQn, hm + nQt + nQn + N, 1.0);
        BPG_HIP(ctx, hipMemcpyAsync(pl->Qt.p, hm, nQt * sizeof(double), hipMemcpyHostToDevice, st));
        if (nQn) BPG_HIP(ctx, hipMemcpyAsync(pl->Qn.p, hm + nQt, nQn * sizeof(double), hipMemcpyHostToDevice, st));
        BPG_HIP(ctx, hipMemcpyAsync(pl->dscale.p, hm + nQt + nQn, NR * sizeof(double), hipMemcpyHostToDevice, st));
        dQt = pl->Qt.p;
        dQn = pl->Qn.p;
        dds = pl->dscale.p;
    }
    {
        bpg::TransformArgs ta{};
        ta.Yc = pl->Yc.p; ta.Qn = dQn; ta.dscale = dds;
        ta.N = N; ta.q = m.q; ta.Vc = Vc; ta.Vpad = Vpad;
        ta.KK = bf ? KK : 0; ta.NS = bf ? 0 : f64 ? NS64 : NS; ta.npy = npy;
        ta.yy = pl->yy.p; ta.Yb = bf ? pl->Yb.p : nullptr; ta.Y32 = bf || f64 || mx ? nullptr : pl->Y32.p;
        ta.Y64 = f64 ? pl->Y64.p : nullptr;
        ta.Ym = mx ? pl->Ym.p : nullptr;
        BPG_HIP(ctx, bpg::launch_transform(ta, st));
        if (f8) BPG_HIP(ctx, bpg::launch_data_f8(pl->Y32.p, NS, N, Vpad, KS, pl->Yf8.p, pl->Ysc.p, st));
    }
    BPG_HIP(ctx, hipEventRecord(pl->ev[1], st));
    // while the transform runs: stage (and check) the permutations, then
    // upload them on the copy stream, in chunks, each uploaded while the
    // next is staged (the staging is free: the last phase's upload finished
    // before its operand build, which its final stream synchronisation
    // waited for)
    const size_t total = (size_t)K * N;
    const size_t chunk = std::max<size_t>((total + 3) / 4, (size_t)1 << 16);
    for (size_t c0 = 0; c0 < total; c0 += chunk) {
        const size_t c1 = std::min(total, c0 + chunk);
        if (m.sign_operand) {
            std::memcpy(pl->h_sign.p + c0, sign + c0, (c1 - c0) * sizeof(float));
            BPG_HIP(ctx, hipMemcpyAsync(pl->sign.p + c0, pl->h_sign.p + c0, (c1 - c0) * sizeof(float),
                                        hipMemcpyHostToDevice, pl->copy));
            continue;
        }
        unsigned short* hp = pl->h_perm.p;
        unsigned short worst = 0;
        for (size_t i = c0; i < c1; i++) {
            hp[i] = perm[i];
            worst = std::max(worst, perm[i]);
        }
        if (worst >= N) {
            // the staged model copies and the chunks in flight must finish
            (void)hipStreamSynchronize(pl->copy);
            (void)hipStreamSynchronize(st);
            BPG_FAIL(ctx, BPG_ERR_ARG, "permutation index out of range");
        }
        BPG_HIP(ctx, hipMemcpyAsync(pl->perm.p + c0, hp + c0, (c1 - c0) * sizeof(unsigned short),
                                    hipMemcpyHostToDevice, pl->copy));
    }
    BPG_HIP(ctx, hipEventRecord(pl->copied, pl->copy));
    BPG_HIP(ctx, hipStreamWaitEvent(st, pl->copied, 0));
    {
        bpg::OperandArgs oa{};
        oa.perm = m.sign_operand ? nullptr : pl->perm.p;
        oa.sign = m.sign_operand ? pl->sign.p : nullptr;
        oa.Qt = dQt;
        oa.N = N; oa.pfit = pfit; oa.K = (int)K; oa.PP = PP; oa.KK = KK; oa.NS = NS;
        if (bf) {
            oa.n_sub = n_sub;
            oa.Gb = pl->Gb.p;
            BPG_HIP(ctx, bpg::launch_operand_bf16(oa, st));
        } else if (f64) {
            // the FP64 kernel builds its operand tiles itself
        } else if (mx) {
            oa.n_sub = n_sub;
            oa.Gm = pl->Gm.p;
            BPG_HIP(ctx, bpg::launch_operand_f32mx(oa, KG, st));
        } else if (f8) {
            oa.n_sub = n_sub;
            BPG_HIP(ctx, bpg::launch_operand_f8(oa, KS, m.sign_operand, pl->Gf8.p, pl->Gsc.p, st));
        } else {
            oa.n_sub = (int)Kpad;
            oa.G32 = pl->G32.p;
            BPG_HIP(ctx, bpg::launch_operand_f32(oa, st));
        }
    }
    BPG_HIP(ctx, hipEventRecord(pl->ev[2], st));
    int launches = 0;
    {
        bpg::ContractArgs ca{};
        ca.Gb = pl->Gb.p; ca.Yb = pl->Yb.p; ca.G32 = pl->G32.p; ca.Y32 = pl->Y32.p; ca.Y64 = pl->Y64.p;
        ca.Gm = pl->Gm.p; ca.Ym = pl->Ym.p;
        ca.Gf8 = pl->Gf8.p; ca.Gsc = pl->Gsc.p; ca.Yf8 = pl->Yf8.p; ca.Ysc = pl->Ysc.p;
        ca.perm = m.sign_operand ? nullptr : pl->perm.p;
        ca.sign = m.sign_operand ? pl->sign.p : nullptr;
        ca.Qt = dQt;
        ca.dof = (double)m.dof;
        ca.yy = pl->yy.p; ca.partial = pl->partial.p;
        ca.stats = want_stats ? pl->stats.p : nullptr;
        ca.bits = cluster ? pl->bits.p : nullptr;
        ca.nwords = pl->nwords;
        ca.K = (int)K; ca.Vc = Vc; ca.Vpad = Vpad; ca.N = N; ca.pfit = pfit; ca.r = m.r; ca.NS = NS;
        ca.ftest = d.test == BPG_FTEST;
        ca.sqrt_dof = (float)std::sqrt((double)m.dof);
        ca.f_scale = (float)((double)m.dof / (double)m.r);
        ca.cdt = d.cdt;
        // with three data fragments every product is exact: the FP32 bound
        ca.tau = f8 ? bpg::BPG_TAU_FP8 : bf && npy == 2 ? bpg::BPG_TAU_BF16 : bpg::BPG_TAU_FP32;
        ca.pinf = INFINITY;
        ca.unsafe = pl->unsafe.p;
        ca.flags = pl->flags.p;
        // the band of the accuracy gate, 1e-4 (1 + |cdt|), around the
        // threshold; none for the FP64 form
        const float band = f64 ? 0.0f : (float)(1e-4 * (1.0 + std::fabs((double)d.cdt)));
        ca.cdt_lo = d.cdt - band;
        ca.cdt_hi = d.cdt + band;
        ca.ties = cluster && !f64 ? pl->ties.p : nullptr;
        ca.samples = scap ? pl->samples.p : nullptr;
        ca.n_samples = scap ? pl->n_samples.p : nullptr;
        ca.sample_cap = scap;
        if (bf) BPG_HIP(ctx, bpg::launch_contract_bf16(ca, Lb, KK, PP, m.sign_operand, st));
        else if (f64) BPG_HIP(ctx, bpg::launch_contract_f64(ca, Lf, st));
        else if (mx) BPG_HIP(ctx, bpg::launch_contract_f32mx(ca, Lm, PP, KG, st));
        else if (f8) BPG_HIP(ctx, bpg::launch_contract_f8(ca, L8, PP, KS, m.sign_operand, st));
        else BPG_HIP(ctx, bpg::launch_contract_f32(ca, Lf, st));
        launches = 1;
        if (!cluster) BPG_HIP(ctx, bpg::launch_reduce_max(pl->partial.p, n_part, (int)K, pl->maxout.p, st));
        if (scap) {
            // the listed samples, recomputed in FP64, into the maxima
            bpg::SampleFixArgs fa{};
            fa.samples = pl->samples.p; fa.n_samples = pl->n_samples.p; fa.cap = scap;
            fa.Yc = pl->Yc.p; fa.Qn = dQn; fa.QB = std::max(QB, 1); fa.q = m.q;
            fa.Qt = dQt; fa.perm = pl->perm.p;
            fa.N = N; fa.pfit = pfit; fa.Vc = Vc; fa.K = (int)K;
            fa.dof = (double)m.dof;
            fa.maxout = pl->maxout.p;
            BPG_HIP(ctx, bpg::launch_fix_samples(fa, st));
        }
    }
    BPG_HIP(ctx, hipEventRecord(pl->ev[3], st));
    // both results are 4-byte words: the cluster maxima are uint sizes /
    // masses, the voxel maxima floats
    unsigned* host = reinterpret_cast<unsigned*>(pl->h_max.p);
    const float* hmax = pl->h_max.p;
    if (cluster) {
        BPG_HIP(ctx, pl->maxsize.ensure(K));
        bpg::CclArgs cc{};
        cc.bits = pl->bits.p;
        cc.nbr = pl->d_nbr.p;
        cc.values = mass ? pl->stats.p : nullptr;
        cc.Vc = Vc; cc.nwords = pl->nwords;
        cc.connectivity = d.connectivity;
        const int rc = bpg::cluster_max(ctx, pl->cl, cc, K, pl->maxsize.p, st);
        if (rc != BPG_OK) return rc;
        BPG_HIP(ctx, hipEventRecord(pl->ev[4], st));
        BPG_HIP(ctx, hipMemcpyAsync(host, pl->maxsize.p, K * sizeof(unsigned), hipMemcpyDeviceToHost, st));
    } else if (tfce && Vc > 0) {
        // threshold-free cluster enhancement (LIB:16756-16762,
        // LIB:18464-18508): the levels 0, 0.2846, ... (float accumulation)
        // up to each map's quantised max; level 0 adds sqrt(size) * 0 * 0
        // and is skipped
        BPG_HIP(ctx, hipMemcpyAsync(pl->h_max.p, pl->maxout.p, K * sizeof(float), hipMemcpyDeviceToHost, st));
        BPG_HIP(ctx, hipStreamSynchronize(st));
        BPG_HIP(ctx, pl->h_nlev.ensure(K));
        int Lmax = 0;
        for (long k = 0; k < K; k++) {
            const float top = pl->h_max.p[k] > -INFINITY ? (float)quantize(pl->h_max.p[k]) / 10000.0f : empty_max();
            int L = 0;
            for (float h = 0.0f; h <= top; h += 0.2846f) L++;
            pl->h_nlev.p[k] = L;
            Lmax = std::max(Lmax, L);
        }
        std::vector<float> levels(std::max(Lmax, 1));
        {
            float h = 0.0f;
            for (int j = 0; j < Lmax; j++, h += 0.2846f) levels[j] = h;
        }
        BPG_HIP(ctx, pl->nlev.ensure(K));
        BPG_HIP(ctx, hipMemcpyAsync(pl->nlev.p, pl->h_nlev.p, K * sizeof(int), hipMemcpyHostToDevice, st));
        const size_t budget = (size_t)8 << 30;
        const size_t per_map = (size_t)Vc * 3 * sizeof(unsigned);
        const long B = std::min<long>(K, (long)std::max<size_t>(1, budget / per_map));
        BPG_HIP(ctx, pl->cl.parent.ensure((size_t)B * Vc));
        BPG_HIP(ctx, pl->cl.sizes.ensure((size_t)B * Vc));
        BPG_HIP(ctx, pl->tfce.ensure((size_t)B * Vc));
        if (pl->sqrt_size.n < (size_t)Vc + 1) {
            // sqrt(n) for every possible component size, correctly rounded
            std::vector<float> tab((size_t)Vc + 1);
            for (size_t n = 0; n <= (size_t)Vc; n++) tab[n] = std::sqrt((float)n);
            BPG_HIP(ctx, pl->sqrt_size.ensure(tab.size()));
            BPG_HIP(ctx, hipMemcpy(pl->sqrt_size.p, tab.data(), tab.size() * sizeof(float), hipMemcpyHostToDevice));
        }
        for (long k0 = 0; k0 < K; k0 += B) {
            const int nb = (int)std::min<long>(B, K - k0);
            BPG_HIP(ctx, hipMemsetAsync(pl->tfce.p, 0, (size_t)nb * Vc * sizeof(float), st));
            bpg::CclArgs cc{};
            cc.bits = pl->bits.p;
            cc.nbr = pl->d_nbr.p;
            cc.parent = pl->cl.parent.p;
            cc.sizes = pl->cl.sizes.p;
            cc.tfce = pl->tfce.p;
            cc.sqrt_size = pl->sqrt_size.p;
            cc.Vc = Vc; cc.nwords = pl->nwords; cc.nmaps = nb;
            cc.connectivity = d.connectivity;
            for (int j = 1; j < Lmax; j++) {
                BPG_HIP(ctx, bpg::launch_level_bits(pl->stats.p + (size_t)k0 * Vc, pl->nlev.p + k0, j, levels[j], nb,
                                                    Vc, Vpad, pl->bits.p, pl->nwords, st));
                cc.level = levels[j];
                BPG_HIP(ctx, bpg::launch_ccl_tfce_level(cc, st));
            }
            BPG_HIP(ctx, bpg::launch_rowmax(pl->tfce.p, nb, Vc, pl->maxout.p + k0, st));
        }
        BPG_HIP(ctx, hipEventRecord(pl->ev[4], st));
        BPG_HIP(ctx, hipMemcpyAsync(pl->h_max.p, pl->maxout.p, K * sizeof(float), hipMemcpyDeviceToHost, st));
    } else {
        BPG_HIP(ctx, hipEventRecord(pl->ev[4], st));
        BPG_HIP(ctx, hipMemcpyAsync(pl->h_max.p, pl->maxout.p, K * sizeof(float), hipMemcpyDeviceToHost, st));
    }
    if (stats)
        BPG_HIP(ctx, hipMemcpyAsync(stats, pl->stats.p, (size_t)K * Vc * sizeof(float), hipMemcpyDeviceToHost, st));
    BPG_HIP(ctx, hipMemcpyAsync(pl->h_unsafe.p, pl->unsafe.p, sizeof(unsigned), hipMemcpyDeviceToHost, st));
    if (scap)
        BPG_HIP(ctx, hipMemcpyAsync(pl->h_nsamples.p, pl->n_samples.p,
                                    bpg::BPG_SAMPLE_SEGS * sizeof(unsigned long long), hipMemcpyDeviceToHost, st));
    if (cluster) BPG_HIP(ctx, hipMemcpyAsync(pl->h_ties.p, pl->ties.p, K, hipMemcpyDeviceToHost, st));
    BPG_HIP(ctx, hipEventRecord(pl->ev[5], st));
    auto& ps = pl->pending;
    ps.active = true;
    ps.map = map;
    ps.fp32_perms = form == BPG_PREC_FP32 || form == BPG_PREC_FP32_VALU ? K : 0;
    ps.fp64_perms = f64 ? K : 0;
    ps.have_times = false;
    ps.perm = perm;
    ps.sign = sign;
    ps.stats = stats;
    ps.K = K;
    ps.cluster = cluster;
    ps.mass = mass;
    ps.tfce = tfce;
    ps.form = form;
    ps.exact = bf && npy == 3;
    ps.pfit = pfit;
    ps.mx = mx || f8;
    ps.PP = PP;
    ps.launches = launches;
    ps.sample_cap = scap;
    ps.fp64_samples = 0;
    ps.t0 = t0;
    return BPG_OK;
}

// The flags of the finished phase ([K] bytes), copied when they are needed:
// only a phase that met ill-conditioned samples reads them.
int fetch_flags(bpg_plan* pl, long K) {
    BPG_HIP(pl->ctx, hipMemcpy(pl->h_flags.p, pl->flags.p, (size_t)K, hipMemcpyDeviceToHost));
    return BPG_OK;
}

// Repeats permutations S of the finished phase (indices into it) in `form`,
// synchronously, and merges their results into the pinned result array, the
// tie flags and the caller's statistics.  On return S holds those of them
// whose repeat met ill-conditioned samples again (never for FP64), and
// final_form records the form each permutation's results now come from.
int repeat_subset(bpg_plan* pl, std::vector<long>& S, int form, std::vector<signed char>& final_form) {
    bpg_ctx* ctx = pl->ctx;
    auto& ps = pl->pending;
    const auto saved = ps;
    const long K = ps.K, Ks = (long)S.size();
    const int N = pl->d.N, Vc = pl->Vc;
    const bool whole = Ks == K;
    std::vector<unsigned> keep(reinterpret_cast<const unsigned*>(pl->h_max.p),
                               reinterpret_cast<const unsigned*>(pl->h_max.p) + K);
    std::vector<unsigned char> keep_ties;
    if (ps.cluster) keep_ties.assign(pl->h_ties.p, pl->h_ties.p + K);
    std::vector<unsigned short> sp;
    std::vector<float> ss, sst;
    if (!whole) {
        if (ps.sign) {
            ss.resize((size_t)Ks * N);
            for (long i = 0; i < Ks; i++) std::copy(ps.sign + S[i] * N, ps.sign + (S[i] + 1) * N, ss.begin() + i * N);
        } else {
            sp.resize((size_t)Ks * N);
            for (long i = 0; i < Ks; i++) std::copy(ps.perm + S[i] * N, ps.perm + (S[i] + 1) * N, sp.begin() + i * N);
        }
        if (ps.stats) sst.resize((size_t)Ks * Vc);
    }
    const unsigned short* rp = whole ? ps.perm : (ps.sign ? nullptr : sp.data());
    const float* rs = whole ? ps.sign : (ps.sign ? ss.data() : nullptr);
    float* rst = whole ? ps.stats : (ps.stats ? sst.data() : nullptr);
    const int rc = phase_enqueue(pl, ps.map, rp, rs, Ks, rst, form);
    if (rc != BPG_OK) return rc;
    BPG_HIP(ctx, hipStreamSynchronize(pl->stream));
    const bool again = form != BPG_PREC_FP64 && *pl->h_unsafe.p > 0;
    if (again && fetch_flags(pl, Ks) != BPG_OK) return BPG_ERR_HIP;
    const unsigned* sub = reinterpret_cast<const unsigned*>(pl->h_max.p);
    std::vector<long> S2;
    for (long i = 0; i < Ks; i++) {
        keep[S[i]] = sub[i];
        if (saved.cluster) keep_ties[S[i]] = pl->h_ties.p[i];
        final_form[S[i]] = (signed char)form;
        if (again && pl->h_flags.p[i]) S2.push_back(S[i]);
    }
    if (!whole && saved.stats)
        for (long i = 0; i < Ks; i++)
            std::copy(sst.begin() + i * Vc, sst.begin() + (i + 1) * Vc, saved.stats + S[i] * Vc);
    std::copy(keep.begin(), keep.end(), reinterpret_cast<unsigned*>(pl->h_max.p));
    if (saved.cluster) std::copy(keep_ties.begin(), keep_ties.end(), pl->h_ties.p);
    ps = saved;
    S.swap(S2);
    return BPG_OK;
}

int phase_finish(bpg_plan* pl, float* maxdist) {
    bpg_ctx* ctx = pl->ctx;
    auto& ps = pl->pending;
    if (!ps.active) BPG_FAIL(ctx, BPG_ERR_ARG, "no permutation phase in flight");
    ps.active = false;
    BPG_HIP(ctx, hipStreamSynchronize(pl->stream));
    if (ps.sample_cap)
        for (int g = 0; g < bpg::BPG_SAMPLE_SEGS; g++)
            ps.fp64_samples += (long)std::min(pl->h_nsamples.p[g], ps.sample_cap);
    const int unsafe_first = ps.form != BPG_PREC_FP64 ? (int)*pl->h_unsafe.p : 0;
    if (unsafe_first > 0 && pl->Vc > 0) {
        // Ill-conditioned samples: SSE = |y*|^2 - |z|^2 is too small a
        // fraction of |y*|^2 for this form's z.  The flagged permutations are
        // repeated in the next form (split-BF16 -> FP32 -> FP64; an exact
        // split-BF16 run goes to FP64 directly), whose results replace
        // theirs, until no sample is flagged; FP64 computes the residual
        // itself and is never flagged.
        if (!ps.have_times) {
            ps.have_times = true;
            ps.t_transform = elapsed(pl->ev[0], pl->ev[1]);
            ps.t_operand = elapsed(pl->ev[1], pl->ev[2]);
            ps.t_contract = elapsed(pl->ev[2], pl->ev[3]);
            ps.t_cluster = elapsed(pl->ev[3], pl->ev[4]);
        }
        std::vector<signed char> final_form(ps.K, (signed char)ps.form);
        std::vector<long> S;
        if (fetch_flags(pl, ps.K) != BPG_OK) return BPG_ERR_HIP;
        for (long k = 0; k < ps.K; k++)
            if (pl->h_flags.p[k]) S.push_back(k);
        int form = ps.form;
        while (!S.empty() && form != BPG_PREC_FP64) {
            form = ((form == BPG_PREC_SPLIT_BF16 && !ps.exact) || form == BPG_PREC_SPLIT_FP8) ? BPG_PREC_FP32
                                                                                              : BPG_PREC_FP64;
            const int rc = repeat_subset(pl, S, form, final_form);
            if (rc != BPG_OK) return rc;
        }
        ps.fp32_perms = ps.fp64_perms = 0;
        for (long k = 0; k < ps.K; k++) {
            ps.fp32_perms += final_form[k] == BPG_PREC_FP32 || final_form[k] == BPG_PREC_FP32_VALU;
            ps.fp64_perms += final_form[k] == BPG_PREC_FP64;
        }
    }
    const long K = ps.K;
    const bool cluster = ps.cluster, mass = ps.mass;
    const int Vc = pl->Vc;
    const unsigned* host = reinterpret_cast<const unsigned*>(pl->h_max.p);
    const float* hmax = pl->h_max.p;
    long ties = 0;
    for (long k = 0; k < K; k++) {
        if (cluster) {
            // CalculateLargestCluster, LIB:18453-18460: extent, or mass / 1e4
            maxdist[k] = mass ? (float)host[k] / 10000.0f : (float)host[k];
            if (Vc > 0) ties += pl->h_ties.p[k] != 0;
        } else if (Vc == 0 || !(hmax[k] > -INFINITY)) {
            maxdist[k] = empty_max();
        } else {
            // trunc(max * 1e4) / 1e4 (KM:299, LIB:10988); trunc is monotone,
            // so it commutes with the max.
            maxdist[k] = (float)((float)quantize(hmax[k]) / 10000.0f);
        }
    }
    pl->last_K.resize(pl->nmaps, 0);
    pl->last_K[ps.map] = K;
    if (cluster) {
        pl->map_ties.resize(pl->nmaps, 0);
        pl->map_ties[ps.map] = ties;
        if (Vc > 0) {
            pl->last_ties.assign(pl->h_ties.p, pl->h_ties.p + K);
        } else {
            pl->last_ties.assign(K, 0);
        }
    } else {
        pl->last_ties.assign(K, 0);
    }
    const auto t1 = std::chrono::steady_clock::now();
    bpg_timing& tm = pl->timing;
    tm.phase_ms = std::chrono::duration<double, std::milli>(t1 - ps.t0).count();
    tm.transform_ms = ps.have_times ? ps.t_transform : elapsed(pl->ev[0], pl->ev[1]);
    tm.operand_ms = ps.have_times ? ps.t_operand : elapsed(pl->ev[1], pl->ev[2]);
    tm.contract_ms = ps.have_times ? ps.t_contract : elapsed(pl->ev[2], pl->ev[3]);
    tm.cluster_ms = !(cluster || ps.tfce) ? 0.0 : ps.have_times ? ps.t_cluster : elapsed(pl->ev[3], pl->ev[4]);
    tm.contract_launches = ps.launches;
    tm.perm_voxels = K * (long)Vc;
    tm.fit_columns = ps.pfit;
    tm.slots_per_perm = ps.form == BPG_PREC_SPLIT_BF16 || ps.mx ? ps.PP : ps.pfit;
    tm.precision_used = ps.form;
    tm.unsafe_waves = unsafe_first;
    tm.fp32_permutations = ps.fp32_perms;
    tm.fp64_permutations = ps.fp64_perms;
    tm.fp64_samples = ps.fp64_samples;
    tm.cdt_tie_maps = cluster ? ties : 0;
    return BPG_OK;
}

}  // namespace

extern "C" int bpg_plan_permutation_phase(bpg_plan* pl, int map, const unsigned short* perm,
                                          const float* sign, long K, float* maxdist, float* stats) {
    if (!pl) return BPG_ERR_ARG;
    if (!maxdist) BPG_FAIL(pl->ctx, BPG_ERR_ARG, "null maxdist");
    if (pl->pending.active) BPG_FAIL(pl->ctx, BPG_ERR_ARG, "a permutation phase of this plan is in flight");
    const int rc = phase_enqueue(pl, map, perm, sign, K, stats);
    if (rc != BPG_OK) return rc;
    return phase_finish(pl, maxdist);
}

extern "C" int bpg_plan_permutation_phase_noalloc(bpg_plan* pl, int map, const unsigned short* d_perm,
                                                  const float* d_sign, long K, float* d_maxdist) {
    if (!pl) return BPG_ERR_ARG;
    bpg_ctx* ctx = pl->ctx;
    const bpg_study_desc& d = pl->d;
    if (!d_maxdist) BPG_FAIL(ctx, BPG_ERR_ARG, "null maxdist");
    if (pl->pending.active) BPG_FAIL(ctx, BPG_ERR_ARG, "a permutation phase of this plan is in flight");
    if (map < 0 || map >= pl->nmaps) BPG_FAIL(ctx, BPG_ERR_ARG, "statistical map index out of range");
    if (K <= 0 || K > (1L << 26)) BPG_FAIL(ctx, BPG_ERR_ARG, "number of permutations out of range");
    if (d.inference != BPG_VOXEL)
        BPG_FAIL(ctx, BPG_ERR_UNSUPPORTED, "the no-allocation phase supports voxel inference (max-T) only");
    if (d.precision != BPG_PREC_SPLIT_BF16 && d.precision != BPG_PREC_FP32 && d.precision != BPG_PREC_FP32_VALU)
        BPG_FAIL(ctx, BPG_ERR_UNSUPPORTED, "the no-allocation phase runs the split-BF16 and FP32 forms only");
    if (!ctx->scratch) BPG_FAIL(ctx, BPG_ERR_HIP, ctx->scratch_err.c_str());
    const bpg::MapModel& m = pl->models[map];
    if (m.sign_operand ? !d_sign : !d_perm) BPG_FAIL(ctx, BPG_ERR_ARG, "missing permutation / sign matrix");
    BPG_HIP(ctx, hipSetDevice(ctx->device));
    hipStream_t st = pl->stream;
    const int N = d.N, Vc = pl->Vc, Vpad = pl->Vpad;
    if (Vc == 0) {
        // empty mask: the reference's empty max for every permutation
        float e = empty_max();
        unsigned bits;
        std::memcpy(&bits, &e, sizeof bits);
        BPG_HIP(ctx, bpg::launch_fill_u32(reinterpret_cast<unsigned*>(d_maxdist), (size_t)K, bits, st));
        BPG_HIP(ctx, hipStreamSynchronize(st));
        return BPG_OK;
    }
    const bool bf = d.precision == BPG_PREC_SPLIT_BF16;
    const int pfit = m.pfit, KK = (N + 15) / 16, NS = f32_rows(N);
    const int npy = bpg::bf16_y_pieces(KK, m.sign_operand);
    const int PP = bf16_slots(pfit);
    const bpg::Bf16Launch Lb =
        bf ? bf16_launch(ctx, K, PP, KK, Vpad, m.sign_operand, d.test == BPG_FTEST, false) : bpg::Bf16Launch{};
    const bool mx = d.precision == BPG_PREC_FP32 && pfit <= 16;
    const int KG = (N + 7) / 8;
    const bpg::Bf16Launch Lm =
        mx ? bpg::plan_contract_tiles((int)K, PP, bpg::f32mx_rows_per_wave(KG), Vpad,
                                      ctx->cus / 8 *
                                          bpg::contract_f32mx_blocks_per_cu(PP, KG, d.test == BPG_FTEST, false))
           : bpg::Bf16Launch{};
    const bpg::F32Launch Lf = bf || mx ? bpg::F32Launch{} : bpg::plan_contract_f32((int)K, pfit, Vpad);
    const int n_part = bf ? Lb.n_vchunks : mx ? Lm.n_vchunks : Lf.n_vchunks;
    const int PT = bf || mx ? 1 : bpg::f32_perm_tile(pfit);
    const long Kpad = (K + PT - 1) / PT * PT;
    // scratch layout, each part 256-byte aligned: operand | chunk maxima |
    // maxima | flags | counters (unsafe waves, bad indices)
    auto up = [](size_t b) { return (b + 255) / 256 * 256; };
    const size_t op_bytes = bf   ? (size_t)Lb.n_sub * KK * 2 * 64 * 16
                            : mx ? (size_t)Lm.n_sub * KG * 64 * 16
                                 : (size_t)Kpad * pfit * NS * sizeof(float);
    const size_t o_part = up(op_bytes), o_max = o_part + up((size_t)n_part * K * sizeof(float));
    const size_t o_flags = o_max + up((size_t)K * sizeof(float));
    const size_t o_cnt = o_flags + up((size_t)K);
    const size_t total = o_cnt + 256;
    if (total > bpg::BPG_SCRATCH_BYTES) {
        ctx->err = "the no-allocation phase needs " + std::to_string(total >> 20) + " MiB of scratch for K = " +
                   std::to_string(K) + "; its device scratch holds " +
                   std::to_string(bpg::BPG_SCRATCH_BYTES >> 20) + " MiB (use fewer permutations per call)";
        return BPG_ERR_UNSUPPORTED;
    }
    unsigned char* sc = static_cast<unsigned char*>(ctx->scratch);
    float* partial = reinterpret_cast<float*>(sc + o_part);
    float* maxout = reinterpret_cast<float*>(sc + o_max);
    unsigned char* flags = sc + o_flags;
    unsigned* cnt = reinterpret_cast<unsigned*>(sc + o_cnt);  // [0] unsafe waves, [1] bad indices
    const double* mq = pl->mdev.p + pl->mdev_off[map];
    const int QB = bpg::transform_width(m.q), NR = (N + 15) / 16 * 16;
    const auto t0 = std::chrono::steady_clock::now();
    BPG_HIP(ctx, hipEventRecord(pl->ev[0], st));
    {
        bpg::TransformArgs ta{};
        ta.Yc = pl->Yc.p;
        ta.Qn = mq + m.Qt.size();
        ta.dscale = mq + m.Qt.size() + (size_t)NR * std::max(QB, 1);
        ta.N = N; ta.q = m.q; ta.Vc = Vc; ta.Vpad = Vpad;
        ta.KK = bf ? KK : 0; ta.NS = bf ? 0 : NS; ta.npy = npy;
        ta.yy = pl->yy.p; ta.Yb = bf ? pl->Yb.p : nullptr; ta.Y32 = bf || mx ? nullptr : pl->Y32.p;
        ta.Ym = mx ? pl->Ym.p : nullptr;
        BPG_HIP(ctx, bpg::launch_transform(ta, st));
    }
    BPG_HIP(ctx, hipEventRecord(pl->ev[1], st));
    // flags (K bytes, rounded up to words) and both counters start at zero
    BPG_HIP(ctx, bpg::launch_fill_u32(reinterpret_cast<unsigned*>(flags), (K + 3) / 4, 0u, st));
    BPG_HIP(ctx, bpg::launch_fill_u32(cnt, 2, 0u, st));
    {
        bpg::OperandArgs oa{};
        oa.perm = m.sign_operand ? nullptr : d_perm;
        oa.sign = m.sign_operand ? d_sign : nullptr;
        oa.Qt = mq;
        oa.N = N; oa.pfit = pfit; oa.K = (int)K; oa.PP = PP; oa.KK = KK; oa.NS = NS;
        oa.bad_index = cnt + 1;
        if (bf) {
            oa.n_sub = Lb.n_sub;
            oa.Gb = reinterpret_cast<unsigned*>(sc);
            BPG_HIP(ctx, bpg::launch_operand_bf16(oa, st));
        } else if (mx) {
            oa.n_sub = Lm.n_sub;
            oa.Gm = reinterpret_cast<float*>(sc);
            BPG_HIP(ctx, bpg::launch_operand_f32mx(oa, KG, st));
        } else {
            oa.n_sub = (int)Kpad;
            oa.G32 = reinterpret_cast<float*>(sc);
            BPG_HIP(ctx, bpg::launch_operand_f32(oa, st));
        }
    }
    BPG_HIP(ctx, hipEventRecord(pl->ev[2], st));
    {
        bpg::ContractArgs ca{};
        ca.Gb = reinterpret_cast<const unsigned*>(sc); ca.Yb = pl->Yb.p;
        ca.G32 = reinterpret_cast<const float*>(sc); ca.Y32 = pl->Y32.p;
        ca.Gm = reinterpret_cast<const float*>(sc); ca.Ym = pl->Ym.p;
        ca.yy = pl->yy.p; ca.partial = partial;
        ca.K = (int)K; ca.Vc = Vc; ca.Vpad = Vpad; ca.N = N; ca.pfit = pfit; ca.r = m.r; ca.NS = NS;
        ca.ftest = d.test == BPG_FTEST;
        ca.sqrt_dof = (float)std::sqrt((double)m.dof);
        ca.f_scale = (float)((double)m.dof / (double)m.r);
        ca.cdt = d.cdt;
        ca.tau = bf && npy == 2 ? bpg::BPG_TAU_BF16 : bpg::BPG_TAU_FP32;
        ca.pinf = INFINITY;
        ca.unsafe = cnt;
        ca.flags = flags;
        ca.cdt_lo = ca.cdt_hi = d.cdt;
        if (bf) BPG_HIP(ctx, bpg::launch_contract_bf16(ca, Lb, KK, PP, m.sign_operand, st));
        else if (mx) BPG_HIP(ctx, bpg::launch_contract_f32mx(ca, Lm, PP, KG, st));
        else BPG_HIP(ctx, bpg::launch_contract_f32(ca, Lf, st));
        BPG_HIP(ctx, bpg::launch_reduce_max(partial, n_part, (int)K, maxout, st));
        BPG_HIP(ctx, bpg::launch_quantize_max(maxout, (int)K, d_maxdist, empty_max(), st));
    }
    BPG_HIP(ctx, hipEventRecord(pl->ev[3], st));
    unsigned hc[2] = {0u, 0u};
    BPG_HIP(ctx, hipMemcpyAsync(hc, cnt, sizeof hc, hipMemcpyDeviceToHost, st));
    BPG_HIP(ctx, hipStreamSynchronize(st));
    bpg_timing& tm = pl->timing;
    tm = bpg_timing{};
    tm.phase_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    tm.transform_ms = elapsed(pl->ev[0], pl->ev[1]);
    tm.operand_ms = elapsed(pl->ev[1], pl->ev[2]);
    tm.contract_ms = elapsed(pl->ev[2], pl->ev[3]);
    tm.contract_launches = 1;
    tm.perm_voxels = K * (long)Vc;
    tm.fit_columns = pfit;
    tm.slots_per_perm = bf || mx ? PP : pfit;
    tm.precision_used = d.precision;
    tm.unsafe_waves = (int)hc[0];
    tm.fp32_permutations = bf ? 0 : K;
    if (hc[1]) BPG_FAIL(ctx, BPG_ERR_ARG, "permutation index out of range (maxdist is not valid)");
    if (hc[0])
        BPG_FAIL(ctx, BPG_ERR_UNSUPPORTED,
                 "ill-conditioned samples: their permutations need repeats in a more accurate form, which "
                 "allocates; run bpg_plan_permutation_phase (maxdist is not valid)");
    return BPG_OK;
}

extern "C" int bpg_plans_permutation_phase(bpg_plan* const* plans, int n, const int* maps,
                                           const unsigned short* const* perms, const float* const* signs,
                                           const long* K, float* const* maxdist) {
    if (!plans || n <= 0 || !maps || !K || !maxdist) return BPG_ERR_ARG;
    for (int i = 0; i < n; i++) {
        if (!plans[i] || !maxdist[i]) return BPG_ERR_ARG;
        for (int j = 0; j < i; j++)
            if (plans[j] == plans[i]) BPG_FAIL(plans[i]->ctx, BPG_ERR_ARG, "a plan appears twice in one batch");
        if (plans[i]->pending.active)
            BPG_FAIL(plans[i]->ctx, BPG_ERR_ARG, "a permutation phase of this plan is in flight");
    }
    // enqueue every study (each plan has its own stream); then enqueue the
    // FP32 repeats of the ill-conditioned ones together, then collect.  The
    // first error is returned after the enqueued work has been waited for.
    int rc = BPG_OK;
    int enq = 0;
    for (; enq < n; enq++) {
        rc = phase_enqueue(plans[enq], maps[enq], perms ? perms[enq] : nullptr, signs ? signs[enq] : nullptr,
                           K[enq], nullptr);
        if (rc != BPG_OK) break;
    }
    for (int i = 0; i < enq && rc == BPG_OK; i++) {
        // studies with mostly ill-conditioned permutations: their FP32
        // repeats are enqueued together; a few flagged permutations are
        // repeated when the study is collected
        bpg_plan* pl = plans[i];
        auto& ps = pl->pending;
        if (hipStreamSynchronize(pl->stream) != hipSuccess || ps.form != BPG_PREC_SPLIT_BF16 || ps.exact ||
            *pl->h_unsafe.p == 0)
            continue;
        if (fetch_flags(pl, ps.K) != BPG_OK) {
            rc = BPG_ERR_HIP;
            continue;
        }
        long nf = 0;
        for (long k = 0; k < ps.K; k++) nf += pl->h_flags.p[k] != 0;
        if (nf * 2 <= ps.K) continue;
        const auto t0 = ps.t0;
        const int r = phase_enqueue(pl, ps.map, ps.perm, ps.sign, ps.K, ps.stats, BPG_PREC_FP32);
        ps.t0 = t0;
        if (r != BPG_OK) rc = r;
    }
    for (int i = 0; i < enq; i++) {
        const int r = phase_finish(plans[i], maxdist[i]);
        if (rc == BPG_OK) rc = r;
    }
    return rc;
}

extern "C" int bpg_plan_last_timing(const bpg_plan* pl, bpg_timing* out) {
    if (!pl || !out) return BPG_ERR_ARG;
    *out = pl->timing;
    return BPG_OK;
}

extern "C" int bpg_plan_last_tie_flags(const bpg_plan* pl, unsigned char* flags) {
    if (!pl || !flags) return BPG_ERR_ARG;
    std::copy(pl->last_ties.begin(), pl->last_ties.end(), flags);
    return BPG_OK;
}

extern "C" int bpg_plan_observed_maps(bpg_plan* pl, float* maps) {
    if (!pl || !maps) return BPG_ERR_ARG;
    return bpg_plan_glm(pl, nullptr, maps, nullptr, nullptr);
}

extern "C" int bpg_plan_glm(bpg_plan* pl, float* betas, float* maps, float* residuals, float* resvar) {
    if (!pl) return BPG_ERR_ARG;
    bpg_ctx* ctx = pl->ctx;
    const bpg::ObsModel& o = pl->obs;
    const int Vc = pl->Vc, N = pl->d.N;
    if (Vc == 0) {
        // empty mask: every output is 0 (out-of-mask voxels are 0)
        if (maps) std::fill(maps, maps + (size_t)pl->nmaps * pl->G, 0.0f);
        if (betas) std::fill(betas, betas + (size_t)o.p * pl->G, 0.0f);
        if (residuals) std::fill(residuals, residuals + (size_t)N * pl->G, 0.0f);
        if (resvar) std::fill(resvar, resvar + pl->G, 0.0f);
        pl->have_obs = true;
        return BPG_OK;
    }
    BPG_HIP(ctx, hipSetDevice(ctx->device));
    hipStream_t st = pl->stream;
    BPG_HIP(ctx, pl->oX.ensure(o.X.size()));
    BPG_HIP(ctx, pl->oXp.ensure(o.Xp.size()));
    BPG_HIP(ctx, pl->oC.ensure(o.C.size()));
    BPG_HIP(ctx, pl->osc.ensure(o.sc.size()));
    BPG_HIP(ctx, pl->d_obs.ensure((size_t)pl->nmaps * std::max(Vc, 1)));
    BPG_HIP(ctx, hipMemcpyAsync(pl->oX.p, o.X.data(), o.X.size() * 8, hipMemcpyHostToDevice, st));
    BPG_HIP(ctx, hipMemcpyAsync(pl->oXp.p, o.Xp.data(), o.Xp.size() * 8, hipMemcpyHostToDevice, st));
    BPG_HIP(ctx, hipMemcpyAsync(pl->oC.p, o.C.data(), o.C.size() * 8, hipMemcpyHostToDevice, st));
    BPG_HIP(ctx, hipMemcpyAsync(pl->osc.p, o.sc.data(), o.sc.size() * 8, hipMemcpyHostToDevice, st));
    bpg::ObsArgs a{};
    a.Yc = pl->Yc.p; a.X = pl->oX.p; a.Xp = pl->oXp.p; a.C = pl->oC.p; a.sc = pl->osc.p;
    a.N = N; a.p = o.p; a.nc = o.nc; a.Vc = Vc; a.ftest = o.ftest; a.nmaps = pl->nmaps;
    a.out = pl->d_obs.p;
    // optional outputs in one scratch block [p + N + 1][Vc]
    const size_t extra = (size_t)(betas ? o.p : 0) + (residuals ? N : 0) + (resvar ? 1 : 0);
    if (extra) BPG_HIP(ctx, pl->glm_out.ensure(extra * std::max(Vc, 1)));
    {
        float* q = pl->glm_out.p;
        if (betas) { a.betas = q; q += (size_t)o.p * Vc; }
        if (residuals) { a.residuals = q; q += (size_t)N * Vc; }
        if (resvar) a.resvar = q;
    }
    BPG_HIP(ctx, bpg::launch_observed(a, st));
    std::vector<float> h((size_t)pl->nmaps * Vc), hx(extra * Vc);
    if (Vc > 0) {
        BPG_HIP(ctx, hipMemcpyAsync(h.data(), pl->d_obs.p, h.size() * sizeof(float), hipMemcpyDeviceToHost, st));
        if (extra)
            BPG_HIP(ctx, hipMemcpyAsync(hx.data(), pl->glm_out.p, hx.size() * sizeof(float), hipMemcpyDeviceToHost,
                                        st));
    }
    BPG_HIP(ctx, hipStreamSynchronize(st));
    // out-of-mask voxels are 0 in every output (K1:51-128, K1:1456-1673)
    auto scatter = [&](float* dst, const float* src, int vols) {
        std::fill(dst, dst + (size_t)vols * pl->G, 0.0f);
        for (int c = 0; c < vols; c++)
            for (int v = 0; v < Vc; v++) dst[(size_t)c * pl->G + pl->comp2grid[v]] = src[(size_t)c * Vc + v];
    };
    if (maps) scatter(maps, h.data(), pl->nmaps);
    const float* q = hx.data();
    if (betas) { scatter(betas, q, o.p); q += (size_t)o.p * Vc; }
    if (residuals) { scatter(residuals, q, N); q += (size_t)N * Vc; }
    if (resvar) scatter(resvar, q, 1);
    pl->have_obs = true;
    return BPG_OK;
}

extern "C" int bpg_plan_pvalues(bpg_plan* pl, const long* K, const float* const* maxdist,
                                float* pvalues, unsigned* labels) {
    return bpg_plan_pvalues_ties(pl, K, maxdist, pvalues, labels, nullptr);
}

extern "C" int bpg_plan_pvalues_ties(bpg_plan* pl, const long* K, const float* const* maxdist,
                                     float* pvalues, unsigned* labels, unsigned* ties) {
    if (!pl || !K || !maxdist || !pvalues) return BPG_ERR_ARG;
    bpg_ctx* ctx = pl->ctx;
    if (!pl->have_obs) BPG_FAIL(ctx, BPG_ERR_ARG, "observed maps have not been computed");
    BPG_HIP(ctx, hipSetDevice(ctx->device));
    hipStream_t st = pl->stream;
    const bpg_study_desc& d = pl->d;
    const int Vc = pl->Vc;
    const bool cluster = d.inference == BPG_CLUSTER_EXTENT || d.inference == BPG_CLUSTER_MASS;
    std::fill(pvalues, pvalues + (size_t)pl->nmaps * pl->G, 0.0f);
    if (labels) std::fill(labels, labels + (size_t)pl->nmaps * pl->G, (unsigned)(pl->G * 3));
    if (ties) std::fill(ties, ties + (size_t)pl->nmaps * pl->G, 0u);
    pl->timing.pvalue_ties = 0;
    if (Vc == 0) return BPG_OK;
    BPG_HIP(ctx, pl->P.ensure(Vc));
    BPG_HIP(ctx, pl->d_pties.ensure(Vc));
    if (cluster) {
        BPG_HIP(ctx, pl->bits.ensure(pl->nwords));
        BPG_HIP(ctx, pl->cl.parent.ensure(Vc));
        BPG_HIP(ctx, pl->cl.sizes.ensure(Vc));
        BPG_HIP(ctx, pl->labels.ensure(Vc));
        BPG_HIP(ctx, pl->maxsize.ensure(1));
    }
    std::vector<float> h(Vc);
    std::vector<unsigned> hl(cluster && labels ? Vc : 0), ht(Vc);
    long total_ties = 0;
    for (int c = 0; c < pl->nmaps; c++) {
        if (K[c] <= 0 || !maxdist[c]) BPG_FAIL(ctx, BPG_ERR_ARG, "empty null distribution");
        std::vector<float> srt(maxdist[c], maxdist[c] + K[c]);
        std::sort(srt.begin(), srt.end());
        BPG_HIP(ctx, pl->sorted.ensure(K[c]));
        BPG_HIP(ctx, hipMemcpyAsync(pl->sorted.p, srt.data(), K[c] * sizeof(float), hipMemcpyHostToDevice, st));
        const float* obs = pl->d_obs.p + (size_t)c * Vc;
        if (cluster) {
            // ClusterizeOpenCL on the observed map (LIB:16822, LIB:18301-18401)
            BPG_HIP(ctx, bpg::launch_threshold_bits(obs, Vc, d.cdt, pl->bits.p, pl->nwords, st));
            // compact roots; 0xFFFFFFFF marks voxels at or below the threshold
            BPG_HIP(ctx, hipMemsetAsync(pl->labels.p, 0xFF, Vc * sizeof(unsigned), st));
            BPG_HIP(ctx, hipMemsetAsync(pl->maxsize.p, 0, sizeof(unsigned), st));
            bpg::CclArgs cc{};
            cc.bits = pl->bits.p; cc.nbr = pl->d_nbr.p;
            cc.parent = pl->cl.parent.p; cc.sizes = pl->cl.sizes.p; cc.maxsize = pl->maxsize.p;
            cc.labels = pl->labels.p;
            cc.values = d.inference == BPG_CLUSTER_MASS ? obs : nullptr;
            cc.Vc = Vc; cc.nwords = pl->nwords; cc.nmaps = 1;
            cc.connectivity = d.connectivity;
            BPG_HIP(ctx, bpg::launch_ccl(cc, st));
        }
        bpg::PvalArgs pa{};
        pa.obs = obs; pa.sorted_max = pl->sorted.p;
        pa.labels = cluster ? pl->labels.p : nullptr;
        pa.sizes = cluster ? pl->cl.sizes.p : nullptr;
        pa.K = (int)K[c]; pa.Vc = Vc; pa.inference = d.inference; pa.cdt = d.cdt; pa.P = pl->P.p;
        pa.ties = pl->d_pties.p;
        // the tied maps of this map's last permutation phase, when that phase
        // had as many permutations as this null
        pa.cluster_ties = cluster && c < (int)pl->map_ties.size() &&
                                  pl->last_K.size() > (size_t)c && pl->last_K[c] == K[c]
                              ? (unsigned)pl->map_ties[c]
                              : 0u;
        BPG_HIP(ctx, bpg::launch_pvalues(pa, st));
        BPG_HIP(ctx, hipMemcpyAsync(h.data(), pl->P.p, Vc * sizeof(float), hipMemcpyDeviceToHost, st));
        BPG_HIP(ctx, hipMemcpyAsync(ht.data(), pl->d_pties.p, Vc * sizeof(unsigned), hipMemcpyDeviceToHost, st));
        if (cluster && labels)
            BPG_HIP(ctx, hipMemcpyAsync(hl.data(), pl->labels.p, Vc * sizeof(unsigned), hipMemcpyDeviceToHost, st));
        BPG_HIP(ctx, hipStreamSynchronize(st));
        for (int v = 0; v < Vc; v++) pvalues[(size_t)c * pl->G + pl->comp2grid[v]] = h[v];
        for (int v = 0; v < Vc; v++) total_ties += ht[v];
        if (ties)
            for (int v = 0; v < Vc; v++) ties[(size_t)c * pl->G + pl->comp2grid[v]] = ht[v];
        // grid labels: the component's minimum linear index (its root's grid
        // index), 3*W*H*D elsewhere
        if (cluster && labels)
            for (int v = 0; v < Vc; v++)
                if (hl[v] != 0xFFFFFFFFu) labels[(size_t)c * pl->G + pl->comp2grid[v]] = pl->comp2grid[hl[v]];
    }
    pl->timing.pvalue_ties = total_ties;
    return BPG_OK;
}

// ---------------------------------------------------------------------------
// autotuner
// ---------------------------------------------------------------------------
extern "C" int bpg_plan_tune(bpg_plan* pl, int map, long K, bpg_tune_result* out) {
    if (!pl || !out) return BPG_ERR_ARG;
    bpg_ctx* ctx = pl->ctx;
    const bpg_study_desc& d = pl->d;
    *out = bpg_tune_result{};
    if (map < 0 || map >= pl->nmaps) BPG_FAIL(ctx, BPG_ERR_ARG, "statistical map index out of range");
    if (K <= 0 || K > (1L << 26)) BPG_FAIL(ctx, BPG_ERR_ARG, "number of permutations out of range");
    if (d.precision != BPG_PREC_SPLIT_BF16)
        BPG_FAIL(ctx, BPG_ERR_UNSUPPORTED, "the autotuner searches the split-BF16 contraction's shapes");
    if (pl->pending.active) BPG_FAIL(ctx, BPG_ERR_ARG, "a permutation phase of this plan is in flight");
    if (pl->Vc == 0) return BPG_OK;
    const bpg::MapModel& m = pl->models[map];
    const int N = d.N, KK = (N + 15) / 16, PP = bf16_slots(m.pfit);
    const bool cluster = d.inference == BPG_CLUSTER_EXTENT || d.inference == BPG_CLUSTER_MASS;
    const bool out_stats = cluster || d.inference == BPG_CLUSTER_MASS || d.inference == BPG_TFCE;
    // seeded permutations of the phase's shape (the grid does not depend on
    // their values)
    std::vector<unsigned short> perm;
    std::vector<float> sign;
    int rc = m.sign_operand ? (sign.resize((size_t)K * N), bpg_gen_sign_matrix(N, K, 0, 20261203, sign.data()))
                            : (perm.resize((size_t)K * N), bpg_gen_perm_correlation(N, K, 0, 20261203, perm.data()));
    if (rc != BPG_OK) {
        // fewer distinct permutations than K (tiny N): repeat the identity
        rc = BPG_OK;
        if (m.sign_operand) std::fill(sign.begin(), sign.end(), 1.0f);
        else
            for (long k = 0; k < K; k++)
                for (int s = 0; s < N; s++) perm[(size_t)k * N + s] = (unsigned short)s;
    }
    std::vector<float> md(K);
    std::vector<bpg::TuneEntry> cands;
    const bool alt = KK <= 8 && !out_stats;
    for (int v = 0; v <= (alt ? 1 : 0); v++)
        for (int cp = -1; cp <= 1; cp++) {
            bpg::TuneEntry e;
            e.variant = v;
            e.chunk_pow = cp;
            cands.push_back(e);
        }
    double def_ms = 0.0, best_ms = 1e30;
    bpg::TuneEntry best;
    ctx->tune_forced = true;
    for (auto& e : cands) {
        ctx->tune_force = e;
        double ms = 1e30;
        for (int rep = 0; rep < 3 && rc == BPG_OK; rep++) {
            rc = bpg_plan_permutation_phase(pl, map, m.sign_operand ? nullptr : perm.data(),
                                            m.sign_operand ? sign.data() : nullptr, K, md.data(), nullptr);
            if (rc == BPG_OK && rep > 0) ms = std::min(ms, pl->timing.contract_ms);
        }
        if (rc != BPG_OK) break;
        e.best_ms = ms;
        if (e.variant == 0 && e.chunk_pow == 0) def_ms = ms;
        if (ms < best_ms) {
            best_ms = ms;
            best = e;
        }
    }
    ctx->tune_forced = false;
    if (rc != BPG_OK) return rc;
    best.default_ms = def_ms;
    best.best_ms = best_ms;
    const int Vpad = pl->Vpad;
    ctx->tune[tune_key(KK, PP, m.sign_operand, d.test == BPG_FTEST, out_stats, Vpad)] = best;
    const bpg::Bf16Launch L = bf16_launch(ctx, K, PP, KK, Vpad, m.sign_operand, d.test == BPG_FTEST, out_stats);
    out->variant = best.variant;
    out->rows_per_wave = bpg::bf16_rows_per_wave_v(KK, m.sign_operand, best.variant);
    out->chunk_pow = best.chunk_pow;
    out->vt_per_chunk = L.vt_per_chunk;
    out->default_ms = def_ms;
    out->best_ms = best_ms;
    out->candidates = (int)cands.size();
    return BPG_OK;
}

namespace {
std::string tune_device_tag(const bpg_ctx* ctx) {
    std::string t = ctx->name + "/" + ctx->arch;
    for (char& c : t)
        if (c == ' ' || c == '\t') c = '_';
    return t;
}
}  // namespace

extern "C" int bpg_tune_save(const bpg_ctx* ctx, const char* path) {
    if (!ctx || !path) return BPG_ERR_ARG;
    // keep the entries of other devices already in the file
    std::vector<std::string> keep;
    const std::string tag = tune_device_tag(ctx);
    if (FILE* f = std::fopen(path, "r")) {
        char line[512];
        while (std::fgets(line, sizeof line, f))
            if (line[0] != '#' && std::strncmp(line, tag.c_str(), tag.size()) != 0) keep.push_back(line);
        std::fclose(f);
    }
    FILE* f = std::fopen(path, "w");
    if (!f) return BPG_ERR_ARG;
    std::fprintf(f, "# bpg split-BF16 tuning: device KK PP sign ftest out log2(Vpad) variant chunk_pow default_ms best_ms\n");
    for (const auto& l : keep) std::fputs(l.c_str(), f);
    for (const auto& [k, e] : ctx->tune)
        std::fprintf(f, "%s %d %d %d %d %d %d %d %d %.6f %.6f\n", tag.c_str(), std::get<0>(k), std::get<1>(k),
                     std::get<2>(k), std::get<3>(k), std::get<4>(k), std::get<5>(k), e.variant, e.chunk_pow,
                     e.default_ms, e.best_ms);
    std::fclose(f);
    return BPG_OK;
}

extern "C" int bpg_tune_load(bpg_ctx* ctx, const char* path) {
    if (!ctx || !path) return BPG_ERR_ARG;
    FILE* f = std::fopen(path, "r");
    if (!f) return 0;
    const std::string tag = tune_device_tag(ctx);
    char line[512], dev[256];
    int n = 0;
    while (std::fgets(line, sizeof line, f)) {
        if (line[0] == '#') continue;
        int kk, pp, sg, ft, ou, vb, va, cp;
        double dm, bm;
        if (std::sscanf(line, "%255s %d %d %d %d %d %d %d %d %lf %lf", dev, &kk, &pp, &sg, &ft, &ou, &vb, &va, &cp,
                        &dm, &bm) != 11)
            continue;
        if (tag != dev || va < 0 || va > 1 || cp < -1 || cp > 1) continue;
        bpg::TuneEntry e;
        e.variant = va;
        e.chunk_pow = cp;
        e.default_ms = dm;
        e.best_ms = bm;
        ctx->tune[bpg::TuneKey(kk, pp, sg, ft, ou, vb)] = e;
        n++;
    }
    std::fclose(f);
    return n;
}

extern "C" int bpg_tune_entries(const bpg_ctx* ctx) { return ctx ? (int)ctx->tune.size() : BPG_ERR_ARG; }

// ---------------------------------------------------------------------------
// batches of studies: one persistent contraction per class, two hipGraphs
// ---------------------------------------------------------------------------

struct bpg_batch {
    bpg_ctx* ctx = nullptr;
    struct Study {
        bpg_plan* plan = nullptr;
        int map = 0;
        long K = 0;
        int N = 0, KK = 0, PP = 0, npy = 2;
        int stage = 1;        // 0: likely to meet ill-conditioned samples (runs first)
        bool sign = false;
        // a sample list (its ill-conditioned samples recomputed in FP64):
        // its per-segment capacity, its entries, its segment counters (at
        // counters() + BPG_SAMPLE_SEGS * excl_slot)
        unsigned long long cap = 0;
        DBuf<unsigned long long> samples;
        int excl_slot = -1;
        bpg::Bf16Launch L{};
        DBuf<double> yy;
        DBuf<unsigned> Yb, Gb;
        DBuf<float> partial;
        size_t in_off = 0;    // its permutations / signs in the staging and device copies
        size_t max_off = 0;   // its maxima in max_all / h_max
        size_t flag_off = 0;  // its flags (bytes, a multiple of 4) in the word buffer / h_flags
        int slot = 0;         // its ill-conditioned count in the word buffer / h_unsafe
        // repeats of its flagged permutations (host copies that live until
        // the repeat has finished)
        std::vector<long> rep;
        std::vector<unsigned short> rperm;
        std::vector<float> rsign, rmax;
    };
    std::vector<Study> st;
    struct Class {
        int KK, PP, sign, variant, stage, excl;
        std::vector<int> members;
        int n_items[bpg::BPG_XCDS] = {};  // per XCD queue
        int grid = 0;
        size_t first = 0;  // index of its first study in the device table
    };
    std::vector<Class> cls;
    DBuf<bpg::BatchStudy> table;
    DBuf<bpg::ReduceJob> jobs;
    // the nuisance transforms, one launch per (stage, nuisance width): the
    // group's studies' arguments at ttab + tab_off, their first blocks at
    // tfirst + tab_off
    struct TGroup {
        int stage, QB, n, blocks;
        size_t tab_off;
    };
    std::vector<TGroup> tgroups;
    DBuf<bpg::TransformArgs> ttab;
    DBuf<int> tfirst;
    // one word buffer cleared by one fill: the sample lists' segment counters
    // (64-bit, two words each), queue heads (8 per class), the studies'
    // ill-conditioned counts, then their flag bytes
    DBuf<unsigned> words;
    size_t n_heads = 0, flag_bytes = 0, n_excl = 0;
    DBuf<unsigned short> perm_all;
    DBuf<float> sign_all, max_all;
    size_t n_perm = 0, n_sign = 0, n_max = 0;
    int max_K = 0;
    // per stage: its studies' ranges of maxima, flag bytes and counts (the
    // offsets are assigned stage by stage), its reduction jobs
    size_t max_end[2] = {}, flag_end[2] = {};
    int slot_end[2] = {}, job_end[2] = {};
    HBuf<unsigned short> h_perm;
    HBuf<float> h_sign, h_max;
    HBuf<unsigned> h_unsafe;
    HBuf<unsigned char> h_flags;
    hipStream_t stream = nullptr;
    // the branches of the graphs: stream and side[0]
    static constexpr int NBR = 2;
    hipStream_t side[NBR - 1] = {};
    hipEvent_t first_done = nullptr;
    hipEvent_t fork[2] = {}, join[2][NBR] = {}, prep[2][NBR] = {};
    hipGraph_t graph[2] = {};
    hipGraphExec_t exec[2] = {};
    double last_ms = 0.0, last_graph_ms = 0.0;
    int last_fallback = 0;

    unsigned long long* counters() const { return reinterpret_cast<unsigned long long*>(words.p); }
    unsigned* heads() const { return words.p + 2 * bpg::BPG_SAMPLE_SEGS * n_excl; }
    unsigned* unsafe() const { return heads() + n_heads; }
    unsigned char* flags() const { return reinterpret_cast<unsigned char*>(unsafe() + st.size()); }

    ~bpg_batch() {
        for (int g = 0; g < 2; g++) {
            if (exec[g]) (void)hipGraphExecDestroy(exec[g]);
            if (graph[g]) (void)hipGraphDestroy(graph[g]);
        }
        if (first_done) (void)hipEventDestroy(first_done);
        for (int g = 0; g < 2; g++) {
            if (fork[g]) (void)hipEventDestroy(fork[g]);
            for (int h = 0; h < NBR; h++) {
                if (join[g][h]) (void)hipEventDestroy(join[g][h]);
                if (prep[g][h]) (void)hipEventDestroy(prep[g][h]);
            }
        }
        for (auto& x : side)
            if (x) (void)hipStreamDestroy(x);
        if (stream) (void)hipStreamDestroy(stream);
        for (auto& s : st) {
            s.yy.release(); s.Yb.release(); s.Gb.release(); s.partial.release(); s.samples.release();
        }
        table.release(); jobs.release(); ttab.release(); tfirst.release(); words.release(); perm_all.release(); sign_all.release(); max_all.release();
        h_perm.release(); h_sign.release(); h_max.release(); h_unsafe.release(); h_flags.release();
    }
};

extern "C" int bpg_batch_create(bpg_plan* const* plans, int n, const int* maps, const long* K, bpg_batch** out) {
    if (!plans || n <= 0 || !maps || !K || !out) return BPG_ERR_ARG;
    *out = nullptr;
    bpg_ctx* ctx = plans[0] ? plans[0]->ctx : nullptr;
    if (!ctx) return BPG_ERR_ARG;
    for (int i = 0; i < n; i++) {
        bpg_plan* pl = plans[i];
        if (!pl || pl->ctx != ctx) BPG_FAIL(ctx, BPG_ERR_ARG, "the plans of a batch must share one context");
        for (int j = 0; j < i; j++)
            if (plans[j] == pl) BPG_FAIL(ctx, BPG_ERR_ARG, "a plan appears twice in one batch");
        const bpg_study_desc& d = pl->d;
        if (d.precision != BPG_PREC_SPLIT_BF16 || d.inference != BPG_VOXEL || d.test == BPG_FTEST || d.N > 128)
            BPG_FAIL(ctx, BPG_ERR_UNSUPPORTED,
                     "a batch runs split-BF16 t tests and group means with voxel inference, up to 128 subjects");
        if (maps[i] < 0 || maps[i] >= pl->nmaps) BPG_FAIL(ctx, BPG_ERR_ARG, "statistical map index out of range");
        if (K[i] <= 0 || K[i] > (1L << 26)) BPG_FAIL(ctx, BPG_ERR_ARG, "number of permutations out of range");
    }
    BPG_HIP(ctx, hipSetDevice(ctx->device));
    bpg_batch* b = new (std::nothrow) bpg_batch();
    if (!b) return BPG_ERR_NOMEM;
    b->ctx = ctx;
    auto fail = [&](int rc) { delete b; return rc; };
#define BPG_BHIP(call)                                                              \
    do {                                                                            \
        const hipError_t e_ = (call);                                               \
        if (e_ != hipSuccess) {                                                     \
            ctx->err = std::string(#call) + ": " + hipGetErrorString(e_);           \
            return fail(e_ == hipErrorOutOfMemory ? BPG_ERR_NOMEM : BPG_ERR_HIP);   \
        }                                                                           \
    } while (0)
    b->st.resize(n);
    for (int i = 0; i < n; i++) {
        auto& s = b->st[i];
        bpg_plan* pl = plans[i];
        const bpg::MapModel& m = pl->models[maps[i]];
        s.plan = pl;
        s.map = maps[i];
        s.K = K[i];
        s.N = pl->d.N;
        s.KK = (s.N + 15) / 16;
        s.PP = bf16_slots(m.pfit);
        s.sign = m.sign_operand;
        s.npy = bpg::bf16_y_pieces(s.KK, s.sign);
        const double tau = s.npy == 2 ? bpg::BPG_TAU_BF16 : bpg::BPG_TAU_FP32;
        // as in a single-study phase (phase_enqueue): a sample list where
        // permutations are likely to meet ill-conditioned samples
        if (!s.sign && pl->d.test == BPG_TTEST && s.npy == 2 && s.KK <= bpg::BPG_EXCL_MAX_KK &&
            m.pfit <= bpg::BPG_FIX_MAX_P && m.q <= 24 && pl->Vc > 0)
            s.cap = sample_list_cap((double)m.dof, m.pfit, pl->Vc, s.K, tau);
        // studies likely to repeat permutations run first (a study with a
        // sample list repeats only what overflows it)
        s.stage = !s.cap && ill_conditioned_chance((double)m.dof, m.pfit, pl->Vc, tau) >= 1e-3 ? 0 : 1;
        s.in_off = s.sign ? b->n_sign : b->n_perm;
        (s.sign ? b->n_sign : b->n_perm) += (size_t)s.K * s.N;
        b->max_K = std::max<int>(b->max_K, (int)s.K);
        if (pl->Vc == 0) continue;
        if (s.cap) {
            s.excl_slot = (int)b->n_excl++;
            BPG_BHIP(s.samples.ensure(s.cap * bpg::BPG_SAMPLE_SEGS));
        }
        s.L = bf16_launch(ctx, s.K, s.PP, s.KK, pl->Vpad, s.sign, false, false, s.cap > 0);
        BPG_BHIP(s.yy.ensure(pl->Vpad));
        BPG_BHIP(s.Yb.ensure((size_t)(pl->Vpad / 32) * s.KK * s.npy * 64 * 4));
        BPG_BHIP(s.Gb.ensure((size_t)s.L.n_sub * s.KK * 2 * 64 * 4));
        BPG_BHIP(s.partial.ensure((size_t)s.L.n_vchunks * s.K));
    }
    // result offsets stage by stage, so each stage's results are one range
    for (int g = 0; g < 2; g++) {
        for (auto& s : b->st) {
            if (s.stage != g) continue;
            s.max_off = b->n_max;
            b->n_max += (size_t)s.K;
            s.flag_off = b->flag_bytes;
            b->flag_bytes += ((size_t)s.K + 3) / 4 * 4;
            s.slot = b->slot_end[1]++;
        }
        b->max_end[g] = b->n_max;
        b->flag_end[g] = b->flag_bytes;
        if (g == 0) b->slot_end[0] = b->slot_end[1];
    }
    // classes: same instantiation (k-steps, slots, operand kind, subtile
    // variant) and stage
    for (int i = 0; i < n; i++) {
        const auto& s = b->st[i];
        if (s.plan->Vc == 0) continue;
        bpg_batch::Class* c = nullptr;
        const int excl = s.cap > 0;
        for (auto& x : b->cls)
            if (x.KK == s.KK && x.PP == s.PP && x.sign == (int)s.sign && x.variant == s.L.variant &&
                x.stage == s.stage && x.excl == excl)
                c = &x;
        if (!c) {
            b->cls.push_back(bpg_batch::Class{s.KK, s.PP, (int)s.sign, s.L.variant, s.stage, excl, {}, {}, 0, 0});
            c = &b->cls.back();
        }
        c->members.push_back(i);
    }
    b->n_heads = b->cls.size() * bpg::BPG_XCDS;
    const size_t n_words = 2 * bpg::BPG_SAMPLE_SEGS * b->n_excl + b->n_heads + (size_t)n + b->flag_bytes / 4;
    BPG_BHIP(b->words.ensure(n_words));
    BPG_BHIP(b->perm_all.ensure(std::max<size_t>(b->n_perm, 1)));
    BPG_BHIP(b->sign_all.ensure(std::max<size_t>(b->n_sign, 1)));
    BPG_BHIP(b->max_all.ensure(std::max<size_t>(b->n_max, 1)));
    BPG_BHIP(b->h_perm.ensure(std::max<size_t>(b->n_perm, 1)));
    BPG_BHIP(b->h_sign.ensure(std::max<size_t>(b->n_sign, 1)));
    BPG_BHIP(b->h_max.ensure(std::max<size_t>(b->n_max, 1)));
    BPG_BHIP(b->h_unsafe.ensure(n));
    BPG_BHIP(b->h_flags.ensure(std::max<size_t>(b->flag_bytes, 1)));
    std::fill(b->h_unsafe.p, b->h_unsafe.p + n, 0u);
    std::vector<bpg::BatchStudy> host;
    for (size_t ci = 0; ci < b->cls.size(); ci++) {
        auto& c = b->cls[ci];
        c.first = host.size();
        for (int i : c.members) {
            auto& s = b->st[i];
            bpg_plan* pl = s.plan;
            const bpg::MapModel& m = pl->models[s.map];
            bpg::BatchStudy bs{};
            bpg::ContractArgs& a = bs.a;
            a.Gb = s.Gb.p; a.Yb = s.Yb.p; a.yy = s.yy.p; a.partial = s.partial.p;
            a.K = (int)s.K; a.Vc = pl->Vc; a.Vpad = pl->Vpad; a.N = s.N; a.pfit = m.pfit; a.r = m.r;
            a.nwords = pl->nwords;
            a.sqrt_dof = (float)std::sqrt((double)m.dof);
            a.f_scale = (float)((double)m.dof / (double)m.r);
            a.cdt = pl->d.cdt;
            a.tau = s.npy == 2 ? bpg::BPG_TAU_BF16 : bpg::BPG_TAU_FP32;
            a.pinf = INFINITY;
            a.unsafe = b->unsafe() + s.slot;
            a.flags = b->flags() + s.flag_off;
            a.cdt_lo = a.cdt_hi = pl->d.cdt;
            if (s.cap) {
                a.samples = s.samples.p;
                a.n_samples = b->counters() + (size_t)bpg::BPG_SAMPLE_SEGS * s.excl_slot;
                a.sample_cap = s.cap;
            }
            bs.L = s.L;
            for (int x = 0; x < bpg::BPG_XCDS; x++) {
                // voxel chunks x, x + 8, ... of this study go to queue x
                const int nvc = s.L.n_vchunks > x ? (s.L.n_vchunks - 1 - x) / bpg::BPG_XCDS + 1 : 0;
                bs.item0[x] = c.n_items[x];
                c.n_items[x] += nvc * s.L.n_mgroups;
            }
            host.push_back(bs);
        }
        int occ = 0, total = 0;
        for (int x = 0; x < bpg::BPG_XCDS; x++) total += c.n_items[x];
        BPG_BHIP(bpg::launch_contract_bf16_persist(bpg::BatchClass{}, c.KK, c.PP, c.sign != 0, c.variant, c.excl != 0,
                                                   0, nullptr, &occ));
        c.grid = std::max(1, std::min(total, ctx->cus * std::max(occ, 1)));
    }
    // the reduction jobs stage by stage
    std::vector<bpg::ReduceJob> rjobs;
    for (int g = 0; g < 2; g++) {
        for (const auto& s : b->st)
            if (s.stage == g && s.plan->Vc > 0)
                rjobs.push_back(bpg::ReduceJob{s.partial.p, b->max_all.p + s.max_off, s.L.n_vchunks, (int)s.K});
        b->job_end[g] = (int)rjobs.size();
    }
    // the transform groups
    {
        std::vector<bpg::TransformArgs> targs;
        std::vector<int> tfirst;
        // (each width's studies in two halves, one per branch)
        for (int g = 0; g < 2; g++)
            for (int QB : {24, 16, 8, 4, 0})
                for (int half = 0; half < 2; half++) {
                bpg_batch::TGroup tg{g, QB, 0, 0, targs.size()};
                int seen = 0;
                for (const auto& s : b->st) {
                    const bpg_plan* pl = s.plan;
                    if (s.stage != g || pl->Vc == 0) continue;
                    const bpg::MapModel& m = pl->models[s.map];
                    if (bpg::transform_width(m.q) != QB) continue;
                    if ((seen++ & 1) != half) continue;
                    const double* mq = pl->mdev.p + pl->mdev_off[s.map];
                    const int NR = (s.N + 15) / 16 * 16;
                    bpg::TransformArgs ta{};
                    ta.Yc = pl->Yc.p;
                    ta.Qn = mq + m.Qt.size();
                    ta.dscale = mq + m.Qt.size() + (size_t)NR * std::max(QB, 1);
                    ta.N = s.N; ta.q = m.q; ta.Vc = pl->Vc; ta.Vpad = pl->Vpad;
                    ta.KK = s.KK; ta.NS = 0; ta.npy = s.npy;
                    ta.yy = s.yy.p; ta.Yb = s.Yb.p;
                    targs.push_back(ta);
                    tfirst.push_back(tg.blocks);
                    tg.blocks += (pl->Vpad + 255) / 256;
                    tg.n++;
                }
                if (tg.n) b->tgroups.push_back(tg);
                }
        BPG_BHIP(b->ttab.ensure(std::max<size_t>(targs.size(), 1)));
        BPG_BHIP(b->tfirst.ensure(std::max<size_t>(tfirst.size(), 1)));
        if (!targs.empty()) {
            BPG_BHIP(hipMemcpy(b->ttab.p, targs.data(), targs.size() * sizeof(bpg::TransformArgs),
                               hipMemcpyHostToDevice));
            BPG_BHIP(hipMemcpy(b->tfirst.p, tfirst.data(), tfirst.size() * sizeof(int), hipMemcpyHostToDevice));
        }
    }
    BPG_BHIP(b->table.ensure(std::max<size_t>(host.size(), 1)));
    if (!host.empty())
        BPG_BHIP(hipMemcpy(b->table.p, host.data(), host.size() * sizeof(bpg::BatchStudy), hipMemcpyHostToDevice));
    BPG_BHIP(b->jobs.ensure(std::max<size_t>(rjobs.size(), 1)));
    if (!rjobs.empty())
        BPG_BHIP(hipMemcpy(b->jobs.p, rjobs.data(), rjobs.size() * sizeof(bpg::ReduceJob), hipMemcpyHostToDevice));
    BPG_BHIP(hipStreamCreate(&b->stream));
    BPG_BHIP(hipEventCreateWithFlags(&b->first_done, hipEventDisableTiming));
    for (auto& x : b->side) BPG_BHIP(hipStreamCreate(&x));
    for (int g = 0; g < 2; g++) {
        BPG_BHIP(hipEventCreateWithFlags(&b->fork[g], hipEventDisableTiming));
        for (int h = 0; h < bpg_batch::NBR; h++) {
            BPG_BHIP(hipEventCreateWithFlags(&b->join[g][h], hipEventDisableTiming));
            BPG_BHIP(hipEventCreateWithFlags(&b->prep[g][h], hipEventDisableTiming));
        }
    }
    // Two graphs.  The first: the upload of all permutations / signs and one
    // fill clearing the sample-list counters, queue heads, ill-conditioned
    // counts and flags (a kernel, not a memset node); then, as the second,
    // the stage's classes, the largest work first and taken in turn by two
    // branches: all members' nuisance transforms and operand builds, then
    // (both branches done) the persistent contractions; one launch of the
    // stage's max reductions; the FP64 recomputation of listed samples; the
    // copies of its maxima, counts and flags to the host.  Stage 0 holds the
    // studies likely to repeat permutations, so that their repeats run
    // beside the second graph.
    std::vector<double> work(b->cls.size(), 0.0);
    for (size_t ci = 0; ci < b->cls.size(); ci++)
        for (int i : b->cls[ci].members) {
            const auto& s = b->st[i];
            work[ci] += (double)s.K * s.plan->Vpad * s.KK * s.PP;
        }
    std::vector<int> order(b->cls.size());
    for (size_t ci = 0; ci < order.size(); ci++) order[ci] = (int)ci;
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return work[x] > work[y]; });
    hipStream_t cs = b->stream;
    for (int g = 0; g < 2; g++) {
        const size_t m0 = g ? b->max_end[0] : 0, f0 = g ? b->flag_end[0] : 0;
        const int s0 = g ? b->slot_end[0] : 0, j0 = g ? b->job_end[0] : 0;
        const bool empty = b->job_end[g] == j0;
        if (g == 1 && empty) break;
        BPG_BHIP(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
        int rc = BPG_OK;
        auto cap = [&](hipError_t e) {
            if (e != hipSuccess && rc == BPG_OK) {
                ctx->err = std::string("graph capture: ") + hipGetErrorString(e);
                rc = BPG_ERR_HIP;
            }
        };
        if (g == 0) {
            if (b->n_perm)
                cap(hipMemcpyAsync(b->perm_all.p, b->h_perm.p, b->n_perm * sizeof(unsigned short),
                                   hipMemcpyHostToDevice, cs));
            if (b->n_sign)
                cap(hipMemcpyAsync(b->sign_all.p, b->h_sign.p, b->n_sign * sizeof(float), hipMemcpyHostToDevice, cs));
            cap(bpg::launch_fill_u32(b->words.p, n_words, 0u, cs));
        }
        // the branches take the stage's work in turn: first the data
        // fragments (one transform launch per nuisance width, so that small
        // studies fill the chip together) and operand fragments of its
        // studies (bandwidth-bound kernels that would wait for CUs behind a
        // persistent contraction), then, once every branch has built its
        // share, the contractions, class by class, so that a class's blocks
        // start in the tail of the one before it
        constexpr int NBR = bpg_batch::NBR;
        auto branch = [&](int t) { return t % NBR == 0 ? cs : b->side[t % NBR - 1]; };
        cap(hipEventRecord(b->fork[g], cs));
        for (int h = 1; h < NBR; h++) cap(hipStreamWaitEvent(b->side[h - 1], b->fork[g], 0));
        int turn = 0;
        for (const auto& tg : b->tgroups)
            if (tg.stage == g && rc == BPG_OK)
                cap(bpg::launch_transform_batch(b->ttab.p + tg.tab_off, b->tfirst.p + tg.tab_off, tg.n, tg.blocks,
                                                tg.QB, branch(turn++)));
        for (int ci : order) {
            if (rc != BPG_OK) break;
            const auto& c = b->cls[ci];
            if (c.stage != g) continue;
            hipStream_t bs = branch(turn++);
            for (int i : c.members) {
                auto& s = b->st[i];
                const bpg::MapModel& m = s.plan->models[s.map];
                bpg::OperandArgs oa{};
                oa.perm = s.sign ? nullptr : b->perm_all.p + s.in_off;
                oa.sign = s.sign ? b->sign_all.p + s.in_off : nullptr;
                oa.Qt = s.plan->mdev.p + s.plan->mdev_off[s.map];
                oa.N = s.N; oa.pfit = m.pfit; oa.K = (int)s.K; oa.PP = s.PP; oa.KK = s.KK;
                oa.n_sub = s.L.n_sub;
                oa.Gb = s.Gb.p;
                cap(bpg::launch_operand_bf16(oa, bs));
            }
        }
        for (int h = 0; h < NBR; h++) cap(hipEventRecord(b->prep[g][h], branch(h)));
        for (int h = 0; h < NBR; h++)
            for (int o = 0; o < NBR; o++)
                if (o != h) cap(hipStreamWaitEvent(branch(h), b->prep[g][o], 0));
        turn = 0;
        for (int ci : order) {
            if (rc != BPG_OK) break;
            const auto& c = b->cls[ci];
            if (c.stage != g) continue;
            hipStream_t bs = branch(turn++);
            bpg::BatchClass bc{};
            bc.studies = b->table.p + c.first;
            bc.n_studies = (int)c.members.size();
            std::copy(c.n_items, c.n_items + bpg::BPG_XCDS, bc.n_items);
            bc.counter = b->heads() + (size_t)ci * bpg::BPG_XCDS;
            cap(bpg::launch_contract_bf16_persist(bc, c.KK, c.PP, c.sign != 0, c.variant, c.excl != 0, c.grid, bs));
        }
        for (int h = 1; h < NBR; h++) {
            cap(hipEventRecord(b->join[g][h], b->side[h - 1]));
            cap(hipStreamWaitEvent(cs, b->join[g][h], 0));
        }
        if (!empty) {
            cap(bpg::launch_reduce_max_jobs(b->jobs.p + j0, b->job_end[g] - j0, b->max_K, cs));
            // the stage's listed samples, recomputed in FP64, into its maxima
            for (auto& s : b->st) {
                if (s.stage != g || !s.cap || s.plan->Vc == 0) continue;
                const bpg::MapModel& m = s.plan->models[s.map];
                bpg::SampleFixArgs fa{};
                fa.samples = s.samples.p;
                fa.n_samples = b->counters() + (size_t)bpg::BPG_SAMPLE_SEGS * s.excl_slot;
                fa.cap = s.cap;
                fa.Yc = s.plan->Yc.p;
                fa.Qn = s.plan->mdev.p + s.plan->mdev_off[s.map] + m.Qt.size();
                fa.QB = std::max(bpg::transform_width(m.q), 0); fa.q = m.q;
                fa.Qt = s.plan->mdev.p + s.plan->mdev_off[s.map];
                fa.perm = b->perm_all.p + s.in_off;
                fa.N = s.N; fa.pfit = m.pfit; fa.Vc = s.plan->Vc; fa.K = (int)s.K;
                fa.dof = (double)m.dof;
                fa.maxout = b->max_all.p + s.max_off;
                cap(bpg::launch_fix_samples(fa, cs));
            }
            cap(hipMemcpyAsync(b->h_max.p + m0, b->max_all.p + m0, (b->max_end[g] - m0) * sizeof(float),
                               hipMemcpyDeviceToHost, cs));
            cap(hipMemcpyAsync(b->h_unsafe.p + s0, b->unsafe() + s0, (size_t)(b->slot_end[g] - s0) * sizeof(unsigned),
                               hipMemcpyDeviceToHost, cs));
            cap(hipMemcpyAsync(b->h_flags.p + f0, b->flags() + f0, b->flag_end[g] - f0, hipMemcpyDeviceToHost, cs));
        }
        const hipError_t ec = hipStreamEndCapture(cs, &b->graph[g]);
        if (rc != BPG_OK) return fail(rc);
        BPG_BHIP(ec);
        BPG_BHIP(hipGraphInstantiate(&b->exec[g], b->graph[g], nullptr, nullptr, 0));
    }
#undef BPG_BHIP
    *out = b;
    return BPG_OK;
}

namespace {

// The flagged permutations of study i of a finished stage, enqueued for a
// repeat through its plan in the next form (split-BF16 -> FP32, an exact
// split -> FP64); the plan cascades further when the repeat is collected.
int batch_enqueue_repeat(bpg_batch* b, int i, const unsigned short* const* perms, const float* const* signs) {
    auto& s = b->st[i];
    s.rep.clear();
    if (s.plan->Vc == 0 || b->h_unsafe.p[s.slot] == 0) return BPG_OK;
    const unsigned char* fl = b->h_flags.p + s.flag_off;
    for (long k = 0; k < s.K; k++)
        if (fl[k]) s.rep.push_back(k);
    if (s.rep.empty()) return BPG_OK;
    const long Ks = (long)s.rep.size();
    const unsigned short* rp = nullptr;
    const float* rs = nullptr;
    if (Ks == s.K) {
        rp = s.sign ? nullptr : perms[i];
        rs = s.sign ? signs[i] : nullptr;
    } else if (s.sign) {
        s.rsign.resize((size_t)Ks * s.N);
        for (long j = 0; j < Ks; j++)
            std::copy(signs[i] + s.rep[j] * s.N, signs[i] + (s.rep[j] + 1) * s.N, s.rsign.begin() + j * s.N);
        rs = s.rsign.data();
    } else {
        s.rperm.resize((size_t)Ks * s.N);
        for (long j = 0; j < Ks; j++)
            std::copy(perms[i] + s.rep[j] * s.N, perms[i] + (s.rep[j] + 1) * s.N, s.rperm.begin() + j * s.N);
        rp = s.rperm.data();
    }
    if (s.plan->pending.active) {
        s.rep.clear();
        BPG_FAIL(b->ctx, BPG_ERR_ARG, "a permutation phase of a batched plan is in flight");
    }
    const int rc = phase_enqueue(s.plan, s.map, rp, rs, Ks, nullptr, s.npy == 3 ? BPG_PREC_FP64 : BPG_PREC_FP32);
    if (rc != BPG_OK) s.rep.clear();
    return rc;
}

}  // namespace

extern "C" int bpg_batch_run(bpg_batch* b, const unsigned short* const* perms, const float* const* signs,
                             float* const* maxdist) {
    if (!b || !maxdist) return BPG_ERR_ARG;
    bpg_ctx* ctx = b->ctx;
    const int n = (int)b->st.size();
    const auto t0 = std::chrono::steady_clock::now();
    for (int i = 0; i < n; i++) {
        const auto& s = b->st[i];
        if (!maxdist[i]) BPG_FAIL(ctx, BPG_ERR_ARG, "null maxdist");
        const void* src = s.sign ? (const void*)(signs ? signs[i] : nullptr) : (const void*)(perms ? perms[i] : nullptr);
        if (!src) BPG_FAIL(ctx, BPG_ERR_ARG, "missing permutation / sign matrix");
        if (s.plan->pending.active) BPG_FAIL(ctx, BPG_ERR_ARG, "a permutation phase of a batched plan is in flight");
        if (s.sign) {
            std::memcpy(b->h_sign.p + s.in_off, signs[i], (size_t)s.K * s.N * sizeof(float));
        } else {
            unsigned short worst = 0;
            for (size_t q = 0; q < (size_t)s.K * s.N; q++) {
                b->h_perm.p[s.in_off + q] = perms[i][q];
                worst = std::max(worst, perms[i][q]);
            }
            if (worst >= s.N) BPG_FAIL(ctx, BPG_ERR_ARG, "permutation index out of range");
        }
    }
    BPG_HIP(ctx, hipSetDevice(ctx->device));
    BPG_HIP(ctx, hipGraphLaunch(b->exec[0], b->stream));
    BPG_HIP(ctx, hipEventRecord(b->first_done, b->stream));
    if (b->exec[1]) BPG_HIP(ctx, hipGraphLaunch(b->exec[1], b->stream));
    // the first stage's repeats are enqueued (on the plans' streams) while
    // the second graph runs; then the second stage's
    int rc = BPG_OK;
    std::vector<int> pending;
    for (int g = 0; g < 2 && rc == BPG_OK; g++) {
        const hipError_t e = g == 0 ? hipEventSynchronize(b->first_done) : hipStreamSynchronize(b->stream);
        if (e != hipSuccess) {
            ctx->err = std::string("batch graph: ") + hipGetErrorString(e);
            rc = BPG_ERR_HIP;
            break;
        }
        if (g == 1)
            b->last_graph_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        for (int i = 0; i < n && rc == BPG_OK; i++) {
            if (b->st[i].stage != g) continue;
            rc = batch_enqueue_repeat(b, i, perms, signs);
            if (rc == BPG_OK && !b->st[i].rep.empty()) pending.push_back(i);
        }
    }
    if (rc != BPG_OK) (void)hipStreamSynchronize(b->stream);
    int fallback = 0;
    for (int i : pending) {
        auto& s = b->st[i];
        s.rmax.resize(s.rep.size());
        const int r = phase_finish(s.plan, s.rmax.data());
        if (rc == BPG_OK) rc = r;
        fallback++;
    }
    if (rc != BPG_OK) return rc;
    for (int i = 0; i < n; i++) {
        const auto& s = b->st[i];
        float* md = maxdist[i];
        if (s.plan->Vc == 0) {
            for (long k = 0; k < s.K; k++) md[k] = empty_max();
            continue;
        }
        const float* hm = b->h_max.p + s.max_off;
        for (long k = 0; k < s.K; k++)
            md[k] = hm[k] > -INFINITY ? (float)((float)quantize(hm[k]) / 10000.0f) : empty_max();
        for (size_t j = 0; j < s.rep.size(); j++) md[s.rep[j]] = s.rmax[j];
    }
    b->last_fallback = fallback;
    b->last_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return BPG_OK;
}

extern "C" int bpg_batch_last_timing(const bpg_batch* b, double* run_ms, double* graph_ms, int* repeated_studies,
                                     int* classes) {
    if (!b) return BPG_ERR_ARG;
    if (run_ms) *run_ms = b->last_ms;
    if (graph_ms) *graph_ms = b->last_graph_ms;
    if (repeated_studies) *repeated_studies = b->last_fallback;
    if (classes) *classes = (int)b->cls.size();
    return BPG_OK;
}

extern "C" void bpg_batch_destroy(bpg_batch* b) {
    if (!b) return;
    (void)hipSetDevice(b->ctx->device);
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    delete b;
}
