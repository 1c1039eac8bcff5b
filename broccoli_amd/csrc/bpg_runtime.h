// bpg_runtime.h -- host runtime shared by the plans of the C ABI: the
// context, growable device / pinned buffers, error macros, the compact
// neighbour table, the batched cluster-size pass and the max quantisation.
#pragma once

#include <hip/hip_runtime.h>

#include <climits>
#include <cmath>
#include <map>
#include <string>
#include <tuple>
#include <vector>

#include "bpg.h"
#include "bpg_kernels.h"

namespace bpg {
// Autotuner entries: per (k-steps, slots per permutation, sign operand,
// F test, statistic output, log2 of the padded voxel count) the subtile
// variant and the chunk-length scaling that ran fastest on this device.
using TuneKey = std::tuple<int, int, int, int, int, int>;
struct TuneEntry {
    int variant = 0, chunk_pow = 0;
    double default_ms = 0.0, best_ms = 0.0;
};
}  // namespace bpg

struct bpg_ctx {
    int device = 0;
    std::map<bpg::TuneKey, bpg::TuneEntry> tune;
    bool tune_forced = false;      // bpg_plan_tune's candidate in force
    bpg::TuneEntry tune_force;
    std::string err;
    std::string name;
    std::string arch;
    size_t lds_per_block = 0;  // shared memory one block may use
    int cus = 256;             // compute units (8 XCDs on gfx950)
    // device address of this GPU's no-allocation scratch (bpg_noalloc.hip),
    // or null with the reason in scratch_err
    void* scratch = nullptr;
    std::string scratch_err;
};

namespace bpg {

// device buffer that only grows
template <class T>
struct DBuf {
    T* p = nullptr;
    size_t n = 0;
    hipError_t ensure(size_t want) {
        if (want <= n && p) return hipSuccess;
        if (p) {
            (void)hipFree(p);
            p = nullptr;
            n = 0;
        }
        if (want == 0) want = 1;
        const hipError_t e = hipMalloc(reinterpret_cast<void**>(&p), want * sizeof(T));
        if (e == hipSuccess) n = want;
        return e;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
};

// pinned host staging: the per-phase uploads and downloads run as true
// asynchronous DMA, and the host fills / reads them while kernels run
template <typename T>
struct HBuf {
    T* p = nullptr;
    size_t n = 0;
    hipError_t ensure(size_t want) {
        if (want <= n && p) return hipSuccess;
        if (p) {
            (void)hipHostFree(p);
            p = nullptr;
            n = 0;
        }
        if (want == 0) want = 1;
        const hipError_t e = hipHostMalloc(reinterpret_cast<void**>(&p), want * sizeof(T), hipHostMallocDefault);
        if (e == hipSuccess) n = want;
        return e;
    }
    void release() {
        if (p) (void)hipHostFree(p);
        p = nullptr;
        n = 0;
    }
};

// (int)(T * 10000.0f) as KM:299, made total (saturating, NaN -> INT_MIN);
// inline, so that the loops over K maxima vectorise
inline int quantize(float t) {
    const float x = t * 10000.0f;
    if (std::isnan(x)) return INT_MIN;
    if (x >= 2147483648.0f) return INT_MAX;
    if (x <= -2147483648.0f) return INT_MIN;
    return (int)x;
}
// The empty-mask value of LIB:10971: the int buffer starts from the bit
// pattern of -1e6f and is returned as (float)max / 10000.0f.
float empty_max();
float elapsed(hipEvent_t a, hipEvent_t b);

// In-mask voxels in linear grid order (mask == 1.0f); grid2comp is -1
// outside the mask.
void compact_mask(const float* mask, size_t G, std::vector<unsigned>& comp2grid, std::vector<int>& grid2comp);

// Compact indices of the 13 lower-index 26-neighbours of every in-mask
// voxel, face neighbours first, -1 outside the mask or the grid: [Vc][13].
void build_neighbours(int W, int H, int D, const std::vector<unsigned>& comp2grid,
                      const std::vector<int>& grid2comp, std::vector<int>& nbr);

// Scratch of the cluster-size pass.
struct ClusterScratch {
    DBuf<unsigned> parent, sizes, overflow;
    void release() {
        parent.release();
        sizes.release();
        overflow.release();
    }
};

// Largest component per map of K bitmask maps (cluster extent, or mass with
// `values`), written to maxsize[K] (uint).  base carries bits (K maps),
// nbr, values (K maps or null), Vc, nwords and connectivity.  Maps run in
// batches whose union-find scratch fits 8 GB; whole maps go through LDS
// when the device's shared memory holds them, and the few with too many
// active voxels take the global-forest path.
int cluster_max(bpg_ctx* ctx, ClusterScratch& s, const CclArgs& base, long K, unsigned* maxsize, hipStream_t st);

}  // namespace bpg

#define BPG_HIP(ctx, call)                                                         \
    do {                                                                           \
        const hipError_t e_ = (call);                                              \
        if (e_ != hipSuccess) {                                                    \
            (ctx)->err = std::string(#call) + ": " + hipGetErrorString(e_);        \
            return e_ == hipErrorOutOfMemory ? BPG_ERR_NOMEM : BPG_ERR_HIP;        \
        }                                                                          \
    } while (0)

#define BPG_FAIL(ctx, code, msg) \
    do {                         \
        (ctx)->err = (msg);      \
        return (code);           \
    } while (0)
